"""CGLS and split Bregman with the reference's class signatures.

Mirrors matlab/classes/+Optimizers/@CGLS (CGLS.m:21-45, solveCGLS.m:1-31) and
@SplitBregman (SplitBregman.m:22-48, solveSplitBregman.m:1-44).  When the
operator is one of this package's projectors (a Projector or a scanner's
bound apply/applyAdjoint) the whole solve runs device-resident in the native
library (fl_cgls / fl_split_bregman: FP64 reductions and stopping tests on
the GPU, optional hipGraph replay and BF16/FP16 storage); otherwise the same
recurrences run over the caller's callables.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .datatypes import ObjectData
from .framelets import FrameletSystem
from .projector import Projector


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _projector_of(A):
    if isinstance(A, Projector):
        return A
    owner = getattr(A, "__self__", None)
    if isinstance(owner, Projector):        # P.forward
        return owner
    return getattr(owner, "_proj", None) if owner is not None else None


def _data(v):
    return v.dataArray if isinstance(v, ObjectData) else v


def _flags(graph=False, matched=False, storage=torch.float32):
    f = 0
    if graph:
        f |= _lib.FL_SOLVE_GRAPH
    if matched:
        f |= _lib.FL_SOLVE_MATCHED
    if storage == torch.bfloat16:
        f |= _lib.FL_SOLVE_LOWP_BF16
    elif storage == torch.float16:
        f |= _lib.FL_SOLVE_LOWP_F16
    return f


def cgls(P: Projector, b: torch.Tensor, u0: torch.Tensor, lam: float = 0.0, iters: int = 20,
         tol: float = 0.0, graph: bool = False, matched: bool = False, storage=torch.float32,
         log: bool = True):
    """Device-resident CGLS on (A^T A + lam I) u = b.  Returns (u, gamma_log, iters_done)."""
    u = u0.detach().float().contiguous().clone()
    b = b.float().contiguous()
    glog = (C.c_double * (iters + 1))() if log else None
    done = C.c_int(0)
    _lib.check(_lib.lib().fl_cgls(P.handle, C.c_void_p(b.data_ptr()), C.c_void_p(u.data_ptr()), float(lam),
                                  int(iters), float(tol), _flags(graph, matched, storage),
                                  glog, C.byref(done), _stream()))
    return u, (np.array(glog[:]) if log else None), done.value


def cgls_f64(P: Projector, b: torch.Tensor, u0: torch.Tensor, lam: float = 0.0, iters: int = 20,
             tol: float = 0.0):
    """CGLS with FP64 vectors over the FP64 projector pair (the accuracy
    reference of the FP32 and mixed-precision solves).  Returns (u, gamma_log, iters_done)."""
    u = u0.detach().to(torch.float64).contiguous().clone()
    b = b.to(torch.float64).contiguous()
    glog = (C.c_double * (iters + 1))()
    done = C.c_int(0)
    _lib.check(_lib.lib().fl_cgls_f64(P.handle, C.c_void_p(b.data_ptr()), C.c_void_p(u.data_ptr()), float(lam),
                                      int(iters), float(tol), glog, C.byref(done), _stream()))
    return u, np.array(glog[:]), done.value


def split_bregman_f64(P: Projector, kind: str, f: torch.Tensor, u0: torch.Tensor, lam: float, mu: float,
                      outer: int, inner: int, history: bool = False):
    """Split Bregman with every value in double (FP64 projector pair, FP64
    one-level framelets): the accuracy reference of the FP32 / BF16 / FP16
    solves.  Returns (u, constraint residual per outer step), and with
    history=True also the iterate after each outer step (outer volumes)."""
    from .framelets import _KIND
    u = u0.detach().to(torch.float64).contiguous().clone()
    f = f.to(torch.float64).contiguous()
    res = (C.c_double * max(outer, 1))()
    hist = torch.empty((max(outer, 1),) + tuple(u.shape), dtype=torch.float64, device=u.device) if history else None
    _lib.check(_lib.lib().fl_split_bregman_f64(P.handle, _KIND[kind], C.c_void_p(f.data_ptr()),
                                               C.c_void_p(u.data_ptr()), float(lam), float(mu), int(outer),
                                               int(inner), res, C.c_void_p(hist.data_ptr() if history else 0),
                                               _stream()))
    return (u, np.array(res[:outer]), hist[:outer]) if history else (u, np.array(res[:outer]))


def split_bregman(P: Projector, kind: str, f: torch.Tensor, u0: torch.Tensor, lam: float, mu: float,
                  outer: int, inner: int, cg_tol: float = 0.0, graph: bool = False, matched: bool = False,
                  storage=torch.float32, level: int = 1, history: bool = False):
    """Device-resident split Bregman with a `level`-level frame.  Returns (u,
    constraint residual per outer step), and with history=True also the
    iterate after each outer step (outer volumes)."""
    from .framelets import _KIND
    u = u0.detach().float().contiguous().clone()
    f = f.float().contiguous()
    res = (C.c_double * max(outer, 1))()
    hist = torch.empty((max(outer, 1),) + tuple(u.shape), dtype=torch.float32, device=u.device) if history else None
    _lib.check(_lib.lib().fl_split_bregman(P.handle, _KIND[kind], int(level), C.c_void_p(f.data_ptr()),
                                           C.c_void_p(u.data_ptr()), float(lam), float(mu), int(outer),
                                           int(inner), float(cg_tol), _flags(graph, matched, storage), res,
                                           C.c_void_p(hist.data_ptr() if history else 0), _stream()))
    return (u, np.array(res[:outer]), hist[:outer]) if history else (u, np.array(res[:outer]))


class CGLS:
    """CGLS(A, At, b, u0, lam, globalIter, convergenceTol) (CGLS.m:21-45)."""

    def __init__(self, A, At, b, u0, lam, globalIter=500, convergenceTol=1e-14, graph=False,
                 storage=torch.float32):
        self.A, self.At, self.b, self.u0, self.lam = A, At, b, u0, lam
        self.globalIter, self.convergenceTol = globalIter, convergenceTol
        self.graph, self.storage = graph, storage
        self.gamma_log = None

    def solveCGLS(self):
        P = _projector_of(self.A)
        b, u0 = _data(self.b), _data(self.u0)
        if P is not None:
            matched = getattr(getattr(self.At, "__self__", None), "adjoint_kind", 0) == 1
            u, self.gamma_log, _ = cgls(P, b, u0, self.lam, self.globalIter, self.convergenceTol,
                                        graph=self.graph, matched=matched, storage=self.storage)
        else:
            u = self._generic(b, u0)
        return ObjectData(self.u0.dim, u, self.u0.L) if isinstance(self.u0, ObjectData) else u

    def _generic(self, b, u0):
        """solveCGLS.m:7-30 over arbitrary callables: the operators are the
        caller's, the vector steps and FP64 dots are the native ones."""
        A, At, lam = self.A, self.At, float(self.lam)
        L = _lib.lib()
        u = u0.detach().float().contiguous().clone()
        b = b.float().contiguous()
        if not (u.is_cuda and b.is_cuda):
            raise TypeError("CGLS vectors must be CUDA (HIP) tensors")
        n = u.numel()
        ptr = lambda t: C.c_void_p(t.data_ptr())

        def normal(v):
            return _data(At(A(v))).float().contiguous()

        r, p = torch.empty_like(u), torch.empty_like(u)
        gold = C.c_double(0.0)
        _lib.check(L.fl_vec_cgls_init(n, ptr(normal(u)), ptr(u), ptr(b), lam, ptr(r), ptr(p), C.byref(gold),
                                      _stream()))
        log = [gold.value]
        for _ in range(self.globalIter):
            q = normal(p)
            pq = C.c_double(0.0)
            _lib.check(L.fl_vec_cgls_pq(n, ptr(q), ptr(p), lam, C.byref(pq), _stream()))
            if pq.value == 0.0:
                break
            gnew = C.c_double(0.0)
            _lib.check(L.fl_vec_cgls_update(n, ptr(u), ptr(r), ptr(p), ptr(q), gold.value / pq.value,
                                            C.byref(gnew), _stream()))
            log.append(gnew.value)
            _lib.check(L.fl_vec_cgls_dir(n, ptr(p), ptr(r), gnew.value / gold.value, _stream()))
            gold = gnew
            if gold.value < self.convergenceTol:
                break
        self.gamma_log = np.array(log)
        return u


class SplitBregman:
    """SplitBregman(A, At, W, Wt, lam, mu, f, u0, globalIter, cgIter, cgTol) (SplitBregman.m:22-48)."""

    def __init__(self, A, At, W, Wt, lam, mu, f, u0, globalIter=1, cgIter=10, cgTol=1e-10, graph=False,
                 storage=torch.float32):
        self.A, self.At, self.W, self.Wt = A, At, W, Wt
        self.lam, self.mu, self.f, self.u0 = lam, mu, f, u0
        self.globalIter, self.cgIter, self.cgTol = globalIter, cgIter, cgTol
        self.graph, self.storage = graph, storage
        self.residuals = None

    def solveSplitBregman(self):
        P = _projector_of(self.A)
        system = getattr(self.W, "__self__", None)
        f, u0 = _data(self.f), _data(self.u0)
        if P is None or not isinstance(system, FrameletSystem):
            raise NotImplementedError("split Bregman runs on this package's projectors with a "
                                      "FrameletSystem (W=system.forwardTransform)")
        matched = getattr(getattr(self.At, "__self__", None), "adjoint_kind", 0) == 1
        u, self.residuals = split_bregman(P, system.type, f, u0, self.lam, self.mu, self.globalIter,
                                          self.cgIter, self.cgTol, graph=self.graph, matched=matched,
                                          storage=self.storage, level=system.level)
        return ObjectData(self.u0.dim, u, self.u0.L) if isinstance(self.u0, ObjectData) else u

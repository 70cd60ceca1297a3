"""Batches of independent 2D fan-beam problems of uneven sizes.

``ProjectorBatch(paras)`` owns an ``fl_batch`` built from a list of FanPara
(framelab_amd.geometry).  Volumes and sinograms of all problems are packed
back to back in one float32 tensor each (``pack`` / ``unpack``); every
operator is one persistent launch over the whole batch, and CGLS / split
Bregman advance all problems in lock step with per-problem FP64 scalars and
stopping tests.  Each problem's result equals what a single ``Projector`` on
the same geometry returns (identical per-ray and per-pixel arithmetic).  The
reference handles one problem per call (matlab/classes/+Operators/
@FanBeamScanner/apply.m); this is the batched form of the same operators.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .framelets import _KIND
from .geometry import FanPara
from .projector import ADJ_EXACT, ADJ_MATCHED, geometry_struct  # noqa: F401
from .solvers import _flags


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t: torch.Tensor):
    return C.c_void_p(t.data_ptr())


class ProjectorBatch:
    """Forward / adjoint / solvers for many fan problems in single launches."""

    def __init__(self, paras):
        paras = list(paras)
        if not paras:
            raise ValueError("empty batch")
        for p in paras:
            if not isinstance(p, FanPara):
                raise TypeError("batches hold 2D fan-beam geometries")
            if p.nt != 1:
                raise ValueError("batched problems have one frame each")
        self.paras = paras
        keep: list = []
        arr = (_lib.FlGeometry * len(paras))()
        for i, p in enumerate(paras):
            arr[i] = geometry_struct(p, keep)
        h = C.c_void_p()
        _lib.check(_lib.lib().fl_batch_create(C.byref(h), len(paras), arr))
        self.handle = h
        P = len(paras)
        vo = (C.c_longlong * (P + 1))()
        so = (C.c_longlong * (P + 1))()
        _lib.check(_lib.lib().fl_batch_layout(h, vo, so))
        self.vol_offsets = list(vo)
        self.sino_offsets = list(so)
        self.vol_shapes = [(p.ny, p.nx) for p in paras]
        self.sino_shapes = [(p.nv, p.nd) for p in paras]

    def __len__(self):
        return len(self.paras)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.lib().fl_batch_destroy(h)
            except Exception:
                pass
            self.handle = None

    # ------------------------------------------------------------ layout --
    @property
    def n_vol(self) -> int:
        return self.vol_offsets[-1]

    @property
    def n_sino(self) -> int:
        return self.sino_offsets[-1]

    def pack(self, tensors, sino: bool = False) -> torch.Tensor:
        """Concatenate per-problem volumes (or sinograms) into the packed layout."""
        shapes = self.sino_shapes if sino else self.vol_shapes
        if len(tensors) != len(shapes):
            raise ValueError("one tensor per problem")
        for t, s in zip(tensors, shapes):
            if tuple(t.shape) != s:
                raise ValueError(f"shape {tuple(t.shape)} != {s}")
        return torch.cat([t.reshape(-1).float() for t in tensors]).contiguous()

    def unpack(self, flat: torch.Tensor, sino: bool = False):
        off = self.sino_offsets if sino else self.vol_offsets
        shapes = self.sino_shapes if sino else self.vol_shapes
        return [flat[off[i]:off[i + 1]].view(shapes[i]) for i in range(len(shapes))]

    def _check(self, t: torch.Tensor, n: int, name: str):
        if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous() or t.numel() != n:
            raise ValueError(f"{name}: contiguous float32 CUDA tensor of {n} elements required")

    # --------------------------------------------------------- operators --
    def forward(self, x: torch.Tensor, out: torch.Tensor | None = None, method: str = "siddon") -> torch.Tensor:
        self._check(x, self.n_vol, "x")
        if out is None:
            out = torch.empty(self.n_sino, device=x.device, dtype=torch.float32)
        self._check(out, self.n_sino, "out")
        m = {"siddon": 0, "gao": 1}[method]
        _lib.check(_lib.lib().fl_batch_forward(self.handle, _ptr(x), _ptr(out), m, _stream()))
        return out

    def adjoint(self, y: torch.Tensor, out: torch.Tensor | None = None, kind: int = ADJ_EXACT) -> torch.Tensor:
        self._check(y, self.n_sino, "y")
        if out is None:
            out = torch.empty(self.n_vol, device=y.device, dtype=torch.float32)
        self._check(out, self.n_vol, "out")
        _lib.check(_lib.lib().fl_batch_adjoint(self.handle, _ptr(y), _ptr(out), int(kind), _stream()))
        return out

    # ----------------------------------------------------------- solvers --
    def cgls(self, b: torch.Tensor, u0: torch.Tensor, lam: float = 0.0, iters: int = 20, tol: float = 0.0,
             graph: bool = False, matched: bool = False):
        """CGLS on (A^T A + lam I) u = b for every problem.  Returns (u, gamma_log[P, iters+1], iters_done[P])."""
        self._check(b, self.n_vol, "b")
        u = u0.detach().float().contiguous().clone()
        self._check(u, self.n_vol, "u0")
        P = len(self)
        glog = (C.c_double * (P * (iters + 1)))()
        done = (C.c_int * P)()
        _lib.check(_lib.lib().fl_batch_cgls(self.handle, _ptr(b), _ptr(u), float(lam), int(iters), float(tol),
                                            _flags(graph, matched), glog, done, _stream()))
        return u, np.array(glog).reshape(P, iters + 1), np.array(done)

    def split_bregman(self, f: torch.Tensor, u0: torch.Tensor, frame: str = "linear", lam: float = 1e-3,
                      mu: float = 1.0, outer: int = 10, inner: int = 3, cg_tol: float = 0.0,
                      graph: bool = False, matched: bool = False):
        """Split Bregman for every problem.  Returns (u, resid[P, outer])."""
        self._check(f, self.n_sino, "f")
        u = u0.detach().float().contiguous().clone()
        self._check(u, self.n_vol, "u0")
        P = len(self)
        res = (C.c_double * max(1, P * outer))()
        _lib.check(_lib.lib().fl_batch_split_bregman(self.handle, _KIND[frame], _ptr(f), _ptr(u), float(lam),
                                                     float(mu), int(outer), int(inner), float(cg_tol),
                                                     _flags(graph, matched), res, _stream()))
        return u, np.array(res)[:P * outer].reshape(P, outer)

    def autotune(self):
        """Pick the forward ray-tile width and persistent grid sizes by timing them on this batch."""
        t, fg, ag = C.c_int(0), C.c_int(0), C.c_int(0)
        _lib.check(_lib.lib().fl_batch_autotune(self.handle, C.byref(t), C.byref(fg), C.byref(ag)))
        return {"fwd_tile": t.value, "fwd_grid": fg.value, "adj_grid": ag.value}

"""Gradients of the reconstruction objective (the hand-written backward passes).

J(x) = 1/2 |A x - b|^2 + lam * sum_{bands != low-pass} psi((W x)_b),
psi(t) = sqrt(t^2 + eps) (smoothed l1).  The data term's backward is the
voxel-driven adjoint (by default the matched Siddon transpose, so the
gradient is the exact transpose of the FP32 forward); the framelet term's
backward is the synthesis transform.  No reference counterpart: the
reference has no gradient API; its operators are the forward passes.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from .framelets import FrameletSystem
from .projector import Projector


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def data_gradient(P: Projector, x: torch.Tensor, b: torch.Tensor, exact_adjoint: bool = False):
    """(g, loss) with g = A^T (A x - b), loss = 0.5 |A x - b|^2 (FP64)."""
    x = x.float().contiguous()
    b = b.float().contiguous()
    g = torch.empty_like(x)
    loss = C.c_double(0.0)
    flags = _lib.FL_GRAD_EXACT_ADJOINT if exact_adjoint else 0
    _lib.check(_lib.lib().fl_grad_data(P.handle, C.c_void_p(x.data_ptr()), C.c_void_p(b.data_ptr()),
                                       C.c_void_p(g.data_ptr()), C.byref(loss), flags, _stream()))
    return g, loss.value


def framelet_gradient(W: FrameletSystem, x: torch.Tensor, lam: float, eps: float = 1e-6,
                      with_lowpass: bool = False, accumulate_into: torch.Tensor | None = None):
    """(g, loss) with g = lam * W^T psi'(W x) (+ accumulate_into), loss = lam * sum psi(W x) over
    every band of every level but the low-pass ones (the top level's too unless with_lowpass)."""
    x = x.float().contiguous()
    g = accumulate_into if accumulate_into is not None else torch.zeros_like(x)
    work = torch.empty((W.level * W.nbands,) + tuple(x.shape), dtype=torch.float32, device=x.device)
    nx, ny, nz = W._sizes(x.shape)
    loss = C.c_double(0.0)
    _lib.check(_lib.lib().fl_grad_framelet(W.dim, W.kind, W.level, nx, ny, nz, C.c_void_p(x.data_ptr()), float(lam),
                                           float(eps), int(with_lowpass), C.c_void_p(g.data_ptr()),
                                           C.c_void_p(work.data_ptr()), C.byref(loss), _stream()))
    return g, loss.value


def objective_gradient(P: Projector, W: FrameletSystem, x: torch.Tensor, b: torch.Tensor, lam: float,
                       eps: float = 1e-6, with_lowpass: bool = False, exact_adjoint: bool = False):
    """(g, J): the full gradient, data term plus framelet term, in two native calls."""
    g, ld = data_gradient(P, x, b, exact_adjoint)
    g, lr = framelet_gradient(W, x, lam, eps, with_lowpass, accumulate_into=g)
    return g, ld + lr

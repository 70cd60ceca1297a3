"""B-spline tight-framelet systems on the GPU.

Mirror of matlab/classes/+Transforms/@FrameletSystem (FrameletSystem.m:1-73):
``FrameletSystem(dim, type, level)`` with ``forwardTransform`` (analysis W,
FraDecMultiLevel{2D,3D}.m) and ``adjointTransform`` (synthesis W^T,
FraRecMultiLevel{2D,3D}.m), and the coefficient container
``FrameletExpansion`` with ``softThreshold`` (softThreshold.m:1-80).

Coefficients are one tensor of shape (level, NB, *volume_shape); band
b = i + NF*j (+ NF^2*k) is the MATLAB cell {i,j(,k)} in linear order.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib

_KIND = {"haar": _lib.FL_FRAME_HAAR, "linear": _lib.FL_FRAME_LINEAR, "cubic": _lib.FL_FRAME_CUBIC}
_DT = {torch.float32: _lib.FL_F32, torch.bfloat16: _lib.FL_BF16, torch.float16: _lib.FL_F16}


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def num_bands(dim: int, kind: str) -> int:
    nb = _lib.lib().fl_framelet_num_bands(dim, _KIND[kind])
    if nb < 0:
        _lib.check(nb)
    return nb


class FrameletSystem:
    """FrameletSystem(dim, type, level) -- FrameletSystem.m:12-30."""

    def __init__(self, dim: int = 2, type: str = "haar", level: int = 1):
        if dim not in (2, 3):
            raise ValueError("dim must be 2 or 3")
        if type not in _KIND:
            raise ValueError("type must be 'haar', 'linear' or 'cubic'")
        if level < 1:
            raise ValueError("level must be >= 1")
        self.dim, self.type, self.level = dim, type, level
        self.kind = _KIND[type]
        self.nbands = num_bands(dim, type)

    def _sizes(self, shape):
        if self.dim == 3:
            nz, ny, nx = shape[-3:]
        else:
            ny, nx = shape[-2:]
            nz = 1
        return nx, ny, nz

    def forwardTransform(self, u, coef_dtype=torch.float32) -> "FrameletExpansion":
        """alpha = W u (FrameletSystem.m:31-49)."""
        from .datatypes import ObjectData
        if isinstance(u, ObjectData):
            u = u.dataArray
        u = u.contiguous()
        nx, ny, nz = self._sizes(u.shape)
        c = torch.empty((self.level, self.nbands) + tuple(u.shape[-self.dim:]), dtype=coef_dtype,
                        device=u.device)
        _lib.check(_lib.lib().fl_framelet_forward(self.dim, self.kind, self.level, nx, ny, nz,
                                                  C.c_void_p(u.data_ptr()), _DT[u.dtype],
                                                  C.c_void_p(c.data_ptr()), _DT[c.dtype], _stream()))
        return FrameletExpansion(self.dim, self, c)

    def adjointTransform(self, alpha, out_dtype=torch.float32) -> torch.Tensor:
        """u = W^T alpha (FrameletSystem.m:51-64); W^T W = I."""
        c = alpha.frameletArray if isinstance(alpha, FrameletExpansion) else alpha
        c = c.contiguous()
        shape = tuple(c.shape[2:])
        nx, ny, nz = self._sizes(shape)
        u = torch.empty(shape, dtype=out_dtype, device=c.device)
        work = None
        if self.level > 1:
            work = torch.empty((2,) + shape, dtype=torch.float32, device=c.device)
        _lib.check(_lib.lib().fl_framelet_adjoint(
            self.dim, self.kind, self.level, nx, ny, nz, C.c_void_p(c.data_ptr()), _DT[c.dtype],
            C.c_void_p(u.data_ptr()), _DT[u.dtype], C.c_void_p(work.data_ptr() if work is not None else 0),
            _stream()))
        return u


class FrameletExpansion:
    """Framelet coefficients (FrameletExpansion.m:1-80) with arithmetic.

    The reference's plus/minus/mtimes read fields that do not exist and its
    softThreshold has no output (softThreshold.m:1); here they return new
    expansions and softThreshold also updates in place."""

    def __init__(self, dim, system: FrameletSystem, frameletArray: torch.Tensor):
        self.dim, self.system, self.frameletArray = dim, system, frameletArray

    def _wrap(self, a):
        return FrameletExpansion(self.dim, self.system, a)

    def __add__(self, o):
        return self._wrap(self.frameletArray + (o.frameletArray if isinstance(o, FrameletExpansion) else o))

    def __sub__(self, o):
        return self._wrap(self.frameletArray - (o.frameletArray if isinstance(o, FrameletExpansion) else o))

    def __rmul__(self, a):
        return self._wrap(a * self.frameletArray)

    __mul__ = __rmul__

    def adjointFrameletTransform(self) -> torch.Tensor:
        """u = W^T alpha with this expansion's own system (FrameletExpansion.m:36-41)."""
        return self.system.adjointTransform(self)

    def softThreshold(self, tau: float) -> "FrameletExpansion":
        """sign(a)*max(|a|-tau,0) on every band except each level's low-pass
        (native kernel, in place; also returned)."""
        a = self.frameletArray
        if not a.is_contiguous():
            raise ValueError("frameletArray must be contiguous")
        level, nb = a.shape[0], a.shape[1]
        band = a[0, 0].numel()
        _lib.check(_lib.lib().fl_framelet_shrink(nb, level, band, C.c_void_p(a.data_ptr()), _DT[a.dtype],
                                                 float(tau), _stream()))
        return self

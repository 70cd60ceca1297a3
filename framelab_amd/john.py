"""John's equation on multi-helix cone-beam data, device-resident.

``JohnOperator`` wraps the native stencil (fl_john_*): D (the discretised
John's equation, classes/+DataTypes/@CTData/applyJohn.m:38-106), its
transpose and D^T D (applyJohnNormal.m), in FP64, on torch tensors laid out
(nz helices, nv views, nb rows, na columns) -- the reference's
[na, nb, nv, nz] array.  ``john_geometry`` derives the constants from a
multi-helix ConeBeamScanner exactly as the reference does (scale-multiplied
single coordinates promoted to double, R = SO + OD, h = vtab/(2 pi rps),
dzeta = phaseShift[1]*h, dtheta = 2 pi rps/fps).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _lib

f32 = np.float32


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def john_geometry(scanner) -> dict:
    """Constants of applyJohn.m:6-47 for a multiHelix ConeBeamScanner with para set."""
    p = scanner.para
    nz = int(scanner.nHelix)
    if nz < 2:
        raise ValueError("Johns equation only applies on multiHelix data!")
    scale = f32(p.scale)
    h = scanner.vtab / (2 * math.pi * scanner.rps)
    return dict(na=int(scanner.na), nb=int(scanner.nb), nv=int(p.nv) // nz, nz=nz,
                R=float(scanner.SO + scanner.OD), SO=float(scanner.SO), h=h,
                dzeta=float(scanner.phaseShift[1]) * h, dtheta=2 * math.pi * scanner.rps / scanner.fps,
                da=float(f32(scale * f32(p.dy_det))), db=float(f32(scale * f32(p.dz_det))),
                a=(scale * np.asarray(p.y_det, f32)).astype(f32).astype(np.float64),
                b=(scale * np.asarray(p.z_det, f32)).astype(f32).astype(np.float64))


class JohnOperator:
    def __init__(self, geom: dict):
        self.geom = dict(geom)
        g = _lib.FlJohnGeometry()
        for k in ("na", "nb", "nv", "nz"):
            setattr(g, k, int(geom[k]))
        for k in ("R", "SO", "h", "dzeta", "dtheta", "da", "db"):
            setattr(g, k, float(geom[k]))
        self._a = np.ascontiguousarray(geom["a"], np.float64)
        self._b = np.ascontiguousarray(geom["b"], np.float64)
        g.a = self._a.ctypes.data_as(C.POINTER(C.c_double))
        g.b = self._b.ctypes.data_as(C.POINTER(C.c_double))
        h = C.c_void_p()
        _lib.check(_lib.lib().fl_john_create(C.byref(h), C.byref(g)))
        self.handle = h
        na, nb, nv, nz = (int(geom[k]) for k in ("na", "nb", "nv", "nz"))
        self.in_shape = (nz, nv, nb, na)
        self.out_shape = (nz - 2 if nz > 2 else 1, nv - 2, nb - 2, na - 2)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.lib().fl_john_destroy(h)
            except Exception:
                pass
            self.handle = None

    def _check(self, t, shape, dtypes, what):
        if not t.is_cuda or t.dtype not in dtypes or not t.is_contiguous() or t.numel() != int(np.prod(shape)):
            raise ValueError(f"{what}: contiguous CUDA tensor of {shape} ({dtypes}) required")

    def apply(self, g: torch.Tensor) -> torch.Tensor:
        """D g (FP64)."""
        self._check(g, self.in_shape, (torch.float32, torch.float64), "g")
        out = torch.empty(self.out_shape, dtype=torch.float64, device=g.device)
        dt = _lib.FL_F32 if g.dtype == torch.float32 else _lib.FL_F64
        _lib.check(_lib.lib().fl_john_apply(self.handle, C.c_void_p(g.data_ptr()), dt,
                                            C.c_void_p(out.data_ptr()), _stream()))
        return out

    def adjoint(self, u: torch.Tensor) -> torch.Tensor:
        """D^T u (FP64)."""
        self._check(u, self.out_shape, (torch.float64,), "u")
        out = torch.empty(self.in_shape, dtype=torch.float64, device=u.device)
        _lib.check(_lib.lib().fl_john_adjoint(self.handle, C.c_void_p(u.data_ptr()),
                                              C.c_void_p(out.data_ptr()), _stream()))
        return out

    def normal(self, g: torch.Tensor) -> torch.Tensor:
        """D^T D g (FP64)."""
        self._check(g, self.in_shape, (torch.float32, torch.float64), "g")
        work = torch.empty(self.out_shape, dtype=torch.float64, device=g.device)
        out = torch.empty(self.in_shape, dtype=torch.float64, device=g.device)
        dt = _lib.FL_F32 if g.dtype == torch.float32 else _lib.FL_F64
        _lib.check(_lib.lib().fl_john_normal(self.handle, C.c_void_p(g.data_ptr()), dt,
                                             C.c_void_p(work.data_ptr()), C.c_void_p(out.data_ptr()),
                                             _stream()))
        return out

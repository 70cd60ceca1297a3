"""Device-resident projector plans over the C ABI.

``Projector(para)`` owns an ``fl_plan`` built from a fan or cone ``para``
(framelab_amd.geometry) and applies the forward projector A and its adjoints
to torch tensors that live on the GPU.  Layouts follow the reference (MATLAB
column-major, x fastest):
  volume   (nt, nz, ny, nx)  or (nz, ny, nx) / (ny, nx)  -- X[it][iz][iy][ix]
  sinogram (nv, nb, na)      or (nv, nd)                 -- y[iv][ib][ia]
Storage dtypes float32 / bfloat16 / float16 are accepted for both; arithmetic
is FP32 inside the kernels.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .geometry import ConePara, FanPara

_DT = {torch.float32: _lib.FL_F32, torch.bfloat16: _lib.FL_BF16, torch.float16: _lib.FL_F16}

ADJ_EXACT = _lib.FL_ADJ_EXACT
ADJ_MATCHED = _lib.FL_ADJ_MATCHED
ADJ_FAN_FB = _lib.FL_ADJ_FAN_FB


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def geometry_struct(p, keep: list) -> _lib.FlGeometry:
    """Fill an fl_geometry from a FanPara / ConePara (arrays kept alive in `keep`)."""
    def f32(a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        keep.append(a)
        return _fptr(a)

    def i32(a):
        a = np.ascontiguousarray(a, dtype=np.int32)
        keep.append(a)
        return _iptr(a)

    g = _lib.FlGeometry()
    g.nx, g.ny, g.nt = p.nx, p.ny, p.nt
    g.SO, g.OD, g.scale = float(p.SO), float(p.OD), float(p.scale)
    g.dy_det, g.y_os = float(p.dy_det), float(p.y_os)
    g.sd_phi, g.y_det = f32(p.sd_phi), f32(p.y_det)
    g.cos_phi, g.sin_phi = f32(p.cos_phi), f32(p.sin_phi)
    g.id_X, g.id_Y, g.Nv = i32(p.id_X), i32(p.id_Y), i32(p.Nv)
    g.tmp_size = int(p.tmp_size)
    if isinstance(p, ConePara):
        g.nz, g.nv, g.na, g.nb = p.nz, p.nv, p.na, p.nb
        g.dz, g.dz_det = float(p.dz), float(p.dz_det)
        g.sd_z, g.z_det = f32(p.sd_z), f32(p.z_det)
    else:
        g.nz, g.nv, g.na, g.nb = 0, p.nv, p.nd, 1
        g.dz, g.dz_det = 1.0, 1.0
        g.cos_det, g.sin_det = f32(p.cos_det), f32(p.sin_det)
        if getattr(p, "cos_det2", None) is not None:
            g.cos_det2, g.sin_det2 = f32(p.cos_det2), f32(p.sin_det2)
    return g


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class Projector:
    """A (forward) and A^T (adjoint) for one scanner geometry on the current GPU."""

    def __init__(self, para):
        self.para = para
        self.cone = isinstance(para, ConePara)
        keep: list = []
        g = geometry_struct(para, keep)
        h = C.c_void_p()
        _lib.check(_lib.lib().fl_plan_create(C.byref(h), C.byref(g)))
        self._h = h
        nvox, nt, ndet, nv = C.c_longlong(), C.c_int(), C.c_longlong(), C.c_int()
        _lib.check(_lib.lib().fl_plan_sizes(h, C.byref(nvox), C.byref(nt), C.byref(ndet), C.byref(nv)))
        self.nvox, self.nt, self.ndet, self.nv = nvox.value, nt.value, ndet.value, nv.value
        if self.cone:
            self.vol_shape = (para.nz, para.ny, para.nx)
            self.sino_shape = (para.nv, para.nb, para.na)
        else:
            self.vol_shape = (para.ny, para.nx)
            self.sino_shape = (para.nv, para.nd)
        if self.nt > 1:
            self.vol_shape = (self.nt,) + self.vol_shape

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().fl_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- helpers -------------------------------------------------------------
    @staticmethod
    def _check_tensor(t: torch.Tensor, n: int, what: str):
        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise TypeError(f"{what} must be a CUDA (HIP) tensor")
        if t.dtype not in _DT:
            raise TypeError(f"{what}: unsupported dtype {t.dtype}")
        if not t.is_contiguous():
            raise ValueError(f"{what} must be contiguous")
        if t.numel() != n:
            raise ValueError(f"{what} has {t.numel()} elements, expected {n}")

    # -- operators -----------------------------------------------------------
    def forward(self, x: torch.Tensor, out: torch.Tensor | None = None, dtype=None,
                method: str = "siddon") -> torch.Tensor:
        """y = A x (Siddon; method='gao' selects the fan column walk, 'fb' the fan
        finite-width-beam strip areas, FP32 only)."""
        self._check_tensor(x, self.nvox * self.nt, "x")
        if out is None:
            out = torch.empty(self.sino_shape, dtype=dtype or x.dtype, device=x.device)
        self._check_tensor(out, self.ndet * self.nv, "out")
        if method == "fb":
            if x.dtype != torch.float32 or out.dtype != torch.float32:
                raise ValueError("the finite-width beam is FP32")
            _lib.check(_lib.lib().fl_forward_fan_fb(self._h, C.c_void_p(x.data_ptr()),
                                                    C.c_void_p(out.data_ptr()), _stream()))
            return out
        if method not in ("siddon", "gao"):
            raise ValueError(f"unknown forward method {method!r}")
        fn = _lib.lib().fl_forward if method == "siddon" else _lib.lib().fl_forward_gao_fan
        _lib.check(fn(self._h, C.c_void_p(x.data_ptr()), _DT[x.dtype], C.c_void_p(out.data_ptr()),
                      _DT[out.dtype], _stream()))
        return out

    def adjoint(self, y: torch.Tensor, out: torch.Tensor | None = None, dtype=None,
                kind: int = ADJ_EXACT) -> torch.Tensor:
        """x = A^T y; kind ADJ_EXACT (reference chords), ADJ_MATCHED (Siddon transpose) or
        ADJ_FAN_FB (fan finite-width-beam strip areas, FP32)."""
        self._check_tensor(y, self.ndet * self.nv, "y")
        if out is None:
            out = torch.empty(self.vol_shape, dtype=dtype or y.dtype, device=y.device)
        self._check_tensor(out, self.nvox * self.nt, "out")
        _lib.check(_lib.lib().fl_adjoint(self._h, C.c_void_p(y.data_ptr()), _DT[y.dtype],
                                         C.c_void_p(out.data_ptr()), _DT[out.dtype], int(kind),
                                         _stream()))
        return out

    def forward_f64(self, x: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """y = A x in FP64 (the Siddon traversal rule in double): the accuracy reference."""
        if x.dtype != torch.float64 or not x.is_cuda or not x.is_contiguous() or x.numel() != self.nvox * self.nt:
            raise ValueError("x must be a contiguous float64 CUDA tensor of the volume's size")
        if out is None:
            out = torch.empty(self.sino_shape, dtype=torch.float64, device=x.device)
        if out.dtype != torch.float64 or not out.is_contiguous() or out.numel() != self.ndet * self.nv:
            raise ValueError("out must be a contiguous float64 tensor of the sinogram's size")
        _lib.check(_lib.lib().fl_forward_f64(self._h, C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()),
                                             _stream()))
        return out

    def adjoint_f64(self, y: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """x = A^T y in FP64 (voxel-driven transpose of forward_f64)."""
        if y.dtype != torch.float64 or not y.is_cuda or not y.is_contiguous() or y.numel() != self.ndet * self.nv:
            raise ValueError("y must be a contiguous float64 CUDA tensor of the sinogram's size")
        if out is None:
            out = torch.empty(self.vol_shape, dtype=torch.float64, device=y.device)
        if out.dtype != torch.float64 or not out.is_contiguous() or out.numel() != self.nvox * self.nt:
            raise ValueError("out must be a contiguous float64 tensor of the volume's size")
        _lib.check(_lib.lib().fl_adjoint_f64(self._h, C.c_void_p(y.data_ptr()), C.c_void_p(out.data_ptr()),
                                             _stream()))
        return out

    def forward_norm(self, y: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """yNorm = y / (|d| * scale) per ray (cone, FP32): the reference gateway's second output."""
        self._check_tensor(y, self.ndet * self.nv, "y")
        if y.dtype != torch.float32:
            raise TypeError(f"forward_norm reads FP32 sinograms, got {y.dtype}")
        if out is None:
            out = torch.empty_like(y)
        self._check_tensor(out, self.ndet * self.nv, "out")
        if out.dtype != torch.float32:
            raise TypeError(f"forward_norm writes FP32, got out of {out.dtype}")
        _lib.check(_lib.lib().fl_forward_norm(self._h, C.c_void_p(y.data_ptr()), C.c_void_p(out.data_ptr()),
                                              _stream()))
        return out

    def autotune(self):
        """Time the cone-forward ray-tile shapes on this geometry and keep the fastest.
        Returns (tile index, milliseconds of one forward with it)."""
        t, ms = C.c_int(0), C.c_float(0.0)
        _lib.check(_lib.lib().fl_plan_autotune(self._h, C.byref(t), C.byref(ms)))
        return t.value, ms.value

    def set_tile(self, tile: int):
        """Select a cone-forward ray-tile shape by index (0..5)."""
        _lib.check(_lib.lib().fl_plan_set_tile(self._h, int(tile)))

    def trace(self, rays, cap: int | None = None):
        """Traversal log (voxel index, length) of (view, ia, ib) rays, for parity tests."""
        rays = np.ascontiguousarray(np.asarray(rays, dtype=np.int32).reshape(-1, 3))
        n = rays.shape[0]
        p = self.para
        if cap is None:
            cap = p.nx + p.ny + (p.nz if self.cone else 0) + 3
        J = np.zeros(n * cap, np.int32)
        S = np.zeros(n * cap, np.float32)
        cnt = np.zeros(n, np.int32)
        _lib.check(_lib.lib().fl_siddon_trace(self._h, n, _iptr(rays), cap, _iptr(J), _fptr(S), _iptr(cnt)))
        return [(J[k * cap:k * cap + min(cnt[k], cap)].copy(), S[k * cap:k * cap + min(cnt[k], cap)].copy())
                for k in range(n)]

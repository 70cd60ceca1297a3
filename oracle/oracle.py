"""CPU oracle for the CT-operator hot path.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg, and only as the reference results the
GPU path is compared with: nothing in the product package (framelab_amd/)
imports it.

Contents
  * ctypes bindings to liboracle.so (ct_oracle.c, this repository's C
    restatement of the reference projectors) and, when it was built, to
    _ref/libframelab_ref.so (the reference's own CPU kernels compiled from
    its sources by oracle/Makefile).
  * numpy FP64 restatements of the framelet transforms, soft threshold, CGLS
    and split Bregman (matlab/classes/+Transforms, +Optimizers, +DataTypes).

Parity pinning per function is listed in each docstring; functions with no
independent check say "parity unpinned".  Parity unpinned against the
reference (no reference values exist for them: MATLAB is absent and no
reference test holds any): the framelet band values (fradec*/frarec*; pinned
only by scipy correlation, W^T W = I and the tight-frame identities), the
CGLS iterates (cgls; pinned by the dense normal-equation solution) and split
Bregman (split_bregman; pinned by objective decrease and by convergence to
the minimiser of its objective, which the tests compute independently from
the dual box-constrained least-squares problem).
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libframelab_ref.so")

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)
_f = C.c_float
_i = C.c_int
_d = C.c_double


def build() -> None:
    """Compile liboracle.so and, where the reference checkout exists, _ref."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load(path):
    if not os.path.exists(path):
        return None
    return C.CDLL(path)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _lib.orc_merge_union.restype = _i
        _lib.orc_merge_union.argtypes = [_i, _fp, _i, _fp, _fp]
        _lib.orc_trig_f32.argtypes = [_i, _fp, _fp, _fp]
        _lib.orc_siddon_fan_f32.argtypes = [_fp, _fp, _f, _f, _f, _i, _i, _i, _fp, _fp, _i, _fp, _ip]
        _lib.orc_siddon_cone_f32.argtypes = [_fp, _fp, _f, _f, _f, _i, _i, _i, _i, _fp, _fp, _fp,
                                             _i, _i, _fp, _fp, _ip, _f]
        _lib.orc_siddon_trace_f32.restype = _i
        _lib.orc_siddon_trace_f32.argtypes = [_i, _ip, _f, _f, _i, _i, _i, _fp, _fp, _fp, _fp, _fp,
                                              _f, _i, _ip, _fp, _ip]
        _lib.orc_siddon_f64.argtypes = [_dp, _dp, _d, _d, _d, _i, _i, _i, _i, _dp, _dp, _dp, _i, _i,
                                        _dp, _dp, _d]
        _lib.orc_siddon_adj_f64.argtypes = _lib.orc_siddon_f64.argtypes
        _lib.orc_find_l.restype = _f
        _lib.orc_find_l.argtypes = [_f] * 5
        _lib.orc_find_l2.restype = _f
        _lib.orc_find_l2.argtypes = [_f] * 8
        _lib.orc_find_l_3d.restype = _f
        _lib.orc_find_l_3d.argtypes = [_f] * 12
        _lib.orc_atx_fan_f32.argtypes = [_fp, _fp, _f, _f, _f, _i, _i, _i, _f, _f, _i, _ip, _ip,
                                         _i, _fp, _fp, _fp, _fp]
        _lib.orc_atx_cone_f32.argtypes = [_fp, _fp, _fp, _fp, _fp, _fp, _fp, _ip, _ip, _f, _f, _f,
                                          _f, _f, _f, _f, _i, _i, _i, _i, _i, _i, _i]
        _lib.orc_chord_box_f64.restype = _d
        _lib.orc_chord_box_f64.argtypes = [_dp, _dp, _dp, _dp, _i]
        _lib.orc_set_threads.argtypes = [_i]
        _lib.orc_find_area.restype = _f
        _lib.orc_find_area.argtypes = [_f] * 8
        _lib.orc_fan_fb_fwd_f32.argtypes = [_fp, _fp, _f, _f, _f, _i, _i, _i, _fp, _i, _fp, _ip, _f]
        _lib.orc_fan_fb_adj_f32.argtypes = [_fp, _fp, _f, _f, _f, _i, _i, _i, _f, _i, _ip, _ip, _i,
                                            _fp, _fp, _fp, _fp]
    return _lib


def ref():
    """The reference's own CPU kernels, or None when oracle/_ref was not built."""
    global _ref
    if _ref is None:
        _ref = _load(REF_PATH)
        if _ref is None:
            return None
        _ref.ref_sort_alpha.restype = _i
        _ref.ref_sort_alpha.argtypes = [_i, _i, _fp, _fp, _fp]
        _ref.ref_Ax_fan_mf_cpu_siddon.argtypes = [_fp, _fp, _f, _f, _f, _i, _i, _i, _fp, _i, _fp, _ip]
        _ref.ref_Ax_cone_mf_cpu_siddon.argtypes = [_fp, _fp, _f, _f, _f, _i, _i, _i, _i, _fp, _fp,
                                                   _i, _i, _fp, _fp, _ip, _f]
        _ref.ref_find_l2.restype = _f
        _ref.ref_find_l2.argtypes = [_f] * 8
        _ref.ref_find_l_3d.restype = _f
        _ref.ref_find_l_3d.argtypes = [_f] * 12
        _ref.ref_Atx_cone_mf_cpu_new.argtypes = [_fp, _fp, _f, _f, _f, _i, _i, _i, _fp, _fp, _fp,
                                                 _f, _f, _f, _f, _i, _ip, _ip, _i, _i, _fp, _fp,
                                                 _i, _i]
        _ref.ref_Ax_fan_mf_cpu_new.argtypes = _ref.ref_Ax_fan_mf_cpu_siddon.argtypes
        _ref.ref_find_area.restype = _f
        _ref.ref_find_area.argtypes = [_f] * 8
        _ref.ref_Ax_fan_mf_cpu_new_fb.argtypes = [_fp, _fp, _f, _f, _f, _i, _i, _i, _fp, _i, _fp, _ip, _f]
        _ref.ref_Atx_fan_mf_cpu_new_fb.argtypes = [_fp, _fp, _f, _f, _f, _i, _i, _fp, _i, _fp, _f, _i,
                                                   _ip, _ip, _i, _i, _fp, _fp, _fp, _fp]
    return _ref


def has_ref() -> bool:
    return ref() is not None


def _pf(a):
    return a.ctypes.data_as(_fp)


def _pi(a):
    return a.ctypes.data_as(_ip)


def _pd(a):
    return a.ctypes.data_as(_dp)


def _c32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ci(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# --------------------------------------------------------------------------
# projectors (C restatement)
# --------------------------------------------------------------------------

def trig_f32(phi):
    """Per-view (cosf, sinf) as the reference's C++ forward computes them."""
    phi = _c32(phi)
    c = np.empty_like(phi)
    s = np.empty_like(phi)
    lib().orc_trig_f32(phi.size, _pf(phi), _pf(c), _pf(s))
    return c, s


def siddon_fan_f32(X, p):
    """2D fan Siddon (Ax_fan_mf_cpu_siddon.cpp:19-93); pinned bit-exact to _ref."""
    X = _c32(X).ravel()
    c, s = trig_f32(p.sd_phi)
    y = np.zeros(p.nv * p.nd, dtype=np.float32)
    lib().orc_siddon_fan_f32(_pf(X), _pf(y), p.SO, p.OD, p.scale, p.nx, p.ny, p.nv, _pf(c),
                             _pf(s), p.nd, _pf(_c32(p.y_det)), _pi(_ci(p.id_X)))
    return y


def siddon_cone_f32(X, p):
    """3D cone Siddon (Ax_cone_mf_cpu_siddon.cpp:19-134); pinned bit-exact to _ref."""
    X = _c32(X).ravel()
    c, s = trig_f32(p.sd_phi)
    y = np.zeros(p.nv * p.na * p.nb, dtype=np.float32)
    lib().orc_siddon_cone_f32(_pf(X), _pf(y), p.SO, p.OD, p.scale, p.nx, p.ny, p.nz, p.nv,
                              _pf(c), _pf(s), _pf(_c32(p.sd_z)), p.na, p.nb, _pf(_c32(p.y_det)),
                              _pf(_c32(p.z_det)), _pi(_ci(p.id_X)), p.dz)
    return y


def siddon_trace(p, rays, cone=True):
    """Per-ray (voxel index, length) lists of the FP32 Siddon walk.

    rays: int array (n,3) of (view, ia, ib).  Returns a list of (J, S)."""
    rays = _ci(np.asarray(rays).reshape(-1, 3))
    n = rays.shape[0]
    nz = p.nz if cone else 0
    cap = n * (p.nx + p.ny + nz + 3)
    J = np.zeros(cap, dtype=np.int32)
    S = np.zeros(cap, dtype=np.float32)
    off = np.zeros(n + 1, dtype=np.int32)
    c, s = trig_f32(p.sd_phi)
    sdz = _c32(p.sd_z) if cone else np.zeros(p.nv, np.float32)
    zdet = _c32(p.z_det) if cone else np.zeros(1, np.float32)
    dz0 = p.dz if cone else np.float32(1.0)
    rc = lib().orc_siddon_trace_f32(n, _pi(rays), p.SO, p.OD, p.nx, p.ny, nz, _pf(c), _pf(s),
                                    _pf(sdz), _pf(_c32(p.y_det)), _pf(zdet), dz0, cap, _pi(J),
                                    _pf(S), _pi(off))
    if rc != 0:
        raise RuntimeError("trace capacity exceeded")
    return [(J[off[k]:off[k + 1]].copy(), S[off[k]:off[k + 1]].copy()) for k in range(n)]


def _geom64(p, cone, trig="f64"):
    """FP64 geometry.  trig='f64': cos/sin of the angle in double; 'f32':
    the float cosf/sinf values the FP32 forward uses, promoted (same rays as
    the FP32 path, exact arithmetic along them)."""
    if trig == "f32":
        c, s = trig_f32(p.sd_phi)
        cphi, sphi = c.astype(np.float64), s.astype(np.float64)
    else:
        phi = np.asarray(p.sd_phi, dtype=np.float64)
        cphi, sphi = np.cos(phi), np.sin(phi)
    phi = np.asarray(p.sd_phi, dtype=np.float64)
    sdz = np.asarray(p.sd_z, dtype=np.float64) if cone else np.zeros(phi.size)
    ydet = np.asarray(p.y_det, dtype=np.float64)
    zdet = np.asarray(p.z_det, dtype=np.float64) if cone else np.zeros(1)
    return (cphi, sphi), sdz, ydet, zdet


def siddon_f64(X, p, cone=True, trig="f64"):
    """FP64 Siddon forward (same traversal rule in double; accuracy oracle).

    Pinned by closed forms (box/disc/ball chords) and brute-force clipping."""
    (cphi, sphi), sdz, ydet, zdet = _geom64(p, cone, trig)
    nz = p.nz if cone else 0
    nb = p.nb if cone else 1
    na = p.na if cone else p.nd
    X = np.ascontiguousarray(X, dtype=np.float64).ravel()
    y = np.zeros(cphi.size * na * nb)
    lib().orc_siddon_f64(_pd(X), _pd(y), float(p.SO), float(p.OD), float(p.scale), p.nx, p.ny,
                         nz, cphi.size, _pd(cphi), _pd(sphi), _pd(sdz), na, nb, _pd(ydet), _pd(zdet),
                         float(p.dz) if cone else 1.0)
    return y


def siddon_adj_f64(y, p, cone=True, trig="f64"):
    """FP64 exact transpose of siddon_f64 (dot-product test to ~1e-15)."""
    (cphi, sphi), sdz, ydet, zdet = _geom64(p, cone, trig)
    nz = p.nz if cone else 0
    nb = p.nb if cone else 1
    na = p.na if cone else p.nd
    y = np.ascontiguousarray(y, dtype=np.float64).ravel()
    X = np.zeros(p.nx * p.ny * max(nz, 1))
    lib().orc_siddon_adj_f64(_pd(X), _pd(y), float(p.SO), float(p.OD), float(p.scale), p.nx,
                             p.ny, nz, cphi.size, _pd(cphi), _pd(sphi), _pd(sdz), na, nb, _pd(ydet),
                             _pd(zdet), float(p.dz) if cone else 1.0)
    return X


def find_l(a, b, c, x, y):
    """find_l.cpp:18-45 (unbuildable here: it includes mex.h); pinned by the
    closed-form square chord and by equality with the 8-argument form."""
    return lib().orc_find_l(a, b, c, x, y)


def find_l2(*args):
    return lib().orc_find_l2(*args)


def find_l_3d(*args):
    return lib().orc_find_l_3d(*args)


def chord_box_f64(p0, p1, c, h):
    p0, p1, c, h = (np.ascontiguousarray(v, dtype=np.float64) for v in (p0, p1, c, h))
    return lib().orc_chord_box_f64(_pd(p0), _pd(p1), _pd(c), _pd(h), p0.size)


def atx_fan_f32(Y, p):
    """Pixel-driven exact adjoint (Atx_fan_mf_cpu_new.cpp:21-63)."""
    Y = _c32(Y).ravel()
    x = np.zeros(p.nx * p.ny * p.nt, dtype=np.float32)
    lib().orc_atx_fan_f32(_pf(x), _pf(Y), p.SO, p.OD, p.scale, p.nx, p.ny, p.nd, p.dy_det,
                          p.y_os, p.nt, _pi(_ci(p.id_Y)), _pi(_ci(p.Nv)), p.tmp_size,
                          _pf(_c32(p.cos_phi)), _pf(_c32(p.sin_phi)), _pf(_c32(p.cos_det)),
                          _pf(_c32(p.sin_det)))
    return x


def find_area(a1, b1, c1, a2, b2, c2, x, y):
    """Strip-area weight of a unit pixel (find_area.cpp:18-69)."""
    return lib().orc_find_area(a1, b1, c1, a2, b2, c2, x, y)


def fan_fb_fwd_f32(X, p):
    """Finite-width-beam fan forward (Ax_fan_mf_cpu_new_fb.cpp:19-123)."""
    X = _c32(X).ravel()
    y = np.zeros(p.nv * p.nd, dtype=np.float32)
    lib().orc_fan_fb_fwd_f32(_pf(X), _pf(y), p.SO, p.OD, p.scale, p.nx, p.ny, p.nv, _pf(_c32(p.sd_phi)),
                             p.nd, _pf(_c32(p.y_det)), _pi(_ci(p.id_X)), p.dy_det)
    return y


def fan_fb_adj_f32(Y, p):
    """Finite-width-beam fan adjoint (Atx_fan_mf_cpu_new_fb.cpp:21-66), edge tables
    p.cos_det2 / p.sin_det2 (nd+1 entries)."""
    Y = _c32(Y).ravel()
    x = np.zeros(p.nx * p.ny * p.nt, dtype=np.float32)
    lib().orc_fan_fb_adj_f32(_pf(x), _pf(Y), p.SO, p.OD, p.scale, p.nx, p.ny, p.nd, p.dy_det, p.nt,
                             _pi(_ci(p.id_Y)), _pi(_ci(p.Nv)), p.tmp_size, _pf(_c32(p.cos_phi)),
                             _pf(_c32(p.sin_phi)), _pf(_c32(p.cos_det2)), _pf(_c32(p.sin_det2)))
    return x


def ref_find_area(*args):
    return ref().ref_find_area(*args)


def ref_fan_fb_fwd(X, p):
    X = _c32(X).ravel().copy()
    y = np.zeros(p.nv * p.nd, dtype=np.float32)
    ref().ref_Ax_fan_mf_cpu_new_fb(_pf(X), _pf(y), p.SO, p.OD, p.scale, p.nx, p.ny, p.nv,
                                   _pf(_c32(p.sd_phi)), p.nd, _pf(_c32(p.y_det)), _pi(_ci(p.id_X)),
                                   p.dy_det)
    return y


def ref_fan_fb_adj(Y, p):
    Y = _c32(Y).ravel().copy()
    x = np.zeros(p.nx * p.ny * p.nt, dtype=np.float32)
    ref().ref_Atx_fan_mf_cpu_new_fb(_pf(x), _pf(Y), p.SO, p.OD, p.scale, p.nx, p.ny, _pf(_c32(p.sd_phi)),
                                    p.nd, _pf(_c32(p.y_det)), p.dy_det, p.nt, _pi(_ci(p.id_Y)),
                                    _pi(_ci(p.Nv)), p.tmp_size, p.nv, _pf(_c32(p.cos_phi)),
                                    _pf(_c32(p.sin_phi)), _pf(_c32(p.cos_det2).copy()),
                                    _pf(_c32(p.sin_det2).copy()))
    return x


def atx_cone_f32(Y, p):
    """Voxel-driven exact adjoint, reference GPU semantics (Atx_cone_mf_gpu_new.cu:41-84)."""
    Y = _c32(Y).ravel()
    x = np.zeros(p.nx * p.ny * p.nz * p.nt, dtype=np.float32)
    lib().orc_atx_cone_f32(_pf(x), _pf(Y), _pf(_c32(p.cos_phi)), _pf(_c32(p.sin_phi)),
                           _pf(_c32(p.sd_z)), _pf(_c32(p.y_det)), _pf(_c32(p.z_det)),
                           _pi(_ci(p.id_Y)), _pi(_ci(p.Nv)), p.SO, p.OD, p.scale, p.dy_det,
                           p.y_os, p.dz_det, p.dz, p.nx, p.ny, p.nz, p.nt, p.na, p.nb,
                           p.tmp_size)
    return x


# --------------------------------------------------------------------------
# the reference's own kernels (oracle/_ref)
# --------------------------------------------------------------------------

def ref_siddon_fan(X, p):
    X = _c32(X).ravel().copy()
    y = np.zeros(p.nv * p.nd, dtype=np.float32)
    ref().ref_Ax_fan_mf_cpu_siddon(_pf(X), _pf(y), p.SO, p.OD, p.scale, p.nx, p.ny, p.nv,
                                   _pf(_c32(p.sd_phi)), p.nd, _pf(_c32(p.y_det)),
                                   _pi(_ci(p.id_X)))
    return y


def ref_siddon_cone(X, p):
    X = _c32(X).ravel().copy()
    y = np.zeros(p.nv * p.na * p.nb, dtype=np.float32)
    ref().ref_Ax_cone_mf_cpu_siddon(_pf(X), _pf(y), p.SO, p.OD, p.scale, p.nx, p.ny, p.nz,
                                    p.nv, _pf(_c32(p.sd_phi)), _pf(_c32(p.sd_z)), p.na, p.nb,
                                    _pf(_c32(p.y_det)), _pf(_c32(p.z_det)), _pi(_ci(p.id_X)),
                                    p.dz)
    return y


def ref_atx_cone_cpu(Y, p):
    """Reference CPU cone adjoint; equals the GPU semantics when dz=1, sd_z=0."""
    Y = _c32(Y).ravel().copy()
    x = np.zeros(p.nx * p.ny * p.nz * p.nt, dtype=np.float32)
    ref().ref_Atx_cone_mf_cpu_new(_pf(x), _pf(Y), p.SO, p.OD, p.scale, p.nx, p.ny, p.nz,
                                  _pf(_c32(p.sd_z)), _pf(_c32(p.y_det)), _pf(_c32(p.z_det)),
                                  p.dy_det, p.y_os, p.dz_det, p.dz, p.nt, _pi(_ci(p.id_Y)),
                                  _pi(_ci(p.Nv)), p.tmp_size, p.nv, _pf(_c32(p.cos_phi)),
                                  _pf(_c32(p.sin_phi)), p.na, p.nb)
    return x


def ref_gao_fan(X, p):
    X = _c32(X).ravel().copy()
    y = np.zeros(p.nv * p.nd, dtype=np.float32)
    ref().ref_Ax_fan_mf_cpu_new(_pf(X), _pf(y), p.SO, p.OD, p.scale, p.nx, p.ny, p.nv,
                                _pf(_c32(p.sd_phi)), p.nd, _pf(_c32(p.y_det)), _pi(_ci(p.id_X)))
    return y


def ref_sort_alpha(a, b):
    a = _c32(a)
    b = _c32(b)
    pad_a = np.concatenate([a, np.full(2, np.inf, np.float32)]).astype(np.float32)
    pad_b = np.concatenate([b, np.full(2, np.inf, np.float32)]).astype(np.float32)
    out = np.zeros(a.size + b.size + 2, dtype=np.float32)
    n = ref().ref_sort_alpha(a.size, b.size, _pf(pad_a), _pf(pad_b), _pf(out))
    return out[:n]


def merge_union(a, b):
    a = _c32(a)
    b = _c32(b)
    out = np.zeros(a.size + b.size, dtype=np.float32)
    n = lib().orc_merge_union(a.size, _pf(a), b.size, _pf(b), _pf(out))
    return out[:n]


# --------------------------------------------------------------------------
# framelets (GenerateFrameletFilter.m, ConvSymAsym{2D,3D}.m, FraDec/FraRec)
# --------------------------------------------------------------------------

def framelet_filters(kind="linear"):
    """Decomposition / reconstruction taps, GenerateFrameletFilter.m:1-30."""
    s2, s6 = math.sqrt(2.0), math.sqrt(6.0)
    if kind == "haar":
        D = [np.array([0, 1, 1]) / 2, -np.array([0, 1, -1]) / 2]
        R = [np.array([1, 1, 0]) / 2, -np.array([-1, 1, 0]) / 2]
    elif kind == "linear":
        D = [np.array([1, 2, 1]) / 4, -np.array([1, 0, -1]) / 4 * s2, np.array([-1, 2, -1]) / 4]
        R = [np.array([1, 2, 1]) / 4, -np.array([-1, 0, 1]) / 4 * s2, np.array([-1, 2, -1]) / 4]
    elif kind == "cubic":
        D = [np.array([1, 4, 6, 4, 1]) / 16, -np.array([1, 2, 0, -2, -1]) / 8,
             np.array([-1, 0, 2, 0, -1]) / 16 * s6, np.array([-1, 2, 0, -2, 1]) / 8,
             -np.array([1, -4, 6, -4, 1]) / 16]
        R = [np.array([1, 4, 6, 4, 1]) / 16, -np.array([-1, -2, 0, 2, 1]) / 8,
             np.array([-1, 0, 2, 0, -1]) / 16 * s6, np.array([1, -2, 0, 2, -1]) / 8,
             -np.array([1, -4, 6, -4, 1]) / 16]
    else:
        raise ValueError(kind)
    return [d.astype(np.float64) for d in D], [r.astype(np.float64) for r in R]


def corr_circ(A, taps, axis, level):
    """imfilter(A, dilated taps, 'circular') along `axis` (ConvSymAsym3D.m:19-24):
    circular correlation, kernel centre floor((len+1)/2), dilation 2^(level-1).
    Pinned against scipy.ndimage.correlate1d(mode='wrap') in the tests."""
    step = 2 ** (level - 1)
    m = len(taps)
    c = (m - 1) // 2
    out = np.zeros_like(A, dtype=np.float64)
    for k, t in enumerate(taps):
        if t == 0:
            continue
        out += t * np.roll(A, -(k - c) * step, axis=axis)
    return out


# Arrays are indexed [z, y, x] / [y, x] (x fastest, MATLAB dim 1 = x).
AX_X3, AX_Y3, AX_Z3 = 2, 1, 0
AX_X2, AX_Y2 = 1, 0


def fradec2d(A, D, level=1):
    """FraDecMultiLevel2D.m:16-53: band(i,j) = corr_y[D_j](corr_x[D_i](A)).
    Returns a list over levels of arrays [nD, nD, ny, nx].  Band values:
    parity unpinned against the reference (pinned by scipy correlation and
    the tight-frame identities)."""
    out = []
    cur = np.asarray(A, dtype=np.float64)
    n = len(D)
    for L in range(1, level + 1):
        bands = np.zeros((n, n) + cur.shape)
        for i in range(n):
            ti = corr_circ(cur, D[i], AX_X2, L)
            for j in range(n):
                bands[i, j] = corr_circ(ti, D[j], AX_Y2, L)
        out.append(bands)
        cur = bands[0, 0]
    return out


def frarec2d(C, R):
    """FraRecMultiLevel2D.m:15-50 (Rec = sum_i corr_x[R_i](sum_j corr_y[R_j](C_ij)))."""
    C = [c.copy() for c in C]
    n = len(R)

    def one(bands, L):
        rec = np.zeros(bands.shape[2:])
        for i in range(n):
            t = np.zeros(bands.shape[2:])
            for j in range(n):
                t += corr_circ(bands[i, j], R[j], AX_Y2, L)
            rec += corr_circ(t, R[i], AX_X2, L)
        return rec

    for L in range(len(C), 1, -1):
        C[L - 2][0, 0] = one(C[L - 1], L)
    return one(C[0], 1)


def fradec3d(A, D, level=1):
    """FraDecMultiLevel3D.m:14-48: band(i,j,k) = corr_z[D_k](corr_x[D_j](corr_y[D_i](A))).
    Returns a list over levels of arrays [nD, nD, nD, nz, ny, nx]."""
    out = []
    cur = np.asarray(A, dtype=np.float64)
    n = len(D)
    for L in range(1, level + 1):
        bands = np.zeros((n, n, n) + cur.shape)
        for i in range(n):
            ti = corr_circ(cur, D[i], AX_Y3, L)
            for j in range(n):
                tj = corr_circ(ti, D[j], AX_X3, L)
                for k in range(n):
                    bands[i, j, k] = corr_circ(tj, D[k], AX_Z3, L)
        out.append(bands)
        cur = bands[0, 0, 0]
    return out


def frarec3d(C, R):
    """FraRecMultiLevel3D.m:14-49."""
    C = [c.copy() for c in C]
    n = len(R)

    def one(bands, L):
        shape = bands.shape[3:]
        rec = np.zeros(shape)
        for i in range(n):
            t1 = np.zeros(shape)
            for j in range(n):
                t2 = np.zeros(shape)
                for k in range(n):
                    t2 += corr_circ(bands[i, j, k], R[k], AX_Z3, L)
                t1 += corr_circ(t2, R[j], AX_X3, L)
            rec += corr_circ(t1, R[i], AX_Y3, L)
        return rec

    for L in range(len(C), 1, -1):
        C[L - 2][(0,) * 3] = one(C[L - 1], L)
    return one(C[0], 1)


def soft_threshold(C, tau, dim):
    """shrinkFramelet{2D,3D} (softThreshold.m:15-80): every band of every
    level except the low-pass band (1,1[,1]) -> sign(a)*max(|a|-tau,0)."""
    out = []
    for bands in C:
        b = np.sign(bands) * np.maximum(np.abs(bands) - tau, 0.0)
        lp = (0,) * dim
        b[lp] = bands[lp]
        out.append(b)
    return out


# --------------------------------------------------------------------------
# solvers (solveCGLS.m, solveSplitBregman.m) -- FP64 numpy
# --------------------------------------------------------------------------

def cgls(A, At, b, u0, lam, iters, tol=0.0, log=None):
    """CG on (At A + lam I) u = b, solveCGLS.m:1-31.  A, At are callables.
    Iterates: parity unpinned against the reference (pinned by the dense
    solution of the regularised normal equations)."""
    Aloc = lambda v: At(A(v)) + lam * v
    r = Aloc(u0) - b
    p = -r
    u = u0.copy()
    gold = float(np.dot(r.ravel(), r.ravel()))
    for _ in range(iters):
        q = Aloc(p)
        pq = float(np.dot(p.ravel(), q.ravel()))
        if pq == 0.0:          # exact convergence (p = 0): nothing left to do
            break
        alpha = gold / pq
        u = u + alpha * p
        r = r + alpha * q
        gnew = float(np.dot(r.ravel(), r.ravel()))
        if log is not None:
            log.append(gnew)
        beta = gnew / gold
        p = -r + beta * p
        gold = gnew
        if gold < tol:
            break
    return u


def split_bregman(A, At, W, Wt, lam, mu, f, u0, outer, cg_iters, cg_tol=0.0, shrink=None):
    """solveSplitBregman.m:16-41 (class form v <- v + (Wu - alpha)).
    W maps a volume to a coefficient array, Wt back; shrink(c, tau) applies
    the soft threshold to everything but the low-pass band.
    Parity unpinned against the reference (its MATLAB cannot run); pinned by
    the objective-decrease check and by convergence to the objective's
    minimiser, computed independently (dual bounded least squares) in the
    tests, at one and two frame levels."""
    alpha = W(u0)
    v = alpha.copy()
    Atf = At(f)
    b = Atf + mu * Wt(alpha - v)
    u = u0
    for _ in range(outer):
        u = cgls(A, At, b, u, mu, cg_iters, cg_tol)
        Wu = W(u)
        alpha = shrink(v + Wu, lam / mu)
        v = v + (Wu - alpha)
        b = Atf + mu * Wt(alpha - v)
    return u


# --------------------------------------------------------------------------
# John's equation (classes/+DataTypes/@CTData/applyJohn.m:38-106): the
# sparse operator assembled as the reference assembles it -- 1D difference
# matrices (first differences (g[i+2]-g[i])/(2 step), second differences
# (g[i]-2g[i+1]+g[i+2])/step^2, a forward difference across two helices),
# Kronecker products in (helix, view, row, column) order, trims to the
# interior, and the coefficient diagonals a, b, a^2 + R^2 -- in FP64.
def john_matrix(na, nb, nv, nz, R, SO, h, dzeta, dtheta, da, db, a, b):
    import scipy.sparse as sp

    def eye(n):
        return sp.identity(n, format="csr", dtype=np.float64)

    def d1(n, step):
        i = np.arange(n - 2)
        return sp.csr_matrix((np.concatenate([-np.ones(n - 2), np.ones(n - 2)]),
                              (np.concatenate([i, i]), np.concatenate([i, i + 2]))), shape=(n - 2, n)) / (2 * step)

    def d2(n, step):
        i = np.arange(n - 2)
        return sp.csr_matrix((np.concatenate([np.ones(n - 2), -2 * np.ones(n - 2), np.ones(n - 2)]),
                              (np.concatenate([i, i, i]), np.concatenate([i, i + 1, i + 2]))),
                             shape=(n - 2, n)) / step ** 2

    def K(z, v, bb, aa):
        return sp.kron(sp.kron(z, v), sp.kron(bb, aa), format="csr")

    def trim(n):
        return eye(n)[1:-1, :]

    Da1, Db1, Dbb1, Dt1 = d1(na, da), d1(nb, db), d2(nb, db), d1(nv, dtheta)
    if nz > 2:
        Dz1, zT, nzd = d1(nz, dzeta), trim(nz), nz - 2
    else:
        Dz1 = sp.csr_matrix(([-1.0, 1.0], ([0, 0], [0, 1])), shape=(1, nz)) / dzeta
        zT, nzd = eye(nz)[0:1, :], nz - 1
    Dtb = K(eye(nz), Dt1, Db1, eye(na))
    Daz = K(Dz1, eye(nv), eye(nb), Da1)
    Dbz = K(Dz1, eye(nv), Db1, eye(na))
    Db = K(eye(nz), eye(nv), Db1, eye(na))
    Dbb = K(eye(nz), eye(nv), Dbb1, eye(na))
    Dab = K(eye(nz), eye(nv), Db1, Da1)
    aT, bT, tT = trim(na), trim(nb), trim(nv)
    azTrim = K(zT, eye(nv - 2), eye(nb - 2), aT)
    btTrim = K(eye(nzd), tT, bT, eye(na - 2))
    atTrim = K(eye(nzd), tT, eye(nb - 2), aT)
    atzTrim = K(zT, tT, eye(nb - 2), aT)
    tzTrim = K(zT, tT, eye(nb - 2), eye(na - 2))
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    A = np.broadcast_to(a[None, None, None, :], (nz, nv, nb - 2, na))
    B = np.broadcast_to(b[1:-1][None, None, :, None], (nz, nv, nb - 2, na))
    A2 = np.broadcast_to((a[1:-1] ** 2 + R ** 2)[None, None, None, :], (nz, nv, nb - 2, na - 2))
    a1, b1, a2 = sp.diags(A.ravel()), sp.diags(B.ravel()), sp.diags(A2.ravel())
    return (R * (azTrim @ Dtb) - R * SO * (btTrim @ Daz) - R * h * (atTrim @ Dbz)
            + 2 * (atzTrim @ a1 @ Db) + atzTrim @ a1 @ b1 @ Dbb + tzTrim @ a2 @ Dab).tocsr()


def john_apply(g, M):
    """D g for g of shape (nz, nv, nb, na) (a fastest)."""
    return M @ np.asarray(g, np.float64).ravel()


def john_normal(g, M):
    """D^T D g (applyJohnNormal.m:155)."""
    return M.T @ (M @ np.asarray(g, np.float64).ravel())


def set_threads(n: int) -> None:
    """OpenMP threads of the restated kernels."""
    lib().orc_set_threads(int(n))

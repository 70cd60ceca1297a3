"""ctypes binding of libframelab_hip.so (the C ABI declared in include/framelab_hip.h).

The library is built in-tree by ``framelab_amd.build.build()``.  Loading fails
loudly when it is missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libframelab_hip.so")

FL_OK = 0
FL_ERR_INVALID = -1
FL_ERR_HIP = -2
FL_ERR_ALLOC = -3
FL_ERR_UNSUPPORTED = -4

FL_F32, FL_BF16, FL_F16, FL_F64 = 0, 1, 2, 3
FL_ADJ_EXACT, FL_ADJ_MATCHED, FL_ADJ_FAN_FB = 0, 1, 2

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class FlGeometry(C.Structure):
    """Mirror of ``fl_geometry`` (include/framelab_hip.h)."""
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("nt", C.c_int),
        ("nv", C.c_int), ("na", C.c_int), ("nb", C.c_int),
        ("SO", C.c_float), ("OD", C.c_float), ("scale", C.c_float), ("dz", C.c_float),
        ("dy_det", C.c_float), ("dz_det", C.c_float), ("y_os", C.c_float),
        ("sd_phi", _fp), ("sd_z", _fp), ("y_det", _fp), ("z_det", _fp),
        ("cos_phi", _fp), ("sin_phi", _fp), ("cos_det", _fp), ("sin_det", _fp),
        ("id_X", _ip), ("id_Y", _ip), ("Nv", _ip),
        ("tmp_size", C.c_int),
        ("slot_view", _ip),
        ("cos_det2", _fp), ("sin_det2", _fp),
    ]


class FlJohnGeometry(C.Structure):
    """Mirror of ``fl_john_geometry`` (include/framelab_hip.h)."""
    _fields_ = [
        ("na", C.c_int), ("nb", C.c_int), ("nv", C.c_int), ("nz", C.c_int),
        ("R", C.c_double), ("SO", C.c_double), ("h", C.c_double), ("dzeta", C.c_double),
        ("dtheta", C.c_double), ("da", C.c_double), ("db", C.c_double),
        ("a", C.POINTER(C.c_double)), ("b", C.POINTER(C.c_double)),
    ]


class FrameLabError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"framelab_hip error {code}: {msg}")
        self.code = code


_lib = None

# every C entry point declared in include/framelab_hip.h, with its signature
SIGNATURES = {
    "fl_get_last_error": (C.c_int, []),
    "fl_last_error_message": (C.c_char_p, []),
    "fl_error_string": (C.c_char_p, [C.c_int]),
    "fl_version": (C.c_int, []),
    "Ax_fan_mf_gpu_siddon": (None, [_fp, _fp, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int,
                                    C.c_int, _fp, C.c_int, _fp, _ip, C.c_int]),
    "Ax_fan_mf_gpu_new": (None, [_fp, _fp, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int,
                                 C.c_int, _fp, C.c_int, _fp, _ip, C.c_int]),
    "Atx_fan_mf_gpu_new": (None, [_fp, _fp, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int, _fp,
                                  C.c_int, _fp, C.c_float, C.c_float, C.c_int, _ip, _ip, C.c_int,
                                  C.c_int, _fp, _fp, _fp, _fp]),
    "Ax_cone_mf_gpu_siddon": (None, [_fp, _fp, _fp, _fp, _fp, _fp, _ip, _ip, C.c_float, C.c_float,
                                     C.c_float, C.c_float, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "Ax_cone_mf_gpu_new": (None, [_fp, _fp, _fp, _fp, _fp, _fp, _fp, _ip, _ip, C.c_float,
                                  C.c_float, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int,
                                  C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "Atx_cone_mf_gpu_new": (None, [_fp, _fp, _fp, _fp, _fp, _fp, _fp, _ip, _ip, C.c_float,
                                   C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                   C.c_float, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_int, C.c_int, C.c_int, C.c_int]),
    "fl_plan_create": (C.c_int, [C.POINTER(_vp), C.POINTER(FlGeometry)]),
    "fl_plan_destroy": (C.c_int, [_vp]),
    "fl_plan_sizes": (C.c_int, [_vp, C.POINTER(C.c_longlong), C.POINTER(C.c_int),
                                C.POINTER(C.c_longlong), C.POINTER(C.c_int)]),
    "fl_forward": (C.c_int, [_vp, _vp, C.c_int, _vp, C.c_int, _vp]),
    "fl_forward_gao_fan": (C.c_int, [_vp, _vp, C.c_int, _vp, C.c_int, _vp]),
    "fl_forward_fan_fb": (C.c_int, [_vp, _vp, _vp, _vp]),
    "Ax_fan_mf_gpu_new_fb": (None, [_fp, _fp, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int,
                                    C.c_int, _fp, C.c_int, _fp, _ip, C.c_int, C.c_float]),
    "Atx_fan_mf_gpu_new_fb": (None, [_fp, _fp, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int, _fp,
                                     C.c_int, _fp, C.c_float, C.c_int, _ip, _ip, C.c_int, C.c_int,
                                     _fp, _fp, _fp, _fp]),
    "fl_adjoint": (C.c_int, [_vp, _vp, C.c_int, _vp, C.c_int, C.c_int, _vp]),
    "fl_siddon_trace": (C.c_int, [_vp, C.c_int, _ip, C.c_int, _ip, _fp, _ip]),
    "fl_forward_f64": (C.c_int, [_vp, _vp, _vp, _vp]),
    "fl_adjoint_f64": (C.c_int, [_vp, _vp, _vp, _vp]),
    "fl_split_bregman_f64": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_double, C.c_double, C.c_int, C.c_int,
                                       C.POINTER(C.c_double), _vp, _vp]),
    "fl_cgls_f64": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_int, C.c_double, C.POINTER(C.c_double),
                              C.POINTER(C.c_int), _vp]),
    "fl_framelet_num_bands": (C.c_int, [C.c_int, C.c_int]),
    "fl_framelet_forward": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _vp, C.c_int,
                                      _vp, C.c_int, _vp]),
    "fl_framelet_adjoint": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _vp, C.c_int,
                                      _vp, C.c_int, _vp, _vp]),
    "fl_cgls": (C.c_int, [_vp, _vp, _vp, C.c_float, C.c_int, C.c_double, C.c_int,
                          C.POINTER(C.c_double), C.POINTER(C.c_int), _vp]),
    "fl_split_bregman": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp, C.c_float, C.c_float, C.c_int, C.c_int,
                                   C.c_double, C.c_int, C.POINTER(C.c_double), _vp, _vp]),
    "fl_vec_cgls_init": (C.c_int, [C.c_longlong, _vp, _vp, _vp, C.c_float, _vp, _vp, C.POINTER(C.c_double), _vp]),
    "fl_vec_cgls_pq": (C.c_int, [C.c_longlong, _vp, _vp, C.c_float, C.POINTER(C.c_double), _vp]),
    "fl_vec_cgls_update": (C.c_int, [C.c_longlong, _vp, _vp, _vp, _vp, C.c_double, C.POINTER(C.c_double), _vp]),
    "fl_vec_cgls_dir": (C.c_int, [C.c_longlong, _vp, _vp, C.c_double, _vp]),
    "fl_framelet_shrink": (C.c_int, [C.c_int, C.c_int, C.c_longlong, _vp, C.c_int, C.c_float, _vp]),
    "fl_grad_data": (C.c_int, [_vp, _vp, _vp, _vp, C.POINTER(C.c_double), C.c_int, _vp]),
    "fl_grad_framelet": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _vp, C.c_float,
                                   C.c_float, C.c_int, _vp, _vp, C.POINTER(C.c_double), _vp]),
    "fl_grad_framelet_noalloc": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _vp, C.c_float,
                                           C.c_float, C.c_int, _vp, _vp, _vp, _vp]),
    "fl_plan_autotune": (C.c_int, [_vp, _ip, _fp]),
    "fl_plan_set_tile": (C.c_int, [_vp, C.c_int]),
    "fl_forward_norm": (C.c_int, [_vp, _vp, _vp, _vp]),
    "fl_john_create": (C.c_int, [C.POINTER(_vp), C.POINTER(FlJohnGeometry)]),
    "fl_john_destroy": (C.c_int, [_vp]),
    "fl_john_sizes": (C.c_int, [_vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "fl_john_apply": (C.c_int, [_vp, _vp, C.c_int, _vp, _vp]),
    "fl_john_adjoint": (C.c_int, [_vp, _vp, _vp, _vp]),
    "fl_john_normal": (C.c_int, [_vp, _vp, C.c_int, _vp, _vp, _vp]),
    "fl_batch_create": (C.c_int, [C.POINTER(_vp), C.c_int, C.POINTER(FlGeometry)]),
    "fl_batch_destroy": (C.c_int, [_vp]),
    "fl_batch_layout": (C.c_int, [_vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "fl_batch_forward": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp]),
    "fl_batch_adjoint": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp]),
    "fl_batch_cgls": (C.c_int, [_vp, _vp, _vp, C.c_float, C.c_int, C.c_double, C.c_int,
                                C.POINTER(C.c_double), C.POINTER(C.c_int), _vp]),
    "fl_batch_split_bregman": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_float, C.c_float, C.c_int, C.c_int,
                                         C.c_double, C.c_int, C.POINTER(C.c_double), _vp]),
    "fl_batch_autotune": (C.c_int, [_vp, _ip, _ip, _ip]),
}

FL_FRAME_HAAR, FL_FRAME_LINEAR, FL_FRAME_CUBIC = 0, 1, 3
FL_SOLVE_GRAPH, FL_SOLVE_MATCHED, FL_SOLVE_LOWP_BF16, FL_SOLVE_LOWP_F16 = 1, 2, 4, 8
FL_GRAD_EXACT_ADJOINT = 1
FL_NOALLOC_PARTIALS = 1 << 21   # doubles in fl_grad_framelet_noalloc's static scratch


def lib():
    """Load libframelab_hip.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "(the HIP extension is required; there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            if not hasattr(L, name):
                continue
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != FL_OK:
        msg = lib().fl_last_error_message().decode(errors="replace")
        raise FrameLabError(rc, msg)


def check_last() -> None:
    """Raise if the last drop-in (void) call on this thread failed."""
    rc = lib().fl_get_last_error()
    check(rc)

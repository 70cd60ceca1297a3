"""ctypes loader of libfl_diag.so (tests/diag/diag.hip): the VALU issue-peak
calibration kernels and the exhaustive chord square-root check.  Test and
profiling infrastructure; the product library does not contain them."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfl_diag.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.fl_check_chord_sqrt.restype = C.c_int
        L.fl_check_chord_sqrt.argtypes = [C.c_void_p, C.c_void_p]
        L.fl_calibrate_valu.restype = C.c_int
        L.fl_calibrate_valu.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]
        _lib = L
    return _lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed with status {rc}")

"""Shared test geometry and data generators."""
import numpy as np

from framelab_amd import geometry as G


def small_fan(nd=128, nv=90, N=64):
    return G.fan_para(nd, nv, N=(N, N), SO=4.0, OD=4.0, Ly=G.full_fan_width())


def small_cone(na=64, nb=48, nv=24, N=32, **kw):
    return G.cone_para(na, nb, nv, N=(N, N, N), SO=4.0, OD=4.0, Ly=G.full_fan_width(),
                       Lz=G.full_cone_height(), **kw)


def bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def first_mismatch(a, b):
    a = np.asarray(a, np.float32).ravel()
    b = np.asarray(b, np.float32).ravel()
    bad = np.nonzero(a.view(np.uint32) != b.view(np.uint32))[0]
    if bad.size == 0:
        return None
    k = bad[0]
    return f"{bad.size} mismatches, first at {k}: {a[k]!r} vs {b[k]!r}"

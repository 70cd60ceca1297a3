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


def dot_gap(y, Y, X, Xt):
    """<y, Y> - <X, Xt> for y = A X and Xt = A^T Y, over the scale of the
    terms, max(sum |y||Y|, sum |X||Xt|) (FP64): the rounding of the FP32
    operators is relative to that scale, not to the dot product itself,
    which random signs can cancel to a small fraction of it."""
    y, Y, X, Xt = (np.asarray(v, np.float64).ravel() for v in (y, Y, X, Xt))
    lhs, rhs = float(np.dot(y, Y)), float(np.dot(X, Xt))
    scale = max(float(np.dot(np.abs(y), np.abs(Y))), float(np.dot(np.abs(X), np.abs(Xt))), 1e-300)
    return abs(lhs - rhs) / scale


def first_mismatch(a, b):
    a = np.asarray(a, np.float32).ravel()
    b = np.asarray(b, np.float32).ravel()
    bad = np.nonzero(a.view(np.uint32) != b.view(np.uint32))[0]
    if bad.size == 0:
        return None
    k = bad[0]
    return f"{bad.size} mismatches, first at {k}: {a[k]!r} vs {b[k]!r}"


def clip_halfplane(poly, a, b, c, sign):
    """Sutherland-Hodgman clip (FP64) of a convex polygon [(x, y), ...] to
    sign*(a*x + b*y - c) >= 0."""
    out = []
    n = len(poly)
    for i in range(n):
        p, q = poly[i], poly[(i + 1) % n]
        fp = sign * (a * p[0] + b * p[1] - c)
        fq = sign * (a * q[0] + b * q[1] - c)
        if fp >= 0:
            out.append(p)
        if (fp >= 0) != (fq >= 0):
            t = fp / (fp - fq)
            out.append((p[0] + t * (q[0] - p[0]), p[1] + t * (q[1] - p[1])))
    return out


def poly_area(poly):
    if len(poly) < 3:
        return 0.0
    s = 0.0
    for i in range(len(poly)):
        x0, y0 = poly[i]
        x1, y1 = poly[(i + 1) % len(poly)]
        s += x0 * y1 - x1 * y0
    return abs(s) / 2


def strip_area(poly, l1, l2):
    """Area (FP64) of `poly` between two lines a*x + b*y = c, i.e. where the
    two line functions have opposite signs."""
    tot = 0.0
    for s in (1, -1):
        q = clip_halfplane(poly, *l1, s)
        q = clip_halfplane(q, *l2, -s)
        tot += poly_area(q)
    return tot

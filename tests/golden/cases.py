"""Seeded inputs of the golden cases (shared by make_golden.py and the tests)."""
import numpy as np

from framelab_amd import geometry as G


def merge_cases():
    rng = np.random.default_rng(100)
    cases = []
    for k in range(20):
        na, nb = int(rng.integers(1, 40)), int(rng.integers(1, 40))
        a = np.sort(rng.random(na).astype(np.float32))
        b = np.sort(rng.random(nb).astype(np.float32))
        # plant exact cross-list ties
        for _ in range(int(rng.integers(0, 5))):
            if na and nb:
                b[int(rng.integers(0, nb))] = a[int(rng.integers(0, na))]
        b = np.sort(b)
        cases.append((np.unique(a), np.unique(b)))
    # first elements tied, last elements tied, fully interleaved
    cases.append((np.array([0.1, 0.2, 0.3], np.float32), np.array([0.1, 0.25, 0.3], np.float32)))
    cases.append((np.array([0.0, 0.5, 1.0], np.float32), np.array([0.5], np.float32)))
    return cases


def fan_cases():
    rng = np.random.default_rng(200)
    out = []
    for k, (N, nd, nv) in enumerate(((16, 32, 8), (32, 64, 12), (24, 40, 9))):
        sd_phi = None
        if k == 0:   # axis-aligned and diagonal views
            sd_phi = (np.arange(nv) * np.pi / 4).astype(np.float32)
        p = G.fan_para(nd, nv, N=(N, N), SO=4.0, OD=4.0, Ly=G.full_fan_width(), sd_phi=sd_phi)
        X = rng.random((N, N)).astype(np.float32)
        out.append((p, X))
    return out


def cone_cases():
    rng = np.random.default_rng(300)
    out = []
    p = G.cone_para(24, 20, 8, N=(16, 16, 16), SO=4.0, OD=4.0, Ly=G.full_fan_width(),
                    Lz=G.full_cone_height())
    out.append((p, rng.random((16, 16, 16)).astype(np.float32),
                rng.standard_normal((8, 20, 24)).astype(np.float32), True))
    p = G.cone_para(20, 16, 0, N=(12, 12, 12), SO=4.0, OD=4.0, Ly=G.full_fan_width(),
                    Lz=G.full_cone_height(), kind="helix", zmax=0.4, rps=1.0, vtab=1.0, fps=20.0)
    out.append((p, rng.random((12, 12, 12)).astype(np.float32),
                rng.standard_normal((p.nv, 16, 20)).astype(np.float32), False))
    p = G.cone_para(20, 18, 6, N=(12, 12, 8), L=(2.0, 2.0, 1.0), SO=4.0, OD=4.0,
                    Ly=G.full_fan_width(), Lz=G.full_cone_height(),
                    sd_phi=(np.arange(6) * np.pi / 2).astype(np.float32))
    out.append((p, rng.random((8, 12, 12)).astype(np.float32),
                rng.standard_normal((6, 18, 20)).astype(np.float32), False))
    return out


def find_l_3d_args():
    rng = np.random.default_rng(400)
    args = [
        # matlab/mex/src2/test_find_l_3d.cpp:9-23 and its backup copy
        (0, 0, 0, 0, 0, 1, 0.25, 0.5, 0.25, 5, 5, 5),
        (0, 0, 0, 0, 0, 1, 0.25, 0.5, 0.25, 0, 0, 0.5),
    ]
    for _ in range(400):
        s = rng.uniform(-6, 6, 3)
        d = rng.uniform(-6, 6, 3)
        if rng.random() < 0.1:
            d[0] = s[0]          # a == 0 branch
        elif rng.random() < 0.1:
            d[1] = s[1]          # b == 0 branch
        c = rng.uniform(-1, 1, 3)
        h = rng.choice([1.0, 1.0, 0.5, 2.0], 3)
        args.append(tuple(np.concatenate([s, d, h, c]).astype(np.float32)))
    return np.array(args, dtype=np.float32)

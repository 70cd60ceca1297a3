"""Seeded sweep over small random geometries: volume sides, panel shapes,
near and far sources, coarse and fine detector pitches, detector offsets,
non-unit voxel height, helical orbits, several image frames, fan and cone.
On each, the GPU forward and exact adjoint equal the oracle bit for bit
(the reference's float semantics), and the forward and matched adjoint
pass the dot test.  The fixed-geometry tests pin the named configurations;
this sweep looks for the edge cases between them."""
import math
import sys

import numpy as np
import pytest
import torch

from helpers import bits_equal, first_mismatch

pytestmark = pytest.mark.gpu

N_CONE, N_FAN = 40, 12


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _cone_case(seed):
    from framelab_amd import geometry as G
    rng = np.random.default_rng(1000 + seed)
    nx = 2 * int(rng.integers(4, 25))
    ny = 2 * int(rng.integers(4, 25))
    nz = 2 * int(rng.integers(4, 21))
    na, nb = int(rng.integers(8, 65)), int(rng.integers(4, 49))
    hz = float(rng.choice([1.0, rng.uniform(0.6, 1.4)]))          # voxel height / width
    L = (2.0, 2.0 * ny / nx, 2.0 * nz / nx * hz)
    R = 0.5 * math.hypot(L[0], L[1])
    SO = R * float(rng.uniform(1.15, 3.0))
    OD = float(rng.uniform(1.0, 6.0))
    half = math.asin(min(0.95, R / SO))
    Ly = 2.0 * (SO + OD) * math.tan(half) * float(rng.uniform(0.5, 1.4))
    Lz = (L[2] * (SO + OD) / (SO - R)) * float(rng.uniform(0.5, 1.4))
    y_os = float(rng.choice([0.0, rng.uniform(-0.4, 0.4) * Ly / na]))
    kw = dict(N=(nx, ny, nz), L=L, SO=SO, OD=OD, Ly=Ly, Lz=Lz, y_os=y_os)
    if rng.random() < 0.3:
        kw.update(kind="helix", zmax=float(rng.uniform(0.2, 0.8) * L[2]), rps=float(rng.uniform(0.5, 2.0)),
                  vtab=1.0, fps=float(rng.integers(6, 16)))
        p = G.cone_para(na, nb, 0, **kw)
    else:
        p = G.cone_para(na, nb, int(rng.integers(3, 21)), **kw)
    if rng.random() < 0.2 and p.nv >= 3:
        vf = np.arange(p.nv) % 3
        rng.shuffle(vf)
        p = G.with_frames(p, vf)
    return p


def _fan_case(seed):
    from framelab_amd import geometry as G
    rng = np.random.default_rng(2000 + seed)
    n = 2 * int(rng.integers(4, 40))
    nd, nv = int(rng.integers(8, 129)), int(rng.integers(3, 41))
    R = math.sqrt(2.0)
    SO = R * float(rng.uniform(1.15, 3.0))
    OD = float(rng.uniform(1.0, 6.0))
    Ly = 2.0 * (SO + OD) * math.tan(math.asin(min(0.95, R / SO))) * float(rng.uniform(0.5, 1.4))
    return G.fan_para(nd, nv, N=(n, n), SO=SO, OD=OD, Ly=Ly)


@pytest.mark.parametrize("seed", range(N_CONE))
def test_random_cone(gpu, oracle, seed):
    from framelab_amd.projector import ADJ_EXACT, ADJ_MATCHED, Projector
    p = _cone_case(seed)
    _say(f"case: random cone {seed}: vol {p.nx}x{p.ny}x{p.nz} dz {float(p.dz):.3f}, panel {p.na}x{p.nb}, "
         f"{p.nv} views, {p.nt} frames, SO {float(p.SO) * float(p.scale):.2f}, helix {np.ptp(p.sd_z) > 0}")
    rng = np.random.default_rng(seed)
    X = rng.random(p.nx * p.ny * p.nz * p.nt).astype(np.float32)
    Y = rng.standard_normal(p.nv * p.na * p.nb).astype(np.float32)
    oracle.set_threads(8)
    P = Projector(p)
    y = P.forward(torch.from_numpy(X.reshape(P.vol_shape)).to(gpu))
    y = y.cpu().numpy().ravel()
    yo = oracle.siddon_cone_f32(X, p)
    assert bits_equal(y, yo), first_mismatch(y, yo)
    Yt = torch.from_numpy(Y).reshape(P.sino_shape).to(gpu)
    x = P.adjoint(Yt, kind=ADJ_EXACT).cpu().numpy().ravel()
    xo = oracle.atx_cone_f32(Y, p)
    assert bits_equal(x, xo), first_mismatch(x, xo)
    xm = P.adjoint(Yt, kind=ADJ_MATCHED).cpu().numpy().ravel()
    lhs = float(np.dot(y.astype(np.float64), Y.astype(np.float64)))
    rhs = float(np.dot(X.astype(np.float64), xm.astype(np.float64)))
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs), (lhs, rhs)


@pytest.mark.parametrize("seed", range(N_FAN))
def test_random_fan(gpu, oracle, seed):
    from framelab_amd.projector import ADJ_EXACT, ADJ_MATCHED, Projector
    p = _fan_case(seed)
    _say(f"case: random fan {seed}: image {p.nx}, {p.nd} bins, {p.nv} views, SO {float(p.SO) * float(p.scale):.2f}")
    rng = np.random.default_rng(seed)
    X = rng.random(p.nx * p.ny).astype(np.float32)
    Y = rng.standard_normal(p.nv * p.nd).astype(np.float32)
    P = Projector(p)
    y = P.forward(torch.from_numpy(X.reshape(P.vol_shape)).to(gpu)).cpu().numpy().ravel()
    yo = oracle.siddon_fan_f32(X, p)
    assert bits_equal(y, yo), first_mismatch(y, yo)
    Yt = torch.from_numpy(Y).reshape(P.sino_shape).to(gpu)
    x = P.adjoint(Yt, kind=ADJ_EXACT).cpu().numpy().ravel()
    xo = oracle.atx_fan_f32(Y, p)
    assert bits_equal(x, xo), first_mismatch(x, xo)
    xm = P.adjoint(Yt, kind=ADJ_MATCHED).cpu().numpy().ravel()
    lhs = float(np.dot(y.astype(np.float64), Y.astype(np.float64)))
    rhs = float(np.dot(X.astype(np.float64), xm.astype(np.float64)))
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs), (lhs, rhs)

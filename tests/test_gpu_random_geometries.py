"""Seeded sweep over small random geometries: volume sides, panel shapes,
near and far sources, coarse and fine detector pitches, detector offsets,
non-unit voxel height, helical orbits, several image frames, fan and cone.
On each, the GPU forward and exact adjoint equal the oracle bit for bit
(the reference's float semantics), and the forward and matched adjoint
pass the dot test (gap over the scale of its terms).  The fixed-geometry tests pin the named configurations;
this sweep looks for the edge cases between them."""
import math
import sys

import numpy as np
import pytest
import torch

from helpers import bits_equal, dot_gap, first_mismatch

pytestmark = pytest.mark.gpu

N_CONE, N_FAN = 300, 100


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / np.linalg.norm(b)


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
    assert dot_gap(y, Y, X, xm) <= 2e-6


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
    assert dot_gap(y, Y, X, xm) <= 2e-6


@pytest.mark.parametrize("seed", range(N_FAN))
def test_random_fan_finite_beam(gpu, oracle, seed):
    """The finite-width-beam pair (version 2: strip areas) on the same random
    fan geometries: forward and adjoint bit-exact against the oracle."""
    from framelab_amd.projector import ADJ_FAN_FB, Projector
    p = _fan_case(seed)
    _say(f"case: random fan, finite beam {seed}: image {p.nx}, {p.nd} bins, {p.nv} views")
    rng = np.random.default_rng(100 + seed)
    X = rng.random(p.nx * p.ny).astype(np.float32)
    Y = rng.standard_normal(p.nv * p.nd).astype(np.float32)
    P = Projector(p)
    y = P.forward(torch.from_numpy(X.reshape(P.vol_shape)).to(gpu), method="fb").cpu().numpy().ravel()
    yo = oracle.fan_fb_fwd_f32(X, p)
    assert bits_equal(y, yo), first_mismatch(y, yo)
    x = P.adjoint(torch.from_numpy(Y).reshape(P.sino_shape).to(gpu), kind=ADJ_FAN_FB).cpu().numpy().ravel()
    xo = oracle.fan_fb_adj_f32(Y, p)
    assert bits_equal(x, xo), first_mismatch(x, xo)


@pytest.mark.parametrize("seed", range(96))
def test_random_framelets(gpu, oracle, seed):
    """Framelet analysis and synthesis on random sizes, odd and tiny ones
    included (the circular taps of the higher levels then wrap more than
    once): bands and the synthesis of arbitrary coefficients against the
    FP64 oracle within 2e-6, and W^T W = I."""
    from framelab_amd.framelets import FrameletSystem
    rng = np.random.default_rng(3000 + seed)
    dim = int(rng.choice([2, 3]))
    kind = str(rng.choice(["haar", "linear", "cubic"]))
    level = int(rng.integers(1, 4))
    shape = tuple(int(v) for v in rng.integers(3, 41 if dim == 3 else 97, dim))
    _say(f"case: random framelet {seed}: {kind} {dim}D level {level} shape {shape}")
    u = rng.standard_normal(shape).astype(np.float32)
    D, R = oracle.framelet_filters(kind)
    nf = len(D)
    dec, rec = (oracle.fradec3d, oracle.frarec3d) if dim == 3 else (oracle.fradec2d, oracle.frarec2d)
    W = FrameletSystem(dim, kind, level)
    c = W.forwardTransform(torch.from_numpy(u).to(gpu)).frameletArray.cpu().numpy()
    ref = dec(u.astype(np.float64), D, level)
    idx = ([(i, j, k) for k in range(nf) for j in range(nf) for i in range(nf)] if dim == 3
           else [(i, j) for j in range(nf) for i in range(nf)])
    for L in range(level):
        want = np.stack([ref[L][t] for t in idx])
        assert _rel(c[L], want) < 2e-6, L
    back = W.adjointTransform(torch.from_numpy(c).to(gpu)).cpu().numpy()
    assert _rel(back, u) < 2e-6
    cr = rng.standard_normal(c.shape).astype(np.float32)
    got = W.adjointTransform(torch.from_numpy(cr).to(gpu)).cpu().numpy()
    perm = (2, 1, 0, 3, 4, 5) if dim == 3 else (1, 0, 2, 3)
    want = rec([cr[L].astype(np.float64).reshape((nf,) * dim + shape).transpose(perm) for L in range(level)], R)
    assert _rel(got, want) < 2e-6

"""Batched fan problems of uneven sizes: every problem's result from the one
persistent launch must equal the single-plan result on the same geometry
(bit-exact for the operators, which run the same per-ray / per-pixel code;
within 1e-4 relative L2 for the solvers, whose FP64 reductions are split
differently), and the single-plan path is itself pinned to the oracle."""
import sys

import numpy as np
import pytest
import torch

from framelab_amd import geometry as G
from helpers import bits_equal, first_mismatch

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _uneven_paras(n, seed, sides=(16, 24, 32, 48, 64)):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        N = int(rng.choice(sides))
        nv = int(rng.integers(12, 61))
        nd = int(rng.integers(24, 129))
        out.append(G.fan_para(nd, nv, N=(N, N), SO=4.0, OD=4.0, Ly=G.full_fan_width()))
    return out


def _phantoms(paras, seed):
    from framelab_amd.phantoms import ellipses_2d
    return [ellipses_2d(p.nx, 6, seed=seed + i).astype(np.float32) for i, p in enumerate(paras)]


def test_batch_operators_bit_exact(gpu, oracle):
    """11 uneven problems (sides 16..64, 12..60 views, 24..128 bins): batched
    Siddon / Gao forward and exact / matched adjoint equal the single-plan
    calls bit for bit, before and after autotuning; problem 0 also equals
    the oracle's Siddon."""
    from framelab_amd.batch import ProjectorBatch
    from framelab_amd.projector import ADJ_EXACT, ADJ_MATCHED, Projector
    paras = _uneven_paras(11, seed=5)
    _say("case: batch operators, sizes " + ", ".join(f"{p.nx}/{p.nv}/{p.nd}" for p in paras))
    B = ProjectorBatch(paras)
    Ps = [Projector(p) for p in paras]
    xs = _phantoms(paras, 100)
    rng = np.random.default_rng(6)
    ys = [rng.standard_normal((p.nv, p.nd)).astype(np.float32) for p in paras]
    x = B.pack([torch.from_numpy(v) for v in xs]).to(gpu)
    y = B.pack([torch.from_numpy(v) for v in ys], sino=True).to(gpu)
    for rnd in range(2):
        if rnd == 1:
            tuned = B.autotune()
            _say(f"  autotuned {tuned}")
            assert tuned["fwd_tile"] in (64, 128, 256)
        for method in ("siddon", "gao"):
            got = B.unpack(B.forward(x, method=method).cpu(), sino=True)
            for i, P in enumerate(Ps):
                xi = torch.from_numpy(xs[i]).to(gpu)
                want = P.forward(xi, method=method).cpu().numpy()
                assert bits_equal(got[i].numpy(), want), (method, i, first_mismatch(got[i].numpy(), want))
        for kind in (ADJ_EXACT, ADJ_MATCHED):
            got = B.unpack(B.adjoint(y, kind=kind).cpu())
            for i, P in enumerate(Ps):
                yi = torch.from_numpy(ys[i]).to(gpu)
                want = P.adjoint(yi, kind=kind).cpu().numpy()
                assert bits_equal(got[i].numpy(), want), (kind, i, first_mismatch(got[i].numpy(), want))
    want0 = oracle.siddon_fan_f32(xs[0], paras[0])
    got0 = B.unpack(B.forward(x).cpu(), sino=True)[0].numpy()
    assert bits_equal(got0, want0.reshape(got0.shape)), first_mismatch(got0, want0)


@pytest.mark.parametrize("graph", [False, True])
def test_batch_cgls_matches_single(gpu, graph):
    """7 uneven problems, lam = 0.01, 8 iterations, matched adjoint: each
    problem's iterate within 1e-4 relative L2 of the single-plan CGLS and its
    gamma log within 1e-6; with a stopping tolerance each problem stops at
    its own iteration."""
    from framelab_amd.batch import ProjectorBatch
    from framelab_amd.projector import Projector
    from framelab_amd.solvers import cgls
    paras = _uneven_paras(7, seed=9)
    _say(f"case: batch CGLS graph={graph}")
    B = ProjectorBatch(paras)
    xs = _phantoms(paras, 200)
    bs, Ps = [], []
    for p, x in zip(paras, xs):
        P = Projector(p)
        Ps.append(P)
        xi = torch.from_numpy(x).to(gpu)
        bs.append(P.adjoint(P.forward(xi), kind=1))
    b = B.pack(bs)
    u0 = torch.zeros_like(b)
    u, glog, done = B.cgls(b, u0, lam=0.01, iters=8, graph=graph, matched=True)
    got = B.unpack(u.cpu())
    for i, P in enumerate(Ps):
        wu, wlog, wdone = cgls(P, bs[i], torch.zeros_like(bs[i]), lam=0.01, iters=8, matched=True)
        rel = _rel(got[i].numpy(), wu.cpu().numpy())
        _say(f"  problem {i} ({paras[i].nx}^2): rel={rel:.2e} iters={done[i]}")
        assert rel < 1e-4
        assert done[i] == wdone == 8
        assert np.allclose(glog[i], wlog, rtol=1e-6, atol=0)
    # per-problem stopping: a tol below every gamma of problem k up to
    # iteration 4 but above some gamma of the others by then
    low = glog[:, :5].min(axis=1)
    k = int(np.argmax(low))
    tol = float(low[k]) * 0.999
    if (low < tol).sum() >= 1:
        u2, glog2, done2 = B.cgls(b, u0, lam=0.01, iters=8, tol=tol, graph=graph, matched=True)
        assert done2[k] > 4
        for i in range(len(paras)):
            if low[i] < tol:
                assert done2[i] <= 4
                assert glog2[i, done2[i]] < tol


def test_batch_split_bregman_matches_single(gpu):
    """5 uneven problems, linear framelet, mu=1, lam=1e-2, 3 outer x 4 inner:
    each problem within 1e-4 relative L2 of the single-plan split Bregman,
    residual logs within 1e-4."""
    from framelab_amd.batch import ProjectorBatch
    from framelab_amd.projector import Projector
    from framelab_amd.solvers import split_bregman
    paras = _uneven_paras(5, seed=13, sides=(16, 32, 48))
    _say("case: batch split Bregman")
    B = ProjectorBatch(paras)
    xs = _phantoms(paras, 300)
    fs, Ps = [], []
    for p, x in zip(paras, xs):
        P = Projector(p)
        Ps.append(P)
        fs.append(P.forward(torch.from_numpy(x).to(gpu)))
    f = B.pack(fs, sino=True)
    u, res = B.split_bregman(f, torch.zeros(B.n_vol, device=gpu), frame="linear", lam=1e-2, mu=1.0, outer=3,
                             inner=4, graph=True, matched=True)
    got = B.unpack(u.cpu())
    for i, P in enumerate(Ps):
        wu, wres = split_bregman(P, "linear", fs[i], torch.zeros(P.vol_shape, device=gpu), 1e-2, 1.0, 3, 4,
                                 graph=True, matched=True)
        rel = _rel(got[i].numpy(), wu.cpu().numpy())
        _say(f"  problem {i}: rel={rel:.2e} res={res[i]} single={wres}")
        assert rel < 1e-4
        assert np.allclose(res[i], wres, rtol=1e-4)


def test_batch_rejects_cone_and_bad_sizes(gpu):
    from framelab_amd.batch import ProjectorBatch
    from helpers import small_cone
    with pytest.raises(TypeError):
        ProjectorBatch([small_cone()])
    with pytest.raises(ValueError):
        ProjectorBatch([])
    B = ProjectorBatch(_uneven_paras(2, seed=1))
    with pytest.raises(ValueError):
        B.forward(torch.zeros(B.n_vol + 1, device=gpu))

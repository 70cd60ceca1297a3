"""GPU checks of the framelet stencils, device-resident solvers and gradients
against the FP64 oracle (tolerances stated per test)."""
import sys

import numpy as np
import pytest
import torch

from helpers import small_cone, small_fan

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("kind", ["haar", "linear", "cubic"])
@pytest.mark.parametrize("dim", [2, 3])
def test_framelet_vs_oracle(gpu, oracle, kind, dim):
    _say(f"case: framelet {kind} {dim}D")
    from framelab_amd.framelets import FrameletSystem
    rng = np.random.default_rng(1)
    shape = (16, 20, 24) if dim == 3 else (32, 40)
    u = rng.standard_normal(shape).astype(np.float32)
    D, R = oracle.framelet_filters(kind)
    # 3D levels 3-4: tiled synthesis with a 4- or 8-voxel halo (haar,
    # linear), the per-voxel gather once the halo outgrows the tiles (cubic);
    # every level below the top reads the reconstruction of the level above
    # in place of its low-pass band
    for level in ((1, 2, 3, 4) if dim == 3 else (1, 2, 3)):
        W = FrameletSystem(dim, kind, level)
        c = W.forwardTransform(torch.from_numpy(u).to(gpu)).frameletArray.cpu().numpy()
        ref = (oracle.fradec3d if dim == 3 else oracle.fradec2d)(u.astype(np.float64), D, level)
        nf = len(D)
        for L in range(level):
            # device band b = i + NF*j (+ NF^2*k): MATLAB linear order of the cell
            want = np.stack([ref[L][(i, j, k) if dim == 3 else (i, j)]
                             for k in (range(nf) if dim == 3 else [0]) for j in range(nf) for i in range(nf)])
            assert _rel(c[L], want) < 2e-6, (kind, dim, level, L)
        rec = W.adjointTransform(torch.from_numpy(c).to(gpu)).cpu().numpy()
        assert _rel(rec, u) < 2e-6                       # W^T W = I
        want_rec = (oracle.frarec3d if dim == 3 else oracle.frarec2d)(ref, R)
        assert _rel(rec, want_rec) < 2e-6


@pytest.mark.parametrize("nz", [20, 6])
def test_framelet_partial_z_blocks(gpu, oracle, nz):
    """3-D linear frame, levels 1-2, with nz not a multiple of the synthesis
    tiles' eight output slices (nz = 20: a partial last block; nz = 6: fewer
    slices than one block's input window, so its wrapped window repeats
    slices): analysis, synthesis and W^T W = I against the FP64 oracle."""
    _say(f"case: framelet linear 3D, nz={nz}")
    from framelab_amd.framelets import FrameletSystem
    rng = np.random.default_rng(3)
    u = rng.standard_normal((nz, 24, 72)).astype(np.float32)
    D, R = oracle.framelet_filters("linear")
    for level in (1, 2):
        W = FrameletSystem(3, "linear", level)
        c = W.forwardTransform(torch.from_numpy(u).to(gpu)).frameletArray.cpu().numpy()
        ref = oracle.fradec3d(u.astype(np.float64), D, level)
        for L in range(level):
            want = np.stack([ref[L][(i, j, k)] for k in range(3) for j in range(3) for i in range(3)])
            assert _rel(c[L], want) < 2e-6, (nz, level, L)
        rec = W.adjointTransform(torch.from_numpy(c).to(gpu)).cpu().numpy()
        assert _rel(rec, u) < 2e-6
        assert _rel(rec, oracle.frarec3d(ref, R)) < 2e-6
        # synthesis of arbitrary coefficients (not a range of W)
        cr = rng.standard_normal(c.shape).astype(np.float32)
        got = W.adjointTransform(torch.from_numpy(cr).to(gpu)).cpu().numpy()
        want = oracle.frarec3d([cr[L].astype(np.float64).reshape((3, 3, 3) + u.shape).transpose(2, 1, 0, 3, 4, 5)
                                for L in range(level)], R)
        assert _rel(got, want) < 2e-6, (nz, level)


@pytest.mark.parametrize("kind", ["haar", "linear", "cubic"])
def test_framelet_full_row_synthesis(gpu, oracle, kind):
    """3-D synthesis on rows 256 and 512 wide (the full-row tiles, whose x
    taps wrap inside the tile; row counts and slice counts not multiples of
    the tiles'), levels 1-2 (dilations 1 and 2): arbitrary coefficients
    against the FP64 oracle within 2e-6, and W^T W = I."""
    from framelab_amd.framelets import FrameletSystem
    rng = np.random.default_rng(11)
    D, R = oracle.framelet_filters(kind)
    nf = len(D)
    for shape in ((7, 19, 256), (5, 9, 512)):
        _say(f"case: framelet {kind} 3D full rows {shape}")
        u = rng.standard_normal(shape).astype(np.float32)
        for level in (1, 2):
            W = FrameletSystem(3, kind, level)
            c = W.forwardTransform(torch.from_numpy(u).to(gpu))
            rec = W.adjointTransform(c).cpu().numpy()
            assert _rel(rec, u) < 2e-6, (kind, shape, level)
            cr = rng.standard_normal(tuple(c.frameletArray.shape)).astype(np.float32)
            got = W.adjointTransform(torch.from_numpy(cr).to(gpu)).cpu().numpy()
            want = oracle.frarec3d([cr[L].astype(np.float64).reshape((nf,) * 3 + shape).transpose(2, 1, 0, 3, 4, 5)
                                    for L in range(level)], R)
            assert _rel(got, want) < 2e-6, (kind, shape, level)


def test_framelet_dot_test(gpu):
    _say("case: framelet adjointness 48^3")
    from framelab_amd.framelets import FrameletSystem
    g = torch.Generator(device="cpu").manual_seed(2)
    W = FrameletSystem(3, "linear", 1)
    u = torch.randn((48, 48, 48), generator=g).to(gpu)
    c = torch.randn((1, 27, 48, 48, 48), generator=g).to(gpu)
    lhs = torch.dot(W.forwardTransform(u).frameletArray.double().ravel(), c.double().ravel()).item()
    rhs = torch.dot(u.double().ravel(), W.adjointTransform(c).double().ravel()).item()
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)


def _oracle_cgls(oracle, p, f, lam, iters, cone):
    A = lambda v: oracle.siddon_f64(v, p, cone=cone)
    At = lambda y: oracle.siddon_adj_f64(y, p, cone=cone)
    u0 = np.zeros(p.nx * p.ny * (p.nz if cone else 1))
    return oracle.cgls(A, At, At(f), u0, lam, iters)


@pytest.mark.parametrize("cone", [True, False])
def test_cgls_vs_fp64_oracle(gpu, oracle, cone):
    """FP32 operators + FP64 reductions, 20 iterations: the reference's FP32
    operator pair within 1e-4 of the same recurrences over the oracle's FP32
    operators; both adjoints within 3e-2 of the FP64 oracle's iterate; graph
    replay bit-identical to the eager loop; the device stopping test ends the
    loop at the right iteration."""
    _say(f"case: CGLS cone={cone}")
    from framelab_amd.phantoms import ellipses_2d, ellipsoids_3d
    from framelab_amd.projector import Projector
    from framelab_amd.solvers import cgls
    p = small_cone(na=48, nb=40, nv=30, N=24) if cone else small_fan(nd=96, nv=60, N=48)
    X = ellipsoids_3d(24, 8, seed=3) if cone else ellipses_2d(48, 6, seed=3)
    f64 = oracle.siddon_f64(X.astype(np.float64), p, cone=cone)
    want = _oracle_cgls(oracle, p, f64, 0.0, 20, cone)
    P = Projector(p)
    f = torch.from_numpy(f64.astype(np.float32)).reshape(P.sino_shape).to(gpu)
    b = P.adjoint(f)
    u0 = torch.zeros(P.vol_shape, device=gpu)
    # (a) the reference's FP32 operator pair (Siddon + exact-chord adjoint):
    # against the same recurrences over the oracle's bit-identical FP32
    # operators with FP64 dots, to 1e-4; it drifts from the FP64 pair because
    # its float chords and Siddon lengths differ by ~1e-5 (not a transpose)
    f32 = f64.astype(np.float32)
    A32 = lambda v: (oracle.siddon_cone_f32 if cone else oracle.siddon_fan_f32)(v.astype(np.float32), p).astype(np.float64)
    At32 = lambda y: (oracle.atx_cone_f32 if cone else oracle.atx_fan_f32)(y.astype(np.float32), p).astype(np.float64)
    want32 = oracle.cgls(A32, At32, At32(f32), np.zeros_like(want), 0.0, 20)
    u, log, done = cgls(P, b, u0, 0.0, 20)
    rel32, rel64 = _rel(u.cpu().numpy(), want32), _rel(u.cpu().numpy(), want)
    _say(f"  exact adjoint: rel vs FP32-operator oracle {rel32:.2e}, vs FP64 oracle {rel64:.2e}")
    assert done == 20 and rel32 < 1e-4 and rel64 < 3e-2
    # (b) the matched pair: same FP32 forward, so the same distance to the
    # FP64 iterate (CG amplifies the forward's ~1e-5 length rounding ~200x
    # on this 24^3, 30-view problem; both are ~7% from the phantom)
    u, log, done = cgls(P, P.adjoint(f, kind=1), u0, 0.0, 20, matched=True)
    rel = _rel(u.cpu().numpy(), want)
    _say(f"  matched adjoint: rel vs FP64 oracle {rel:.2e}")
    assert done == 20 and rel < 3e-2
    ug, logg, _ = cgls(P, b, u0, 0.0, 20, graph=True)
    ue, loge, _ = cgls(P, b, u0, 0.0, 20, graph=False)
    assert torch.equal(ug, ue) and np.array_equal(logg, loge)
    # the device stopping test ends the loop: tol between gamma_5 and gamma_4
    tol = 0.5 * (loge[4] + loge[5])
    us, logs, done = cgls(P, b, u0, 0.0, 20, tol=tol, graph=True)
    assert done == 5


def test_cgls_low_precision_storage(gpu, oracle):
    _say("case: CGLS bf16/fp16 storage")
    from framelab_amd.phantoms import ellipsoids_3d
    from framelab_amd.projector import Projector
    from framelab_amd.solvers import cgls
    p = small_cone(na=48, nb=40, nv=30, N=24)
    X = ellipsoids_3d(24, 8, seed=3)
    f64 = oracle.siddon_f64(X.astype(np.float64), p)
    want = _oracle_cgls(oracle, p, f64, 0.0, 20, True)
    P = Projector(p)
    f = torch.from_numpy(f64.astype(np.float32)).reshape(P.sino_shape).to(gpu)
    b = P.adjoint(f)
    u0 = torch.zeros(P.vol_shape, device=gpu)
    u32, _, _ = cgls(P, b, u0, 0.0, 20)
    for dt, tol in ((torch.float16, 5e-2), (torch.bfloat16, 2e-1)):
        u, _, _ = cgls(P, b, u0, 0.0, 20, storage=dt)
        rel = _rel(u.cpu().numpy(), want)
        rel32 = _rel(u.cpu().numpy(), u32.cpu().numpy())
        _say(f"  {dt} rel vs FP64 oracle={rel:.2e} vs FP32 path={rel32:.2e}")
        assert rel < tol


def test_fp32_solve_after_low_precision_solve(gpu):
    """An FP32-storage solve on a plan whose workspace an FP16 solve has
    already set up (its two-byte copy of p exists) equals, bit for bit, the
    same solve on a fresh plan: eager and graph-replayed, exact and matched
    adjoint.  Regression test: the FP32 iteration once wrote four-byte
    values into that two-byte buffer."""
    _say("case: FP32 CGLS after an FP16 CGLS on one plan")
    from framelab_amd.phantoms import ellipsoids_3d
    from framelab_amd.projector import Projector
    from framelab_amd.solvers import cgls
    p = small_cone(na=48, nb=40, nv=20, N=24)
    x = torch.from_numpy(ellipsoids_3d(24, 6, seed=4)).to(gpu)
    P, Q = Projector(p), Projector(p)
    b = P.adjoint(P.forward(x))
    u0 = torch.zeros(P.vol_shape, device=gpu)
    cgls(P, b, u0, 0.0, 3, storage=torch.float16, graph=True)
    for graph in (False, True):
        for matched in (False, True):
            u, g, _ = cgls(P, b, u0, 0.0, 5, graph=graph, matched=matched)
            w, h, _ = cgls(Q, b, u0, 0.0, 5, graph=graph, matched=matched)
            assert torch.equal(u, w) and np.array_equal(g, h), (graph, matched)


def test_cgls_graph_reuse_across_calls(gpu):
    """A graph captured by a short solve is reused by a longer one (the log
    buffer grows in between) and by a solve on another iterate tensor: every
    graph-replayed result and gamma log equals the eager one bit for bit."""
    _say("case: CGLS graph reuse")
    from framelab_amd.phantoms import ellipses_2d
    from framelab_amd.projector import Projector
    from framelab_amd.solvers import cgls
    p = small_fan(nd=96, nv=40, N=32)
    P = Projector(p)
    x = torch.from_numpy(ellipses_2d(32, 5, seed=8).astype(np.float32)).to(gpu)
    b = P.adjoint(P.forward(x))
    u0 = torch.zeros(P.vol_shape, device=gpu)
    cgls(P, b, u0, 0.01, 1, graph=True)
    for iters in (6, 9):
        ug, lg, dg = cgls(P, b, u0 + 0.0, 0.01, iters, graph=True)
        ue, le, de = cgls(P, b, u0, 0.01, iters, graph=False)
        assert dg == de == iters
        assert np.array_equal(lg, le), (lg, le)
        assert torch.equal(ug, ue)


def test_split_bregman_vs_oracle(gpu, oracle):
    """Fan 48^2, linear framelet, mu=1, lam=1e-2, 4 outer x 5 inner: relative
    L2 difference to the FP64 oracle below 2e-3."""
    _say("case: split Bregman fan 48^2")
    from framelab_amd.phantoms import ellipses_2d
    from framelab_amd.projector import Projector
    from framelab_amd.solvers import split_bregman
    p = small_fan(nd=96, nv=60, N=48)
    X = ellipses_2d(48, 6, seed=4).astype(np.float64)
    f64 = oracle.siddon_f64(X, p, cone=False)
    D, R = oracle.framelet_filters("linear")
    A = lambda v: oracle.siddon_f64(v.ravel(), p, cone=False)
    At = lambda y: oracle.siddon_adj_f64(y, p, cone=False).reshape(48, 48)
    W = lambda u: oracle.fradec2d(u.reshape(48, 48), D, 1)[0]
    Wt = lambda c: oracle.frarec2d([c], R)
    shrink = lambda c, t: oracle.soft_threshold([c], t, 2)[0]
    lam, mu = 1e-2, 1.0
    want = oracle.split_bregman(A, At, W, Wt, lam, mu, f64, np.zeros((48, 48)), 4, 5, shrink=shrink)
    P = Projector(p)
    f = torch.from_numpy(f64.astype(np.float32)).reshape(P.sino_shape).to(gpu)
    u, res = split_bregman(P, "linear", f, torch.zeros(P.vol_shape, device=gpu), lam, mu, 4, 5,
                           matched=True)
    rel = _rel(u.cpu().numpy(), want)
    _say(f"  rel={rel:.2e} resid={res}")
    assert rel < 2e-3
    assert np.all(np.isfinite(res))


@pytest.mark.parametrize("geom", ["fan", "cone"])
def test_split_bregman_multilevel_vs_oracle(gpu, oracle, geom):
    """Split Bregman with a two-level linear frame (FrameletSystem level 2:
    level 2 analyses level 1's low-pass band at dilation 2; shrink keeps
    every level's low-pass band; W^T synthesises level 2 into level 1's
    low-pass band) against the FP64 oracle with the same multi-level frame:
    fan 48^2 (4 outer x 5 inner) and cone 24^3 (3 x 3), matched adjoint,
    relative L2 below 2e-3."""
    _say(f"case: split Bregman, 2-level frame, {geom}")
    from framelab_amd import geometry as G
    from framelab_amd.framelets import FrameletSystem
    from framelab_amd.phantoms import ellipses_2d, ellipsoids_3d
    from framelab_amd.projector import Projector
    from framelab_amd.solvers import SplitBregman, split_bregman
    cone = geom == "cone"
    if cone:
        p = G.cone_para(40, 32, 30, N=(24, 24, 24), SO=4.0, OD=4.0, Ly=G.full_fan_width(), Lz=G.full_cone_height())
        X = ellipsoids_3d(24, count=5, seed=6).astype(np.float64)
        shape, outer, inner = (24, 24, 24), 3, 3
        dec, rec = oracle.fradec3d, oracle.frarec3d
    else:
        p = small_fan(nd=96, nv=60, N=48)
        X = ellipses_2d(48, 6, seed=4).astype(np.float64)
        shape, outer, inner = (48, 48), 4, 5
        dec, rec = oracle.fradec2d, oracle.frarec2d
    f64 = oracle.siddon_f64(X, p, cone=cone)
    D, R = oracle.framelet_filters("linear")
    dim = len(shape)
    A = lambda v: oracle.siddon_f64(v.ravel(), p, cone=cone)
    At = lambda y: oracle.siddon_adj_f64(y, p, cone=cone).reshape(shape)
    W = lambda u: np.stack(dec(u.reshape(shape), D, 2))
    Wt = lambda c: rec(list(c), R)
    shrink = lambda c, t: np.stack(oracle.soft_threshold(list(c), t, dim))
    lam, mu = 1e-2, 1.0
    want = oracle.split_bregman(A, At, W, Wt, lam, mu, f64, np.zeros(shape), outer, inner, shrink=shrink)
    P = Projector(p)
    system = FrameletSystem(dim, "linear", 2)
    f = torch.from_numpy(f64.astype(np.float32)).reshape(P.sino_shape).to(gpu)
    u, res = split_bregman(P, "linear", f, torch.zeros(P.vol_shape, device=gpu), lam, mu, outer, inner,
                           matched=True, level=system.level)
    rel = _rel(u.cpu().numpy(), want)
    _say(f"  rel={rel:.2e} resid={res}")
    assert rel < 2e-3
    assert np.all(np.isfinite(res))
    # the class mirror passes the system's level through (exact adjoint here)
    sb = SplitBregman(P.forward, P.adjoint, system.forwardTransform, system.adjointTransform, lam, mu, f,
                      torch.zeros(P.vol_shape, device=gpu), globalIter=outer, cgIter=inner, cgTol=0.0)
    ue = sb.solveSplitBregman()
    ue2, _ = split_bregman(P, "linear", f, torch.zeros(P.vol_shape, device=gpu), lam, mu, outer, inner, level=2)
    assert torch.equal(ue, ue2)


def test_framelet_gradient_multilevel(gpu, oracle):
    """Smoothed-l1 gradient over a two-level 3-D linear frame (16^3): the
    GPU gradient equals the FP64 chain rule -- psi' on every high-pass band
    of both levels, synthesised with level 2 replacing level 1's low-pass
    band -- within 1e-5, its loss within 1e-5, and central differences of
    the FP64 objective agree along 6 directions (|g||d| scale, 1e-4)."""
    _say("case: framelet gradient, 2 levels")
    from framelab_amd.framelets import FrameletSystem
    from framelab_amd.gradients import framelet_gradient
    rng = np.random.default_rng(8)
    x = rng.random((16, 16, 16))
    lam, eps = 1.0, 1e-4
    D, R = oracle.framelet_filters("linear")

    def parts(v):
        C = oracle.fradec3d(v.reshape(16, 16, 16), D, 2)
        loss, psi = 0.0, []
        for c in C:
            r = np.sqrt(c ** 2 + eps)
            r[0, 0, 0] = 0.0
            loss += r.sum()
            g = c / np.sqrt(c ** 2 + eps)
            g[0, 0, 0] = 0.0
            psi.append(g)
        return lam * loss, lam * oracle.frarec3d(psi, R).ravel()

    J64, g64 = parts(x)
    W = FrameletSystem(3, "linear", 2)
    g, J = framelet_gradient(W, torch.from_numpy(x.astype(np.float32)).to(gpu), lam, eps)
    g = g.cpu().numpy().astype(np.float64).ravel()
    _say(f"  grad rel {_rel(g, g64):.2e}, loss rel {abs(J - J64) / J64:.2e}")
    assert _rel(g, g64) < 1e-5
    assert abs(J - J64) < 1e-5 * J64
    h = 1e-4
    for k in range(6):
        d = rng.standard_normal((16, 16, 16))
        fd = (parts(x + h * d)[0] - parts(x - h * d)[0]) / (2 * h)
        assert abs(fd - g64 @ d.ravel()) <= 1e-4 * np.linalg.norm(g64) * np.linalg.norm(d)


def test_gradients_fd(gpu, oracle):
    """Gradient of J = 0.5|Ax-b|^2 + lam*sum_hp sqrt((Wx)^2+eps), lam=1e-3,
    eps=1e-6, 16^3 cone, central differences with h=1e-3 along 16 seeded
    Gaussian directions d.  A random direction's inner product cancels by
    ~sqrt(N) and psi's curvature (1/sqrt(eps)) makes the O(h^2) term visible,
    so differences are measured on the scale |g||d|:
      (1) FP64 analytic gradient vs FD of the FP64 oracle objective <= 1e-6;
      (2) the framelet term alone (FD of sum psi(Wx)) <= 1e-4;
      (3) GPU gradient within 1e-4 relative L2 of the FP64 gradient;
      (4) GPU gradient vs FD <= 1e-5."""
    _say("case: gradient finite differences")
    from framelab_amd.framelets import FrameletSystem
    from framelab_amd.gradients import data_gradient, framelet_gradient, objective_gradient
    from framelab_amd.projector import Projector
    p = small_cone(na=40, nb=32, nv=20, N=16)
    rng = np.random.default_rng(5)
    x = rng.random((16, 16, 16))
    b = oracle.siddon_f64(rng.random((16, 16, 16)), p) + 0.01 * rng.standard_normal(p.nv * p.nb * p.na)
    lam, eps = 1e-3, 1e-6
    D, R = oracle.framelet_filters("linear")

    def J(v):
        r = oracle.siddon_f64(v, p) - b
        wb = oracle.fradec3d(v.reshape(16, 16, 16), D, 1)[0]
        hp = np.sqrt(wb ** 2 + eps)
        hp[0, 0, 0] = 0.0
        return 0.5 * r @ r + lam * hp.sum()

    def grad64(v):
        g = oracle.siddon_adj_f64(oracle.siddon_f64(v, p) - b, p)
        wb = oracle.fradec3d(v.reshape(16, 16, 16), D, 1)[0]
        psi = wb / np.sqrt(wb ** 2 + eps)
        psi[0, 0, 0] = 0.0
        return g + lam * oracle.frarec3d([psi], R).ravel()

    g64 = grad64(x)
    P = Projector(p)
    W = FrameletSystem(3, "linear", 1)
    xt = torch.from_numpy(x.astype(np.float32)).to(gpu)
    bt = torch.from_numpy(b.astype(np.float32)).reshape(P.sino_shape).to(gpu)
    g, Jg = objective_gradient(P, W, xt, bt, lam, eps)
    g = g.cpu().numpy().astype(np.float64).ravel()
    assert abs(Jg - J(x)) < 1e-4 * abs(J(x))
    assert _rel(g, g64) < 1e-4
    def Jr(v):
        wb = oracle.fradec3d(v.reshape(16, 16, 16), D, 1)[0]
        hp = np.sqrt(wb ** 2 + eps)
        hp[0, 0, 0] = 0.0
        return hp.sum()

    wbx = oracle.fradec3d(x, D, 1)[0]
    psi_x = wbx / np.sqrt(wbx ** 2 + eps)
    psi_x[0, 0, 0] = 0.0
    gr64 = oracle.frarec3d([psi_x], R).ravel()
    h = 1e-3
    worst = [0.0, 0.0, 0.0]
    for k in range(16):
        d = rng.standard_normal((16, 16, 16))
        dv = d.ravel()
        nd = np.linalg.norm(dv)
        fd = (J(x + h * d) - J(x - h * d)) / (2 * h)
        fdr = (Jr(x + h * d) - Jr(x - h * d)) / (2 * h)
        worst[0] = max(worst[0], abs(fd - g64 @ dv) / (np.linalg.norm(g64) * nd))
        worst[1] = max(worst[1], abs(fdr - gr64 @ dv) / (np.linalg.norm(gr64) * nd))
        worst[2] = max(worst[2], abs(fd - g @ dv) / (np.linalg.norm(g) * nd))
    _say(f"  GPU gradient rel L2 vs FP64 {_rel(g, g64):.2e}; FD errors on |g||d| scale: "
         f"FP64 {worst[0]:.1e}, framelet term {worst[1]:.1e}, GPU {worst[2]:.1e}")
    assert worst[0] <= 1e-6 and worst[1] <= 1e-4 and worst[2] <= 1e-5
    # the pieces separately against their FP64 closed forms
    gd, ld = data_gradient(P, xt, bt)
    r = oracle.siddon_f64(x, p) - b
    assert _rel(gd.cpu().numpy(), oracle.siddon_adj_f64(r, p)) < 1e-4
    assert abs(ld - 0.5 * r @ r) < 1e-4 * (0.5 * r @ r)
    gr, lr = framelet_gradient(W, xt, 1.0, eps)
    wb = oracle.fradec3d(x, D, 1)[0]
    psi = wb / np.sqrt(wb ** 2 + eps)
    psi[0, 0, 0] = 0.0
    assert _rel(gr.cpu().numpy(), oracle.frarec3d([psi], R)) < 1e-5


def test_class_mirror_native_paths(gpu, oracle):
    """CGLS.solveCGLS over caller-supplied callables (not one of this
    package's projectors) runs its vector steps and FP64 dots in the native
    kernels: with lambdas around a projector it reproduces the plan solver's
    eager iterates and residuals bit for bit.  FrameletExpansion.softThreshold
    (native, in place) equals the oracle's shrink with every level's
    low-pass band kept."""
    _say("case: class mirror native paths")
    from framelab_amd.framelets import FrameletSystem
    from framelab_amd.projector import Projector
    from framelab_amd.solvers import CGLS, cgls
    P = Projector(small_cone(na=40, nb=32, nv=16, N=16))
    g = torch.Generator(device="cpu").manual_seed(9)
    b = P.adjoint(P.forward(torch.rand(P.vol_shape, generator=g).to(gpu)))
    u0 = torch.zeros_like(b)
    want, wlog, _ = cgls(P, b, u0, 0.05, 7, graph=False)
    solver = CGLS(lambda v: P.forward(v), lambda y: P.adjoint(y), b, u0, 0.05, 7, 0.0)
    got = solver.solveCGLS()
    assert torch.equal(got, want)
    assert np.array_equal(solver.gamma_log, wlog)
    W = FrameletSystem(3, "linear", 2)
    x = torch.randn((12, 12, 12), generator=g).to(gpu)
    e = W.forwardTransform(x)
    c = e.frameletArray.double().cpu().numpy()
    e.softThreshold(0.3)
    got = e.frameletArray.double().cpu().numpy()
    want = np.stack(oracle.soft_threshold(list(c), 0.3, 1))   # band axis 1: band 0 per level kept
    assert np.allclose(got, want, rtol=0, atol=1e-7)
    assert np.array_equal(got[:, 0], c[:, 0])


def test_scanner_classes(gpu, oracle):
    """The reference's object API: apply / applyAdjoint / CGLS.solveCGLS."""
    _say("case: scanner classes")
    from framelab_amd.datatypes import ObjectData
    from framelab_amd.phantoms import ellipsoids_3d
    from framelab_amd.scanners import ConeBeamScanner, FanBeamScanner
    from framelab_amd.solvers import CGLS
    cb = ConeBeamScanner("circle", 48, 40, 24, SO=4.0, OD=4.0)
    X = ellipsoids_3d(24, 6, seed=6)
    obj = ObjectData(3, torch.from_numpy(X).to(gpu), (2.0, 2.0, 2.0))
    y = cb.apply(obj)
    assert y.dataArray.shape == (24, 40, 48)
    yo = oracle.siddon_cone_f32(X, cb.para)
    assert np.array_equal(y.dataArray.cpu().numpy().ravel(), yo)
    back = cb.applyAdjoint(y, [24, 24, 24])
    assert np.array_equal(back.dataArray.cpu().numpy().ravel(), oracle.atx_cone_f32(yo, cb.para))
    solver = CGLS(cb.apply, cb.applyAdjoint, back, ObjectData(3, torch.zeros_like(obj.dataArray), obj.L),
                  0.0, 5)
    u = solver.solveCGLS()
    assert u.dataArray.shape == (24, 24, 24) and solver.gamma_log[-1] < solver.gamma_log[0]
    fb = FanBeamScanner(64, 30, SO=4.0, OD=4.0, version=0)
    o2 = ObjectData(2, torch.rand((32, 32), device=gpu), (2.0, 2.0))
    y2 = fb.apply(o2)
    assert np.array_equal(y2.dataArray.cpu().numpy().ravel(),
                          oracle.siddon_fan_f32(o2.dataArray.cpu().numpy(), fb.para))


def test_data_type_framelet_methods(gpu, oracle):
    """ObjectData.frameletTransform / adjointFrameletTransform,
    FrameletExpansion.adjointFrameletTransform and CTData.frameletTransform
    (ObjectData.m:58-69, FrameletExpansion.m:36-41, CTData.m:26-30): the
    bands equal the oracle's analysis and the synthesis restores the array
    (W^T W = I)."""
    _say("case: data-type framelet methods")
    from framelab_amd.datatypes import CTData, ObjectData
    from framelab_amd.framelets import FrameletSystem
    g = torch.Generator(device="cpu").manual_seed(21)
    x = torch.randn((10, 12, 14), generator=g)
    W = FrameletSystem(3, "linear", 2)
    obj = ObjectData(3, x.to(gpu), (2.0, 2.0, 2.0))
    alpha = obj.frameletTransform(W)
    D, _ = oracle.framelet_filters("linear")
    ref = oracle.fradec3d(x.double().numpy(), D, 2)
    nf = len(D)
    idx = [(i, j, k) for k in range(nf) for j in range(nf) for i in range(nf)]
    for L in range(2):
        assert _rel(alpha.frameletArray[L].cpu().numpy(), np.stack([ref[L][t] for t in idx])) < 2e-6
    back = alpha.adjointFrameletTransform()
    assert _rel(back.cpu().numpy(), x.numpy()) < 2e-6
    obj2 = ObjectData(3, torch.zeros_like(obj.dataArray), obj.L)
    obj2.adjointFrameletTransform(W, alpha)
    assert torch.equal(obj2.dataArray, back)
    y = torch.randn((6, 8, 9), generator=g).to(gpu)
    ct = CTData(None, y, None, None)
    assert torch.equal(ct.frameletTransform(W).frameletArray, W.forwardTransform(y).frameletArray)

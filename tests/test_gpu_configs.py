"""The benchmark's 2-D and differentiation configurations at their real
sizes: C1 (fan 256^2, 360 views x 512 bins, 20 CGLS iterations on the
noiseless sinogram) and C3 (cone 128^3, 180 views, 256x256 panel, Poisson
data at 1e5 photons, lam = 1e-3 linear framelet, psi = sqrt(t^2 + 1e-6))."""
import sys

import numpy as np
import pytest
import torch

from helpers import bits_equal, first_mismatch

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_c1_fan_full_size(gpu, oracle):
    """Siddon forward bit-exact against the reference CPU kernel and the
    oracle; exact adjoint bit-exact against the oracle; 20 CGLS iterations
    within 2e-4 relative L2 of the same recurrence over the oracle's FP32
    operators (FP64 dots)."""
    from framelab_amd import geometry as G
    from framelab_amd.phantoms import ellipses_2d
    from framelab_amd.projector import Projector
    from framelab_amd.solvers import cgls
    _say("case: C1 fan 256^2, 360 x 512, 20 CGLS")
    p = G.fan_para(512, 360, N=(256, 256), SO=4.0, OD=4.0, Ly=G.full_fan_width())
    X = ellipses_2d(256, 10, seed=1)
    P = Projector(p)
    y = P.forward(torch.from_numpy(X).to(gpu)).cpu().numpy().ravel()
    yo = oracle.siddon_fan_f32(X, p)
    assert bits_equal(y, yo), first_mismatch(y, yo)
    if oracle.has_ref():
        assert bits_equal(y, oracle.ref_siddon_fan(X, p))
    b = P.adjoint(torch.from_numpy(y).reshape(P.sino_shape).to(gpu))
    bo = oracle.atx_fan_f32(yo, p)
    assert bits_equal(b.cpu().numpy().ravel(), bo), first_mismatch(b.cpu().numpy(), bo)
    oracle.set_threads(16)
    want = oracle.cgls(lambda v: oracle.siddon_fan_f32(v.astype(np.float32), p).astype(np.float64),
                       lambda s: oracle.atx_fan_f32(s.astype(np.float32), p).astype(np.float64),
                       bo.astype(np.float64), np.zeros(bo.size), 0.0, 20)
    oracle.set_threads(1)
    u, _, done = cgls(P, b, torch.zeros_like(b), 0.0, 20, graph=True)
    rel = _rel(u.cpu().numpy(), want)
    _say(f"  CGLS rel L2 vs oracle recurrence {rel:.2e}, vs phantom {_rel(u.cpu().numpy(), X):.3f}")
    assert done == 20 and rel < 2e-4


def test_c3_gradient_full_size(gpu, oracle):
    """At the full C3 size on the GPU: J = 0.5|Ax-b|^2 + lam sum psi(Wx)
    (high-pass bands), its gradient against central differences (h = 1e-3,
    16 Gaussian directions) on the |g||d| scale; projector and framelet dot
    tests; and, on an 18-view subset of the same geometry, the GPU gradient
    against the FP64 oracle gradient."""
    from framelab_amd import geometry as G
    from framelab_amd.framelets import FrameletSystem
    from framelab_amd.gradients import objective_gradient
    from framelab_amd.phantoms import ellipsoids_3d
    from framelab_amd.projector import ADJ_MATCHED, Projector
    _say("case: C3 cone 128^3, 180 views, 256^2, gradient")
    N, nv = 128, 180
    kw = dict(N=(N, N, N), SO=4.0, OD=4.0, Ly=G.full_fan_width(), Lz=G.full_cone_height())
    p = G.cone_para(256, 256, nv, **kw)
    X = ellipsoids_3d(N, count=20, seed=3)
    P = Projector(p)
    xt = torch.from_numpy(X).to(gpu)
    ax = P.forward(xt).double().cpu().numpy()
    rng = np.random.default_rng(7)
    counts = np.maximum(rng.poisson(1e5 * np.exp(-ax)), 1)
    b = (-np.log(counts / 1e5)).astype(np.float32)
    bt = torch.from_numpy(b).to(gpu)
    lam, eps = 1e-3, 1e-6
    W = FrameletSystem(3, "linear", 1)
    x0 = (X + 0.05 * rng.standard_normal(X.shape)).astype(np.float32)
    x0t = torch.from_numpy(x0).to(gpu)
    g, J0 = objective_gradient(P, W, x0t, bt, lam, eps)
    g = g.double().cpu().numpy().ravel()
    h = 1e-3
    worst = 0.0
    for k in range(16):
        d = rng.standard_normal(X.shape).astype(np.float32)
        dt = torch.from_numpy(d).to(gpu)
        _, Jp = objective_gradient(P, W, x0t + h * dt, bt, lam, eps)
        _, Jm = objective_gradient(P, W, x0t - h * dt, bt, lam, eps)
        fd = (Jp - Jm) / (2 * h)
        worst = max(worst, abs(fd - g @ d.ravel().astype(np.float64)) / (np.linalg.norm(g) * np.linalg.norm(d)))
    _say(f"  FD error on the |g||d| scale: {worst:.2e}")
    assert worst <= 1e-4
    # dot tests at full size
    yv = torch.randn(P.sino_shape, device=gpu)
    xv = torch.randn(P.vol_shape, device=gpu)
    lhs = float(torch.dot(P.forward(xv).double().ravel(), yv.double().ravel()))
    rhs = float(torch.dot(xv.double().ravel(), P.adjoint(yv, kind=ADJ_MATCHED).double().ravel()))
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
    c = W.forwardTransform(xv).frameletArray
    cv = torch.randn_like(c)
    lhs = float(torch.dot(c.double().ravel(), cv.double().ravel()))
    rhs = float(torch.dot(xv.double().ravel(), W.adjointTransform(cv).double().ravel()))
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
    # FP64 oracle gradient on every tenth view
    sub = G.cone_para(256, 256, 18, sd_phi=p.sd_phi[::10], **kw)
    Ps = Projector(sub)
    bs = b.reshape(nv, 256, 256)[::10].copy()
    gs, _ = objective_gradient(Ps, W, x0t, torch.from_numpy(bs).to(gpu), lam, eps)
    oracle.set_threads(16)
    r = oracle.siddon_f64(x0.astype(np.float64), sub) - bs.astype(np.float64).ravel()
    gd = oracle.siddon_adj_f64(r, sub)
    oracle.set_threads(1)
    D, R = oracle.framelet_filters("linear")
    wb = oracle.fradec3d(x0.astype(np.float64), D, 1)[0]
    psi = wb / np.sqrt(wb ** 2 + eps)
    psi[0, 0, 0] = 0.0
    g64 = gd + lam * oracle.frarec3d([psi], R).ravel()
    rel = _rel(gs.double().cpu().numpy(), g64)
    _say(f"  gradient rel L2 vs FP64 oracle (18 views): {rel:.2e}")
    assert rel < 1e-4

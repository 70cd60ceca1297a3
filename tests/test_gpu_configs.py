"""The 2-D, differentiation and precision workloads at their real sizes:
fan 256^2 with 360 views x 512 bins and 20 CGLS iterations on the
noiseless sinogram; cone 128^3 with 180 views, a 256x256 panel, Poisson
data at 1e5 photons, lam = 1e-3 linear framelet, psi = sqrt(t^2 + 1e-6);
cone 512^3 with 720 views, a 1024x768 panel, BF16/FP16 storage and split
Bregman with the 27-band framelet."""
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
    # dot tests at full size, seeded; uniform [0, 1) vectors keep <Ax, y>
    # free of cancellation, so the relative difference measures the pair
    gen = torch.Generator(device=gpu).manual_seed(13)
    yv = torch.rand(P.sino_shape, device=gpu, generator=gen)
    xv = torch.rand(P.vol_shape, device=gpu, generator=gen)
    lhs = float(torch.dot(P.forward(xv).double().ravel(), yv.double().ravel()))
    rhs = float(torch.dot(xv.double().ravel(), P.adjoint(yv, kind=ADJ_MATCHED).double().ravel()))
    _say(f"  projector dot test rel {abs(lhs - rhs) / abs(lhs):.2e}")
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)
    c = W.forwardTransform(xv).frameletArray
    cv = torch.rand(c.shape, device=gpu, generator=gen)
    lhs = float(torch.dot(c.double().ravel(), cv.double().ravel()))
    rhs = float(torch.dot(xv.double().ravel(), W.adjointTransform(cv).double().ravel()))
    _say(f"  framelet dot test rel {abs(lhs - rhs) / abs(lhs):.2e}")
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


@pytest.mark.timeout(600)
def test_c4_precision_full_size(gpu, oracle):
    """The precision configuration at its full size: 512^3, 720 views over
    2 pi, 1024 x 768 panel.  (1) The forward of the BF16- and FP16-stored
    phantom and the matched adjoint of the BF16/FP16-stored sinogram (FP32
    accumulation) against the FP64 oracle on a two-view subset of the same
    geometry, each error stated.  (2) The matched pair's dot test at full
    size.  (3) Split Bregman with the 27-band linear framelet (mu = 1,
    lam = 1e-3) storing the CG vectors and the framelet coefficients in BF16
    and in FP16, against the FP32-storage run after the same 2 outer x 3
    inner iterations (the FP32 run's own parity with the oracle is pinned at
    the smaller sizes; an FP64 solve at this size is beyond a test's time)."""
    from framelab_amd import geometry as G
    from framelab_amd.phantoms import ellipsoids_3d_device
    from framelab_amd.projector import ADJ_MATCHED, Projector
    from framelab_amd.solvers import split_bregman
    _say("case: C4 cone 512^3, 720 views, 1024 x 768, BF16/FP16 storage")
    N, nv, na, nb = 512, 720, 1024, 768
    kw = dict(N=(N, N, N), SO=4.0, OD=4.0, Ly=G.full_fan_width(), Lz=G.full_cone_height())
    p = G.cone_para(na, nb, nv, **kw)
    x = ellipsoids_3d_device(N, count=20, seed=4, device=gpu)
    # (1) storage error against the FP64 oracle, views 0 and nv/2
    sub = G.cone_para(na, nb, 2, sd_phi=p.sd_phi[::nv // 2], **kw)
    Ps = Projector(sub)
    X64 = x.double().cpu().numpy()
    oracle.set_threads(16)
    y64 = oracle.siddon_f64(X64, sub)
    errs = {}
    for dt, name in ((torch.float32, "fp32"), (torch.bfloat16, "bf16"), (torch.float16, "fp16")):
        yl = Ps.forward(x.to(dt))
        errs[name] = _rel(yl.double().cpu().numpy(), y64)
    ys = torch.from_numpy(y64.astype(np.float32)).reshape(Ps.sino_shape).to(gpu)
    b64 = oracle.siddon_adj_f64(y64, sub)
    oracle.set_threads(1)
    for dt, name in ((torch.float32, "fp32"), (torch.bfloat16, "bf16"), (torch.float16, "fp16")):
        bl = Ps.adjoint(ys.to(dt), kind=ADJ_MATCHED, out=torch.empty(Ps.vol_shape, dtype=dt, device=gpu))
        errs["adj_" + name] = _rel(bl.double().cpu().numpy(), b64)
    _say("  rel L2 vs FP64 oracle (2 views): " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    # the adjoint's voxel values are sums of a few single intersection
    # lengths, each an FP32 difference of ray parameters near 1/2 (relative
    # error up to ~3e-5); the forward's ray sums average those errors out
    assert errs["fp32"] < 1e-5 and errs["adj_fp32"] < 5e-5
    assert errs["fp16"] < 1e-3 and errs["adj_fp16"] < 1e-3
    assert errs["bf16"] < 5e-3 and errs["adj_bf16"] < 5e-3
    del X64, y64, b64, ys
    # (2) dot test of the matched pair at full size, with uniform [0, 1)
    # vectors: positive terms, so |<Ax, y>| is not a cancellation remnant
    # and the relative difference measures the operators, not the draw
    P = Projector(p)
    g = torch.Generator(device=gpu).manual_seed(11)
    xv = torch.rand(P.vol_shape, device=gpu, generator=g)
    yv = torch.rand(P.sino_shape, device=gpu, generator=g)
    lhs = float(torch.dot(P.forward(xv).double().ravel(), yv.double().ravel()))
    rhs = float(torch.dot(xv.double().ravel(), P.adjoint(yv, kind=ADJ_MATCHED).double().ravel()))
    _say(f"  dot test rel {abs(lhs - rhs) / abs(lhs):.2e}")
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)
    del xv, yv
    # (3) split Bregman, 2 outer x 3 inner, FP32 vs BF16 / FP16 storage
    f = P.forward(x)
    u0 = torch.zeros(P.vol_shape, device=gpu)
    runs = {}
    for dt, name in ((torch.float32, "fp32"), (torch.bfloat16, "bf16"), (torch.float16, "fp16")):
        u, res = split_bregman(P, "linear", f, u0, 1e-3, 1.0, 2, 3, graph=True, matched=True, storage=dt)
        torch.cuda.synchronize()
        runs[name] = (u, res)
        _say(f"  split Bregman {name}: residuals {np.array2string(res, precision=3)}")
    u32 = runs["fp32"][0].double()
    rel = {k: float((runs[k][0].double() - u32).norm() / u32.norm()) for k in ("bf16", "fp16")}
    relx = float((u32 - x.double()).norm() / x.double().norm())
    _say(f"  rel L2 vs FP32 run: bf16 {rel['bf16']:.2e}, fp16 {rel['fp16']:.2e}; FP32 run vs phantom {relx:.3f}")
    assert np.all(np.isfinite(runs["fp32"][1]))
    assert rel["fp16"] < 2e-4 and rel["bf16"] < 1e-3

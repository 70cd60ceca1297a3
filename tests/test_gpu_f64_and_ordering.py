"""The FP64 projector pair and FP64 CGLS against the CPU oracle; the
frame-list sinogram order of the cone drop-ins; ordering of plan work
issued on several streams; argument checks of forward_norm and the drop-in
yNorm output."""
import ctypes as C
import sys

import numpy as np
import pytest
import torch

from helpers import bits_equal, first_mismatch, small_cone, small_fan

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _cone_dz():
    from framelab_amd import geometry as G
    return G.cone_para(48, 40, 20, N=(32, 32, 24), L=(2.0, 2.0, 1.2), SO=4.0, OD=4.0,
                       Ly=G.full_fan_width(), Lz=G.full_cone_height())


def _geom(kind):
    from framelab_amd import geometry as G
    if kind == "circle":
        return small_cone(), True
    if kind == "helix":
        return small_cone(kind="helix", zmax=0.5, rps=1.0, vtab=1.0, fps=24.0), True
    if kind == "dz":
        return _cone_dz(), True
    if kind == "frames":
        base = small_cone(nv=24)
        return G.with_frames(base, np.arange(base.nv) % 3), True
    if kind == "c128":
        return G.cone_para(128, 96, 12, N=(128, 128, 128), SO=4.0, OD=4.0, Ly=G.full_fan_width(),
                           Lz=G.full_cone_height()), True
    return small_fan(), False


def _frames_sum_oracle(fn, arr, p, per_frame_in, cone):
    """The oracle's FP64 pair works on one frame: apply it per frame (views
    of frame f only, a one-frame geometry of those angles) and assemble."""
    from framelab_amd import geometry as G
    out = []
    for f in range(p.nt):
        views = np.nonzero(np.asarray(p.id_X) == f)[0]
        sub = G.cone_para(p.na, p.nb, int(views.size), N=(p.nx, p.ny, p.nz), SO=4.0, OD=4.0,
                          Ly=G.full_fan_width(), Lz=G.full_cone_height(), sd_phi=p.sd_phi[views])
        assert np.array_equal(sub.y_det, p.y_det) and np.array_equal(sub.z_det, p.z_det)
        out.append((views, fn(per_frame_in(arr, f, views), sub, cone=cone)))
    return out


@pytest.mark.parametrize("kind", ["circle", "helix", "dz", "frames", "fan", "c128"])
def test_f64_pair_vs_oracle(gpu, oracle, kind):
    """forward_f64 / adjoint_f64 equal the oracle's FP64 Siddon and its
    exact transpose (oracle/ct_oracle.c: orc_siddon_f64, orc_siddon_adj_f64)
    to FP64 rounding (relative max-norm 1e-12), up to 128^3."""
    from framelab_amd.projector import Projector
    p, cone = _geom(kind)
    _say(f"case: FP64 pair vs oracle ({kind})")
    P = Projector(p)
    rng = np.random.default_rng(41)
    X = rng.random(P.vol_shape)
    Y = rng.standard_normal(P.sino_shape)
    y = P.forward_f64(torch.from_numpy(X).to(gpu)).cpu().numpy().reshape(p.nv, -1)
    x = P.adjoint_f64(torch.from_numpy(Y).to(gpu)).cpu().numpy()
    if p.nt == 1:
        yo = oracle.siddon_f64(X, p, cone=cone).reshape(p.nv, -1)
        xo = oracle.siddon_adj_f64(Y, p, cone=cone).reshape(x.shape)
    else:
        nd = Y.reshape(p.nv, -1).shape[1]
        yo = np.zeros_like(y)
        for views, yf in _frames_sum_oracle(oracle.siddon_f64, X, p, lambda a, f, v: a[f], cone):
            yo[views] = yf.reshape(views.size, nd)
        xo = np.zeros_like(x)
        for f, (views, xf) in enumerate(_frames_sum_oracle(
                oracle.siddon_adj_f64, Y.reshape(p.nv, -1), p, lambda a, f, v: a[v], cone)):
            xo[f] = xf.reshape(x.shape[1:])
    ey = np.abs(y - yo).max() / np.abs(yo).max()
    ex = np.abs(x - xo).max() / np.abs(xo).max()
    _say(f"  forward rel {ey:.2e}, adjoint rel {ex:.2e}")
    assert ey <= 1e-12 and ex <= 1e-12


def test_f64_dot(gpu):
    """<A x, y> = <x, A^T y> for the FP64 pair to 1e-13 (Gaussian x, y)."""
    from framelab_amd.projector import Projector
    _say("case: FP64 pair dot test")
    for p in (small_cone(), _cone_dz(), small_fan()):
        P = Projector(p)
        g = torch.Generator(device="cpu").manual_seed(42)
        x = torch.randn(P.vol_shape, generator=g, dtype=torch.float64).to(gpu)
        y = torch.randn(P.sino_shape, generator=g, dtype=torch.float64).to(gpu)
        lhs = torch.dot(P.forward_f64(x).ravel(), y.ravel()).item()
        rhs = torch.dot(x.ravel(), P.adjoint_f64(y).ravel()).item()
        _say(f"  dot rel {abs(lhs - rhs) / abs(lhs):.2e}")
        assert abs(lhs - rhs) <= 1e-13 * abs(lhs)


def test_cgls_f64_vs_oracle(gpu, oracle):
    """fl_cgls_f64 (FP64 vectors, FP64 pair) equals the oracle's FP64 CGLS
    recurrence over the oracle's FP64 pair: 12 iterations on a 24^3 cone
    problem within 1e-9 relative, and the same residual norms r.r."""
    from framelab_amd import geometry as G
    from framelab_amd.projector import Projector
    from framelab_amd.solvers import cgls_f64
    _say("case: FP64 CGLS vs oracle (24^3)")
    p = G.cone_para(40, 32, 18, N=(24, 24, 24), SO=4.0, OD=4.0, Ly=G.full_fan_width(),
                    Lz=G.full_cone_height())
    P = Projector(p)
    rng = np.random.default_rng(43)
    X = rng.random((24, 24, 24))
    A = lambda v: oracle.siddon_f64(v, p)
    At = lambda w: oracle.siddon_adj_f64(w, p)
    b = At(A(X))
    log = []
    want = oracle.cgls(A, At, b, np.zeros_like(b), 0.0, 12, log=log)
    u, glog, done = cgls_f64(P, torch.from_numpy(b.reshape(P.vol_shape)).to(gpu),
                             torch.zeros(P.vol_shape, dtype=torch.float64, device=gpu), iters=12)
    got = u.cpu().numpy().ravel()
    rel = np.linalg.norm(got - want) / np.linalg.norm(want)
    _say(f"  rel {rel:.2e}, iters {done}")
    assert done == 12 and rel <= 1e-9
    assert np.allclose(glog[1:], np.array(log), rtol=1e-8, atol=0.0)


def test_cone_dropin_frame_list_order(gpu, oracle):
    """The cone drop-ins order the sinogram by frame lists, as the reference
    GPU wrappers do (Ax_cone_mf_gpu_new.cu:500-517, Atx_cone_mf_gpu_new.cu:
    140-150): with nt = 3 frames and views dealt round-robin, slot k holds
    the k-th listed view.  The forward writes, and the adjoint reads, that
    order: both equal the plan results permuted accordingly, bit for bit,
    and the drop-in pair passes the dot test."""
    from framelab_amd import _lib, geometry as G
    from framelab_amd.projector import Projector
    _say("case: cone drop-ins, 3 frames round-robin, frame-list order")
    base = small_cone(nv=24)
    p = G.with_frames(base, np.arange(base.nv) % 3)
    order = np.concatenate([np.asarray(p.id_Y).reshape(p.nt, p.tmp_size)[f, :p.Nv[f]] for f in range(p.nt)])
    assert not np.array_equal(order, np.arange(p.nv))
    L = _lib.lib()
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
    F = lambda *v: [float(t) for t in v]
    rng = np.random.default_rng(44)
    X = rng.random((p.nt, p.nz, p.ny, p.nx), dtype=np.float32)
    nd = p.na * p.nb
    y = np.zeros(p.nv * nd, np.float32)
    L.Ax_cone_mf_gpu_siddon(fp(X), fp(y), fp(p.sd_phi), fp(p.sd_z), fp(p.y_det), fp(p.z_det),
                            ip(p.id_Y), ip(p.Nv), *F(p.SO, p.OD, p.scale, p.dz), p.nx, p.ny, p.nz, p.nt,
                            p.na, p.nb, p.nv, p.tmp_size, 4)
    _lib.check_last()
    P = Projector(p)
    yv = P.forward(torch.from_numpy(X).to(gpu)).cpu().numpy().reshape(p.nv, nd)   # view order
    assert bits_equal(y.reshape(p.nv, nd), yv[order])
    Ys = rng.standard_normal(p.nv * nd).astype(np.float32)                         # slot order
    x = np.zeros_like(X).ravel()
    L.Atx_cone_mf_gpu_new(fp(x), fp(Ys), fp(p.cos_phi), fp(p.sin_phi), fp(p.sd_z), fp(p.y_det),
                          fp(p.z_det), ip(p.id_Y), ip(p.Nv), *F(p.SO, p.OD, p.scale, p.dy_det, p.y_os,
                          p.dz_det, p.dz), p.nx, p.ny, p.nz, p.nt, p.na, p.nb, p.nv, p.tmp_size, 4)
    _lib.check_last()
    Yv = np.empty_like(Ys.reshape(p.nv, nd))
    Yv[order] = Ys.reshape(p.nv, nd)
    xv = P.adjoint(torch.from_numpy(Yv.reshape(P.sino_shape)).to(gpu)).cpu().numpy().ravel()
    assert bits_equal(x, xv), first_mismatch(x, xv)
    assert bits_equal(xv, oracle.atx_cone_f32(Yv, p))
    lhs = float(np.dot(y.astype(np.float64), Ys))
    rhs = float(np.dot(X.ravel().astype(np.float64), x))
    _say(f"  drop-in pair dot rel {abs(lhs - rhs) / abs(lhs):.2e}")
    assert abs(lhs - rhs) <= 1e-4 * abs(lhs)


def test_plan_slot_view_adjoint(gpu):
    """A plan with a permuted slot_view: the adjoint reads each view from the
    slot the forward wrote it to, so the matched pair stays a transpose."""
    from framelab_amd import _lib
    from framelab_amd.projector import ADJ_MATCHED, Projector, geometry_struct
    _say("case: plan with permuted output slots")
    p = small_cone(nv=16)
    perm = np.random.default_rng(45).permutation(p.nv).astype(np.int32)
    keep = []
    g = geometry_struct(p, keep)
    g.slot_view = perm.ctypes.data_as(C.POINTER(C.c_int))
    h = C.c_void_p()
    _lib.check(_lib.lib().fl_plan_create(C.byref(h), C.byref(g)))
    P0 = Projector(p)
    try:
        st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        x = torch.rand(P0.vol_shape, device=gpu)
        y = torch.empty(P0.sino_shape, device=gpu)
        _lib.check(_lib.lib().fl_forward(h, C.c_void_p(x.data_ptr()), 0, C.c_void_p(y.data_ptr()), 0, st))
        ref = P0.forward(x)
        assert torch.equal(y, ref[torch.from_numpy(perm).long().to(gpu)])
        for kind in (0, ADJ_MATCHED):
            back = torch.empty_like(x)
            _lib.check(_lib.lib().fl_adjoint(h, C.c_void_p(y.data_ptr()), 0, C.c_void_p(back.data_ptr()), 0,
                                             kind, st))
            assert torch.equal(back, P0.adjoint(ref, kind=kind))
    finally:
        _lib.lib().fl_plan_destroy(h)


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _st(s):
    return C.c_void_p(s.cuda_stream)


def test_graph_cgls_and_forward_on_two_streams(gpu):
    """A graph-replayed CGLS on one stream and a forward of the same plan on
    another (both write the plan's brick buffer), issued without host
    synchronisation in between, give the serial results."""
    from framelab_amd import _lib
    from framelab_amd.projector import Projector
    _say("case: graph CGLS and a forward on two streams")
    L = _lib.lib()
    P = Projector(small_cone(nv=20))
    g = torch.Generator(device="cpu").manual_seed(46)
    x = torch.rand(P.vol_shape, generator=g).to(gpu)
    x2 = torch.rand(P.vol_shape, generator=g).to(gpu)
    b = P.adjoint(P.forward(x))
    s0, s1, s2 = torch.cuda.current_stream(), torch.cuda.Stream(), torch.cuda.Stream()
    u_ref = torch.zeros_like(b)
    _lib.check(L.fl_cgls(P.handle, _ptr(b), _ptr(u_ref), 0.0, 6, 0.0, _lib.FL_SOLVE_GRAPH, None, None, _st(s0)))
    y_ref = torch.empty(P.sino_shape, device=gpu)
    _lib.check(L.fl_forward(P.handle, _ptr(x2), 0, _ptr(y_ref), 0, _st(s0)))
    torch.cuda.synchronize()
    for _ in range(3):
        u = torch.zeros_like(b)
        y = torch.empty_like(y_ref)
        torch.cuda.synchronize()
        _lib.check(L.fl_cgls(P.handle, _ptr(b), _ptr(u), 0.0, 6, 0.0, _lib.FL_SOLVE_GRAPH, None, None, _st(s1)))
        _lib.check(L.fl_forward(P.handle, _ptr(x2), 0, _ptr(y), 0, _st(s2)))
        torch.cuda.synchronize()
        assert torch.equal(u, u_ref)
        assert torch.equal(y, y_ref)


def test_data_gradients_on_two_streams(gpu):
    """Two data gradients of one plan issued on two streams without host
    synchronisation (they share the plan's residual workspace) equal the
    serial ones."""
    from framelab_amd import _lib
    from framelab_amd.projector import Projector
    _say("case: data gradients on two streams")
    L = _lib.lib()
    P = Projector(small_cone(nv=20))
    g = torch.Generator(device="cpu").manual_seed(47)
    xa, xb = (torch.rand(P.vol_shape, generator=g).to(gpu) for _ in range(2))
    b = P.forward(torch.rand(P.vol_shape, generator=g).to(gpu))
    s0, s1, s2 = torch.cuda.current_stream(), torch.cuda.Stream(), torch.cuda.Stream()
    ga, gb = torch.empty_like(xa), torch.empty_like(xb)
    _lib.check(L.fl_grad_data(P.handle, _ptr(xa), _ptr(b), _ptr(ga), None, 0, _st(s0)))
    _lib.check(L.fl_grad_data(P.handle, _ptr(xb), _ptr(b), _ptr(gb), None, 0, _st(s0)))
    torch.cuda.synchronize()
    for _ in range(3):
        g1, g2 = torch.empty_like(xa), torch.empty_like(xb)
        torch.cuda.synchronize()
        _lib.check(L.fl_grad_data(P.handle, _ptr(xa), _ptr(b), _ptr(g1), None, 0, _st(s1)))
        _lib.check(L.fl_grad_data(P.handle, _ptr(xb), _ptr(b), _ptr(g2), None, 0, _st(s2)))
        torch.cuda.synchronize()
        assert torch.equal(g1, ga) and torch.equal(g2, gb)


def test_forward_norm_checks_and_dropin_ynorm(gpu, oracle):
    """forward_norm refuses non-FP32 tensors (its kernel reads and writes
    FP32); the drop-in Ax_cone_mf_gpu_new's second output equals the plan's
    forward_norm bit for bit and y / (|d| scale) with |d| the FP64 ray length
    (Ax_cone_mf.cpp:55-57)."""
    from framelab_amd import _lib
    from framelab_amd.projector import Projector
    _say("case: yNorm checks")
    p = small_cone()
    P = Projector(p)
    with pytest.raises(TypeError):
        P.forward_norm(torch.zeros(P.sino_shape, dtype=torch.bfloat16, device=gpu))
    with pytest.raises(TypeError):
        P.forward_norm(torch.zeros(P.sino_shape, device=gpu),
                       out=torch.zeros(P.sino_shape, dtype=torch.float16, device=gpu))
    L = _lib.lib()
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
    rng = np.random.default_rng(48)
    X = rng.random(p.nx * p.ny * p.nz, dtype=np.float32)
    y = np.zeros(p.nv * p.na * p.nb, np.float32)
    yn = np.zeros_like(y)
    L.Ax_cone_mf_gpu_new(fp(X), fp(y), fp(yn), fp(p.sd_phi), fp(p.sd_z), fp(p.y_det), fp(p.z_det),
                         ip(p.id_Y), ip(p.Nv), float(p.SO), float(p.OD), float(p.scale), float(p.dz), p.nx,
                         p.ny, p.nz, 1, p.na, p.nb, p.nv, p.tmp_size, 4)
    _lib.check_last()
    yt = P.forward(torch.from_numpy(X.reshape(P.vol_shape)).to(gpu))
    assert bits_equal(y, yt.cpu().numpy().ravel())
    assert bits_equal(yn, P.forward_norm(yt).cpu().numpy().ravel())
    c, s = oracle.trig_f32(p.sd_phi)
    c, s = c.astype(np.float64), s.astype(np.float64)
    ya, zb = np.asarray(p.y_det, np.float64), np.asarray(p.z_det, np.float64)
    SO, OD = float(p.SO), float(p.OD)
    dx = (c[:, None, None] * OD - s[:, None, None] * ya[None, None, :]) + SO * c[:, None, None]
    dy = (s[:, None, None] * OD + c[:, None, None] * ya[None, None, :]) + SO * s[:, None, None]
    dz = zb[None, :, None]
    L64 = np.sqrt(dx ** 2 + dy ** 2 + dz ** 2)
    want = y.reshape(L64.shape) / (L64 * float(p.scale))
    rel = np.abs(yn.reshape(L64.shape) - want).max() / np.abs(want).max()
    _say(f"  drop-in yNorm vs closed form: max rel {rel:.2e}")
    assert rel <= 1e-6


@pytest.mark.parametrize("geom", ["fan", "cone"])
def test_split_bregman_f64_vs_oracle(gpu, oracle, geom):
    """fl_split_bregman_f64 (FP64 pair, FP64 one-level framelets) equals the
    oracle's FP64 split Bregman (solveSplitBregman.m:16-41 over the oracle's
    FP64 Siddon pair and numpy framelets): fan 48^2 with the 9-band linear
    frame (4 outer x 5 inner), cone 24^3 with the 27-band frame (3 x 3),
    within 1e-9 relative L2, residual logs within 1e-8."""
    from framelab_amd import geometry as G
    from framelab_amd.phantoms import ellipses_2d, ellipsoids_3d
    from framelab_amd.projector import Projector
    from framelab_amd.solvers import split_bregman_f64
    _say(f"case: FP64 split Bregman vs oracle ({geom})")
    cone = geom == "cone"
    if cone:
        p = G.cone_para(40, 32, 30, N=(24, 24, 24), SO=4.0, OD=4.0, Ly=G.full_fan_width(), Lz=G.full_cone_height())
        X = ellipsoids_3d(24, count=5, seed=6).astype(np.float64)
        shape, outer, inner, dec, rec, dim = (24, 24, 24), 3, 3, oracle.fradec3d, oracle.frarec3d, 3
    else:
        p = small_fan(nd=96, nv=60, N=48)
        X = ellipses_2d(48, 6, seed=4).astype(np.float64)
        shape, outer, inner, dec, rec, dim = (48, 48), 4, 5, oracle.fradec2d, oracle.frarec2d, 2
    f64 = oracle.siddon_f64(X, p, cone=cone)
    D, R = oracle.framelet_filters("linear")
    A = lambda v: oracle.siddon_f64(v.ravel(), p, cone=cone)
    At = lambda y: oracle.siddon_adj_f64(y, p, cone=cone).reshape(shape)
    W = lambda u: np.stack(dec(u.reshape(shape), D, 1))
    Wt = lambda c: rec(list(c), R)
    shrink = lambda c, t: np.stack(oracle.soft_threshold(list(c), t, dim))
    lam, mu = 1e-2, 1.0
    want = oracle.split_bregman(A, At, W, Wt, lam, mu, f64, np.zeros(shape), outer, inner, shrink=shrink)
    P = Projector(p)
    f = torch.from_numpy(f64).reshape(P.sino_shape).to(gpu)
    u, res = split_bregman_f64(P, "linear", f, torch.zeros(P.vol_shape, dtype=torch.float64, device=gpu), lam, mu,
                               outer, inner)
    got = u.cpu().numpy().ravel()
    rel = np.linalg.norm(got - want.ravel()) / np.linalg.norm(want)
    _say(f"  rel {rel:.2e}, residuals {res}")
    assert rel <= 1e-9
    assert np.all(np.isfinite(res)) and res[-1] < res[0]


def test_split_bregman_iterate_history(gpu):
    """The per-outer-step iterate history of both split-Bregman entry points:
    entry k equals, bit for bit, a fresh solve with k+1 outer steps, and the
    last entry equals the returned iterate (FP32 graph-replayed with the
    matched adjoint and FP16 storage, and FP64)."""
    from framelab_amd.phantoms import ellipses_2d
    from framelab_amd.projector import Projector
    from framelab_amd.solvers import split_bregman, split_bregman_f64
    _say("case: split Bregman iterate history")
    p = small_fan(nd=96, nv=60, N=48)
    P = Projector(p)
    x = torch.from_numpy(ellipses_2d(48, 6, seed=4).astype(np.float32)).to(gpu)
    f = P.forward(x)
    u0 = torch.zeros(P.vol_shape, device=gpu)
    outer, inner, lam, mu = 4, 3, 1e-2, 1.0
    kw = dict(graph=True, matched=True, storage=torch.float16)
    u, res, hist = split_bregman(P, "linear", f, u0, lam, mu, outer, inner, history=True, **kw)
    assert hist.shape == (outer,) + tuple(P.vol_shape)
    assert torch.equal(hist[-1], u)
    for k in range(outer - 1):
        uk, _ = split_bregman(P, "linear", f, u0, lam, mu, k + 1, inner, **kw)
        assert torch.equal(hist[k], uk), k
    u64, res64, h64 = split_bregman_f64(P, "linear", f.double(), u0.double(), lam, mu, outer, inner, history=True)
    assert torch.equal(h64[-1], u64)
    uk64, _ = split_bregman_f64(P, "linear", f.double(), u0.double(), lam, mu, 2, inner)
    assert torch.equal(h64[1], uk64)

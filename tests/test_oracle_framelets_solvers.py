"""Pin the numpy oracle of the framelet transforms and solvers (CPU).

The reference's framelet arithmetic is MATLAB imfilter (absent) and its
solvers cannot run (see the split-Bregman notes in oracle/oracle.py), so the
oracle is pinned by the library routine imfilter(...,'circular') reduces to
(scipy.ndimage.correlate1d, mode='wrap'), by the identities the reference's
test/testFramelet.m checks (W^T W = I), by tight-frame and adjointness
identities, and by a dense least-squares solve for CGLS.
"""
import numpy as np
import pytest
from scipy.ndimage import correlate1d


@pytest.mark.parametrize("kind", ["haar", "linear", "cubic"])
@pytest.mark.parametrize("level", [1, 2, 3])
def test_corr_circ_is_imfilter_circular(oracle, kind, level):
    rng = np.random.default_rng(1)
    A = rng.standard_normal((12, 14, 16))
    D, _ = oracle.framelet_filters(kind)
    step = 2 ** (level - 1)
    for taps in D:
        ker = np.zeros(step * (len(taps) - 1) + 1)
        ker[::step] = taps
        for axis in range(3):
            want = correlate1d(A, ker, axis=axis, mode="wrap", origin=0)
            assert np.allclose(oracle.corr_circ(A, taps, axis, level), want, atol=1e-13)


@pytest.mark.parametrize("kind", ["haar", "linear", "cubic"])
def test_perfect_reconstruction_and_tight_frame(oracle, kind):
    """test/testFramelet.m:17-20: ||W^T W u - u|| / ||u|| ~ 0 (2D and 3D)."""
    rng = np.random.default_rng(2)
    D, R = oracle.framelet_filters(kind)
    u2 = rng.standard_normal((16, 20))
    u3 = rng.standard_normal((8, 10, 12))
    for level in (1, 2):
        C2 = oracle.fradec2d(u2, D, level)
        assert np.linalg.norm(oracle.frarec2d(C2, R) - u2) < 1e-12 * np.linalg.norm(u2)
        C3 = oracle.fradec3d(u3, D, level)
        assert np.linalg.norm(oracle.frarec3d(C3, R) - u3) < 1e-12 * np.linalg.norm(u3)
    # tight frame: |W u| = |u| for one level
    assert abs(np.linalg.norm(oracle.fradec2d(u2, D, 1)[0]) - np.linalg.norm(u2)) < 1e-12 * np.linalg.norm(u2)
    assert abs(np.linalg.norm(oracle.fradec3d(u3, D, 1)[0]) - np.linalg.norm(u3)) < 1e-12 * np.linalg.norm(u3)


def test_synthesis_is_adjoint_and_constant_input(oracle):
    rng = np.random.default_rng(3)
    D, R = oracle.framelet_filters("linear")
    u = rng.standard_normal((6, 8, 10))
    c = rng.standard_normal((3, 3, 3, 6, 8, 10))
    lhs = np.sum(oracle.fradec3d(u, D, 1)[0] * c)
    rhs = np.sum(u * oracle.frarec3d([c], R))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # a constant volume has only a low-pass response (filters 2.. sum to zero)
    bands = oracle.fradec3d(np.full((6, 6, 6), 2.5), D, 1)[0]
    assert np.allclose(bands[0, 0, 0], 2.5)
    bands[0, 0, 0] = 0
    assert np.max(np.abs(bands)) < 1e-14
    # band order: band (i,j,k) applies D_i along y, D_j along x, D_k along z
    ramp = np.tile(np.arange(10.0), (6, 8, 1))            # varies along x only
    b = oracle.fradec3d(ramp, D, 1)[0]
    assert np.max(np.abs(b[0, 1, 0][:, :, 2:-2])) > 0.1   # x-derivative band responds
    assert np.max(np.abs(b[1, 0, 0])) < 1e-13             # y-derivative band does not


def test_soft_threshold_keeps_lowpass(oracle):
    C = [np.array([[[1.5, -0.2]], [[-3.0, 0.7]]])]       # 2 bands of shape (1,2)
    out = oracle.soft_threshold(C, 1.0, 1)[0]
    assert np.allclose(out[0], C[0][0])
    assert np.allclose(out[1], [[-2.0, 0.0]])


def test_cgls_matches_lstsq(oracle):
    """CG on the normal equations converges to the regularised LS solution."""
    rng = np.random.default_rng(4)
    M = rng.standard_normal((60, 25))
    f = rng.standard_normal(60)
    lam = 0.3
    A = lambda v: M @ v
    At = lambda y: M.T @ y
    b = At(f)
    u = oracle.cgls(A, At, b, np.zeros(25), lam, 60)
    want = np.linalg.solve(M.T @ M + lam * np.eye(25), b)
    assert np.linalg.norm(u - want) < 1e-10 * np.linalg.norm(want)
    log = []
    oracle.cgls(A, At, b, np.zeros(25), lam, 5, log=log)
    assert all(log[k + 1] < log[0] for k in range(4))


def test_split_bregman_denoising_fixed_point(oracle):
    """With A = I the split-Bregman iterates decrease the objective
    0.5|u-f|^2 + lam |W_hp u|_1 and end near a fixed point alpha = shrink(v+Wu)."""
    rng = np.random.default_rng(5)
    D, R = oracle.framelet_filters("linear")
    f = np.zeros((24, 24))
    f[6:18, 6:18] = 1.0
    f += 0.1 * rng.standard_normal(f.shape)
    W = lambda u: oracle.fradec2d(u, D, 1)[0]
    Wt = lambda c: oracle.frarec2d([c], R)
    shrink = lambda c, t: oracle.soft_threshold([c], t, 2)[0]
    lam, mu = 0.05, 1.0
    obj = lambda u: 0.5 * np.sum((u - f) ** 2) + lam * np.sum(np.abs(W(u)[1:])) + lam * np.sum(np.abs(W(u)[0, 1:]))
    I = lambda v: v
    u1 = oracle.split_bregman(I, I, W, Wt, lam, mu, f, f.copy(), 1, 20, shrink=shrink)
    u8 = oracle.split_bregman(I, I, W, Wt, lam, mu, f, f.copy(), 8, 20, shrink=shrink)
    assert obj(u8) < obj(f)
    assert obj(u8) <= obj(u1) + 1e-9


def _highpass_matrix(oracle, kind, shape, level):
    """The analysis rows the split-Bregman objective penalises, as a dense
    matrix: every band but the low-pass one of every level (level L analyses
    the low-pass band of level L-1), built column by column from fradec2d."""
    D, _ = oracle.framelet_filters(kind)
    n = shape[0] * shape[1]
    cols = []
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        C = oracle.fradec2d(e.reshape(shape), D, level)
        rows = []
        for c in C:
            flat = c.reshape(-1, n)
            rows.append(flat[1:])                 # band (0, 0) is the low-pass
        cols.append(np.concatenate(rows).ravel())
    return np.stack(cols, axis=1)


@pytest.mark.parametrize("level", [1, 2])
def test_split_bregman_reaches_the_minimiser(oracle, level):
    """With A = I, split Bregman (solveSplitBregman.m:16-41, class form) must
    converge to the minimiser of 0.5|u - f|^2 + lam |B u|_1, B = the
    high-pass analysis rows of the `level`-level linear frame (its fixed
    point: alpha = shrink(v + W u) with W u = alpha on those rows).  The
    minimiser is computed independently from the dual problem
    min_{|p|_inf <= lam} 0.5|f - B^T p|^2, a box-constrained least-squares
    problem (scipy's lsq_linear, bounded-variable LS), u* = f - B^T p*, and
    checked for optimality against random perturbations.  A dropped
    Bregman term, a wrong sign in v's update or a shrunk low-pass band moves
    the iterate off u*."""
    from scipy.optimize import lsq_linear
    rng = np.random.default_rng(6)
    shape = (8, 8)
    f = np.zeros(shape)
    f[2:6, 1:5] = 1.0
    f += 0.2 * rng.standard_normal(shape)
    lam, mu = 0.08, 4.0
    B = _highpass_matrix(oracle, "linear", shape, level)
    sol = lsq_linear(B.T, f.ravel(), bounds=(-lam, lam), method="bvls", tol=1e-15, max_iter=10000)
    ustar = f.ravel() - B.T @ sol.x
    obj = lambda v: 0.5 * np.sum((v - f.ravel()) ** 2) + lam * np.abs(B @ v).sum()
    for _ in range(20):
        assert obj(ustar) <= obj(ustar + 1e-3 * rng.standard_normal(ustar.size)) + 1e-12
    D, R = oracle.framelet_filters("linear")
    W = lambda u: np.stack(oracle.fradec2d(u.reshape(shape), D, level))
    Wt = lambda c: oracle.frarec2d(list(c), R)
    shrink = lambda c, t: np.stack(oracle.soft_threshold(list(c), t, 2))
    I = lambda v: v
    u = oracle.split_bregman(I, I, W, Wt, lam, mu, f, f.copy(), 3000, 2, shrink=shrink)
    err = np.linalg.norm(u.ravel() - ustar) / np.linalg.norm(ustar)
    assert err < 1e-9, err

"""John's equation oracle (the reference's sparse assembly restated with
scipy) pinned by closed forms: on polynomials whose differences are exact,
D g equals R g_tb - R SO g_az - R h g_bz + 2 a g_b + a b g_bb + (a^2+R^2) g_ab
evaluated analytically; D^T D is symmetric positive semi-definite."""
import numpy as np
import pytest

P = dict(R=8.0, SO=4.0, h=0.1591549, dzeta=0.2, dtheta=0.05, da=0.03, db=0.04)


def _grid(na, nb, nv, nz):
    a = -0.6 + P["da"] * np.arange(na)
    b = -0.3 + P["db"] * np.arange(nb)
    th = P["dtheta"] * np.arange(nv)
    ze = P["dzeta"] * np.arange(nz)
    Z, T, B, A = np.meshgrid(ze, th, b, a, indexing="ij")
    return a, b, A, B, T, Z


@pytest.mark.parametrize("nz", [2, 4])
def test_john_oracle_polynomial_closed_form(oracle, nz):
    na, nb, nv = 9, 8, 7
    a, b, A, B, T, Z = _grid(na, nb, nv, nz)
    # f = 1.5 a b + 0.7 b t + 0.9 a z + 1.1 b z - 0.4 b^2 + 0.3 a^2 b + 0.2 t b^2
    f = 1.5 * A * B + 0.7 * B * T + 0.9 * A * Z + 1.1 * B * Z - 0.4 * B ** 2 + 0.3 * A ** 2 * B + 0.2 * T * B ** 2
    f_tb = 0.7 + 0.4 * B
    f_az = 0.9
    f_bz = 1.1
    f_b = 1.5 * A + 0.7 * T + 1.1 * Z - 0.8 * B + 0.3 * A ** 2 + 0.4 * T * B
    f_bb = -0.8 + 0.4 * T
    f_ab = 1.5 + 0.6 * A
    R, SO, h = P["R"], P["SO"], P["h"]
    want = R * f_tb - R * SO * f_az - R * h * f_bz + 2 * A * f_b + A * B * f_bb + (A ** 2 + R ** 2) * f_ab
    zi = slice(1, -1) if nz > 2 else slice(0, 1)
    want = want[zi, 1:-1, 1:-1, 1:-1]
    M = oracle.john_matrix(na, nb, nv, nz, R, SO, h, P["dzeta"], P["dtheta"], P["da"], P["db"], a, b)
    got = oracle.john_apply(f, M).reshape(want.shape)
    assert np.allclose(got, want, rtol=1e-11, atol=1e-10 * np.abs(want).max())


def test_john_oracle_normal_symmetric(oracle):
    na, nb, nv, nz = 8, 7, 6, 3
    a, b, *_ = _grid(na, nb, nv, nz)
    M = oracle.john_matrix(na, nb, nv, nz, P["R"], P["SO"], P["h"], P["dzeta"], P["dtheta"], P["da"], P["db"], a, b)
    rng = np.random.default_rng(0)
    x, y = rng.standard_normal((2, nz * nv * nb * na))
    Nx, Ny = oracle.john_normal(x, M), oracle.john_normal(y, M)
    assert abs(np.dot(Nx, y) - np.dot(x, Ny)) <= 1e-10 * abs(np.dot(Nx, y))
    assert np.dot(Nx, x) >= 0

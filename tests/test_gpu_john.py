"""John's equation operators on the GPU against the FP64 sparse oracle (the
reference's assembly restated), and on a multi-helix scan of a smooth
object, where the equation nearly holds."""
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _scanner_geom(na=20, nb=14, nvh=12, nz=2):
    """John constants of a real multiHelix geometry (detector coordinates from setPara)."""
    from framelab_amd import geometry as G
    from framelab_amd.john import john_geometry

    class S:   # the scanner fields applyJohn.m reads
        pass
    s = S()
    s.na, s.nb, s.nHelix, s.SO, s.OD = na, nb, nz, 4.0, 4.0
    s.vtab, s.rps, s.fps = 1.0, 1.0, 16.0
    s.phaseShift = tuple(0.4 * k for k in range(nz))
    s.para = G.cone_para(na, nb, 0, N=(16, 16, 16), SO=4.0, OD=4.0, Ly=G.full_fan_width(),
                         Lz=G.full_cone_height(), kind="multiHelix", zmax=0.5, rps=1.0, vtab=1.0,
                         fps=16.0, nHelix=nz, phaseShift=s.phaseShift)
    g = john_geometry(s)
    g["nv"] = nvh          # any view count per helix for the operator test
    return g


@pytest.mark.parametrize("nz", [2, 3, 4])
def test_john_vs_oracle(gpu, oracle, nz):
    """D g (FP32 and FP64 input), D^T u and D^T D g within 1e-12 of the oracle,
    relative to |D| |g| elementwise."""
    from framelab_amd.john import JohnOperator
    _say(f"case: John operator nz={nz}")
    geom = _scanner_geom(nz=nz)
    J = JohnOperator(geom)
    M = oracle.john_matrix(geom["na"], geom["nb"], geom["nv"], geom["nz"], geom["R"], geom["SO"], geom["h"],
                           geom["dzeta"], geom["dtheta"], geom["da"], geom["db"], geom["a"], geom["b"])
    Mabs = abs(M)
    rng = np.random.default_rng(20 + nz)
    g = rng.standard_normal(J.in_shape).astype(np.float32)
    for dt in (torch.float32, torch.float64):
        got = J.apply(torch.from_numpy(g).to(gpu, dt)).cpu().numpy().ravel()
        want = oracle.john_apply(g, M)
        bound = Mabs @ np.abs(g.astype(np.float64)).ravel()
        assert np.all(np.abs(got - want) <= 1e-12 * bound + 1e-300), np.max(np.abs(got - want) / bound)
    u = rng.standard_normal(J.out_shape)
    got = J.adjoint(torch.from_numpy(u).to(gpu)).cpu().numpy().ravel()
    want = M.T @ u.ravel()
    bound = Mabs.T @ np.abs(u).ravel()
    assert np.all(np.abs(got - want) <= 1e-12 * bound + 1e-300)
    got = J.normal(torch.from_numpy(g).to(gpu)).cpu().numpy().ravel()
    want = oracle.john_normal(g, M)
    bound = Mabs.T @ (Mabs @ np.abs(g.astype(np.float64)).ravel())
    assert np.all(np.abs(got - want) <= 1e-11 * bound + 1e-300)


def _john_terms(g, geom):
    """The six terms of the discretised equation separately (numpy, FP64),
    two-helix forward difference across helices; each (1, nv-2, nb-2, na-2)."""
    g = g.astype(np.float64)
    R, SO, h = geom["R"], geom["SO"], geom["h"]
    da, db, dt, dz = geom["da"], geom["db"], geom["dtheta"], geom["dzeta"]
    a = geom["a"][None, None, 1:-1]
    b = geom["b"][None, 1:-1, None]
    c = lambda x: x[1:-1, 1:-1, 1:-1]          # (t, b, a) interior of one helix
    g0, g1 = g[0], g[1]
    gtb = (g0[2:, 2:, 1:-1] - g0[2:, :-2, 1:-1] - g0[:-2, 2:, 1:-1] + g0[:-2, :-2, 1:-1]) / (4 * db * dt)
    dzg = (g1 - g0) / dz
    gaz = (dzg[1:-1, 1:-1, 2:] - dzg[1:-1, 1:-1, :-2]) / (2 * da)
    gbz = (dzg[1:-1, 2:, 1:-1] - dzg[1:-1, :-2, 1:-1]) / (2 * db)
    gb = (g0[1:-1, 2:, 1:-1] - g0[1:-1, :-2, 1:-1]) / (2 * db)
    gbb = (g0[1:-1, 2:, 1:-1] - 2 * c(g0) + g0[1:-1, :-2, 1:-1]) / db ** 2
    gab = (g0[1:-1, 2:, 2:] - g0[1:-1, 2:, :-2] - g0[1:-1, :-2, 2:] + g0[1:-1, :-2, :-2]) / (4 * da * db)
    return [R * gtb, -R * SO * gaz, -R * h * gbz, 2 * a * gb, a * b * gbb, (a ** 2 + R ** 2) * gab]


def test_john_on_multihelix_scan(gpu, oracle):
    """End to end on a two-helix scan of a smooth object: the scanner's
    normalised output feeds CTData.applyJohn, whose result equals the sum of
    the six terms formed separately with numpy differences.  (Whether the
    terms cancel depends on the reference's sign and unit conventions for
    the helix and view coordinates, which the operator reproduces as they
    are; the residual ratio is reported, not asserted.)"""
    from framelab_amd.datatypes import ObjectData
    from framelab_amd.scanners import ConeBeamScanner
    from framelab_amd.john import JohnOperator, john_geometry
    from framelab_amd import geometry as G
    _say("case: John residual on a two-helix scan")
    N = 48
    c = (np.arange(N) + 0.5) / N * 2 - 1
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    vol = np.exp(-((X / 0.5) ** 2 + (Y / 0.4) ** 2 + (Z / 0.45) ** 2)).astype(np.float32)
    sc = ConeBeamScanner("multiHelix", 64, 48, 0, zmax=0.4, rps=1.0, vtab=1.0, fps=90.0, nHelix=2,
                         phaseShift=(0.0, 0.1), SO=4.0, OD=4.0, Ly=G.full_fan_width(), Lz=G.full_cone_height())
    d = sc.apply(ObjectData(3, torch.from_numpy(vol).to(gpu), (2.0, 2.0, 2.0)))
    g = d.dataArrayNorm.cpu().numpy()
    geom = john_geometry(sc)
    D = d.applyJohn().cpu().numpy()[0]
    terms = _john_terms(g, geom)
    assert np.allclose(D, sum(terms), rtol=0, atol=1e-9 * max(np.abs(t).max() for t in terms))
    r = np.linalg.norm(D) / max(np.linalg.norm(t) for t in terms)
    gi = g.copy()
    gi[1] = np.roll(g[1], 3, axis=1)
    J = JohnOperator(geom)
    Di = J.apply(torch.from_numpy(gi).to(gpu)).cpu().numpy()[0]
    ri = np.linalg.norm(Di) / max(np.linalg.norm(t) for t in _john_terms(gi, geom))
    _say(f"  ||D g|| / max term: scan {r:.3e}, second helix shifted {ri:.3e}")
    assert np.isfinite(r) and np.isfinite(ri)
    # normal operator on the same data: D^T D g is the adjoint stencil of D g
    N2 = d.applyJohnNormal().cpu().numpy()
    assert np.allclose(N2, J.adjoint(torch.from_numpy(D[None]).to(gpu)).cpu().numpy(), rtol=0,
                       atol=1e-12 * np.abs(N2).max())

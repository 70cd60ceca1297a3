"""Analytic projections of a ball (the reference's test/testBall.m:23-117 for
circular, helical and multi-helix scans, closed form utils/coneBeamSphere.m:
10-16).  The line from the source at height zs to the detector point (a, b),
b measured from the source height, meets the unit ball in a chord of length
2 sqrt(((Rs R - b zs)^2 - (a^2+b^2+R^2)(Rs^2+zs^2-1)) / (a^2+b^2+R^2)),
R = Rs + Rd, a quadratic in the line parameter solved in closed form, and
the normalised projection divides it by |xs - xd| = sqrt(a^2+b^2+R^2).  A
voxelised ball (centres inside) projects to it up to the staircase error,
which halves when the voxels do."""
import numpy as np
import pytest
import torch

SO, OD = 4.0, 4.0


def _para(kind, N, na=96, nb=80):
    from framelab_amd import geometry as G
    kw = dict(N=(N, N, N), L=(2.2, 2.2, 2.2), SO=SO, OD=OD, Ly=5.0, Lz=9.0, kind=kind)
    if kind == "circle":
        return G.cone_para(na, nb, 24, **kw)
    kw.update(zmax=1.0, rps=1.0, vtab=1.0, fps=30.0)
    if kind == "multiHelix":
        kw.update(nHelix=2, phaseShift=(0.0, np.pi))
    return G.cone_para(na, nb, 0, **kw)


def _ball(N):
    c = -1.1 + (np.arange(N) + 0.5) * 2.2 / N
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    return (X ** 2 + Y ** 2 + Z ** 2 <= 1.0).astype(np.float32)


def _closed_form(p, normalized):
    s = float(p.scale)
    Rs, Rd = float(p.SO) * s, float(p.OD) * s
    R = Rs + Rd
    a = p.y_det.astype(np.float64)[None, None, :] * s
    b = p.z_det.astype(np.float64)[None, :, None] * s
    zs = p.sd_z.astype(np.float64)[:, None, None] * s
    q = a ** 2 + b ** 2 + R ** 2
    disc = np.maximum((Rs * R - b * zs) ** 2 - q * (Rs ** 2 + zs ** 2 - 1.0), 0.0)
    return 2 * np.sqrt(disc) / q if normalized else 2 * np.sqrt(disc / q)


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


@pytest.mark.parametrize("kind", ["circle", "helix"])
def test_ball_oracle(oracle, kind):
    """The oracle's FP32 Siddon forward of a 48^3 voxelised ball against the
    closed form (pins the oracle's helical source trajectory)."""
    p = _para(kind, 48, na=64, nb=56)
    y = oracle.siddon_cone_f32(_ball(48), p).reshape(p.nv, p.nb, p.na)
    err = _rel(y, _closed_form(p, False))
    assert err < 4e-2, err


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["circle", "helix", "multiHelix"])
def test_ball_gpu(gpu, kind):
    """GPU forward and normalised forward (yNorm) of voxelised balls at 64^3
    and 128^3 against the closed forms; the error falls with the voxel size."""
    from framelab_amd.projector import Projector
    print(f"case: analytic ball, {kind}", flush=True)
    errs = []
    for N in (64, 128):
        p = _para(kind, N)
        P = Projector(p)
        y = P.forward(torch.from_numpy(_ball(N)).to(gpu))
        yn = P.forward_norm(y)
        e = _rel(y.double().cpu().numpy().reshape(p.nv, p.nb, p.na), _closed_form(p, False))
        en = _rel(yn.double().cpu().numpy().reshape(p.nv, p.nb, p.na), _closed_form(p, True))
        print(f"  N={N}: rel L2 {e:.3e}, normalised {en:.3e}", flush=True)
        assert abs(e - en) < 0.2 * e + 1e-6
        errs.append(e)
    assert errs[0] < 3e-2 and errs[1] < 0.7 * errs[0], errs

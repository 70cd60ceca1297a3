"""Edge cases of the projector pair on the GPU: the smallest volumes and
panels (2x2x2 voxels, 1x1 and 3x2 cells, a single view), a 4-D scan with a
frame that no view projects, and geometries the plan rejects with a status
and a message instead of launching.  Forward and exact adjoint bit-exact
against the oracle; the matched adjoint passes the dot test; a frame
without views back-projects to exact zeros."""
import math
import sys

import numpy as np
import pytest
import torch

from helpers import bits_equal, dot_gap, first_mismatch

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _pair_checks(P, p, X, Y, oracle, cone=True):
    from framelab_amd.projector import ADJ_EXACT, ADJ_MATCHED
    gpu = torch.device("cuda:0")
    y = P.forward(torch.from_numpy(X.reshape(P.vol_shape)).to(gpu)).cpu().numpy().ravel()
    yo = oracle.siddon_cone_f32(X, p) if cone else oracle.siddon_fan_f32(X, p)
    assert bits_equal(y, yo), first_mismatch(y, yo)
    Yt = torch.from_numpy(Y).reshape(P.sino_shape).to(gpu)
    x = P.adjoint(Yt, kind=ADJ_EXACT).cpu().numpy().ravel()
    xo = oracle.atx_cone_f32(Y, p) if cone else oracle.atx_fan_f32(Y, p)
    assert bits_equal(x, xo), first_mismatch(x, xo)
    xm = P.adjoint(Yt, kind=ADJ_MATCHED).cpu().numpy().ravel()
    assert dot_gap(y, Y, X, xm) <= 2e-6
    return y, x, xm


@pytest.mark.parametrize("na,nb", [(1, 1), (3, 2), (5, 7)])
def test_tiny_cone(gpu, oracle, na, nb):
    """2x2x2 voxels, one to a few cells, one and then three views."""
    from framelab_amd import geometry as G
    from framelab_amd.projector import Projector
    for nv in (1, 3):
        _say(f"case: tiny cone 2^3, panel {na}x{nb}, {nv} views")
        p = G.cone_para(na, nb, nv, N=(2, 2, 2), SO=4.0, OD=4.0, Ly=2.0 * na / max(na, 2), Lz=2.0 * nb / max(nb, 2))
        rng = np.random.default_rng(na * 10 + nb + nv)
        X = rng.random(8).astype(np.float32)
        Y = rng.standard_normal(nv * na * nb).astype(np.float32)
        y, _, _ = _pair_checks(Projector(p), p, X, Y, oracle)
        assert np.all(np.isfinite(y))


@pytest.mark.parametrize("nd", [1, 4])
def test_tiny_fan(gpu, oracle, nd):
    """A 2x2 image, one or four bins, one and then five views."""
    from framelab_amd import geometry as G
    from framelab_amd.projector import Projector
    for nv in (1, 5):
        _say(f"case: tiny fan 2^2, {nd} bins, {nv} views")
        p = G.fan_para(nd, nv, N=(2, 2), SO=4.0, OD=4.0, Ly=2.0)
        rng = np.random.default_rng(nd * 7 + nv)
        X = rng.random(4).astype(np.float32)
        Y = rng.standard_normal(nv * nd).astype(np.float32)
        _pair_checks(Projector(p), p, X, Y, oracle, cone=False)


def test_frame_without_views(gpu, oracle):
    """Three frames, views alternating between frames 0 and 2: frame 1 has
    no view (Nv = 0).  The forward reads each view's own frame; the exact
    and matched adjoints leave frame 1 at exact zeros and match the oracle
    on the others."""
    from framelab_amd import geometry as G
    from framelab_amd.projector import ADJ_MATCHED, Projector
    from helpers import small_cone
    _say("case: cone 16^3, 3 frames, frame 1 without views")
    p = G.with_frames(small_cone(na=24, nb=20, nv=12, N=16), (np.arange(12) % 2) * 2)
    assert p.nt == 3 and int(p.Nv[1]) == 0
    rng = np.random.default_rng(5)
    X = rng.random(16 ** 3 * 3).astype(np.float32)
    Y = rng.standard_normal(12 * 24 * 20).astype(np.float32)
    P = Projector(p)
    _, x, xm = _pair_checks(P, p, X, Y, oracle)
    n = 16 ** 3
    assert not np.any(x[n:2 * n]) and not np.any(xm[n:2 * n])
    assert np.any(x[:n]) and np.any(x[2 * n:])


def test_plan_rejects_bad_geometry(gpu):
    """Zero sizes, odd sides, a fan panel with rows, a source inside the
    volume and a view index out of range come back as errors naming the
    problem; nothing is launched."""
    import dataclasses
    from framelab_amd import _lib, geometry as G
    from framelab_amd.projector import Projector
    from helpers import small_cone, small_fan
    base = small_cone(na=8, nb=6, nv=4, N=8)
    bad = [
        (dataclasses.replace(base, y_det=np.zeros(0, np.float32)), "non-positive"),
        (dataclasses.replace(base, nx=7), "even"),
        (dataclasses.replace(base, SO=np.float32(1.0)), "outside the volume"),
        (dataclasses.replace(base, id_Y=np.array([0, 1, 2, 9], dtype=np.int32)), "outside"),
    ]
    for p, words in bad:
        _say(f"case: rejected plan ({words})")
        with pytest.raises(_lib.FrameLabError) as e:
            Projector(p)
        assert words in str(e.value), str(e.value)
    fan = small_fan(nd=8, nv=4, N=8)
    with pytest.raises(_lib.FrameLabError):
        Projector(dataclasses.replace(fan, nx=0))

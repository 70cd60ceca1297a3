"""The oracle against the reference's own CPU kernels (oracle/_ref, compiled
from the reference sources in this container) at the headline cone
configuration's full volume and panel size: 256^3, 512 x 384 cells, two of
the 360 views.  Skipped where oracle/_ref was not built."""
import numpy as np
import pytest

from helpers import bits_equal, first_mismatch


def test_cone_oracle_equals_reference_at_full_size(oracle):
    if not oracle.has_ref():
        pytest.skip("oracle/_ref not built (reference checkout absent)")
    from framelab_amd import geometry as G
    from framelab_amd.phantoms import ellipsoids_3d
    p = G.cone_para(512, 384, 360, N=(256, 256, 256), SO=4.0, OD=4.0, Ly=G.full_fan_width(),
                    Lz=G.full_cone_height())
    sub = G.cone_para(512, 384, 2, N=(256, 256, 256), SO=4.0, OD=4.0, Ly=G.full_fan_width(),
                      Lz=G.full_cone_height(), sd_phi=p.sd_phi[[7, 131]])
    X = ellipsoids_3d(256, count=20, seed=2)
    oracle.set_threads(8)
    y = oracle.siddon_cone_f32(X, sub)
    yr = oracle.ref_siddon_cone(X, sub)
    assert bits_equal(y, yr), first_mismatch(y, yr)
    Y = np.random.default_rng(53).standard_normal(y.size).astype(np.float32)
    x = oracle.atx_cone_f32(Y, sub)
    xr = oracle.ref_atx_cone_cpu(Y, sub)
    assert bits_equal(x, xr), first_mismatch(x, xr)

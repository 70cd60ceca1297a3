"""Host-side checks of the data-type mirror (ObjectData.m, CTData.m): the
constructor defaults, arithmetic, dot and norm, the setter and
applyOperator, on CPU tensors (these methods launch no kernel)."""
import math

import numpy as np
import pytest
import torch

from framelab_amd.datatypes import CTData, ObjectData


def test_object_defaults_follow_reference_constructor():
    o = ObjectData(3)                                    # ObjectData.m:21-33
    assert tuple(o.dataArray.shape) == (128, 128, 128) and float(o.dataArray.abs().sum()) == 0.0
    assert o.N == (128, 128, 128) and o.L == (10.0, 10.0, 10.0)
    assert o.Dx == (10.0 / 128,) * 3
    o2 = ObjectData(2, np.ones((6, 4), np.float32))      # ObjectData.m:34-44: L = 10, N from the array
    assert o2.N == (4, 6) and o2.L == (10.0, 10.0) and o2.Dx == (2.5, 10.0 / 6)
    o3 = ObjectData(3, torch.zeros((4, 6, 8)), (2.0, 3.0, 1.0))
    assert o3.N == (8, 6, 4) and o3.Dx == (0.25, 0.5, 0.25)
    with pytest.raises(ValueError):
        ObjectData(4)


def test_object_arithmetic_dot_norm():
    g = torch.Generator().manual_seed(3)
    a = ObjectData(2, torch.randn((5, 7), generator=g), (2.0, 2.0))
    b = ObjectData(2, torch.randn((5, 7), generator=g), (2.0, 2.0))
    A, B = a.dataArray.double().numpy(), b.dataArray.double().numpy()
    assert np.array_equal((a + b).dataArray.numpy(), (a.dataArray + b.dataArray).numpy())
    assert np.array_equal((a - b).dataArray.numpy(), (a.dataArray - b.dataArray).numpy())
    assert np.array_equal((-a).dataArray.numpy(), -a.dataArray.numpy())
    assert np.array_equal((2.5 * a).dataArray.numpy(), (2.5 * a.dataArray).numpy())
    assert (a + b).L == a.L and (3 * a).dim == 2
    assert math.isclose(a.objdot(b), float((A * B).sum()), rel_tol=1e-12)
    assert math.isclose(a.norm(), float(np.sqrt((A * A).sum())), rel_tol=1e-12)
    # <a + b, a - b> = |a|^2 - |b|^2
    assert math.isclose((a + b).objdot(a - b), a.norm() ** 2 - b.norm() ** 2, rel_tol=1e-6, abs_tol=1e-6)
    with pytest.raises(ValueError):
        a + ObjectData(2, torch.zeros((5, 7)), (2.0, 3.0))     # ObjectData.m:104-106
    with pytest.raises(ValueError):
        a - ObjectData(3, torch.zeros((5, 7, 1)), (2.0, 2.0, 2.0))


def test_object_setter_and_apply_operator():
    a = ObjectData(2, torch.ones((4, 4)), (2.0, 2.0))
    a.updateDataArray(np.full((4, 4), 3.0, np.float32))
    assert a.dataArray.dtype == torch.float32 and float(a.dataArray.sum()) == 48.0 and a.L == (2.0, 2.0)

    class Twice:
        def apply(self, obj):
            return 2 * obj

    out = a.applyOperator(Twice())
    assert isinstance(out, ObjectData) and float(out.dataArray.sum()) == 96.0


def test_ct_data_arithmetic():
    g = torch.Generator().manual_seed(4)
    y1, n1 = torch.randn((3, 4, 5), generator=g), torch.randn((3, 4, 5), generator=g)
    y2, n2 = torch.randn((3, 4, 5), generator=g), torch.randn((3, 4, 5), generator=g)
    a, b = CTData("scanner", y1, n1, (2.0, 2.0)), CTData("scanner", y2, n2, (2.0, 2.0))
    s = a + b                                            # CTData.m:44-54
    assert s.scanner == "scanner" and s.L == (2.0, 2.0)
    assert torch.equal(s.dataArray, y1 + y2) and torch.equal(s.dataArrayNorm, n1 + n2)
    s = a + 1.5                                          # scalar to both arrays
    assert torch.equal(s.dataArray, y1 + 1.5) and torch.equal(s.dataArrayNorm, n1 + 1.5)
    d = a - b                                            # CTData.m:56-63
    assert torch.equal(d.dataArray, y1 - y2) and torch.equal(d.dataArrayNorm, n1 - n2)
    m = -a                                               # CTData.m:65-70
    assert torch.equal(m.dataArray, -y1) and torch.equal(m.dataArrayNorm, -n1)
    t = 0.5 * a                                          # CTData.m:77-81
    assert torch.equal(t.dataArray, 0.5 * y1) and torch.equal(t.dataArrayNorm, 0.5 * n1)
    assert math.isclose(a.objdot(b), float((y1.double() * y2.double()).sum()), rel_tol=1e-12)
    assert math.isclose(a.norm(), float(y1.double().norm()), rel_tol=1e-12)
    plain = CTData("scanner", y1, None, None)            # a circular scan has no normalised copy
    assert (plain + plain).dataArrayNorm is None and (-plain).dataArrayNorm is None
    with pytest.raises(ValueError):
        a + CTData("scanner", torch.zeros((3, 4, 6)), None, None)
    with pytest.raises(ValueError):
        a - CTData("scanner", torch.zeros((3, 4, 6)), None, None)

"""Vector-space data types used by the operators and solvers.

Mirrors matlab/classes/+DataTypes/ObjectData.m:1-129 (image/volume with its
physical extent; objdot/norm in FP64 here, single precision in the
reference) and @CTData/CTData.m:1-87 (sinogram plus its scanner).  The data
lives in a GPU tensor; shapes follow the projector layouts ([nz,]ny,nx and
nv,[nb,]na).
"""
from __future__ import annotations

import numpy as np
import torch


class ObjectData:
    def __init__(self, dim: int, dataArray, L=None):
        if not isinstance(dataArray, torch.Tensor):
            dataArray = torch.as_tensor(np.ascontiguousarray(dataArray, dtype=np.float32))
        self.dim = dim
        self.dataArray = dataArray
        shape = tuple(dataArray.shape[-dim:])
        self.N = tuple(reversed(shape))                 # (nx, ny[, nz]) as MATLAB N
        self.L = tuple(L) if L is not None else tuple(2.0 for _ in range(dim))
        self.Dx = tuple(l / n for l, n in zip(self.L, self.N))

    def _like(self, a):
        return ObjectData(self.dim, a, self.L)

    def _check(self, o):
        if self.dim != o.dim or self.L != o.L:
            raise ValueError("Either dimensions or sizes are wrong!")

    def __add__(self, o):
        self._check(o)
        return self._like(self.dataArray + o.dataArray)

    def __sub__(self, o):
        self._check(o)
        return self._like(self.dataArray - o.dataArray)

    def __neg__(self):
        return self._like(-self.dataArray)

    def __rmul__(self, a):
        return self._like(a * self.dataArray)

    __mul__ = __rmul__

    def objdot(self, o) -> float:
        return float(torch.dot(self.dataArray.reshape(-1).double(), o.dataArray.reshape(-1).double()))


class CTData:
    """Sinogram data with the scanner that produced it (CTData.m:1-87)."""

    def __init__(self, scanner, dataArray, dataArrayNorm=None, L=None):
        self.scanner = scanner
        self.dataArray = dataArray
        self.dataArrayNorm = dataArrayNorm
        self.L = L
        self.para = getattr(scanner, "para", None)

    def objdot(self, o) -> float:
        return float(torch.dot(self.dataArray.reshape(-1).double(), o.dataArray.reshape(-1).double()))

    def _john_input(self, smoothed):
        """The normalised multi-helix sinogram (nHelix, nv, nb, na); with
        `smoothed`, each helix replaced by the low-pass band of its one-level
        3D linear framelet expansion (applyJohn.m:28-36)."""
        from .framelets import FrameletSystem
        g = self.dataArrayNorm
        if g is None:
            raise ValueError("CTData has no normalised data (dataArrayNorm)")
        g = g.contiguous()
        if smoothed:
            fs = FrameletSystem(3, "linear", 1)
            g = torch.stack([fs.forwardTransform(g[i].contiguous()).frameletArray[0, 0] for i in range(g.shape[0])])
        return g.contiguous()

    def applyJohn(self, smoothed: bool = False) -> torch.Tensor:
        """Discretised John's equation D g of a multiHelix scan (applyJohn.m:1-106);
        FP64, shape (nHelix-2 or 1, nv-2, nb-2, na-2)."""
        from .john import JohnOperator, john_geometry
        return JohnOperator(john_geometry(self.scanner)).apply(self._john_input(smoothed))

    def applyJohnNormal(self, smoothed: bool = False) -> torch.Tensor:
        """D^T D g (applyJohnNormal.m:1-156); FP64, shape (nHelix, nv, nb, na)."""
        from .john import JohnOperator, john_geometry
        return JohnOperator(john_geometry(self.scanner)).normal(self._john_input(smoothed))

    def norm(self) -> float:
        return float(np.sqrt(self.objdot(self)))

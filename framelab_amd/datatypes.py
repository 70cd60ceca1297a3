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
    """Image or volume with its physical extent (ObjectData.m:1-129).

    ``ObjectData(dim)`` is a zero 128^dim array of side 10 and
    ``ObjectData(dim, array)`` takes side 10 per axis, as the reference's
    constructor does (ObjectData.m:13-52)."""

    def __init__(self, dim: int, dataArray=None, L=None):
        if dim not in (2, 3):
            raise ValueError("dim must be 2 or 3")
        if dataArray is None:
            dataArray = torch.zeros((128,) * dim, dtype=torch.float32)
        elif not isinstance(dataArray, torch.Tensor):
            dataArray = torch.as_tensor(np.ascontiguousarray(dataArray, dtype=np.float32))
        self.dim = dim
        self.meshtype = "cartesian"
        self.dataArray = dataArray
        shape = tuple(dataArray.shape[-dim:])
        self.N = tuple(reversed(shape))                 # (nx, ny[, nz]) as MATLAB N
        self.L = tuple(float(v) for v in L) if L is not None else tuple(10.0 for _ in range(dim))
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
        """<self, o> (ObjectData.m:118-121), accumulated in FP64."""
        return float(torch.dot(self.dataArray.reshape(-1).double(), o.dataArray.reshape(-1).double()))

    def norm(self) -> float:
        """sqrt(<self, self>) (ObjectData.m:123-125)."""
        return float(np.sqrt(self.objdot(self)))

    def frameletTransform(self, sys):
        """alpha = W u for the FrameletSystem ``sys`` (ObjectData.m:58-62)."""
        return sys.forwardTransform(self.dataArray)

    def adjointFrameletTransform(self, sys, alpha) -> None:
        """Replace the data array by W^T alpha (ObjectData.m:64-69).  The
        reference passes ``sys`` on to FrameletExpansion.adjointFrameletTransform,
        which takes no system argument; here ``sys`` performs the synthesis."""
        self.dataArray = sys.adjointTransform(alpha)

    def applyOperator(self, oper):
        """oper.apply(self) (ObjectData.m:71-78): a scanner gives CTData."""
        return oper.apply(self)

    def updateDataArray(self, A) -> None:
        """Setter of the data array (ObjectData.m:80-83); the extent L stays."""
        if not isinstance(A, torch.Tensor):
            A = torch.as_tensor(np.ascontiguousarray(A, dtype=np.float32))
        self.dataArray = A


class CTData:
    """Sinogram data with the scanner that produced it (CTData.m:1-87)."""

    def __init__(self, scanner, dataArray, dataArrayNorm=None, L=None):
        self.scanner = scanner
        self.dataArray = dataArray
        self.dataArrayNorm = dataArrayNorm
        self.L = L
        self.para = getattr(scanner, "para", None)

    def _like(self, a, an):
        return CTData(self.scanner, a, an, self.L)

    @staticmethod
    def _norm_op(op, a, b):
        return None if a is None or b is None else op(a, b)

    def __add__(self, o):
        """Sum of two CTData, or CTData plus a scalar or array added to both
        the data and its normalised copy (CTData.m:44-54)."""
        if isinstance(o, CTData):
            if tuple(self.dataArray.shape) != tuple(o.dataArray.shape):
                raise ValueError("Array dimension mismatch!")
            return self._like(self.dataArray + o.dataArray,
                              self._norm_op(torch.add, self.dataArrayNorm, o.dataArrayNorm))
        return self._like(self.dataArray + o, None if self.dataArrayNorm is None else self.dataArrayNorm + o)

    def __sub__(self, o):
        """Difference of two CTData (CTData.m:56-63)."""
        if tuple(self.dataArray.shape) != tuple(o.dataArray.shape):
            raise ValueError("Array dimension mismatch!")
        return self._like(self.dataArray - o.dataArray,
                          self._norm_op(torch.sub, self.dataArrayNorm, o.dataArrayNorm))

    def __neg__(self):
        """Negated data and normalised data (CTData.m:65-70)."""
        return self._like(-self.dataArray, None if self.dataArrayNorm is None else -self.dataArrayNorm)

    def __rmul__(self, a):
        """alpha * CTData scales the data and its normalised copy (CTData.m:77-81)."""
        return self._like(a * self.dataArray, None if self.dataArrayNorm is None else a * self.dataArrayNorm)

    __mul__ = __rmul__

    def objdot(self, o) -> float:
        """<self, o> over the data arrays (CTData.m:72-75), accumulated in FP64."""
        return float(torch.dot(self.dataArray.reshape(-1).double(), o.dataArray.reshape(-1).double()))

    def frameletTransform(self, sys):
        """Framelet analysis of the sinogram array (CTData.m:26-30)."""
        return sys.forwardTransform(self.dataArray)

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
        """sqrt(<self, self>) over the data array."""
        return float(np.sqrt(self.objdot(self)))

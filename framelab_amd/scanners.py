"""Scanner operator objects with the reference's call signatures.

Mirrors matlab/classes/+Operators/@FanBeamScanner and @ConeBeamScanner:
``apply(object) -> CTData`` (apply.m) and ``applyAdjoint(data[, N]) ->
ObjectData`` (applyAdjoint.m), with ``setPara`` building the ``para`` tables
(setPara.m) and the work running on the GPU through the C ABI.

Differences from the reference, all additive: the fan geometry (SO, OD, Ly)
is a constructor option (the reference hard-codes it privately,
FanBeamScanner.m:6-8); ``adjoint_kind`` selects the exact-chord adjoint
(default, the reference's) or the matched Siddon transpose; ``storage``
selects FP32/BF16/FP16 sinogram storage.
"""
from __future__ import annotations

import torch

from . import geometry as G
from .datatypes import CTData, ObjectData
from .projector import ADJ_EXACT, ADJ_FAN_FB, Projector


class _Scanner:
    verbose = False
    adjoint_kind = ADJ_EXACT

    def _projector(self, key, make_para):
        if getattr(self, "_pkey", None) != key:
            self.para = make_para()
            self._proj = Projector(self.para)
            self._pkey = key
        return self._proj


class FanBeamScanner(_Scanner):
    """FanBeamScanner(nd, nv) (FanBeamScanner.m:21-27)."""

    def __init__(self, nd: int = 256, nv: int = 256, SO: float = 5.0, OD: float = 5.0, Ly: float = 6.0,
                 version: int = 1):
        self.scanType = "2DFan"
        self.nd, self.nv = nd, nv
        self.SO, self.OD, self.Ly = SO, OD, Ly
        self.y_os = 0.0
        self.version = version       # 1: Gao column walk (class default), 0: Siddon, 2: finite-width beam
        self.para = None
        self._pkey = None

    def setPara(self, obj: ObjectData):
        key = (obj.N, obj.L)
        return self._projector(key, lambda: G.fan_para(self.nd, self.nv, N=obj.N, L=obj.L, SO=self.SO,
                                                       OD=self.OD, Ly=self.Ly))

    def apply(self, obj: ObjectData) -> CTData:
        P = self.setPara(obj)
        x = obj.dataArray.to("cuda").float().contiguous()
        y = P.forward(x, method={0: "siddon", 1: "gao", 2: "fb"}[self.version])
        return CTData(self, y, None, obj.L)

    def applyAdjoint(self, data: CTData) -> ObjectData:
        if self._pkey is None:
            raise RuntimeError("applyAdjoint before apply: the scanner has no geometry yet")
        kind = ADJ_FAN_FB if self.version == 2 else self.adjoint_kind
        x = self._proj.adjoint(data.dataArray.contiguous(), kind=kind)
        return ObjectData(2, x, self._pkey[1])


class ConeBeamScanner(_Scanner):
    """ConeBeamScanner(type, na, nb, nv, zmax, rps, vtab, fps, nHelix, phaseShift)
    (ConeBeamScanner.m:33-74)."""

    def __init__(self, type: str, na: int, nb: int, nv: int, zmax=None, rps=3.0, vtab=2.0, fps=30.0,
                 nHelix=1, phaseShift=(0.0,), SO=5.0, OD=5.0, Ly=6.0, Lz=6.0):
        if type not in ("circle", "helix", "multiHelix"):
            raise ValueError("circle, helix and multiHelix are only valid types")
        if type == "circle" and zmax is not None:
            raise ValueError("zmax, rps, vtab, fps are for helical scan only")
        if type != "circle" and zmax is None:
            raise ValueError("helix and multiHelix require more inputs")
        self.type, self.na, self.nb, self.nv = type, na, nb, nv
        self.zmax, self.rps, self.vtab, self.fps = zmax, rps, vtab, fps
        self.nHelix, self.phaseShift = nHelix, tuple(phaseShift)
        self.SO, self.OD, self.Ly, self.Lz = SO, OD, Ly, Lz
        self.GPU = 1
        self.para = None
        self._pkey = None

    def setPara(self, obj: ObjectData):
        key = (obj.N, obj.L)

        def mk():
            kw = {}
            if self.type != "circle":
                kw = dict(zmax=self.zmax, rps=self.rps, vtab=self.vtab, fps=self.fps, nHelix=self.nHelix,
                          phaseShift=self.phaseShift)
            return G.cone_para(self.na, self.nb, self.nv, N=obj.N, L=obj.L, SO=self.SO, OD=self.OD,
                               Ly=self.Ly, Lz=self.Lz, kind=self.type, **kw)
        return self._projector(key, mk)

    def apply(self, obj: ObjectData) -> CTData:
        """Forward scan and its normalised form yNorm = y/(|d| scale) (ConeBeamScanner/apply.m:1-57);
        multiHelix data are split per helix, (nHelix, nv, nb, na)."""
        P = self.setPara(obj)
        x = obj.dataArray.to("cuda").float().contiguous()
        y = P.forward(x)
        yn = P.forward_norm(y)
        if self.type == "multiHelix":
            shape = (self.nHelix, self.para.nv // self.nHelix, self.nb, self.na)
            y, yn = y.reshape(shape), yn.reshape(shape)
        return CTData(self, y, yn, obj.L)

    def applyAdjoint(self, data: CTData, N=None) -> ObjectData:
        if self._pkey is None:
            raise RuntimeError("applyAdjoint before apply: the scanner has no geometry yet")
        if N is not None and tuple(int(n) for n in N) != self._pkey[0]:
            raise ValueError("adjoint size N must match the forward object's size")
        x = self._proj.adjoint(data.dataArray.contiguous(), kind=self.adjoint_kind)
        return ObjectData(3, x, self._pkey[1])

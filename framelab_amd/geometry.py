"""Scanner geometry -> the ``para`` parameter set consumed by the projectors.

Host-side mirror of the reference's ``setPara`` methods:
  * matlab/classes/+Operators/@FanBeamScanner/setPara.m:1-56
  * matlab/classes/+Operators/@ConeBeamScanner/setPara.m:1-83

All lengths are rescaled to voxel units (``scale = Dx(1)``) and every table
is float32, exactly as the MATLAB code stores them (``single``); integer
tables are int32 (MATLAB ``uint32`` reinterpreted as ``int*`` by the mex
gateways, matlab/mex/src/Ax_cone_mf.cpp:44-48).  The arithmetic is done in
float32 step by step so the tables match what MATLAB's single-precision
expressions produce.
"""
from __future__ import annotations

from dataclasses import dataclass, field
import math

import numpy as np

f32 = np.float32


def _frame_tables(nv: int):
    """Single-frame bookkeeping (ConeBeamScanner/setPara.m:59-80,
    FanBeamScanner/setPara.m:32-52)."""
    id_X = np.zeros(nv, dtype=np.int32)
    id_Y = np.arange(nv, dtype=np.int32)
    Nv = np.array([nv], dtype=np.int32)
    return id_X, id_Y, Nv, nv


def with_frames(p, view_frame):
    """Copy of a fan/cone para with several image frames (4D data): view iv
    projects frame view_frame[iv].  The bookkeeping is setPara's for nt
    frames (FanBeamScanner/setPara.m:32-52, ConeBeamScanner/setPara.m:59-80):
    id_X per view, and per frame the list of its views in order, stored as
    the columns of a tmp_size x nt matrix (id_Y) with counts Nv."""
    import dataclasses
    vf = np.asarray(view_frame, dtype=np.int32)
    if vf.size != p.nv or vf.min() < 0:
        raise ValueError("one non-negative frame index per view")
    nt = int(vf.max()) + 1
    lists = [np.nonzero(vf == f)[0].astype(np.int32) for f in range(nt)]
    Nv = np.array([len(v) for v in lists], dtype=np.int32)
    tmp = int(Nv.max())
    id_Y = np.zeros((nt, tmp), dtype=np.int32)
    for f, v in enumerate(lists):
        id_Y[f, :v.size] = v
    return dataclasses.replace(p, nt=nt, id_X=vf.copy(), id_Y=id_Y.ravel(), Nv=Nv, tmp_size=tmp)


@dataclass
class FanPara:
    """Fields of the fan-beam ``para`` struct (FanBeamScanner/setPara.m)."""
    SO: np.float32
    OD: np.float32
    scale: np.float32
    nx: int
    ny: int
    nt: int
    dy_det: np.float32
    y_os: np.float32
    sd_phi: np.ndarray
    y_det: np.ndarray
    cos_phi: np.ndarray
    sin_phi: np.ndarray
    cos_det: np.ndarray
    sin_det: np.ndarray
    id_X: np.ndarray
    id_Y: np.ndarray
    Nv: np.ndarray
    nv: int
    tmp_size: int
    version: int = 1
    GPU: int = 1
    # detector-cell edge angles for the finite-width beam (version 2): edge e
    # of nd+1 bounds cells e-1 and e
    cos_det2: np.ndarray | None = None
    sin_det2: np.ndarray | None = None

    @property
    def nd(self) -> int:
        return int(self.y_det.size)


@dataclass
class ConePara:
    """Fields of the cone-beam ``para`` struct (ConeBeamScanner/setPara.m)."""
    SO: np.float32
    OD: np.float32
    scale: np.float32
    nx: int
    ny: int
    nz: int
    nt: int
    dy_det: np.float32
    dz_det: np.float32
    y_os: np.float32
    dz: np.float32
    sd_phi: np.ndarray
    sd_z: np.ndarray
    y_det: np.ndarray
    z_det: np.ndarray
    cos_phi: np.ndarray
    sin_phi: np.ndarray
    id_X: np.ndarray
    id_Y: np.ndarray
    Nv: np.ndarray
    tmp_size: int
    nv_block: int = 4
    version: int = 1
    GPU: int = 1

    @property
    def na(self) -> int:
        return int(self.y_det.size)

    @property
    def nb(self) -> int:
        return int(self.z_det.size)

    @property
    def nv(self) -> int:
        return int(self.sd_phi.size)


def _cos_sin_single(phi32: np.ndarray):
    """MATLAB ``cos``/``sin`` of a single array: evaluated and rounded to single."""
    p = phi32.astype(np.float64)
    return np.cos(p).astype(f32), np.sin(p).astype(f32)


def _det_coords(n: int, pitch: np.float32, offset: np.float32, scale: np.float32):
    """single(((-n/2:n/2-1)+0.5)*pitch+offset)/scale (FanBeamScanner/setPara.m:23)."""
    k = (np.arange(-(n // 2), n - n // 2, dtype=np.float64) + 0.5).astype(f32)
    return ((k * f32(pitch)).astype(f32) + f32(offset)).astype(f32) / f32(scale)


def fan_para(nd: int, nv: int, N=(256, 256), L=(2.0, 2.0), SO=5.0, OD=5.0, Ly=6.0,
             y_os=0.0, sd_phi: np.ndarray | None = None) -> FanPara:
    """FanBeamScanner(nd,nv).setPara(object) for an image of N pixels spanning L.

    SO/OD/Ly default to the reference's private defaults (FanBeamScanner.m:6-8);
    they are constructor options here so other geometries can be expressed.
    """
    nx, ny = int(N[0]), int(N[1])
    scale = f32(f32(L[0]) / f32(nx))                     # ObjectData Dx = L./N
    SOs, ODs = f32(SO) / scale, f32(OD) / scale
    dy_det = f32(f32(Ly) / f32(nd))
    if sd_phi is None:
        sd_phi = (2.0 * math.pi / nv * np.arange(nv, dtype=np.float64)).astype(f32)
    else:
        sd_phi = np.asarray(sd_phi, dtype=f32)
    y_det = _det_coords(nd, dy_det, f32(y_os), scale)
    cos_phi, sin_phi = _cos_sin_single(sd_phi)
    ang = np.arctan2(y_det.astype(np.float64), np.float64(f32(SOs + ODs))).astype(f32)
    cos_det, sin_det = _cos_sin_single(ang)
    # cell edges y_det[e] - dy/2 (and the last cell's upper edge), the
    # strips of the finite-width forward (Ax_fan_mf_cpu_new_fb.cpp:38-47)
    dyp = f32(dy_det / scale)
    edges = np.concatenate([(y_det - f32(dyp / f32(2))).astype(f32),
                            [f32(y_det[-1] + f32(dyp / f32(2)))]]).astype(f32)
    ang2 = np.arctan2(edges.astype(np.float64), np.float64(f32(SOs + ODs))).astype(f32)
    cos_det2, sin_det2 = _cos_sin_single(ang2)
    id_X, id_Y, Nv, tmp = _frame_tables(nv)
    return FanPara(SO=f32(SOs), OD=f32(ODs), scale=scale, nx=nx, ny=ny, nt=1,
                   dy_det=dyp, y_os=f32(0.0), sd_phi=sd_phi, y_det=y_det,
                   cos_phi=cos_phi, sin_phi=sin_phi, cos_det=cos_det, sin_det=sin_det,
                   id_X=id_X, id_Y=id_Y, Nv=Nv, nv=nv, tmp_size=tmp,
                   cos_det2=cos_det2, sin_det2=sin_det2)


def cone_para(na: int, nb: int, nv: int, N=(256, 256, 256), L=(2.0, 2.0, 2.0),
              SO=5.0, OD=5.0, Ly=6.0, Lz=6.0, y_os=0.0, kind: str = "circle",
              zmax=2.0, rps=3.0, vtab=2.0, fps=30.0, nHelix=1, phaseShift=(0.0,),
              sd_phi: np.ndarray | None = None) -> ConePara:
    """ConeBeamScanner(kind,na,nb,nv,...).setPara(object) (ConeBeamScanner/setPara.m:1-83).

    ``kind`` is 'circle', 'helix' or 'multiHelix' (ConeBeamScanner.m:38-40);
    helical trajectories derive nv from fps*zmax/vtab as the constructor does
    (ConeBeamScanner.m:53-58).
    """
    if kind not in ("circle", "helix", "multiHelix"):
        raise ValueError("circle, helix and multiHelix are only valid types")
    nx, ny, nz = (int(v) for v in N)
    scale = f32(f32(L[0]) / f32(nx))
    SOs, ODs = f32(SO) / scale, f32(OD) / scale
    dy_det = f32(f32(Ly) / f32(na))
    dz_det = f32(f32(Lz) / f32(nb))
    dy_p = f32(dy_det / scale)
    dz_vox = f32(f32(f32(L[2]) / f32(nz)) / scale)
    if kind == "circle":
        if sd_phi is None:
            sd_phi = (2.0 * math.pi / nv * np.arange(nv, dtype=np.float64)).astype(f32)
        else:
            sd_phi = np.asarray(sd_phi, dtype=f32)
        sd_z = np.zeros(sd_phi.size, dtype=f32)
    else:
        phiMax = rps * zmax * 2 * math.pi / vtab
        dphi = 2 * math.pi * rps / fps
        base = np.arange(-phiMax / 2, phiMax / 2 + 1e-12 * phiMax, dphi)
        h = vtab / (2 * math.pi * rps)
        helices = nHelix if kind == "multiHelix" else 1
        shifts = list(phaseShift) + [0.0] * max(0, helices - len(phaseShift))
        phis, zs = [], []
        for i in range(helices):
            shift = shifts[i] if kind == "multiHelix" else 0.0
            phis.append((base + shift).astype(f32))
            zs.append((h * base).astype(f32) / scale)
        sd_phi = np.concatenate(phis).astype(f32)
        sd_z = np.concatenate(zs).astype(f32)
    y_det = _det_coords(na, dy_det, f32(y_os), scale)
    z_det = _det_coords(nb, dz_det, f32(0.0), scale)
    cos_phi, sin_phi = _cos_sin_single(sd_phi)
    id_X, id_Y, Nv, tmp = _frame_tables(sd_phi.size)
    return ConePara(SO=f32(SOs), OD=f32(ODs), scale=scale, nx=nx, ny=ny, nz=nz, nt=1,
                    dy_det=dy_p, dz_det=f32(dz_det / scale), y_os=f32(f32(y_os) / dy_p),
                    dz=dz_vox, sd_phi=sd_phi, sd_z=sd_z, y_det=y_det, z_det=z_det,
                    cos_phi=cos_phi, sin_phi=sin_phi, id_X=id_X, id_Y=id_Y, Nv=Nv,
                    tmp_size=tmp)


# Detector extents that make the fan cover the image's circumscribed circle
# (radius sqrt(2) for [-1,1]^2) with SO=4, SD=8, and the cone cover z=+-1 at the
# nearest approach SO-sqrt(2).  These define the benchmark geometries.
def full_fan_width(SO: float = 4.0, SD: float = 8.0, r: float = math.sqrt(2.0)) -> float:
    half = math.asin(r / SO)
    return 2.0 * SD * math.tan(half)


def full_cone_height(SO: float = 4.0, SD: float = 8.0, r: float = math.sqrt(2.0),
                     zhalf: float = 1.0) -> float:
    return 2.0 * zhalf * SD / (SO - r)

"""CPU-side checks of the C ABI library and the host-side para builders."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

from framelab_amd import geometry as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = []
    for hdr in os.listdir(os.path.join(ROOT, "include")):
        if not hdr.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"typedef struct[^{;]*\{.*?\}[^;]*;", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M):
            if m.group(1) not in ("if", "while", "for", "defined"):
                names.append(m.group(1))
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    from framelab_amd import _lib
    L = _lib.lib()
    declared = _declared_functions()
    assert "Ax_cone_mf_gpu_siddon" in declared and "fl_forward" in declared
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, missing
    # the ctypes table covers the header
    assert set(declared) <= set(_lib.SIGNATURES), set(declared) - set(_lib.SIGNATURES)


REF_MEX = "/root/reference/matlab/mex/src"
REF_HEADERS = ("Ax_fan_mf.h", "Atx_fan_mf.h", "Atx_fan_mf_cpu.h", "Ax_cone_mf.h", "Atx_cone_mf.h")


@pytest.mark.skipif(not os.path.isdir(REF_MEX), reason="reference checkout absent")
def test_dropins_match_the_gateway_headers(tmp_path):
    """The mex gateways link against the drop-ins unchanged only if every
    GPU function their headers declare extern "C" is exported here with the
    same parameter list.  One translation unit includes this library's
    header and the gateways' headers: C-linkage functions cannot be
    overloaded, so any differing declaration is a compile error."""
    import subprocess
    gpu_funcs = set()
    for h in REF_HEADERS:
        gpu_funcs |= set(re.findall(r'extern\s+"C"\s+void\s+(\w+)\s*\(', open(os.path.join(REF_MEX, h)).read()))
    assert len(gpu_funcs) == 8, gpu_funcs
    assert gpu_funcs <= set(_declared_functions()), gpu_funcs - set(_declared_functions())
    tu = tmp_path / "gateway_headers.cpp"
    tu.write_text('#include "framelab_hip.h"\n' + "".join(f'#include "{h}"\n' for h in REF_HEADERS))
    r = subprocess.run(["g++", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), "-I", REF_MEX, str(tu)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_error_strings_without_gpu():
    from framelab_amd import _lib
    L = _lib.lib()
    assert L.fl_error_string(-1) == b"invalid argument"
    assert L.fl_version() >= 100
    # invalid geometry is rejected before any device call
    g = _lib.FlGeometry()
    g.nx = 31
    h = C.c_void_p()
    rc = L.fl_plan_create(C.byref(h), C.byref(g))
    assert rc == _lib.FL_ERR_INVALID and not h.value
    assert b"non-positive" in L.fl_last_error_message() or b"even" in L.fl_last_error_message()


def test_batch_rejects_non_fan_without_gpu():
    """Batch geometry checks run before any device call."""
    from framelab_amd import _lib
    L = _lib.lib()
    gs = (_lib.FlGeometry * 2)()
    for g in gs:
        g.nx = g.ny = 16
        g.nt, g.nv, g.na, g.nb = 1, 8, 16, 1
    gs[1].nz = 16          # a cone geometry
    h = C.c_void_p()
    assert L.fl_batch_create(C.byref(h), 2, gs) == _lib.FL_ERR_INVALID and not h.value
    assert b"problem 1" in L.fl_last_error_message()
    assert L.fl_batch_create(C.byref(h), 0, gs) == _lib.FL_ERR_INVALID


def test_fan_para_matches_setPara_formulas():
    """FanBeamScanner/setPara.m:5-30 with the reference defaults SO=OD=5, Ly=6."""
    p = G.fan_para(256, 180, N=(128, 128), L=(2.0, 2.0))
    scale = np.float32(2.0) / np.float32(128)
    assert p.scale == scale
    assert p.SO == np.float32(5.0) / scale and p.OD == np.float32(5.0) / scale
    dy = np.float32(6.0) / np.float32(256)
    assert p.dy_det == np.float32(dy / scale)
    d = np.diff(p.y_det.astype(np.float64))
    assert np.allclose(d, float(dy) / float(scale), rtol=1e-6)
    assert np.allclose(p.y_det + p.y_det[::-1], 0, atol=1e-4)          # symmetric, y_os = 0
    assert np.allclose(p.sd_phi, 2 * np.pi / 180 * np.arange(180), rtol=1e-6)
    ang = np.arctan2(p.y_det.astype(np.float64), float(p.SO + p.OD))
    assert np.allclose(p.cos_det, np.cos(ang), atol=1e-7) and np.allclose(p.sin_det, np.sin(ang), atol=1e-7)
    assert p.id_X.tolist() == [0] * 180 and p.id_Y.tolist() == list(range(180)) and p.Nv.tolist() == [180]


def test_cone_para_matches_setPara_formulas():
    """ConeBeamScanner/setPara.m:5-56 (circle, and helix z-advance)."""
    p = G.cone_para(64, 48, 30, N=(32, 32, 16), L=(2.0, 2.0, 1.0))
    assert p.dz == np.float32(1.0)            # isotropic voxels in voxel units
    assert np.allclose(np.diff(p.z_det.astype(np.float64)), float(p.dz_det), rtol=1e-5)
    assert np.all(p.sd_z == 0) and p.nv == 30
    h = G.cone_para(64, 48, 0, N=(32, 32, 32), kind="helix", zmax=2.0, rps=3.0, vtab=2.0, fps=30.0)
    assert h.nv == int(math.floor(30 * 2.0 / 2.0)) + 1          # ConeBeamScanner.m:55
    # z advances vtab/(2*pi*rps) per radian, in voxel units
    slope = np.polyfit(h.sd_phi.astype(np.float64), h.sd_z.astype(np.float64), 1)[0]
    assert abs(slope - 2.0 / (2 * np.pi * 3.0) / float(h.scale)) < 1e-3 * abs(slope)
    with pytest.raises(ValueError):
        G.cone_para(8, 8, 8, kind="doubleHelix")


def test_benchmark_detector_extents():
    """Full fan covers the circumscribed circle r=sqrt(2): half-angle asin(sqrt2/4)."""
    assert abs(G.full_fan_width() - 6.047432) < 1e-5
    assert abs(G.full_cone_height() - 6.187673) < 1e-5


def test_matlab_class_fixes_shipped():
    """The MATLAB class methods that replace the reference's defective ones
    are shipped (framelab_amd/matlab/classes, copied over the reference's
    matlab/classes): their signatures return what the callers use --
    softThreshold returns the shrunk expansion and solveSplitBregman keeps
    it; FanBeamScanner.applyAdjoint takes the parameters from data.scanner;
    plus/minus/mtimes build results from the class's own properties.  (No
    MATLAB here: checked as text.)"""
    import os
    root = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "framelab_amd", "matlab",
                        "classes")
    want = {
        "+Operators/@FanBeamScanner/applyAdjoint.m": ("function Aty = applyAdjoint(this, data)", "data.scanner.para"),
        "+DataTypes/@FrameletExpansion/plus.m": ("function c = plus(a, b)", "a.frameletSystem"),
        "+DataTypes/@FrameletExpansion/minus.m": ("function c = minus(a, b)", "a.frameletSystem"),
        "+DataTypes/@FrameletExpansion/mtimes.m": ("function c = mtimes(a, b)", "frameletSystem"),
        "+DataTypes/@FrameletExpansion/softThreshold.m": ("function obj = softThreshold(obj, tau)", "max(abs(a) - tau, 0)"),
        "+DataTypes/@FrameletExpansion/private/combineBands.m": ("function C = combineBands(A, B, op)", "cellfun"),
        "+Optimizers/@SplitBregman/solveSplitBregman.m": ("function u = solveSplitBregman(obj)",
                                                          "alpha = alpha.softThreshold(obj.lam / obj.mu);"),
    }
    for rel, (sig, body) in want.items():
        with open(os.path.join(root, rel)) as fh:
            text = fh.read()
        assert text.startswith(sig), rel
        assert body in text, rel
        code = "\n".join(l for l in text.splitlines() if not l.lstrip().startswith("%"))
        assert "nLevel" not in code and "frameletType" not in code

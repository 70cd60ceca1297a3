"""The bound a cone plan uses to skip the reference's detector window in the
exact adjoint (Atx_cone_mf_gpu_new.cu:64-73): wherever it clears an axis,
every voxel's shadow window (0.05-cell margin) must lie inside the
reference's window, for every view.  Checked by brute force in FP64 over
random geometries drawn so that every term of the bound matters: near and
far sources, fine pitches, voxel heights, helical source heights reaching
several volume heights, and panels shifted against the windows' grid.
Weakening any term (the rotated square's extent, the nearest depth, the
helix's z range, the grid-offset tolerance) lets a cut through.  The bound
is restated here from the plan's (csrc/capi.hip: cone_ref_window_bits);
the GPU parity tests cover the kernel on both sides of it."""
import math

import numpy as np


def ref_window_bits(nx, ny, nz, dz, SO, OD, na, nb, y_det, z_det, dy_det, dz_det, y_os, cphi, sphi, sdz):
    ex = max(0.5 * (abs(c) + abs(s)) for c, s in zip(cphi, sphi)) + 1e-6
    SD = SO + OD
    rc = 0.5 * math.hypot(nx - 1, ny - 1)
    dmin = SO - rc - 1.0
    if not (dmin > 0 and ex < 1):
        return 3
    pa = (y_det[-1] - y_det[0]) / (na - 1) if na > 1 else dy_det
    pb = (z_det[-1] - z_det[0]) / (nb - 1) if nb > 1 else dz_det
    bits = 0
    geo = (1 - ex) * SO - ex - rc * math.hypot(1 - ex, ex)
    slack = abs(1 / pa - 1 / dy_det) * (rc + 1) * SD / dmin
    off = -y_det[0] / pa + y_os - na // 2
    if not (pa > 0 and dy_det > 0 and geo > 1e-3 * SO and abs(off) + slack < 0.8):
        bits |= 1
    hz, d = dz / 2, math.sqrt((1 + dz * dz) / 2)
    zlo, zhi = (0.5 - nz // 2) * dz, (nz - 0.5 - nz // 2) * dz
    zmax = max(max(abs(zlo - s), abs(zhi - s)) for s in sdz)
    geo = (d - hz) * dmin - ex * (zmax + d)
    slack = abs(1 / pb - 1 / dz_det) * (zmax + d) * SD / dmin
    off = -z_det[0] / pb - nb // 2
    if not (pb > 0 and dz_det > 0 and geo > 1e-3 * SO and abs(off) + slack < 0.8):
        bits |= 2
    return bits


def cuts(nx, ny, nz, dz, SO, OD, na, nb, y_det, z_det, dy_det, dz_det, y_os, cphi, sphi, sdz, margin=0.05):
    """Voxel-views where the shadow window leaves the reference window,
    per axis (columns, rows)."""
    SD = SO + OD
    pa = (y_det[-1] - y_det[0]) / (na - 1)
    pb = (z_det[-1] - z_det[0]) / (nb - 1)
    X, Y = np.meshgrid(np.arange(nx) + 0.5 - nx // 2, np.arange(ny) + 0.5 - ny // 2, indexing="ij")
    zc = (np.arange(nz) + 0.5 - nz // 2) * dz
    hz, d = dz / 2, math.sqrt((1 + dz * dz) / 2)
    ccut = rcut = 0
    for c, s, sz in zip(cphi, sphi, sdz):
        xr, yr = c * X + s * Y, -s * X + c * Y
        D = xr + SO
        t = SD / (D * dy_det)
        ahi = np.ceil((yr + 1) * t - y_os + na // 2)
        alo = np.floor((yr - 1) * t - y_os + na // 2)
        us = []
        for k in range(4):
            xk = xr + (0.5 * c if k & 1 else -0.5 * c) + (0.5 * s if k & 2 else -0.5 * s)
            yk = yr + (-0.5 * s if k & 1 else 0.5 * s) + (0.5 * c if k & 2 else -0.5 * c)
            us.append(yk * SD / (xk + SO))
        us = np.array(us)
        slo = np.maximum(np.ceil((us.min(0) - y_det[0]) / pa - margin), 0)
        shi = np.minimum(np.floor((us.max(0) - y_det[0]) / pa + margin), na - 1)
        ccut += int(((slo <= shi) & ((slo < np.maximum(alo, 0)) | (shi > np.minimum(ahi, na - 1)))).sum())
        ex = 0.5 * (abs(c) + abs(s))
        rn, rf = SD / (xr - ex + SO), SD / (xr + ex + SO)
        tz = SD / (D * dz_det)
        for z in zc:
            zr = z - sz
            zl, zh = zr - hz, zr + hz
            vmax = np.where(zh > 0, zh * rn, zh * rf)
            vmin = np.where(zl < 0, zl * rn, zl * rf)
            blo = np.maximum(np.ceil((vmin - z_det[0]) / pb - margin), 0)
            bhi = np.minimum(np.floor((vmax - z_det[0]) / pb + margin), nb - 1)
            rhi = np.ceil((zr + d) * tz + nb // 2)
            rlo = np.floor((zr - d) * tz + nb // 2)
            rcut += int(((blo <= bhi) & ((blo < np.maximum(rlo, 0)) | (bhi > np.minimum(rhi, nb - 1)))).sum())
    return ccut, rcut


def _geometry(rng, SO_over_R, shift=0.0, helix=None):
    n = int(rng.choice([12, 16, 20]))
    nz = int(rng.choice([8, 12, 16]))
    dz = float(rng.choice([0.5, 0.8, 1.0, 1.5]))
    rc = 0.5 * math.hypot(n, n)
    SO = SO_over_R * rc
    OD = SO * float(rng.uniform(0.5, 1.5))
    SD = SO + OD
    # pitches of 2-8 cells per voxel at the centre's magnification (fine
    # enough for cell rounding not to hide a cut), panels covering the
    # volume's fan and its height at the nearest approach
    mag = SD / SO
    dy_det = mag / float(rng.uniform(2.0, 8.0))
    dz_det = mag * dz / float(rng.uniform(2.0, 8.0))
    half = math.asin(min(0.99, (rc + 1) / SO))
    na = int(math.ceil(2 * SD * math.tan(half) / dy_det)) + 2
    nv = 24
    if helix is None:
        helix = float(rng.uniform(0.2, 1.0)) if rng.random() < 0.4 else 0.0
    sdz = np.linspace(-1, 1, nv) * nz * dz * helix   # helical source heights
    # rows cover the volume's height plus the helix's reach
    nb = int(math.ceil(2 * (nz * dz / 2 + nz * dz * helix + 1) * SD / max(SO - rc, 1.0) / dz_det)) + 2
    y_os = float(rng.uniform(-0.3, 0.3))
    y_det = ((np.arange(na) - na / 2) + 0.5 + y_os) * dy_det
    z_det = ((np.arange(nb) - nb / 2) + 0.5) * dz_det
    # panels shifted against the grid the reference's windows assume
    y_det = y_det + shift * dy_det
    z_det = z_det - shift * dz_det
    phi = 2 * math.pi * np.arange(nv) / nv + float(rng.uniform(0, 0.1))
    return dict(nx=n, ny=n, nz=nz, dz=dz, SO=SO, OD=OD, na=na, nb=nb, y_det=y_det, z_det=z_det,
                dy_det=dy_det, dz_det=dz_det, y_os=y_os, cphi=np.cos(phi), sphi=np.sin(phi), sdz=sdz)


def test_bound_implies_no_cut():
    rng = np.random.default_rng(7)
    cleared = {1: 0, 2: 0}
    for trial in range(180):
        # sources near the volume (where the windows come close), far
        # sources, far sources with panels shifted 1-3 cells, and far
        # sources on a helix reaching 1-4 volume heights
        group = trial % 4
        if group == 0:
            g = _geometry(rng, float(rng.uniform(1.05, 3.0)))
        elif group == 1:
            g = _geometry(rng, float(rng.uniform(3.0, 6.0)))
        elif group == 2:
            g = _geometry(rng, float(rng.uniform(3.0, 6.0)), float(rng.choice([-1, 1]) * rng.uniform(1.0, 3.0)))
        else:
            g = _geometry(rng, float(rng.uniform(2.0, 6.0)), helix=float(rng.uniform(1.0, 4.0)))
        bits = ref_window_bits(**g)
        ccut, rcut = cuts(**g)
        if not bits & 1:
            cleared[1] += 1
            assert ccut == 0, (trial, g["SO"], ccut)
        if not bits & 2:
            cleared[2] += 1
            assert rcut == 0, (trial, g["SO"], rcut)
    # both axes were cleared in a good share of the draws, so the checks ran
    assert cleared[1] >= 10 and cleared[2] >= 10, cleared


def test_bound_keeps_the_window_where_it_cuts():
    """Near sources the reference window does cut the shadow (so the plan
    must keep it): the bound sets the bits there."""
    rng = np.random.default_rng(11)
    seen = 0
    for _ in range(20):
        g = _geometry(rng, 1.05)
        ccut, rcut = cuts(**g)
        bits = ref_window_bits(**g)
        if ccut:
            assert bits & 1
        if rcut:
            assert bits & 2
        seen += bool(ccut or rcut)
    assert seen > 0

"""Seeded synthetic phantoms (additive random ellipses / ellipsoids).

Rasterised at pixel/voxel centres -1 + (i + 1/2) * 2/N on [-1,1]^d, in the
projector's axis order (arrays indexed [z, y, x] / [y, x], x fastest), so no
meshgrid axis swap applies.  Centres are uniform in the disk/ball of radius
0.8, semi-axes U[0.05, 0.4], orientation uniform (angle U[0, pi) in 2D, a
Haar-random rotation in 3D) and densities U[0.1, 1.0].
"""
from __future__ import annotations

import numpy as np


def _haar_rotation(rng: np.random.Generator) -> np.ndarray:
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def _uniform_ball(rng, dim, radius):
    v = rng.standard_normal(dim)
    v /= np.linalg.norm(v)
    return v * radius * rng.random() ** (1.0 / dim)


def ellipses_2d(n: int, count: int = 10, seed: int = 0) -> np.ndarray:
    """n x n float32 image of `count` additive random ellipses."""
    rng = np.random.default_rng(seed)
    c = -1.0 + (np.arange(n) + 0.5) * 2.0 / n
    Y, X = np.meshgrid(c, c, indexing="ij")
    img = np.zeros((n, n), dtype=np.float64)
    for _ in range(count):
        x0, y0 = _uniform_ball(rng, 2, 0.8)
        a, b = rng.uniform(0.05, 0.4, 2)
        th = rng.uniform(0.0, np.pi)
        rho = rng.uniform(0.1, 1.0)
        u = (X - x0) * np.cos(th) + (Y - y0) * np.sin(th)
        v = -(X - x0) * np.sin(th) + (Y - y0) * np.cos(th)
        img[(u / a) ** 2 + (v / b) ** 2 <= 1.0] += rho
    return img.astype(np.float32)


def _ellipsoid_params(count: int, seed: int):
    """(centre, semi-axes, rotation, density) of `count` seeded ellipsoids."""
    rng = np.random.default_rng(seed)
    params = []
    for _ in range(count):
        ctr = _uniform_ball(rng, 3, 0.8)
        ax = rng.uniform(0.05, 0.4, 3)
        R = _haar_rotation(rng)
        rho = rng.uniform(0.1, 1.0)
        params.append((ctr, ax, R, rho))
    return params


def ellipsoids_3d(n, count: int = 20, seed: int = 0) -> np.ndarray:
    """float32 volume [nz, ny, nx] of `count` additive random ellipsoids.
    n is an int (cube) or a (nx, ny, nz) triple."""
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    params = _ellipsoid_params(count, seed)
    cx = -1.0 + (np.arange(nx) + 0.5) * 2.0 / nx
    cy = -1.0 + (np.arange(ny) + 0.5) * 2.0 / ny
    cz = -1.0 + (np.arange(nz) + 0.5) * 2.0 / nz
    vol = np.zeros((nz, ny, nx), dtype=np.float32)
    Yg, Xg = np.meshgrid(cy, cx, indexing="ij")
    for iz in range(nz):
        P = np.stack([Xg, Yg, np.full_like(Xg, cz[iz])], -1)   # [ny, nx, 3]
        sl = np.zeros((ny, nx), dtype=np.float64)
        for ctr, ax, R, rho in params:
            if abs(cz[iz] - ctr[2]) > ax.max():
                continue
            q = (P - ctr) @ R          # coordinates in the ellipsoid frame
            sl[((q / ax) ** 2).sum(-1) <= 1.0] += rho
        vol[iz] = sl
    return vol


def ellipsoids_3d_device(n: int, count: int = 20, seed: int = 0, device="cuda"):
    """The ellipsoids of ``ellipsoids_3d(n, count, seed)`` rasterised on a
    torch device, z-slab by z-slab, for volumes too large for the host loop
    (512^3 and up).  FP64 coordinates as on the host; voxels exactly on an
    ellipsoid's surface may round to the other side of it, so the volume is
    a phantom of the same family, not the host's bit for bit."""
    import torch
    params = _ellipsoid_params(count, seed)
    c = -1.0 + (torch.arange(n, dtype=torch.float64, device=device) + 0.5) * 2.0 / n
    vol = torch.zeros((n, n, n), dtype=torch.float32, device=device)
    slab = max(1, (1 << 24) // (n * n))
    Yg, Xg = torch.meshgrid(c, c, indexing="ij")
    for z0 in range(0, n, slab):
        zs = c[z0:z0 + slab]
        P = torch.stack([Xg.expand(zs.numel(), n, n), Yg.expand(zs.numel(), n, n),
                         zs[:, None, None].expand(zs.numel(), n, n)], -1)
        acc = torch.zeros((zs.numel(), n, n), dtype=torch.float64, device=device)
        for ctr, ax, R, rho in params:
            q = (P - torch.from_numpy(ctr).to(device)) @ torch.from_numpy(R).to(device)
            inside = ((q / torch.from_numpy(ax).to(device)) ** 2).sum(-1) <= 1.0
            acc += inside.to(torch.float64) * rho
        vol[z0:z0 + slab] = acc.to(torch.float32)
    return vol

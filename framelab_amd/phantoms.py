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


def ellipsoids_3d(n, count: int = 20, seed: int = 0) -> np.ndarray:
    """float32 volume [nz, ny, nx] of `count` additive random ellipsoids.
    n is an int (cube) or a (nx, ny, nz) triple."""
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    rng = np.random.default_rng(seed)
    params = []
    for _ in range(count):
        ctr = _uniform_ball(rng, 3, 0.8)
        ax = rng.uniform(0.05, 0.4, 3)
        R = _haar_rotation(rng)
        rho = rng.uniform(0.1, 1.0)
        params.append((ctr, ax, R, rho))
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

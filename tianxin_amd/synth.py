"""Seeded synthetic inputs for the assembly workloads.

One counter-based generator (SplitMix64 -> uniform double in [0,1); Box-Muller
-> normal) so every run regenerates the same mesh coordinates, coefficients and
solution vectors from a seed.  The same arrays are handed to the GPU path and to
the CPU checker, so no generator needs to be reproduced in a second language.
"""
from __future__ import annotations

import math

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n outputs of SplitMix64 for counters offset..offset+n-1 (uint64)."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
    u = (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


def normal(seed: int, n: int, sigma: float = 1.0) -> np.ndarray:
    m = (n + 1) // 2
    u = uniform(seed, 2 * m)
    u1 = 1.0 - u[0::2]          # (0, 1]
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.empty(2 * m)
    z[0::2] = r * np.cos(2.0 * math.pi * u2)
    z[1::2] = r * np.sin(2.0 * math.pi * u2)
    return sigma * z[:n]


def brick_coords(nx: int, ny: int, nz: int, *, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0),
                 jitter: float = 0.0, seed: int = 0) -> np.ndarray:
    """Vertex coordinates [(nx+1)(ny+1)(nz+1)][3], node id = i + (nx+1)(j + (ny+1)k).

    Coordinates follow CubeHexMeshFactory's lattice (x0 + i*dx).  With
    ``jitter`` > 0 every interior vertex moves by a seeded uniform offset in
    [-jitter*h, +jitter*h] per axis; boundary vertices stay on the box faces.
    """
    h = [(hi[d] - lo[d]) / n for d, n in enumerate((nx, ny, nz))]
    i = np.arange(nx + 1, dtype=np.float64)
    j = np.arange(ny + 1, dtype=np.float64)
    k = np.arange(nz + 1, dtype=np.float64)
    X = np.empty(((nz + 1), (ny + 1), (nx + 1), 3))
    X[..., 0] = lo[0] + i[None, None, :] * h[0]
    X[..., 1] = lo[1] + j[None, :, None] * h[1]
    X[..., 2] = lo[2] + k[:, None, None] * h[2]
    if jitter > 0.0:
        nn = X.shape[0] * X.shape[1] * X.shape[2]
        off = uniform(seed, 3 * nn, -1.0, 1.0).reshape(X.shape)
        inner = np.zeros(X.shape[:3], bool)
        inner[1:-1, 1:-1, 1:-1] = True
        for d in range(3):
            X[..., d] += np.where(inner, off[..., d] * jitter * h[d], 0.0)
    return X.reshape(-1, 3)


def brick_hex8_conn(nx: int, ny: int, nz: int, shuffle_seed=None) -> np.ndarray:
    """Hex8 connectivity [nx*ny*nz][8] of the brick whose vertices brick_coords numbers.

    Cells in x-fastest order, vertices of a cell in the order (0,0,0) (1,0,0) (1,1,0)
    (0,1,0) (0,0,1) (1,0,1) (1,1,1) (0,1,1) (CubeHexMeshFactory's element order).  With
    ``shuffle_seed`` the cells come in a seeded random order.
    """
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    c = (i + (nx + 1) * (j + (ny + 1) * k)).transpose(2, 1, 0).reshape(-1).astype(np.int64)
    sx, sxy = 1, (nx + 1) * (ny + 1)
    off = np.array([0, sx, sx + nx + 1, nx + 1, sxy, sxy + sx, sxy + sx + nx + 1, sxy + nx + 1], np.int64)
    conn = c[:, None] + off[None, :]
    if shuffle_seed is not None:
        conn = conn[np.argsort(splitmix64(shuffle_seed, conn.shape[0]), kind="stable")]
    return conn


def lognormal(seed: int, n: int, sigma: float) -> np.ndarray:
    """exp(g), g ~ N(0, sigma^2) iid."""
    return np.exp(normal(seed, n, sigma))


def q2_positions(X, nx, ny, nz):
    """Physical position of every Q2 lattice point I + (2nx+1)(J + (2ny+1)K): the
    trilinear map of its cell evaluated at the half-cell offsets."""
    px, py, pz = 2 * nx + 1, 2 * ny + 1, 2 * nz + 1
    I, J, K = np.meshgrid(np.arange(px), np.arange(py), np.arange(pz), indexing="ij")
    I, J, K = I.ravel(order="F"), J.ravel(order="F"), K.ravel(order="F")
    cx, cy, cz = np.minimum(I // 2, nx - 1), np.minimum(J // 2, ny - 1), np.minimum(K // 2, nz - 1)
    lx, ly, lz = (I - 2 * cx) / 2.0, (J - 2 * cy) / 2.0, (K - 2 * cz) / 2.0
    out = np.zeros((I.size, 3))
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                w = (lx if dx else 1 - lx) * (ly if dy else 1 - ly) * (lz if dz else 1 - lz)
                vid = (cx + dx) + (nx + 1) * ((cy + dy) + (ny + 1) * (cz + dz))
                out += w[:, None] * X[vid]
    order = I + px * (J + py * K)
    res = np.zeros_like(out)
    res[order] = out
    return res

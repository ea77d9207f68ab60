"""Q2 lattice positions on a brick of trilinear hex cells (test helper)."""
import numpy as np


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

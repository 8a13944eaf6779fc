"""Square disk lattice of the reference solver, built on the host.

Mirrors `IceDEM.__init__` (src/tensordem/dem.py:105-129): disks of radius r
on a grid of spacing 2r, x centred, y >= 0, row-major order; the edge columns
are fixed when `fix_edges`; every pair closer than 2r*sqrt(2)*(1+1e-10) is a
bond.  On this grid those are exactly the 8 grid neighbours, so bonds are
enumerated from the stencil instead of from all N(N-1)/2 pairs, and listed in
the reference's `torch.triu_indices` order (ascending lower index, then upper
index).
"""
from __future__ import annotations

import math

import numpy as np

# direction codes 4..7 point to the partners with a larger index:
# +1 (right), +nx-1 (up-left), +nx (up), +nx+1 (up-right)
FORWARD = ((1, 0), (-1, 1), (0, 1), (1, 1))


def positions(nx: int, ny: int, radius: float) -> np.ndarray:
    x = np.arange(nx, dtype=np.float64)
    y = np.arange(ny, dtype=np.float64)
    yy, xx = np.meshgrid(y, x, indexing="ij")
    return np.stack((xx.ravel() - (nx - 1) / 2, yy.ravel()), axis=1) * (2 * radius)


def fixed_mask(nx: int, ny: int, fix_edges: bool) -> np.ndarray:
    ix = np.tile(np.arange(nx), ny)
    if not fix_edges:
        return np.zeros(nx * ny, dtype=bool)
    return (ix == 0) | (ix == nx - 1)


def bonds(nx: int, ny: int, radius: float):
    """Return (i, j, rest_length, fwd_bond) in the reference's bond order.

    fwd_bond[p] is the index of the first bond whose lower end is p.
    """
    pos = positions(nx, ny, radius)
    iy, ix = np.divmod(np.arange(nx * ny), nx)
    cols_i, cols_j, counts = [], [], np.zeros(nx * ny, dtype=np.int64)
    for dx, dy in FORWARD:
        ok = (ix + dx >= 0) & (ix + dx < nx) & (iy + dy < ny)
        p = np.nonzero(ok)[0]
        cols_i.append(p)
        cols_j.append(p + dy * nx + dx)
        counts += ok
    i = np.concatenate(cols_i)
    j = np.concatenate(cols_j)
    order = np.lexsort((j, i))
    i, j = i[order], j[order]
    d = pos[j] - pos[i]
    rest = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1])
    cutoff = 2 * radius * math.sqrt(2) * (1 + 1e-10)
    assert bool((rest <= cutoff).all())
    fwd = np.concatenate(([0], np.cumsum(counts)[:-1])).astype(np.int32)
    return i.astype(np.int64), j.astype(np.int64), rest, fwd

"""Ensembles of independent plates stepped as one batch.

A batch is any list of `PlateSpec`; one kernel launch advances every member
(`PlateEngine`).  `random_ensemble` draws the parameter sweep of an
indentation study: lattice sizes log-uniform in nx and uniform in ny,
indenter speeds uniform, and per-disk Weibull break strains seeded per
member (the reference's strength scatter, ice_bond_viscoelastic.py:300-322,
drawn per disk; a bond breaks at its weaker disk).
"""
from __future__ import annotations

import math
from dataclasses import replace

import numpy as np

from .config import DEMConfig
from .engine import PlateSpec, default_tile
from .strength import weibull_strengths

TILE_WIDTHS = (16, 32, 64, 128, 256)


def random_ensemble(members: int = 512, seed: int = 0, *, nx_range=(16, 256), ny_range=(4, 32),
                    speed_range=(0.1, 1.0), weibull_shape: float = 5.0,
                    mean_axial: float = 0.01, mean_shear: float = 0.02, max_degree: int = 8,
                    **config_kw) -> list[PlateSpec]:
    """`members` plates: nx = round(exp U[ln nx0, ln nx1]), ny uniform on
    [ny0, ny1], tool speed U[v0, v1], per-disk Weibull(shape) axial / shear
    strains with the given means (member k draws with seeds 2k+1, 2k+2 offset
    by 1000003*seed), dt bounded by `max_degree` partners per disk."""
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(members):
        nx = int(round(math.exp(rng.uniform(math.log(nx_range[0]), math.log(nx_range[1])))))
        ny = int(rng.integers(ny_range[0], ny_range[1] + 1))
        speed = float(rng.uniform(*speed_range))
        cfg = DEMConfig(nx=nx, ny=ny, tool_speed=speed, breaking_strain=mean_axial,
                        shear_breaking_strain=mean_shear, **config_kw)
        base = 1000003 * seed + 2 * k
        n = nx * ny
        et = weibull_strengths(n, weibull_shape, seed=base + 1, mean=mean_axial)
        es = weibull_strengths(n, weibull_shape, seed=base + 2, mean=mean_shear)
        specs.append(PlateSpec(cfg, dt=cfg.degree_bound_dt(max_degree), eps_t_disk=et,
                               eps_s_disk=es))
    return specs


def tiles(nx: int, ny: int, tile_x: int) -> int:
    ty = 256 // tile_x
    return -(-nx // tile_x) * -(-ny // ty)


def tile_for(nx: int, ny: int, policy) -> int:
    """Tile width for one member: an integer is used as given; "default" is
    the single-plate rule; "min_tiles" launches the fewest 256-disk tiles
    (fewest idle lanes), preferring the squarer tile (less halo) on ties;
    "min_tiles_128" does the same with tiles at most 128 wide (a batch of
    such tiles can stage candidate forces in LDS)."""
    if isinstance(policy, int):
        return policy
    if policy == "default":
        return default_tile(nx)
    if policy in ("min_tiles", "min_tiles_128"):
        widths = TILE_WIDTHS if policy == "min_tiles" else tuple(w for w in TILE_WIDTHS if w <= 128)
        return min(widths, key=lambda tx: (tiles(nx, ny, tx), abs(math.log(tx / (256 // tx)))))
    raise ValueError(f"unknown tile policy {policy!r}")


def with_tiles(specs: list[PlateSpec], policy) -> list[PlateSpec]:
    return [replace(s, tile_x=tile_for(s.config.nx, s.config.ny, policy)) for s in specs]


def total_disks(specs: list[PlateSpec]) -> int:
    return sum(s.config.nx * s.config.ny for s in specs)

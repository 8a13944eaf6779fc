"""Ensemble construction and tile policies (host side)."""
import math

import numpy as np

from tensordem_amd.ensemble import random_ensemble, tile_for, tiles, total_disks, with_tiles


def test_tile_policies():
    assert tiles(20, 4, tile_for(20, 4, "min_tiles")) == 1  # one tile instead of two 16x16
    assert tiles(20, 4, 64) == 1 and tiles(20, 4, 16) == 2
    assert tile_for(100, 10, 32) == 32
    assert tile_for(100, 10, "default") == 32 and tile_for(30, 10, "default") == 16
    for nx in (16, 37, 100, 256):
        for ny in (4, 9, 32):
            best = tile_for(nx, ny, "min_tiles")
            assert all(tiles(nx, ny, best) <= tiles(nx, ny, t) for t in (16, 32, 64, 128, 256))


def test_random_ensemble_distribution_and_seeds():
    a = random_ensemble(512, seed=0)
    b = random_ensemble(512, seed=0)
    c = random_ensemble(512, seed=1)
    nx = np.array([s.config.nx for s in a])
    ny = np.array([s.config.ny for s in a])
    v = np.array([s.config.tool_speed for s in a])
    assert nx.min() >= 16 and nx.max() <= 256 and ny.min() >= 4 and ny.max() <= 32
    # log-uniform nx: mean (256 - 16) / ln 16 = 86.6; ln nx uniform on [ln 16, ln 256]
    assert abs(nx.mean() - 240 / math.log(16)) < 8
    assert abs(np.log(nx).mean() - 0.5 * (math.log(16) + math.log(256))) < 0.1
    assert 0.1 <= v.min() and v.max() <= 1.0
    assert 5e5 < total_disks(a) < 9e5
    assert all(np.array_equal(x.eps_t_disk, y.eps_t_disk) for x, y in zip(a, b))
    assert not np.array_equal(a[0].eps_t_disk, c[0].eps_t_disk)
    # per-member streams differ, and axial / shear draws differ
    assert not np.array_equal(a[0].eps_t_disk[:16], a[1].eps_t_disk[:16])
    all_t = np.concatenate([s.eps_t_disk for s in a])
    assert abs(all_t.mean() - 0.01) < 1e-4
    assert [s.tile_x for s in with_tiles(a[:3], 64)] == [64, 64, 64]

"""Weibull strength scatter and per-bond thresholds in the plate oracle."""
import math

import numpy as np
import pytest

from oracle.plate import OraclePlate
from tensordem_amd.config import DEMConfig
from tensordem_amd.strength import weibull_scale, weibull_strengths


def test_weibull_quantiles_follow_the_closed_form():
    m, mean = 5.0, 0.01
    s = weibull_strengths(200_000, m, seed=3, mean=mean)
    scale = mean / math.gamma(1 + 1 / m)
    # CDF 1 - exp(-(f/scale)^m): check several quantiles and the mean
    for q in (0.05, 0.25, 0.5, 0.75, 0.95):
        exact = scale * (-math.log(1 - q)) ** (1 / m)
        assert np.quantile(s, q) == pytest.approx(exact, rel=4e-3)
    sd = scale * math.sqrt(math.gamma(1 + 2 / m) - math.gamma(1 + 1 / m) ** 2)
    assert abs(s.mean() - mean) < 4 * sd / math.sqrt(len(s))


def test_weibull_median_parameterisation_matches_the_reference_form():
    # the reference scales by the median: F(median) = 1/2
    scale = weibull_scale(5.0, median=0.02)
    assert 1 - math.exp(-(0.02 / scale) ** 5) == pytest.approx(0.5, rel=1e-14)
    a = weibull_strengths(1000, 5.0, seed=11, median=0.02)
    b = weibull_strengths(1000, 5.0, seed=11, median=0.02)
    c = weibull_strengths(1000, 5.0, seed=12, median=0.02)
    assert np.array_equal(a, b) and not np.array_equal(a, c)


def test_uniform_per_disk_thresholds_reproduce_the_scalar_run():
    cfg = DEMConfig(nx=9, ny=4, tool_speed=0.4)
    a = OraclePlate(cfg)
    b = OraclePlate(cfg, eps_t_disk=np.full(36, cfg.breaking_strain),
                    eps_s_disk=np.full(36, cfg.shear_breaking_strain))
    for _ in range(600):
        a.step()
        b.step()
    assert a.broken_bonds > 0
    assert np.array_equal(a.positions, b.positions) and np.array_equal(a.alive, b.alive)


def test_a_bond_breaks_at_its_weaker_disk():
    """Stretch a free 4x3 plate uniformly: every bond carries the same axial
    strain (1 %).  Disk 5 is made weak (threshold 0.5 %), all others strong
    (5 %): exactly the bonds touching disk 5 break, and nothing else."""
    cfg = DEMConfig(nx=4, ny=3, fix_edges=False, drag=0.0, tool_speed=1e-9, tool_gap=100.0)
    et = np.full(12, 0.05)
    et[5] = 0.005
    sim = OraclePlate(cfg, eps_t_disk=et, eps_s_disk=np.full(12, 1.0))
    sim.positions *= 1.01
    sim.forces()
    broken = {(int(sim.bi[b]), int(sim.bj[b])) for b in np.nonzero(sim.alive == 0)[0]}
    touching = {(int(i), int(j)) for i, j in zip(sim.bi, sim.bj) if 5 in (i, j)}
    assert broken == touching and len(touching) == 8
    # the reverse: the weak disk's partners keep their bonds to everyone else
    assert (sim.alive == 0).sum() == 8

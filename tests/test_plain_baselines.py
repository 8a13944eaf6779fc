"""The eager PyTorch FP64 paths that bench.py uses as the speedup denominator
compute the same physics as the FP64 oracle (run here on the CPU device).

A denominator that skipped work or broke different bonds would inflate the
reported speedup, so each is held to the oracle: the plate path to the
reference's 21 golden break events, the bed path to the oracle's positions
and contact-pair set.
"""
import csv

import numpy as np
import pytest
import torch

from bench import PlainBedFP64, PlainFP64
from conftest import GOLDEN
from oracle.bed import OracleBed
from tensordem_amd.bedconfig import BedConfig, column_packing
from tensordem_amd.config import DEMConfig


def test_plain_plate_breaks_the_golden_bonds():
    kw = dict(nx=7, ny=3, tool_speed=0.2)
    dt0 = DEMConfig(**kw).recommended_dt
    cfg = DEMConfig(**kw, dt=dt0, device="cpu")
    p = PlainFP64(cfg, dt0, device="cpu", pinv=False)
    order = []
    for _ in range(345):
        before = p.alive.clone()
        p.step()
        for b in torch.nonzero(before & ~p.alive).flatten().tolist():
            order.append((int(p.i[b]), int(p.j[b])))
    with open(GOLDEN / "dem2d_fracture_events.csv") as fh:
        gold = [(int(g["pair_i"]), int(g["pair_j"])) for g in csv.DictReader(fh)]
    assert sorted(order) == sorted(gold) and len(order) == 21


def test_plain_bed_follows_the_oracle():
    x, r = column_packing(300, 0.01, seed=7)
    cfg = BedConfig()
    v0 = np.random.default_rng(7).normal(scale=0.3, size=x.shape)
    ref = OracleBed(cfg, x, r)
    ref.velocities[:] = v0
    p = PlainBedFP64(cfg, x, r, ref.dt, skin=0.004, v=v0, device="cpu")
    seen = 0
    for k in range(1, 2001):
        ref.step()
        p.step()
        if k % 100 == 0:
            got = p.contact_pairs().numpy()
            exp = ref.contacts().astype(np.int64)
            assert np.array_equal(got[np.lexsort((got[:, 1], got[:, 0]))], exp), k
            seen += len(exp)
    assert seen > 20 and p.rebuilds > 2
    # same law, different summation order (index_add_): FP64 round-off only
    assert np.abs(p.x.numpy() - ref.positions).max() < 1e-9
    assert p.acc[0].item() == pytest.approx(ref.acc[0], rel=1e-6)
    assert p.acc[5].item() == pytest.approx(ref.acc[5], rel=1e-6)


def test_plain_plate_with_weibull_strengths_matches_the_oracle():
    from oracle.plate import OraclePlate
    from tensordem_amd.strength import weibull_strengths
    cfg = DEMConfig(nx=12, ny=5, tool_speed=0.6, breaking_strain=0.01, shear_breaking_strain=0.02,
                    device="cpu")
    dt = cfg.degree_bound_dt()
    et = weibull_strengths(60, 5.0, seed=1, mean=0.01)
    es = weibull_strengths(60, 5.0, seed=2, mean=0.02)
    ref = OraclePlate(cfg, dt=dt, eps_t_disk=et, eps_s_disk=es)
    p = PlainFP64(cfg, dt, device="cpu", pinv=False, eps_t_disk=et, eps_s_disk=es)
    for _ in range(1500):
        ref.step()
        p.step()
    assert ref.broken_bonds > 5
    assert np.array_equal(~p.alive.numpy(), ref.alive == 0)
    assert np.abs(p.x.numpy() - ref.positions).max() < 1e-9

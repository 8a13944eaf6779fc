"""The rigid-disk engine at the benchmark's size: 1 048 576 disks on a
jittered grid in a walled box (the largest neighbours overlap from the start,
gravity, rolling resistance), stepped on the device and by the FP64 oracle
(oracle/disk2d.c, restating granular2d/engine.py:399-627) from the same
state; positions, velocities, spins and angles must be bit-identical after
every step, and the state digest equal."""
import numpy as np
import pytest

from oracle.disk import OracleDisk
from tensordem_amd.granular2d import DiskConfig, DiskSystem, jittered_box

pytestmark = pytest.mark.gpu

STEPS = 6


def test_million_disk_box_matches_oracle_step_by_step():
    pos, r, vel, walls = jittered_box(1024, 0.01, seed=0)
    cfg = DiskConfig(mu=0.5, mu_rolling=0.05, gravity=(0.0, -9.81))
    dev = DiskSystem(pos, r, cfg, walls=walls)
    ref = OracleDisk(pos, r, cfg, walls=walls)
    dev.vel = vel
    ref.vel[:] = vel
    dev.enable_energy_ledger()
    ref.enable_energy_ledger()
    for k in range(STEPS):
        dev.step(1)
        ref.step(1)
        for name in ("pos", "vel", "omega", "theta"):
            assert np.array_equal(getattr(dev, name).cpu().numpy(), getattr(ref, name)), \
                f"step {k + 1}: {name}"
    assert dev.digest() == ref.digest()
    counts = dev.out[10:13].tolist()
    assert counts == ref.last[10:13].tolist()
    assert counts[0] > 50000 and counts[1] > 0          # pair and wall contacts
    for key in ("normal_dashpot", "friction", "rolling"):
        a, b = dev.energy_ledger[key], ref.energy_ledger[key]
        assert abs(a - b) <= 1e-12 * max(abs(a), abs(b)) + 1e-300, key

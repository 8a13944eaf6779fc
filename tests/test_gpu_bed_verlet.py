"""Velocity-Verlet beds (the floe bed's integrator, scale_ceiling/kernel.py:
410-421) on the device against the FP64 oracle's velocity-Verlet, whose step
order is pinned to the reference's step sequence in test_oracle_bed.py.

FP64: positions, half-step velocities, full-step velocities (the readout),
contact sets bit-identical at every checkpoint, over calls of mixed length
and through list rebuilds, a moving indenter and friction; ledger to 1e-11
relative.  A restart continues bit for bit, graph replays equal eager steps.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle.bed import OracleBed
from tensordem_amd import _lib
from tensordem_amd.bed import BedEngine, BedSpec
from tensordem_amd.bedconfig import BedConfig, column_packing

pytestmark = pytest.mark.gpu

R0 = 0.01


def pressed_column(n, seed, steps=30000):
    """A column released 0.3 s ago (collapsing, about one contact per disk;
    an Euler run), then a velocity-Verlet law with the right wall just
    touching the rightmost disk and pushing in, and perturbed velocities."""
    x, r = column_packing(n, R0, seed=seed)
    e = BedEngine([BedSpec(BedConfig(), x, r)])
    e.step(steps)
    assert e.status()[0] == 0
    x, v = e.x.cpu().numpy(), e.v.cpu().numpy()
    cfg = BedConfig(integrator="verlet", right_wall=True,
                    right_wall_x=float((x[:, 0] + r).max()) - 1e-4, right_wall_speed=-0.05,
                    friction=0.4, wall_friction=0.3)
    v = v + np.random.default_rng(seed).normal(scale=0.05, size=x.shape)
    return cfg, x, r, v


def oracle_pairs(ref):
    n = len(ref.radii)
    k = ref.hist_keys
    return np.stack((k // n, k % n), axis=1) if len(k) else np.zeros((0, 2), dtype=np.int64)


def test_verlet_fp64_bit_exact_over_mixed_calls():
    cfg, x, r, v = pressed_column(500, seed=7)
    ref = OracleBed(cfg, x, r)
    ref.velocities[:] = v
    eng = BedEngine([BedSpec(cfg, x, r, dt=ref.dt, velocities=v)])
    # the readout of a fresh state is the given velocity
    assert np.array_equal(eng.full_velocities().cpu().numpy(), v)
    calls = [1, 1, 5, 2, 16, 3, 40, 1, 64, 7] * 12
    done, rebuilds0 = 0, eng.read_members()[0].rebuilds
    for k, c in enumerate(calls):
        eng.step(c)
        for _ in range(c):
            ref.step()
        done += c
        if k % 10 == 9:
            assert eng.status()[0] == 0, f"after {done} steps"
            assert np.array_equal(eng.x.cpu().numpy(), ref.positions), f"after {done} steps"
            assert np.array_equal(eng.v.cpu().numpy(), ref.velocities), f"after {done} steps"
            assert np.array_equal(eng.full_velocities().cpu().numpy(), ref.full_velocity())
            got = eng.contact_pairs() - eng.read_members()[0].first
            assert np.array_equal(got, oracle_pairs(ref)), f"after {done} steps"
    m = eng.read_members()[0]
    assert m.rebuilds > rebuilds0 and m.step_count == done and m.verlet_start == 0
    assert len(ref.hist_keys) > 200
    assert ref.acc[6] != 0                       # the indenter touched the column
    assert eng.status()[0] == 0
    acc = eng.accounting[0].cpu().numpy()
    for i in range(6):
        assert acc[i] == pytest.approx(ref.acc[i], rel=1e-11, abs=1e-15), i
    assert acc[_lib.BED_WALL_WORK] == pytest.approx(ref.acc[6], rel=1e-11, abs=1e-15)
    # energies: the kinetic term of the full-step velocity, as the reference's
    e = eng.energies()[0].cpu().numpy()
    oe = ref.energies()
    assert e[0] == pytest.approx(oe["kinetic_J"], rel=1e-12)
    assert e[1] == pytest.approx(oe["potential_J"], rel=1e-12)
    assert e[2] == pytest.approx(oe["elastic_normal_J"], rel=1e-10)
    # reading energies or velocities changes nothing: one more step still matches
    eng.step(1)
    ref.step()
    assert np.array_equal(eng.x.cpu().numpy(), ref.positions)
    assert np.array_equal(eng.v.cpu().numpy(), ref.velocities)


def test_verlet_restart_and_graph_replays_continue_bit_for_bit():
    cfg, x, r, v = pressed_column(400, seed=11)
    a = BedEngine([BedSpec(cfg, x, r, velocities=v)])
    a.step(301)
    state = a.snapshot()
    assert state["verlet_start"] == 0
    b = BedEngine.from_snapshot(state)
    c = BedEngine.from_snapshot(state)
    a.step(400)
    b.step(400)
    c.graph(50)
    c.run_graph(50, repeats=4)
    torch.cuda.synchronize()
    for e in (b, c):
        assert np.array_equal(e.x.cpu().numpy(), a.x.cpu().numpy())
        assert np.array_equal(e.v.cpu().numpy(), a.v.cpu().numpy())
        assert np.array_equal(e.full_velocities().cpu().numpy(),
                              a.full_velocities().cpu().numpy())
        assert np.array_equal(e.accounting.cpu().numpy(), a.accounting.cpu().numpy())
    assert a.status()[0] == 0


def test_verlet_fp32_forces_and_euler_readout():
    cfg, x, r, v = pressed_column(300, seed=13)
    e32 = BedEngine([BedSpec(cfg, x, r, velocities=v)], fp32=True)
    e64 = BedEngine([BedSpec(cfg, x, r, velocities=v)])
    e32.step(200)
    e64.step(200)
    f32, f64 = e32.full_velocities().cpu().numpy(), e64.full_velocities().cpu().numpy()
    assert e32.status()[0] == 0 and np.isfinite(f32).all()
    # FP32 forces perturb the trajectory at the 1e-7 relative level per force
    assert np.abs(e32.x.cpu().numpy() - e64.x.cpu().numpy()).max() < 1e-6
    assert np.abs(f32 - f64).max() < 1e-2 * np.abs(f64).max()
    # an Euler bed's readout is its velocity
    eu = BedEngine([BedSpec(BedConfig(), x, r, velocities=v)])
    eu.step(10)
    assert np.array_equal(eu.full_velocities().cpu().numpy(), eu.v.cpu().numpy())


def test_verlet_member_needs_the_batch_flag():
    """Setup rejects a velocity-Verlet member in a batch whose arrays do not
    say so; a batch that drops the flag later steps it by Euler and says so
    in the status bits."""
    cfg, x, r, v = pressed_column(120, seed=17, steps=2000)
    eng = BedEngine([BedSpec(cfg, x, r, velocities=v)])
    eng._arrays.verlet = 0
    assert eng.lib.tdm_bed_setup(ctypes.byref(eng._arrays), eng.stream) != 0
    eng._verlet = False
    eng._refresh()
    eng.step(3)
    assert eng.status()[0] & _lib.ST_INTEGRATOR

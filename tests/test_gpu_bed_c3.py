"""The 1M-disk column collapse at its full size against the FP64 oracle.

The column (1 048 576 disks, radii U[0.8, 1.2] x 0.01 m, seed 0) is advanced
40 000 FP64 steps on the GPU, into the collapse; its bed restart state (with
every pair and wall spring history) then starts both the HIP engines and the
FP64 oracle (`oracle/bed2d.c`, its own cell grid and sorted-key history
merge), which runs a short horizon on the host.

* FP64 forces: after every one of the K steps the contact-pair set and its
  spring histories are identical to the oracle's, and positions and
  velocities are bit-identical.
* FP32 forces (FP64 contact decisions, state and histories): the first
  step's contact-pair set equals the oracle's (decided on the same state),
  and after K steps positions stay within 1e-9 m (1e-7 of the disk radius)
  of the oracle's.
* The floe bed's velocity-Verlet (scale_ceiling/kernel.py:410-421) from the
  same state, FP64 forces: contact pairs, springs, positions, half-step
  velocities and the full-step readout bit-identical to the oracle's, over
  single steps and a multi-step call.
"""
import numpy as np
import pytest
import torch

from oracle.bed import OracleBed
from tensordem_amd.bed import BedEngine, BedSpec
from tensordem_amd.bedconfig import BedConfig, column_packing

pytestmark = pytest.mark.gpu

N = 1 << 20
ADVANCE = 40000
K = 20


@pytest.fixture(scope="module")
def state():
    x, r = column_packing(N, 0.01, seed=0)
    eng = BedEngine([BedSpec(BedConfig(), x, r)], fp32=False)
    eng.run_graph(500, repeats=ADVANCE // 1000)
    torch.cuda.synchronize()
    assert int(eng.status()[0]) == 0
    st = eng.snapshot()
    eng.close()
    return st


def _oracle(st) -> OracleBed:
    cfg = BedConfig(**st["config"])
    ref = OracleBed(cfg, st["positions"], st["radii"], dt=st["dt"])
    ref.velocities[:] = st["velocities"]
    n = len(st["radii"])
    pairs = np.asarray(st["pairs"], np.int64)
    ref.hist_keys = np.ascontiguousarray(pairs[:, 0] * n + pairs[:, 1])
    ref.hist_xi = np.ascontiguousarray(st["pair_xi"], np.float64)
    act = np.asarray(st["wall_active"], np.int64)
    wx = np.asarray(st["wall_xi"], np.float64)
    ref.xi_floor = np.ascontiguousarray(np.where((act & 1)[:, None] == 1, wx[:, 0:2], 0.0))
    ref.xi_left = np.ascontiguousarray(np.where((act & 2)[:, None] == 2, wx[:, 2:4], 0.0))
    return ref


def _keys(snap) -> np.ndarray:
    n = len(snap["radii"])
    p = np.asarray(snap["pairs"], np.int64)
    return p[:, 0] * n + p[:, 1]


def test_c3_fp64_steps_bit_identical_to_oracle(state):
    ref = _oracle(state)
    eng = BedEngine.from_snapshot(state, fp32=False)
    print(f"column at step {state['step_count']}: {len(state['pairs'])} pair contacts", flush=True)
    for k in range(1, K + 1):
        eng.step(1)
        ref.step()
        snap = eng.snapshot()
        keys = _keys(snap)
        assert np.array_equal(keys, ref.hist_keys), f"contact-pair set differs at step {k}"
        assert np.array_equal(snap["pair_xi"], ref.hist_xi), f"spring histories differ at step {k}"
        assert np.array_equal(snap["positions"], ref.positions), f"positions differ at step {k}"
        assert np.array_equal(snap["velocities"], ref.velocities), f"velocities differ at step {k}"
        print(f"fp64 step {k}: {len(keys)} contacts, identical", flush=True)
    eng.close()


def test_c3_fp32_contact_set_and_positions(state):
    ref = _oracle(state)
    eng = BedEngine.from_snapshot(state, fp32=True)
    eng.step(1)
    ref.step()
    snap = eng.snapshot()
    assert np.array_equal(_keys(snap), ref.hist_keys), "first-step contact-pair set differs"
    eng.step(K - 1)
    for _ in range(K - 1):
        ref.step()
    snap = eng.snapshot()
    err = float(np.abs(snap["positions"] - ref.positions).max())
    verr = float(np.abs(snap["velocities"] - ref.velocities).max())
    print(f"fp32 after {K} steps: max |dx| = {err:.3e} m, max |dv| = {verr:.3e} m/s, "
          f"contacts {len(snap['pairs'])} vs {len(ref.hist_keys)}", flush=True)
    assert err <= 1e-9
    eng.close()


def test_c3_verlet_fp64_bit_identical_to_oracle(state):
    vstate = dict(state, config=dict(state["config"], integrator="verlet"), verlet_start=1)
    ref = _oracle(vstate)
    assert ref.integrator == "verlet" and ref.verlet_start
    eng = BedEngine.from_snapshot(vstate, fp32=False)

    def same(tag):
        snap = eng.snapshot()
        assert np.array_equal(_keys(snap), ref.hist_keys), f"contact-pair set differs {tag}"
        assert np.array_equal(snap["pair_xi"], ref.hist_xi), f"spring histories differ {tag}"
        assert np.array_equal(snap["positions"], ref.positions), f"positions differ {tag}"
        assert np.array_equal(snap["velocities"], ref.velocities), f"velocities differ {tag}"
        return snap

    for k in range(1, 6):
        eng.step(1)
        ref.step()
        snap = same(f"at step {k}")
        print(f"verlet step {k}: {len(snap['pairs'])} contacts, identical", flush=True)
    eng.step(6)
    for _ in range(6):
        ref.step()
    same("after a 6-step call")
    assert np.array_equal(eng.full_velocities().cpu().numpy(), ref.full_velocity()), \
        "full-step velocities differ"
    assert int(eng.status()[0]) == 0
    eng.close()

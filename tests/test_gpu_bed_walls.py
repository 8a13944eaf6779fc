"""The bed's right wall on the device (gate at rest, then an indenter driven
into the column) against the FP64 oracle: positions, velocities, contact
pairs and wall springs bit-identical step by step, ledger (with the wall's
work) to 1e-11; graph replays and restarts with a moving wall; and the
settled release state of a deposited column."""
import numpy as np
import pytest
import torch

from oracle.bed import OracleBed
from tensordem_amd.bed import BedEngine, BedSpec, settled_column
from tensordem_amd.bedconfig import BedConfig, column_deposit

pytestmark = pytest.mark.gpu

R0 = 0.01


def oracle_pairs(ref):
    n = len(ref.radii)
    k = ref.hist_keys
    return np.stack((k // n, k % n), axis=1) if len(k) else np.zeros((0, 2), dtype=np.int64)


def gated(n, seed):
    x, r, width = column_deposit(n, R0, seed=seed)
    cfg = BedConfig(right_wall=True, right_wall_x=width, right_wall_speed=0.0)
    return cfg, x, r, width


def same(eng, ref, what):
    assert np.array_equal(eng.x.cpu().numpy(), ref.positions), what
    assert np.array_equal(eng.v.cpu().numpy(), ref.velocities), what
    got = eng.contact_pairs() - eng.read_members()[0].first
    assert np.array_equal(got, oracle_pairs(ref)), what
    xw = eng.xi_wall.cpu().numpy()
    for k, arr in enumerate((ref.xi_floor, ref.xi_left, ref.xi_right)):
        assert np.array_equal(xw[:, 2 * k:2 * k + 2], arr), (what, k)


def test_gate_then_indenter_bit_exact_with_oracle():
    cfg, x, r, width = gated(800, 4)
    ref = OracleBed(cfg, x, r)
    eng = BedEngine([BedSpec(cfg, x, r, dt=ref.dt)])
    for k in range(1, 2001):
        eng.step(1)
        ref.step()
        if k % 500 == 0:
            same(eng, ref, f"gate, step {k}")
    assert ref.last[7] > 0                                   # disks lean on the gate
    # the gate becomes an indenter pushing into the column at 0.5 m/s
    eng.set_right_wall(speed=-0.5)
    ref.set_right_wall(speed=-0.5)
    assert eng.right_wall_position() == ref.right_wall_position
    for k in range(1, 3001):
        eng.step(1)
        ref.step()
        if k % 500 == 0:
            same(eng, ref, f"indenter, step {k}")
    assert eng.status()[0] == 0
    acc = eng.accounting[0].cpu().numpy()
    assert ref.acc[6] > 0                                    # the indenter did work
    for i in range(7):
        idx = 8 if i == 6 else i
        assert acc[idx] == pytest.approx(ref.acc[i], rel=1e-11, abs=1e-15), i
    e = eng.energies()[0].cpu().numpy()
    assert e[2] == pytest.approx(ref.energies()["elastic_normal_J"], rel=1e-10)


@pytest.mark.parametrize("fp32", [False, True])
def test_moving_wall_graph_replays_and_restart(fp32):
    cfg, x, r, width = gated(1500, 6)
    cfg = BedConfig(right_wall=True, right_wall_x=width, right_wall_speed=-0.3)
    eager = BedEngine([BedSpec(cfg, x, r)], fp32=fp32)
    graph = BedEngine([BedSpec(cfg, x, r)], fp32=fp32)
    for rep in range(3):
        eager.step(400)
        graph.run_graph(200)
        torch.cuda.synchronize()
        _ = eager.energies()
        assert torch.equal(eager.x, graph.x) and torch.equal(eager.v, graph.v), rep
        assert torch.equal(eager.xi_wall, graph.xi_wall), rep
    state = graph.snapshot()
    assert state["config"]["right_wall"] and state["config"]["right_wall_speed"] == -0.3
    back = BedEngine.from_snapshot(state, fp32=fp32)
    assert back.right_wall_position() == graph.right_wall_position()
    graph.step(600)
    back.step(600)
    assert torch.equal(graph.x, back.x) and torch.equal(graph.v, back.v)
    assert torch.equal(graph.xi_wall, back.xi_wall)
    assert np.array_equal(graph.contact_pairs(), back.contact_pairs())


def test_settled_release_state():
    # settling ends once the kinetic energy is below 1e-6 of m g H (checked
    # every 10 000 steps); the step count that takes depends on the FP32
    # rounding of the forces, so the cap leaves room
    state, info = settled_column(20000, R0, seed=1, max_steps=150000)
    print(info)
    assert info["kinetic_rel"] < 1e-6
    n = len(state["radii"])
    assert state["step_count"] == 0 and not state["config"]["right_wall"]
    assert not (np.asarray(state["wall_active"]) & 4).any()
    pairs_per_disk = info["contact_pairs"] / n
    assert 1.0 < pairs_per_disk < 2.5
    # the released column starts collapsing: its front moves out past the gate
    eng = BedEngine.from_snapshot(state, fp32=True)
    front0 = float((eng.x[:, 0] + eng.radius).max())
    eng.run_graph(1000, repeats=5)
    assert eng.status()[0] == 0
    assert float((eng.x[:, 0] + eng.radius).max()) > front0 + 2 * R0

"""HIP Hertz-Mindlin bed against the FP64 bed oracle on the same seeded columns.

FP64: positions, velocities and the contact-pair set bit-identical to the
oracle at every checkpoint (same operand order, no fused multiply-add);
ledger sums to 1e-12 relative.  FP32 forces: the contact-pair set is decided
in FP64 and must equal the oracle's set at the same positions.
"""
import numpy as np
import pytest
import torch

from oracle.bed import OracleBed
from tensordem_amd.bed import BedEngine, BedSpec
from tensordem_amd.bedconfig import BedConfig, column_packing

pytestmark = pytest.mark.gpu

R0 = 0.01


def column(n, seed, **kw):
    x, r = column_packing(n, R0, seed=seed)
    return BedConfig(**kw), x, r


def oracle_pairs(ref: OracleBed) -> np.ndarray:
    n = len(ref.radii)
    k = ref.hist_keys
    return np.stack((k // n, k % n), axis=1) if len(k) else np.zeros((0, 2), dtype=np.int64)


def shake(n, seed, scale=0.3):
    return np.random.default_rng(seed).normal(scale=scale, size=(n, 2))


def collapsing(n, seed, steps=30000):
    """A column released 0.3 s ago: collapsing, with a dense contact network
    (about one contact per disk).  Both sides of a comparison start here with
    empty spring histories."""
    cfg, x, r = column(n, seed)
    e = BedEngine([BedSpec(cfg, x, r)])
    e.step(steps)
    assert e.status()[0] == 0
    return cfg, e.x.cpu().numpy(), r, e.v.cpu().numpy()


def test_column_fp64_bit_exact_with_contact_sets():
    cfg, x, r, v = collapsing(600, seed=3)
    v0 = v + shake(len(r), 3, scale=0.05)
    ref = OracleBed(cfg, x, r)
    ref.velocities[:] = v0
    eng = BedEngine([BedSpec(cfg, x, r, dt=ref.dt, velocities=v0)])
    rebuild_seen = False
    for k in range(1, 3001):
        eng.step(1)
        ref.step()
        if k % 250 == 0:
            xg = eng.x.cpu().numpy()
            assert np.array_equal(xg, ref.positions), f"step {k}: |dx| {np.abs(xg - ref.positions).max():.3e}"
            assert np.array_equal(eng.v.cpu().numpy(), ref.velocities)
            got = eng.contact_pairs() - eng.read_members()[0].first
            assert np.array_equal(got, oracle_pairs(ref)), f"step {k} contact sets differ"
            rebuild_seen |= eng.read_members()[0].list_parity == 1
    assert len(ref.hist_keys) > 400
    assert rebuild_seen
    assert eng.status()[0] == 0
    acc = eng.accounting[0].cpu().numpy()
    for i in range(6):
        assert acc[i] == pytest.approx(ref.acc[i], rel=1e-11, abs=1e-15), i
    e = eng.energies()[0].cpu().numpy()
    oe = ref.energies()
    assert e[0] == pytest.approx(oe["kinetic_J"], rel=1e-12)
    assert e[1] == pytest.approx(oe["potential_J"], rel=1e-12)
    assert e[2] == pytest.approx(oe["elastic_normal_J"], rel=1e-10)


def test_fp32_contact_sets_equal_fp64_decisions():
    cfg, x, r, v = collapsing(800, seed=5)
    v0 = v + shake(len(r), 5, scale=0.05)
    eng = BedEngine([BedSpec(cfg, x, r, velocities=v0)], fp32=True)
    probe = OracleBed(cfg, x, r)
    seen = 0
    for k in range(1, 2001):
        before = eng.x.cpu().numpy().copy() if k % 200 == 0 else None
        eng.step(1)
        if before is not None:
            # pairs in contact during step k were decided at the positions
            # entering step k: compare with the oracle's FP64 set there
            probe.positions[:] = before
            exp = probe.contacts().astype(np.int64)
            assert np.array_equal(eng.contact_pairs(), exp), f"step {k}"
            seen += len(exp)
    assert seen > 3000
    # FP32 vs FP64 trajectories: stated divergence after 2000 steps
    ref = OracleBed(cfg, x, r)
    ref.velocities[:] = v0
    for _ in range(2000):
        ref.step()
    err = np.abs(eng.x.cpu().numpy() - ref.positions).max()
    print(f"bed fp32: max |x - x_fp64| after 2000 steps = {err:.3e} m")
    assert err < 1e-3 * R0


def test_bed_batch_members_equal_alone():
    specs = []
    for seed, n in ((11, 300), (12, 700), (13, 257)):
        cfg, x, r = column(n, seed, restitution=0.3 + 0.1 * (seed - 11))
        specs.append(BedSpec(cfg, x, r, velocities=shake(n, seed)))
    batch = BedEngine(specs)
    batch.step(900)
    for k, s in enumerate(specs):
        solo = BedEngine([s])
        solo.step(900)
        m = batch.read_members()[k]
        assert np.array_equal(batch.x[m.first:m.first + m.n].cpu().numpy(), solo.x.cpu().numpy())
        assert np.array_equal(batch.accounting[k].cpu().numpy(), solo.accounting[0].cpu().numpy())


def test_bed_graph_replay_equals_eager():
    cfg, x, r = column(500, seed=21)
    v0 = shake(len(r), 21)
    eager = BedEngine([BedSpec(cfg, x, r, velocities=v0)])
    graph = BedEngine([BedSpec(cfg, x, r, velocities=v0)])
    for rep in range(3):
        eager.step(200)
        graph.run_graph(100)
        torch.cuda.synchronize()
        assert np.array_equal(eager.x.cpu().numpy(), graph.x.cpu().numpy()), rep
    assert graph.read_members()[0].step_count == 600


def test_nonfinite_state_sets_the_status_bit():
    """A NaN velocity propagates to NaN positions: the disk is binned into the
    grid's first cell, never compares as in contact, and the member's status
    carries TDM_ST_NONFINITE (the facade's FloatingPointError)."""
    cfg, x, r = column(300, seed=31)
    v = np.zeros_like(x)
    v[17] = [np.nan, 0.0]
    eng = BedEngine([BedSpec(cfg, x, r, velocities=v)])
    eng.step(4)
    assert eng.status()[0] & 1
    assert np.isnan(eng.x.cpu().numpy()[17]).any()
    assert np.isfinite(np.delete(eng.x.cpu().numpy(), 17, axis=0)).all()


def test_list_and_cell_overflow_are_flagged_not_written_past():
    cfg, x, r, v = collapsing(600, seed=33, steps=20000)
    small_lists = BedEngine([BedSpec(cfg, x, r, velocities=v)], cand_slots=2)
    small_lists.step(50)
    assert small_lists.status()[0] & 4          # TDM_ST_CAND_OVERFLOW
    small_cells = BedEngine([BedSpec(cfg, x, r, velocities=v)], cell_cap=1)
    small_cells.step(50)
    assert small_cells.status()[0] & 8          # TDM_ST_CELL_OVERFLOW
    ok = BedEngine([BedSpec(cfg, x, r, velocities=v)])
    ok.step(50)
    assert ok.status()[0] == 0


def test_members_of_very_different_sizes_equal_alone():
    """Ragged batch: a 1-disk bed, a 255-disk and a 3000-disk column."""
    specs = [BedSpec(BedConfig(), np.array([[0.05, 0.02]]), np.array([0.01]),
                     velocities=np.array([[0.3, -0.2]]))]
    for seed, n in ((41, 255), (42, 3000)):
        cfg, x, r = column(n, seed)
        specs.append(BedSpec(cfg, x, r, velocities=shake(n, seed)))
    batch = BedEngine(specs)
    batch.step(1200)
    for k, s in enumerate(specs):
        solo = BedEngine([s])
        solo.step(1200)
        m = batch.read_members()[k]
        assert np.array_equal(batch.x[m.first:m.first + m.n].cpu().numpy(), solo.x.cpu().numpy())
        assert np.array_equal(batch.v[m.first:m.first + m.n].cpu().numpy(), solo.v.cpu().numpy())
    assert not any(batch.status())


def test_bed_restart_continues_bit_for_bit(tmp_path):
    """Snapshot member 1 of a two-bed batch mid-collapse (active spring
    histories, floor and wall springs), write it to disk, restore it as a
    one-bed engine and run both on: positions, velocities, ledger and the
    contact-pair set stay identical."""
    from tensordem_amd.bed import load_bed_restart, save_bed_restart
    cfg0, x0, r0 = column(300, seed=51)
    cfg, x, r, v = collapsing(700, seed=52, steps=20000)
    batch = BedEngine([BedSpec(cfg0, x0, r0, velocities=shake(300, 51)),
                       BedSpec(cfg, x, r, velocities=v + shake(700, 52, scale=0.05))])
    batch.step(777)
    snap = batch.snapshot(1)
    assert len(snap["pairs"]) > 300 and snap["wall_active"].any()
    assert np.abs(snap["pair_xi"]).max() > 0.0
    m = batch.read_members()[1]
    assert np.array_equal(snap["pairs"] + m.first, batch.contact_pairs()[
        (batch.contact_pairs()[:, 0] >= m.first)])
    save_bed_restart(snap, tmp_path / "bed.npz")
    back = BedEngine.from_snapshot(load_bed_restart(tmp_path / "bed.npz"))
    assert back.read_members()[0].step_count == 777
    for k in range(3):
        batch.step(500)
        back.step(500)
        sl = slice(m.first, m.first + m.n)
        assert np.array_equal(batch.x[sl].cpu().numpy(), back.x.cpu().numpy()), k
        assert np.array_equal(batch.v[sl].cpu().numpy(), back.v.cpu().numpy()), k
        assert np.array_equal(batch.accounting[1].cpu().numpy(), back.accounting[0].cpu().numpy())
    assert np.array_equal(batch.snapshot(1)["pair_xi"], back.snapshot()["pair_xi"])
    assert back.read_members()[0].time == batch.read_members()[1].time


def test_polydisperse_bed_sizes_its_hash_and_matches_oracle():
    """The floe bed's widest radius spread (polydispersity 1: radii in
    [0.5, 1.5] r0, a 3:1 ratio): the engine sizes its hash cells by the
    packing bound and stays bit-identical to the oracle, with no overflow."""
    x, r = column_packing(800, R0, polydispersity=0.5, seed=21)
    cfg = BedConfig()
    pre = BedEngine([BedSpec(cfg, x, r)])
    pre.step(40000)                        # fallen and collapsing: dense contacts
    assert pre.status()[0] == 0 and pre.cell_cap == 22
    x, v = pre.x.cpu().numpy(), pre.v.cpu().numpy()
    ref = OracleBed(cfg, x, r)
    ref.velocities[:] = v
    eng = BedEngine([BedSpec(cfg, x, r, dt=ref.dt, velocities=v)])
    for k in range(1, 17):
        eng.step(250)
        for _ in range(250):
            ref.step()
        assert eng.status()[0] == 0, f"step {250 * k}"
        assert np.array_equal(eng.x.cpu().numpy(), ref.positions), f"step {250 * k}"
        assert np.array_equal(eng.v.cpu().numpy(), ref.velocities), f"step {250 * k}"
    got = eng.contact_pairs() - eng.read_members()[0].first
    assert np.array_equal(got, oracle_pairs(ref))
    assert len(ref.hist_keys) > 300

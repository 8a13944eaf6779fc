"""Disk assembly with rotation on the device (`DiskSystem`,
tensordem_amd/csrc/disk.hip) against the FP64 oracle (oracle/disk2d.c), which
restates granular2d/engine.py:399-596.

Positions, velocities, spins and angles must be bit-identical to the oracle
after every checked step, so the state digest (engine.py:289-295) matches
too; ledger and energy sums are added in a different order than the oracle
adds them and must agree to 1e-12 relative (1e-9 for the wall and hull
reaction sums, whose terms cancel)."""
import math

import numpy as np
import pytest
import torch

from oracle.disk import OracleDisk
from tensordem_amd.granular2d import DiskConfig, DiskSystem, box_walls, ship_segments

pytestmark = pytest.mark.gpu


def _packing(n, seed, radius=0.010, aspect=1.0):
    """n disks of radius U[0.9, 1.1] * radius on a jittered square grid of
    pitch 2.1 * radius (the largest neighbours overlap slightly), in a w x h
    box with w / h close to aspect."""
    rng = np.random.default_rng(seed)
    r = radius * rng.uniform(0.9, 1.1, size=n)
    nx = max(1, int(math.ceil(math.sqrt(n * aspect))))
    ny = int(math.ceil(n / nx))
    s = 2.1 * radius
    ix, iy = np.arange(n) % nx, np.arange(n) // nx
    slack = np.maximum(0.5 * s - r, 0.0)[:, None]
    pos = (np.stack([ix, iy], 1) + 0.5) * s + rng.uniform(-1, 1, size=(n, 2)) * slack
    return pos, r, nx * s, ny * s


def _pair(cfg, pos, r, **kw):
    dev = DiskSystem(pos, r, cfg, **kw)
    ref = OracleDisk(pos, r, cfg, **kw)
    return dev, ref


def _same_state(dev, ref, what):
    for name in ("pos", "vel", "omega", "theta"):
        got = getattr(dev, name).cpu().numpy()
        want = getattr(ref, name)
        assert np.array_equal(got, want), f"{what}: {name} differs, max |d| = " \
            f"{np.abs(got - want).max():.3e}"


def _close(a, b, rel, what):
    assert abs(a - b) <= rel * max(abs(a), abs(b)) + 1e-300, f"{what}: {a} vs {b}"


def _ledgers_agree(dev, ref):
    for k in ("normal_dashpot", "friction", "rolling"):
        _close(dev.energy_ledger[k], ref.energy_ledger[k], 1e-12, k)
    de, re_ = dev.energies(), ref.energies()
    for k in ("elastic_normal", "elastic_tangential", "elastic_rolling", "kinetic"):
        _close(de[k], re_[k], 1e-12, k)


def _advance_both(dev, ref, chunks, every):
    for c in range(chunks):
        dev.step(every)
        ref.step(every)
        _same_state(dev, ref, f"after {(c + 1) * every} steps")


def test_box_with_hull_gravity_and_rolling_matches_oracle():
    pos, r, w, h = _packing(400, 1)
    cfg = DiskConfig(mu=0.4, mu_rolling=0.05, zeta=0.2, gravity=(0.0, -9.81), skin=0.003)
    hull = ship_segments(0.3 * w, 0.2 * h, (0.35 * w, 0.5 * h), bow=0.1 * w)
    dev, ref = _pair(cfg, pos, r, walls=box_walls(w, h), hull=hull)
    v = np.random.default_rng(2).normal(scale=0.2, size=pos.shape)
    dev.vel = v
    ref.vel[:] = v
    dev.enable_energy_ledger()
    ref.enable_energy_ledger()
    _advance_both(dev, ref, 6, 150)
    assert dev.digest() == ref.digest()
    assert dev.time == ref.time and dev.steps == ref.steps == 900
    _ledgers_agree(dev, ref)
    assert ref.last[10] > 50 and ref.last[11] > 0 and ref.last[12] > 0  # pair, wall, hull
    assert np.allclose(dev.wall_force, ref.wall_force, rtol=1e-9, atol=1e-12)
    assert np.allclose(dev.hull_force, ref.hull_force, rtol=1e-9, atol=1e-12)
    assert dev.rebuilds > 1


def test_doubly_periodic_with_drag_and_decay_matches_oracle():
    pos, r, w, h = _packing(500, 3)
    cfg = DiskConfig(mu=0.3, mu_rolling=0.02, drift_velocity=(0.05, -0.02), drag_time=0.05,
                     damping_time=0.5, skin=0.002)
    dev, ref = _pair(cfg, pos, r, periodic_x=w, periodic_y=h)
    v = np.random.default_rng(4).normal(scale=0.1, size=pos.shape)
    dev.vel = v
    ref.vel[:] = v
    _advance_both(dev, ref, 4, 200)
    assert bool(((dev.pos[:, 0] >= 0) & (dev.pos[:, 0] <= w)).all())


def test_tiny_periodic_cell_matches_oracle():
    # fewer than three cells along each periodic axis
    pos = np.array([[0.01, 0.01], [0.028, 0.012], [0.005, 0.03], [0.03, 0.032]])
    r = np.array([0.01, 0.011, 0.009, 0.012])
    cfg = DiskConfig(mu=0.5, mu_rolling=0.1, skin=0.001)
    dev, ref = _pair(cfg, pos, r, periodic_x=0.045, periodic_y=0.05)
    _advance_both(dev, ref, 3, 300)


def test_moving_walls_match_oracle_and_drag_the_disk():
    cfg = DiskConfig(kn=1.0e5, kt_ratio=0.5, zeta=0.3, mu=0.5, mu_rolling=0.0, density=2500.0,
                     thickness=0.01, gravity=(0.0, -9.81))
    walls = box_walls(0.2, 0.1, left=False, right=False)
    dev, ref = _pair(cfg, [[0.1, 0.0051]], [0.005], walls=walls,
                     wall_velocity=[[0.5, 0.0], [0.5, 0.0]])
    _advance_both(dev, ref, 3, 1000)
    assert dev.vel[0, 0].item() > 0.02


def test_contact_records_match_oracle_and_inspection_changes_nothing():
    pos, r, w, h = _packing(300, 5)
    cfg = DiskConfig(mu=0.5, mu_rolling=0.05, gravity=(0.0, -9.81), skin=0.003)
    dev, ref = _pair(cfg, pos, r, walls=box_walls(w, h))
    v = np.random.default_rng(6).normal(scale=0.2, size=pos.shape)
    dev.vel = v
    ref.vel[:] = v
    _advance_both(dev, ref, 2, 100)
    for _ in range(2):
        got, want = dev.contact_forces(), ref.contact_forces()
        assert len(want["i"]) > 30 and (want["kind"] == 1).any()
        for k, v in want.items():
            assert np.array_equal(got[k], v), k
    _advance_both(dev, ref, 2, 100)
    # once enabled, `contacts` holds the records of the last step of a call
    got = dev.contacts
    for k, v in ref.contacts.items():
        assert np.array_equal(got[k], v), k


def test_caller_edits_between_steps_match_oracle():
    pos, r, w, h = _packing(200, 7)
    cfg = DiskConfig(mu=0.4, mu_rolling=0.05, gravity=(0.0, -9.81))
    dev, ref = _pair(cfg, pos, r, walls=box_walls(w, h))
    _advance_both(dev, ref, 1, 100)
    dev.vel[3, 0] = 0.7
    ref.vel[3, 0] = 0.7
    dev.omega[5] = -3.0
    ref.omega[5] = -3.0
    _advance_both(dev, ref, 1, 50)
    shift = np.array([0.0, 0.2 * h])
    dev.pos[10] += torch.as_tensor(shift, device="cuda")
    ref.pos[10] += shift
    new = ref.pos.copy()
    dev.pos = new
    _advance_both(dev, ref, 2, 100)


def test_many_blocks_with_rebuilds_match_oracle():
    pos, r, w, h = _packing(6000, 9, radius=0.004, aspect=3.0)
    cfg = DiskConfig(mu=0.4, mu_rolling=0.03, gravity=(0.0, -9.81), skin=0.0005)
    dev, ref = _pair(cfg, pos, r, walls=box_walls(w, h))
    v = np.random.default_rng(10).normal(scale=0.3, size=pos.shape)
    dev.vel = v
    ref.vel[:] = v
    dev.enable_energy_ledger()
    ref.enable_energy_ledger()
    _advance_both(dev, ref, 2, 150)
    _ledgers_agree(dev, ref)
    assert dev.rebuilds >= 3


def test_reference_sliding_disk_case_on_device():
    radius = 0.01
    cfg = DiskConfig(kn=3.0e4, mu=0.3, mu_rolling=0.0, zeta=0.0, gravity=(0.0, -9.81),
                     skin=0.004)
    s = DiskSystem([[0.1, radius - 1.0e-6]], [radius], cfg,
                   walls=box_walls(1.0, 0.2, right=False, top=False))
    s.vel[0, 0] = 0.5
    s.enable_energy_ledger()
    s.step(3000)
    analytic = float(s.mass[0]) * 0.5 ** 2 / 6.0
    assert s.dissipated_work() == pytest.approx(analytic, rel=0.02)
    assert s.vel[0, 0].item() == pytest.approx(2.0 * 0.5 / 3.0, rel=0.02)


def test_graph_replays_match_eager_steps():
    pos, r, w, h = _packing(800, 11)
    cfg = DiskConfig(mu=0.4, mu_rolling=0.05, gravity=(0.0, -9.81), skin=0.002)
    eager = DiskSystem(pos, r, cfg, walls=box_walls(w, h))
    graph = DiskSystem(pos, r, cfg, walls=box_walls(w, h))
    for rep in range(3):
        eager.step(40)
        graph.run_graph(20)
        torch.cuda.synchronize()
        _ = eager.kinetic_energy()         # eager work between the replays
        for name in ("pos", "vel", "omega", "theta"):
            assert torch.equal(getattr(eager, name), getattr(graph, name)), (rep, name)
    assert eager.digest() == graph.digest()
    # replays write no records: `contacts` then evaluates the current state
    before = graph.contact_forces()
    graph.run_graph(1)
    now = graph.contacts
    again = graph.contact_forces()
    assert len(now["i"]) > 0
    for k, v in again.items():
        assert np.array_equal(now[k], v), k
    assert not np.array_equal(before["f"], now["f"])


def test_crowded_disks_with_many_contacts_match_oracle():
    # a hexagonal cluster squeezed so the inner disks touch six neighbours
    # (more rolling torques per disk than the kernel keeps in registers)
    # plus a floor and a side wall
    rows, cols, r = 7, 7, 0.01
    pitch = 1.9 * r
    pos = np.array([[(c + 0.5 * (k % 2)) * pitch + 0.03, k * pitch * math.sqrt(3) / 2 + 0.0095]
                    for k in range(rows) for c in range(cols)])
    rad = np.full(len(pos), r)
    cfg = DiskConfig(mu=0.6, mu_rolling=0.2, gravity=(0.0, -9.81), dt=1e-6)
    walls = np.array([[[-1.0, 0.0], [1.0, 0.0]], [[0.03 - 0.0095, -1.0], [0.03 - 0.0095, 1.0]]])
    dev, ref = _pair(cfg, pos, rad, walls=walls)
    w = np.random.default_rng(12).normal(scale=2.0, size=len(pos))
    dev.omega = w
    ref.omega[:] = w
    dev.enable_energy_ledger()
    ref.enable_energy_ledger()
    _advance_both(dev, ref, 3, 100)
    _ledgers_agree(dev, ref)
    got, c = dev.contact_forces(), ref.contact_forces()
    for k, v in c.items():
        assert np.array_equal(got[k], v), k
    per_disk = np.bincount(np.r_[c["i"], c["j"][c["j"] >= 0]], minlength=len(pos))
    assert per_disk.max() >= 6


def test_doubly_periodic_wrap_and_cell_area_match_reference_case():
    # tests/test_granular2d_energy_minimization.py:231-241: two large disks in
    # a 2 x 2 cell drifting at (10, -7) stay inside [0, 2) on both axes
    pos = np.array([[0.5, 0.5], [1.5, 1.5]])
    cfg = DiskConfig(kn=1.0, mu=0.0, density=1.0, thickness=1.0)
    dev, ref = _pair(cfg, pos, np.array([0.2, 0.2]), periodic_x=2.0, periodic_y=2.0)
    assert dev.domain_area() == pytest.approx(4.0)
    assert dev.n_periodic_axes == 2
    dev.set_uniform_velocity((10.0, -7.0))
    ref.vel[:] = (10.0, -7.0)
    _advance_both(dev, ref, 3, 1)
    p = dev.pos.cpu().numpy()
    assert np.all((p >= 0.0) & (p < 2.0))


def test_periodic_shear_channel_with_moving_walls_matches_oracle():
    # periodic along x, floor and lid walls sliding in opposite directions:
    # the minimum image, the boundary springs and the wall velocities at once
    pos, r, w, h = _packing(600, 13, aspect=4.0)
    cfg = DiskConfig(mu=0.5, mu_rolling=0.05, zeta=0.3, skin=0.002)
    walls = box_walls(w, h, left=False, right=False)
    dev, ref = _pair(cfg, pos, r, walls=walls, periodic_x=w,
                     wall_velocity=[[-0.3, 0.0], [0.3, 0.0]])
    dev.enable_energy_ledger()
    ref.enable_energy_ledger()
    _advance_both(dev, ref, 4, 250)
    _ledgers_agree(dev, ref)
    assert ref.last[11] > 0                                   # wall contacts
    assert np.allclose(dev.wall_force, ref.wall_force, rtol=1e-9, atol=1e-12)
    p = dev.pos.cpu().numpy()
    assert np.all((p[:, 0] >= 0.0) & (p[:, 0] < w))


def test_walls_moved_in_place_and_wall_velocity_set_later_match_oracle():
    # granular_identifiability/targets.py:241-246: a plate is translated by
    # editing DiskSystem.walls in place between steps, after wall_velocity
    # was assigned to the constructed system
    pos, r, w, h = _packing(300, 21)
    cfg = DiskConfig(mu=0.5, mu_rolling=0.05, zeta=0.2, gravity=(0.0, -9.81), skin=0.002)
    walls = box_walls(w, h + 0.004)
    dev = DiskSystem(pos, r, cfg, walls=walls.copy())
    ref = OracleDisk(pos, r, cfg, walls=walls.copy())
    speed = 0.2
    wv = np.zeros((len(walls), 2))
    wv[1] = (0.0, -speed)
    dev.wall_velocity = wv
    ref.wall_velocity = wv.copy()
    for k in range(6):
        for _ in range(40):
            dev.walls[1, :, 1] -= speed * dev.dt
            ref.walls[1, :, 1] -= speed * ref.dt
            dev.step(1)
            ref.step(1)
        _same_state(dev, ref, f"lid lowered for {(k + 1) * 40} steps")
    assert ref.last[11] > 0                                    # the lid touches disks
    assert np.allclose(dev.wall_force, ref.wall_force, rtol=1e-9, atol=1e-12)
    # graph replays pick the moved lid up too (a cached graph holds old boxes)
    g0 = dev.wall_force.copy()
    dev.run_graph(10)
    for _ in range(20):
        ref.step(1)
    dev.walls[1, :, 1] -= 20 * speed * dev.dt
    ref.walls[1, :, 1] -= 20 * speed * ref.dt
    dev.run_graph(10)
    for _ in range(20):
        ref.step(1)
    _same_state(dev, ref, "graph replays around a lid move")
    # the velocity of a wall, edited in place
    dev.wall_velocity[1] = (0.0, 0.0)
    ref.wall_velocity[1] = (0.0, 0.0)
    dev.step(30)
    ref.step(30)
    _same_state(dev, ref, "lid stopped")
    assert not np.array_equal(g0, dev.wall_force)
    with pytest.raises(ValueError, match="fixed at construction"):
        dev.walls = dev.walls[:2].copy()
        dev.step(1)


def test_boundary_spring_rank_passes_on_like_the_oracle():
    # two disks on a floor; the first is lifted off and the second inherits
    # the spring of rank 0 (engine.py:432-456)
    r, dt = 0.01, 1e-5
    cfg = DiskConfig(kn=3.0e4, zeta=0.0, mu=10.0, mu_rolling=0.0, dt=dt)
    floor = np.array([[[-1.0, 0.0], [1.0, 0.0]]])
    pos = np.array([[-0.2, r - 1e-4], [0.2, r - 2e-4]])
    dev, ref = _pair(cfg, pos, np.array([r, r]), walls=floor)
    v = np.array([[0.3, 0.0], [0.1, 0.0]])
    dev.vel = v
    ref.vel[:] = v
    dev.step(1)
    ref.step(1)
    _same_state(dev, ref, "both on the floor")
    # reading the records once turns them on for every later step, so
    # `contacts` then holds the last step's evaluation, as in the reference
    assert dev.contacts["i"].tolist() == ref.contacts["i"].tolist() == [0, 1]
    lifted = np.array([-0.2, 0.5])
    p = dev.pos.cpu().numpy()
    p[0] = lifted
    dev.pos = p
    ref.pos[0] = lifted
    dev.step(3)
    ref.step(3)
    _same_state(dev, ref, "after the lift-off")
    got, want = dev.contacts, ref.contacts
    for k, val in want.items():
        assert np.array_equal(got[k], val), k
    # inherited 0.3 dt plus three steps of about 0.1 dt; a spring kept per
    # disk would hold about 0.4 dt
    assert got["ft"][0] > cfg.kt_ratio * cfg.kn * 0.5 * dt


def test_graph_replays_raise_on_candidate_overflow_and_contacts_keep_energies():
    pos, r, w, h = _packing(400, 5)
    cfg = DiskConfig(mu=0.4, gravity=(0.0, -9.81), skin=0.002)
    sysd = DiskSystem(pos, r, cfg, walls=box_walls(w, h), cand_slots=1)
    with pytest.raises(RuntimeError, match="cand_slots"):
        sysd.run_graph(2)
    ok = DiskSystem(pos, r, cfg, walls=box_walls(w, h))
    ok.step(10)
    ok.run_graph(5)
    e0, wf0 = ok.energies(), ok.wall_force.copy()
    rec = ok.contacts                      # evaluated on the current state
    assert len(rec["i"]) > 0
    assert ok.energies() == e0
    assert np.array_equal(ok.wall_force, wf0)

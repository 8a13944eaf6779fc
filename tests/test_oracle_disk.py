"""Pins of the FP64 disk-assembly oracle (oracle/disk2d.c) against what the
reference's own tests and the mathematics fix:

* tests/test_granular2d_column_collapse.py:55-80 -- a disk launched on a
  frictional floor leaves rolling at 2 v0 / 3 after dissipating m v0^2 / 6;
* tests/test_granular2d_column_collapse.py:155-171 -- a head-on collision
  releases with the clipped-oscillator restitution;
* tests/test_granular2d_column_collapse.py:173-183 and :195-208 -- sticking
  and elastic rolling store energy without slip loss; contact inspection
  leaves the trajectory unchanged;
* tests/test_granular2d_force_chain_jamming.py:95-135, 207-218 -- pair
  repulsion kn*overlap, equal and opposite; wall reaction; momentum;
* tests/test_granular2d_wall_velocity.py -- a moving wall drags a disk, a
  static one does not;
* brute force: the contact set equals the all-pairs overlap filter under the
  minimum image; free fall matches the semi-implicit Euler closed form; the
  velocity decay is exp(-dt/T) per step; a periodic pair across the seam
  feels the force of the same pair moved into the interior.

* the boundary-spring rank rule (engine.py:432-456: a boundary contact's
  spring is keyed n*n + rank, ranks over the wall contacts by disk then
  segment and then the hull contacts): two disks on a floor, the first lifts
  off, and the second inherits its spring -- the stored value and the
  record's |f_t| against the hand-derived u dt + v_c dt; and a hull contact
  of a lower-index disk ranking after a wall contact."""
import math

import numpy as np
import pytest

from oracle.disk import OracleDisk
from tensordem_amd.granular2d import DiskConfig, box_walls


def test_sliding_disk_dissipation_matches_analytic_friction_work():
    radius = 0.01
    cfg = DiskConfig(kn=3.0e4, mu=0.3, mu_rolling=0.0, zeta=0.0, gravity=(0.0, -9.81),
                     skin=0.004)
    s = OracleDisk([[0.1, radius - 1.0e-6]], [radius], cfg,
                   walls=box_walls(1.0, 0.2, right=False, top=False))
    s.vel[0, 0] = 0.5
    s.enable_energy_ledger()
    start = s.energies()
    s.step(3000)
    end = s.energies()
    analytic = s.mass[0] * 0.5 ** 2 / 6.0
    assert s.dissipated_work() == pytest.approx(analytic, rel=0.02)
    assert s.vel[0, 0] == pytest.approx(2.0 * 0.5 / 3.0, rel=0.02)
    residual = ((start["kinetic"] + start["potential"] + start["elastic_normal"])
                - (end["kinetic"] + end["potential"] + end["elastic_normal"])
                - s.dissipated_work())
    assert abs(residual) < 0.01 * analytic


def test_head_on_dashpot_ledger_matches_clipped_restitution():
    zeta, radius = 0.2, 0.01
    cfg = DiskConfig(zeta=zeta, mu=0.0, mu_rolling=0.0, dt=1e-6)
    s = OracleDisk([[-radius, 0.0], [radius, 0.0]], [radius, radius], cfg)
    s.vel[:, 0] = [0.15, -0.15]
    initial = s.kinetic_energy()
    s.enable_energy_ledger()
    period = 2 * math.pi * math.sqrt(s.mass[0] / (2 * cfg.kn))
    s.step(math.ceil(period / s.dt))
    e = math.exp(-2 * zeta * math.acos(zeta) / math.sqrt(1 - zeta * zeta))
    assert s.dissipated_work() == pytest.approx(initial * (1 - e**2), rel=0.004)
    assert s.kinetic_energy() == pytest.approx(initial * e**2, rel=0.01)


def test_sticking_and_elastic_rolling_store_energy_without_slip_loss():
    cfg = DiskConfig(mu=1.0, mu_rolling=1.0, zeta=0.0, dt=1e-6)
    s = OracleDisk([[0.0, 0.0099]], [0.01], cfg, walls=[[[-0.1, 0.0], [0.1, 0.0]]])
    s.vel[0, 0] = 0.01
    s.omega[0] = 0.3
    ledger = s.enable_energy_ledger()
    s.step()
    en = s.energies()
    assert en["elastic_tangential"] > 0
    assert en["elastic_rolling"] > 0
    assert ledger["friction"] == ledger["rolling"] == 0.0


def test_contact_inspection_preserves_sliding_and_rolling_history():
    def integrate(probes):
        cfg = DiskConfig(mu=1.0, mu_rolling=1.0, zeta=0.0, dt=1e-6)
        s = OracleDisk([[0.0, 0.0099]], [0.01], cfg, walls=[[[-0.1, 0.0], [0.1, 0.0]]])
        s.vel[0, 0] = 0.01
        s.omega[0] = 0.3
        s.enable_energy_ledger()
        for _ in range(20):
            for _ in range(probes):
                s.contact_forces()
            s.step()
        return np.r_[s.pos.ravel(), s.vel.ravel(), s.omega, s.dissipated_work()]
    np.testing.assert_array_equal(integrate(3), integrate(0))


def test_pair_contact_is_repulsive_and_equal_opposite():
    s = OracleDisk(np.array([[0.3, 0.2], [0.3195, 0.2]]), np.array([0.01, 0.01]), DiskConfig())
    c = s.contact_forces()
    assert c["kind"].tolist() == [0]
    f = c["f_on_i"][0]
    assert f[0] < 0.0
    assert np.allclose(c["f_on_j"][0], -f)
    assert np.isclose(np.linalg.norm(f), s.config.kn * 5.0e-4, rtol=2e-2)
    s2 = OracleDisk(np.array([[0.3, 0.2], [0.321, 0.2]]), np.array([0.01, 0.01]), DiskConfig())
    assert len(s2.contact_forces()["i"]) == 0


def test_wall_contact_reaction_is_recorded():
    s = OracleDisk(np.array([[0.02, 0.0095]]), np.array([0.01]), DiskConfig(mu=0.3),
                   walls=box_walls(0.1, 0.1))
    c = s.contact_forces()
    wall = c["kind"] == 1
    assert wall.any()
    assert np.isclose(c["f_on_i"][wall, 1].sum(), -s.wall_force[1])
    assert s.wall_force[1] < 0.0


def _packing(n, phi, seed, radius=0.010):
    """Non-overlapping random disks in a square of packing fraction phi."""
    rng = np.random.default_rng(seed)
    r = np.where(rng.random(n) < 0.5, radius, 1.4 * radius)
    side = math.sqrt(float(np.sum(math.pi * r**2)) / phi)
    pos = np.zeros((n, 2))
    k = 0
    while k < n:
        p = rng.uniform(r[k], side - r[k], size=2)
        if k == 0 or np.all(np.linalg.norm(pos[:k] - p, axis=1) >= r[:k] + r[k]):
            pos[k] = p
            k += 1
    return pos, r, side


def test_momentum_is_conserved_without_external_forces():
    pos, r, _ = _packing(80, 0.42, 6)
    s = OracleDisk(pos, r, DiskConfig(mu=0.3, mu_rolling=0.02, skin=0.003))
    s.vel = 0.12 * np.random.default_rng(1).standard_normal((80, 2))
    p0 = s.momentum()
    s.step(1500)
    scale = float(np.sum(s.mass * np.linalg.norm(s.vel, axis=1))) + 1e-30
    assert float(np.linalg.norm(s.momentum() - p0)) / scale < 1e-10


def test_moving_wall_drags_a_disk_and_static_wall_does_not():
    cfg = DiskConfig(kn=1.0e5, kt_ratio=0.5, zeta=0.3, mu=0.5, mu_rolling=0.0, density=2500.0,
                     thickness=0.01, gravity=(0.0, -9.81))
    walls = box_walls(0.2, 0.1, left=False, right=False)
    moving = OracleDisk([[0.1, 0.0051]], [0.005], cfg, walls=walls,
                        wall_velocity=[[0.5, 0.0], [0.5, 0.0]])
    moving.step(3000)
    assert moving.vel[0, 0] > 0.02
    static = OracleDisk([[0.1, 0.0051]], [0.005], cfg, walls=walls)
    static.step(3000)
    assert abs(float(static.vel[0, 0])) < 1e-9


@pytest.mark.parametrize("periodic", [None, (0.5, None), (0.37, 0.41)])
def test_contact_set_equals_all_pairs_filter(periodic):
    rng = np.random.default_rng(4)
    n = 300
    Lx, Ly = (periodic or (None, None))
    w = Lx or 0.5
    h = Ly or 0.5
    pos = rng.uniform(0, 1, size=(n, 2)) * [w, h]
    r = rng.uniform(0.008, 0.02, size=n)
    s = OracleDisk(pos, r, DiskConfig(), periodic_x=Lx, periodic_y=Ly)
    c = s.contact_forces()
    got = set(zip(c["i"].tolist(), c["j"].tolist()))
    want = set()
    for i in range(n):
        d = pos[i + 1:] - pos[i]
        if Lx:
            d[:, 0] -= Lx * np.round(d[:, 0] / Lx)
        if Ly:
            d[:, 1] -= Ly * np.round(d[:, 1] / Ly)
        js = np.nonzero(r[i] + r[i + 1:] - np.linalg.norm(d, axis=1) > 0)[0] + i + 1
        want |= {(i, int(j)) for j in js}
    assert got == want and len(want) > 50
    keys = c["i"] * n + c["j"]
    assert np.all(np.diff(keys) > 0)          # ascending key order


def test_free_fall_and_decay_closed_forms():
    g, dt, k = -9.81, 1e-4, 500
    s = OracleDisk([[0.0, 1.0]], [0.01], DiskConfig(gravity=(0.0, g), dt=dt))
    s.step(k)
    assert s.vel[0, 1] == pytest.approx(k * g * dt, rel=1e-12)
    assert s.pos[0, 1] == pytest.approx(1.0 + g * dt * dt * k * (k + 1) / 2, rel=1e-12)
    T = 0.01
    d = OracleDisk([[0.0, 0.0]], [0.01], DiskConfig(damping_time=T, dt=dt))
    d.vel[0] = [1.0, 0.0]
    d.omega[0] = 2.0
    d.step(k)
    assert d.vel[0, 0] == pytest.approx(math.exp(-dt / T) ** k, rel=1e-12)
    assert d.omega[0] == pytest.approx(2.0 * math.exp(-dt / T) ** k, rel=1e-12)


def test_periodic_seam_pair_matches_interior_pair():
    L, r = 0.2, 0.01
    cfg = DiskConfig(mu=0.4, mu_rolling=0.05, dt=1e-6)
    seam = OracleDisk([[0.004, 0.1], [L - 0.0145, 0.1]], [r, r], cfg, periodic_x=L)
    inner = OracleDisk([[0.104, 0.1], [0.1 - 0.0145, 0.1]], [r, r], cfg, periodic_x=L)
    for s in (seam, inner):
        s.vel[0] = [-0.05, 0.02]
        s.omega[:] = [1.0, -2.0]
    seam.step(200)
    inner.step(200)
    dv = seam.vel - inner.vel
    assert np.abs(dv).max() < 1e-9 * np.abs(inner.vel).max() + 1e-15
    assert np.allclose(seam.omega, inner.omega, rtol=1e-9, atol=1e-12)
    assert np.all((seam.pos[:, 0] >= 0.0) & (seam.pos[:, 0] <= L))


def test_digest_is_reproducible_and_seed_sensitive():
    def digest(seed):
        pos, r, side = _packing(60, 0.40, seed)
        s = OracleDisk(pos, r, DiskConfig(mu=0.4, mu_rolling=0.02, gravity=(0.0, -9.81),
                                          skin=0.003), walls=box_walls(side, side))
        s.step(300)
        return s.digest()
    assert digest(11) == digest(11)
    assert digest(11) != digest(12)


# --- boundary springs are keyed by rank among the step's boundary contacts ---
# (engine.py:432-456: keys n*n + rank, ranks over wall contacts ordered by
# disk then segment, then hull contacts; the spring of rank k is whatever
# rank k held at the previous step)

def _two_on_floor(u, w, dt=1e-5):
    r = 0.01
    cfg = DiskConfig(kn=3.0e4, zeta=0.0, mu=10.0, mu_rolling=0.0, dt=dt)
    floor = np.array([[[-1.0, 0.0], [1.0, 0.0]]])
    s = OracleDisk([[-0.2, r - 1e-4], [0.2, r - 2e-4]], [r, r], cfg, walls=floor)
    s.vel[0] = [u, 0.0]
    s.vel[1] = [w, 0.0]
    return s, cfg, r


def test_boundary_spring_passes_to_the_next_contact_of_its_rank():
    u, w, dt = 0.3, -0.1, 1e-5
    s, cfg, r = _two_on_floor(u, w, dt)
    n = 2
    s.step()
    # both disks on the floor: rank 0 is disk 0, rank 1 is disk 1
    assert s.st_key.tolist() == [n * n, n * n + 1]
    assert s.st_xi[0] == pytest.approx([u * dt, 0.0], rel=1e-15)
    assert s.st_xi[1] == pytest.approx([w * dt, 0.0], rel=1e-15)
    # disk 0 leaves the floor; disk 1 becomes rank 0 and inherits disk 0's
    # spring: xi = u dt + v_c dt, v_c the slip speed of disk 1's contact point
    # (centre speed plus r omega: the lever to the floor is (0, -r))
    s.pos[0] = [-0.2, 0.5]
    k_t = cfg.kt_ratio * cfg.kn
    omega1 = dt * (r * (-k_t * w * dt)) / s.inertia[1]   # torque of the first step
    assert s.omega[1] == pytest.approx(omega1, rel=1e-12)
    v_c = float(s.vel[1, 0]) + r * float(s.omega[1])
    s.step()
    assert s.st_key.tolist() == [n * n]
    expect = u * dt + v_c * dt
    assert s.st_xi[0, 0] == pytest.approx(expect, rel=1e-13)
    assert abs(s.st_xi[0, 1]) < 1e-12 * abs(expect)   # normal (0, 1) up to round-off
    rec = s.contacts
    assert rec["i"].tolist() == [1] and rec["kind"].tolist() == [1]
    assert rec["ft"][0] == pytest.approx(k_t * abs(expect), rel=1e-13)
    # planted: a spring keyed by disk would have kept only disk 1's history
    assert abs(s.st_xi[0, 0] - (w * dt + v_c * dt)) > 1e-3 * abs(expect)


def test_wall_contacts_rank_before_hull_contacts():
    """Disk 0 touches a hull segment, disk 1 the floor: the wall contact of
    disk 1 takes rank 0 although disk 0 has the lower index."""
    r, dt, u, w = 0.01, 1e-5, 0.2, 0.05
    cfg = DiskConfig(kn=3.0e4, zeta=0.0, mu=10.0, mu_rolling=0.0, dt=dt)
    floor = np.array([[[-1.0, 0.0], [1.0, 0.0]]])
    hull = np.array([[[-1.0, 1.0], [1.0, 1.0]]])
    s = OracleDisk([[-0.2, 1.0 - r + 1e-4], [0.2, r - 1e-4]], [r, r], cfg, walls=floor,
                   hull=hull)
    s.vel[0] = [u, 0.0]
    s.vel[1] = [w, 0.0]
    s.step()
    assert s.st_key.tolist() == [4, 5]
    assert s.st_xi[0] == pytest.approx([w * dt, 0.0], rel=1e-15)   # floor, disk 1
    assert s.st_xi[1] == pytest.approx([u * dt, 0.0], rel=1e-15)   # hull, disk 0

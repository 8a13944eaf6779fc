"""The FP64 bed oracle against closed forms of its contact law.

No reference output exists for a 2-D Hertz-Mindlin bed, so the law is pinned
by textbook results: Hertz force and potential, Mindlin stick stiffness,
Coulomb sliding, restitution, momentum conservation, brute-force contacts.
"""
import math

import numpy as np
import pytest

from oracle.bed import OracleBed, BS_DASH, BS_FRIC, BS_NPAIRS
from oracle.plate import pairs_within
from tensordem_amd.bedconfig import BedConfig, column_packing

R = 0.01


def free(**kw):
    base = dict(gravity=0.0, floor=False, back_wall=False)
    base.update(kw)
    return BedConfig(**base)


def test_static_hertz_force_and_potential():
    cfg = free()
    for delta in (1e-6, 3e-5, 2e-4):
        x = np.array([[0.0, 0.0], [2 * R - delta, 0.0]])
        b = OracleBed(cfg, x, np.array([R, R]), dt=1e-7)
        b.step(integrate=False)
        r_eff = R / 2
        expect = (4.0 / 3.0) * cfg.e_star * math.sqrt(r_eff) * delta ** 1.5
        assert b.force[1, 0] == pytest.approx(expect, rel=1e-12)
        assert b.force[0, 0] == -b.force[1, 0]
        e = b.energies()["elastic_normal_J"]
        assert e == pytest.approx((8.0 / 15.0) * cfg.e_star * math.sqrt(r_eff) * delta ** 2.5,
                                  rel=1e-12)


def test_hertz_exponent_from_two_overlaps():
    """log F vs log delta has slope 3/2 (a dropped or squared term would not)."""
    cfg = free()
    f = []
    for delta in (1e-5, 4e-5):
        b = OracleBed(cfg, np.array([[0.0, 0.0], [2 * R - delta, 0.0]]), np.array([R, R]), dt=1e-7)
        b.step(integrate=False)
        f.append(b.force[1, 0])
    assert math.log(f[1] / f[0]) / math.log(4.0) == pytest.approx(1.5, rel=1e-12)


def test_mindlin_stick_stiffness():
    """A disk pressed on the floor and moved tangentially for one step stores
    xi = v dt; the tangential force is s_t xi with s_t = 8 G* sqrt(r d)."""
    cfg = BedConfig(gravity=0.0, back_wall=False, friction=0.5, wall_friction=10.0)
    delta, vt, dt = 1e-4, 1e-3, 1e-6
    b = OracleBed(cfg, np.array([[0.0, R - delta]]), np.array([R]), dt=dt)
    b.velocities[0] = [vt, 0.0]
    b.step(integrate=False)
    s_t = 8 * cfg.g_star * math.sqrt(R * delta)
    assert b.xi_floor[0, 0] == pytest.approx(vt * dt, rel=1e-15)
    assert b.force[0, 0] == pytest.approx(-s_t * vt * dt, rel=1e-12)
    assert 2 * cfg.g_star * 4 / (2 * cfg.e_star) == pytest.approx(
        2 * (1 - cfg.poisson) / (2 - cfg.poisson), rel=1e-12)


def test_coulomb_cap_while_sliding():
    cfg = BedConfig(back_wall=False, wall_friction=0.3)
    m = cfg.rho_pi_t * R * R
    # resting overlap under gravity: (4/3) E* sqrt(r) d^1.5 = m g
    d = (m * cfg.gravity / ((4.0 / 3.0) * cfg.e_star * math.sqrt(R))) ** (2.0 / 3.0)
    b = OracleBed(cfg, np.array([[0.0, R - d]]), np.array([R]), dt=2e-6)
    b.velocities[0] = [0.5, 0.0]
    for _ in range(200):
        b.step()
    b.step(integrate=False)
    fn = b.force[0, 1] + m * cfg.gravity
    ft = abs(b.force[0, 0])
    assert b.velocities[0, 0] > 0.1  # still sliding
    assert ft == pytest.approx(0.3 * fn, rel=1e-9)


def _impact_restitution_rk4(cfg, v0, r, dt=1e-8):
    """Independent 1-DOF impact: m d'' = -max(K d^1.5 + C sqrt(2 E* sqrt(r d) m) d', 0)
    with K = (4/3) E* sqrt(r), C = 2 sqrt(5/6) beta, integrated by RK4."""
    m = cfg.rho_pi_t * r * r
    K = (4.0 / 3.0) * cfg.e_star * math.sqrt(r)
    C = 2.0 * math.sqrt(5.0 / 6.0) * cfg.beta

    def acc(d, dd):
        if d <= 0:
            return 0.0
        sn = 2.0 * cfg.e_star * math.sqrt(r * d)
        f = K * d ** 1.5 + C * math.sqrt(sn * m) * dd
        return -max(f, 0.0) / m

    d, dd = 0.0, v0
    while True:
        k1 = (dd, acc(d, dd))
        k2 = (dd + 0.5 * dt * k1[1], acc(d + 0.5 * dt * k1[0], dd + 0.5 * dt * k1[1]))
        k3 = (dd + 0.5 * dt * k2[1], acc(d + 0.5 * dt * k2[0], dd + 0.5 * dt * k2[1]))
        k4 = (dd + dt * k3[1], acc(d + dt * k3[0], dd + dt * k3[1]))
        d += dt * (k1[0] + 2 * k2[0] + 2 * k3[0] + k4[0]) / 6
        dd += dt * (k1[1] + 2 * k2[1] + 2 * k3[1] + k4[1]) / 6
        if d <= 0 and dd < 0:
            return -dd / v0


def test_restitution_of_a_floor_impact():
    """The restitution realised by the viscous term (with the non-cohesive clamp,
    which raises it above the nominal e) equals an independent RK4 integration
    of the same impact to 1 %, and is within 0.06 of the nominal value."""
    for e in (0.5, 0.8):
        cfg = BedConfig(gravity=0.0, back_wall=False, restitution=e, wall_friction=0.0)
        b = OracleBed(cfg, np.array([[0.0, R + 1e-5]]), np.array([R]), dt=2e-7)
        b.velocities[0] = [0.0, -0.2]
        for _ in range(20000):
            b.step()
            if b.positions[0, 1] > R + 2e-5 and b.velocities[0, 1] > 0:
                break
        e_meas = b.velocities[0, 1] / 0.2
        e_ode = _impact_restitution_rk4(cfg, 0.2, R)
        assert e_meas == pytest.approx(e_ode, rel=0.01), (e, e_meas, e_ode)
        assert abs(e_meas - e) < 0.06


def test_momentum_conserved_in_a_free_cluster():
    rng = np.random.default_rng(4)
    x, r = column_packing(300, R, seed=4, gap=-0.05)
    cfg = free()
    b = OracleBed(cfg, x, r)
    b.velocities[:] = rng.normal(scale=0.05, size=b.velocities.shape)
    m = b.masses[:, None]
    p0 = (m * b.velocities).sum(0)
    for _ in range(300):
        b.step()
    assert b.last[BS_NPAIRS] > 0
    assert np.abs((m * b.velocities).sum(0) - p0).max() < 1e-13 * np.abs(m * b.velocities).sum()


def test_contact_set_equals_brute_force():
    x, r = column_packing(400, R, seed=9, gap=-0.08)
    b = OracleBed(free(), x, r)
    got = {tuple(p) for p in b.contacts().tolist()}
    # brute force with the per-pair cutoff r_i + r_j
    cand = pairs_within(x, 2 * r.max())
    exp = {(int(i), int(j)) for i, j in cand
           if (r[i] + r[j]) - math.sqrt(((x[j] - x[i]) ** 2).sum()) > 0}
    assert got == exp and len(got) > 50


def test_history_resets_on_separation():
    cfg = free(friction=0.5)
    x = np.array([[0.0, 0.0], [2 * R - 1e-5, 0.0]])
    b = OracleBed(cfg, x, np.array([R, R]), dt=1e-6)
    b.velocities[1] = [0.0, 0.01]
    b.step()
    assert len(b.hist_keys) == 1 and np.abs(b.hist_xi).max() > 0
    b.positions[1, 0] = 2 * R + 1e-4  # pulled apart
    b.step()
    assert len(b.hist_keys) == 0


def test_column_packing_is_overlap_free_and_tall():
    x, r = column_packing(2000, R, seed=1)
    assert (x[:, 1] - r).min() > 0 and (x[:, 0] - r).min() > 0
    d = pairs_within(x, 2.4 * R * 1.2)
    for i, j in d:
        assert math.sqrt(((x[j] - x[i]) ** 2).sum()) >= r[i] + r[j]
    h = x[:, 1].max() - x[:, 1].min()
    w = x[:, 0].max() - x[:, 0].min()
    assert h / w == pytest.approx(2.0, rel=0.1)


# --- pair-pair branch: R* = r_i r_j / (r_i + r_j), m* = m_i m_j / (m_i + m_j),
# forces lo -> hi (scale_ceiling/kernel.py:293-322 restated in 2-D) ----------

def _pair(delta, vi, vj, r=R, cfg=None, dt=1e-7):
    cfg = cfg or free(friction=0.5)
    x = np.array([[0.0, 0.0], [2 * r - delta, 0.0]])
    b = OracleBed(cfg, x, np.array([r, r]), dt=dt)
    b.velocities[0] = vi
    b.velocities[1] = vj
    return cfg, b


def test_pair_stick_stiffness_uses_half_radius():
    """Disk 0 slides along +y against disk 1 for one step: xi = v_t dt and the
    tangential force on 1 is +s_t xi with s_t = 8 G* sqrt(R* d), R* = r/2."""
    delta, vt, dt = 2e-5, 1e-3, 1e-7
    cfg, b = _pair(delta, [0.0, vt], [0.0, 0.0], dt=dt)
    b.step(integrate=False)
    s_t = 8 * cfg.g_star * math.sqrt(0.5 * R * delta)
    assert b.hist_keys.tolist() == [1]          # key i * n + j of the pair (0, 1)
    assert b.hist_xi[0] == pytest.approx([0.0, vt * dt], rel=1e-15)
    assert b.force[1, 1] == pytest.approx(s_t * vt * dt, rel=1e-12)
    assert b.force[0, 1] == -b.force[1, 1]
    # planted mistakes: the wall's R* = r, or no tangential force on the partner
    assert b.force[1, 1] != pytest.approx(8 * cfg.g_star * math.sqrt(R * delta) * vt * dt,
                                          rel=1e-3)
    assert b.force[1, 1] > 0


def test_pair_normal_damping_uses_reduced_mass():
    """Approaching at v_n: f_n = (2/3) s_n d + C sqrt(s_n m*) v_n with m* = m/2,
    and the dashpot power (f_n - f_el) v_n = C sqrt(s_n m*) v_n^2 >= 0."""
    delta, vn = 3e-5, 0.02
    cfg, b = _pair(delta, [vn, 0.0], [0.0, 0.0], cfg=free())
    b.step(integrate=False)
    m = cfg.rho_pi_t * R * R
    s_n = 2 * cfg.e_star * math.sqrt(0.5 * R * delta)
    f_el = (2.0 / 3.0) * s_n * delta
    visc = cfg.damping_coefficient * math.sqrt(s_n * 0.5 * m) * vn
    assert b.force[1, 0] == pytest.approx(f_el + visc, rel=1e-12)
    assert b.force[0, 0] == -b.force[1, 0]
    assert b.last[BS_DASH] == pytest.approx(visc * vn, rel=1e-12) and b.last[BS_DASH] > 0
    # planted: full mass instead of the reduced mass
    wrong = f_el + cfg.damping_coefficient * math.sqrt(s_n * m) * vn
    assert b.force[1, 0] != pytest.approx(wrong, rel=1e-4)
    # separating: the viscous term lowers f_n, never below zero
    cfg, b = _pair(delta, [-50.0, 0.0], [0.0, 0.0], cfg=free())
    b.step(integrate=False)
    assert b.force[1, 0] == 0.0 and b.force[0, 0] == 0.0


def test_pair_coulomb_cap_and_friction_power():
    """Held at constant overlap and sliding at v_t, the spring grows by v_t dt
    per step until s_t |xi| = mu f_n; from then on the friction power is
    mu f_n v_t (energy absorbed, >= 0) and the force on each disk opposes
    the relative slip."""
    delta, vt, dt, mu = 2e-5, 0.3, 1e-6, 0.5
    cfg, b = _pair(delta, [0.0, vt], [0.0, 0.0], cfg=free(friction=mu), dt=dt)
    s_n = 2 * cfg.e_star * math.sqrt(0.5 * R * delta)
    f_n = (2.0 / 3.0) * s_n * delta
    s_t = 8 * cfg.g_star * math.sqrt(0.5 * R * delta)
    steps_to_cap = mu * f_n / (s_t * vt * dt)
    assert 2 < steps_to_cap < 200
    for k in range(int(steps_to_cap) + 3):
        b.step(integrate=False)
    assert s_t * abs(b.hist_xi[0, 1]) == pytest.approx(mu * f_n, rel=1e-12)
    assert b.force[0, 1] == pytest.approx(-mu * f_n, rel=1e-12)   # on the sliding disk
    assert b.force[1, 1] == pytest.approx(mu * f_n, rel=1e-12)
    assert b.last[BS_FRIC] == pytest.approx(mu * f_n * vt, rel=1e-12)
    # unequal radii: R* and m* are the harmonic forms
    cfg = free(friction=mu)
    r0, r1 = 0.8 * R, 1.2 * R
    b = OracleBed(cfg, np.array([[0.0, 0.0], [r0 + r1 - delta, 0.0]]), np.array([r0, r1]),
                  dt=dt)
    b.velocities[0] = [0.0, 1e-4]
    b.step(integrate=False)
    reff = r0 * r1 / (r0 + r1)
    assert b.force[1, 1] == pytest.approx(8 * cfg.g_star * math.sqrt(reff * delta) * 1e-4 * dt,
                                          rel=1e-12)


# --- the right wall: indenter (moving) or gate (at rest), law on the velocity
# relative to the wall (scale_ceiling/kernel.py:338-401) ---------------------

def _right(delta, v, wall_v=0.0, dt=1e-7, **kw):
    base = dict(gravity=0.0, floor=False, back_wall=False, right_wall=True, right_wall_x=1.0,
                right_wall_speed=wall_v)
    base.update(kw)
    cfg = BedConfig(**base)
    b = OracleBed(cfg, np.array([[1.0 - (R - delta), 0.5]]), np.array([R]), dt=dt)
    b.velocities[0] = v
    return cfg, b


def test_right_wall_static_hertz_and_galilean_invariance():
    delta = 4e-5
    cfg, b = _right(delta, [0.0, 0.0])
    b.step(integrate=False)
    delta = R - (1.0 - b.positions[0, 0])      # the overlap as stored, to the last bit
    expect = (4.0 / 3.0) * cfg.e_star * math.sqrt(R) * delta ** 1.5
    assert b.force[0, 0] == pytest.approx(-expect, rel=1e-12)
    assert b.force[0, 1] == 0.0
    static = b.force.copy()
    # disk and wall moving together: no viscous term, the same force bit for bit
    cfg, b = _right(delta, [-0.7, 0.0], wall_v=-0.7)
    b.step(integrate=False)
    assert np.array_equal(b.force, static)
    # a disk carried along at the wall's speed sees no work from it either
    # (the work is F . v_wall with F on the disk, here the static force)
    assert b.last[8] == pytest.approx(static[0, 0] * -0.7, rel=1e-15)


def test_right_wall_approach_damping_and_slip():
    delta, U = 4e-5, 0.3
    cfg, b = _right(delta, [0.0, 0.0], wall_v=-U, restitution=0.5)
    b.step(integrate=False)
    delta = R - (1.0 - b.positions[0, 0])
    m = cfg.rho_pi_t * R * R
    s_n = 2 * cfg.e_star * math.sqrt(R * delta)
    f_el = (2.0 / 3.0) * s_n * delta
    assert b.force[0, 0] == pytest.approx(-(f_el + cfg.damping_coefficient *
                                            math.sqrt(s_n * m) * U), rel=1e-12)
    # sliding along the wall at w: xi = w dt, force -s_t xi on the disk
    w, dt = 2e-3, 1e-7
    cfg, b = _right(delta, [0.0, w], dt=dt, wall_friction=10.0)
    b.step(integrate=False)
    delta = R - (1.0 - b.positions[0, 0])
    s_t = 8 * cfg.g_star * math.sqrt(R * delta)
    assert b.xi_right[0] == pytest.approx([0.0, w * dt], rel=1e-15)
    assert b.force[0, 1] == pytest.approx(-s_t * w * dt, rel=1e-12)


def test_right_wall_moves_with_the_clock():
    """A free disk 2 mm left of the contact distance and a wall closing at
    0.4 m/s: the first step in contact is the first whose start time puts
    the wall within r of the disk centre."""
    U, dt, gap = 0.4, 1e-6, 2e-3
    cfg = BedConfig(gravity=0.0, floor=False, back_wall=False, right_wall=True,
                    right_wall_x=1.0, right_wall_speed=-U)
    b = OracleBed(cfg, np.array([[1.0 - R - gap, 0.5]]), np.array([R]), dt=dt)
    first = None
    for k in range(10000):
        b.step()
        if b.last[7] > 0:
            first = k
            break
    t_k = [0.0]
    for _ in range(first):
        t_k.append(t_k[-1] + dt)
    assert (1.0 - U * t_k[-1]) - (1.0 - R - gap) < R
    assert (1.0 - U * t_k[-2]) - (1.0 - R - gap) >= R
    assert first == pytest.approx(gap / (U * dt), abs=1)


@pytest.mark.parametrize("e", [1.0, 0.5])
def test_disk_struck_by_a_moving_indenter(e):
    """A wall at constant speed U strikes a resting free disk.  In the wall's
    frame this is the floor impact: the disk leaves at U (1 + e_real) with
    e_real the restitution of the same impact (RK4 reference, 1 %); the wall
    impulse in the ledger is m v_final and its work is the kinetic energy
    gained plus what the wall's dashpot absorbed (to the discretisation of
    the impact)."""
    U, dt = 0.2, 2e-7
    cfg = BedConfig(gravity=0.0, floor=False, back_wall=False, right_wall=True,
                    right_wall_x=1.0, right_wall_speed=-U, restitution=e, wall_friction=0.0)
    b = OracleBed(cfg, np.array([[1.0 - R - 1e-6, 0.5]]), np.array([R]), dt=dt)
    touched = False
    for _ in range(200000):
        b.step()
        touched = touched or b.last[7] > 0
        if touched and b.last[7] == 0:
            break
    m = cfg.rho_pi_t * R * R
    v = -b.velocities[0, 0]
    e_real = 1.0 if e == 1.0 else _impact_restitution_rk4(cfg, U, R)
    assert v == pytest.approx(U * (1.0 + e_real), rel=0.01)
    assert b.acc[4] == pytest.approx(-m * v, rel=1e-12)        # wall impulse on the disk
    # work of the wall = kinetic energy gained + energy its dashpot absorbed
    assert b.acc[6] == pytest.approx(0.5 * m * v * v + b.acc[2], rel=0.002)
    if e < 1.0:
        assert b.acc[2] > 0.1 * b.acc[6]
    if e == 1.0:
        assert b.acc[2] == 0.0 and abs(b.acc[6] - 0.5 * m * v * v) < 2e-3 * 0.5 * m * v * v


def test_deposited_column_is_loose_overlap_free_and_supported():
    from tensordem_amd.bedconfig import column_deposit
    x, r, width = column_deposit(20000, R, seed=2)
    assert (x[:, 0] - r).min() >= 0 and (x[:, 0] + r).max() <= width and (x[:, 1] - r).min() >= 0
    pr = pairs_within(x, 2.4 * R * 1.0001)
    d = np.sqrt(((x[pr[:, 1]] - x[pr[:, 0]]) ** 2).sum(1))
    gap = d - (r[pr[:, 0]] + r[pr[:, 1]])
    assert gap.min() > -1e-12 * R                       # no overlaps beyond round-off
    touch = np.abs(gap) < 1e-12 * R
    supported = np.zeros(len(r), bool)
    supported[pr[touch, 1]] = True                       # resting on a lower-index disk
    supported |= np.abs(x[:, 1] - r) < 1e-15
    assert supported.all()
    h = (x[:, 1] + r).max()
    phi = math.pi * (r * r).sum() / (width * h)
    assert 0.73 < phi < 0.78
    assert h / width == pytest.approx(2.0, rel=0.05)


# ---- velocity-Verlet (the floe bed's integrator, scale_ceiling/kernel.py:410-421)

def test_verlet_free_flight_under_gravity_is_exact():
    """Constant acceleration: kick-drift-kick reproduces x0 + v0 t + g t^2 / 2
    and v0 + g t (to rounding); semi-implicit Euler is off by g dt t / 2."""
    g, dt, steps = 9.81, 1e-3, 400
    x0, v0 = np.array([[0.3, 2.0]]), np.array([[0.7, 1.5]])
    t = steps * dt
    want_x = x0 + v0 * t + np.array([[0.0, -0.5 * g * t * t]])
    want_v = v0 + np.array([[0.0, -g * t]])
    out = {}
    for integ in ("verlet", "euler"):
        b = OracleBed(free(gravity=g, integrator=integ), x0, np.array([R]), dt=dt)
        b.velocities[:] = v0
        for _ in range(steps):
            b.step()
        out[integ] = (b.positions.copy(), b.full_velocity())
    xv, vv = out["verlet"]
    assert np.allclose(xv, want_x, rtol=0, atol=1e-13)
    assert np.allclose(vv, want_v, rtol=0, atol=1e-12)
    xe, _ = out["euler"]
    assert abs(xe[0, 1] - want_x[0, 1]) == pytest.approx(0.5 * g * dt * t, rel=1e-6)


def _reference_floe_steps(cfg, x, r, v, dt, steps):
    """kernel.py:410-421 literally: a force evaluation (springs committed)
    when no force is cached, v += 0.5 dt F / m, x += dt v, a force
    evaluation at the new positions, v += 0.5 dt F / m.  The force engine is
    the oracle's own evaluation (its Euler instance, never integrated)."""
    f = OracleBed(BedConfig(**{**cfg.__dict__, "integrator": "euler"}), x, r, dt=dt)
    f.velocities[:] = v
    inv_m = 1.0 / f.masses
    force = None
    for _ in range(steps):
        if force is None:
            f.step(integrate=False)
            force = f.force.copy()
        f.velocities += ((0.5 * dt) * force) * inv_m[:, None]
        f.positions += dt * f.velocities
        f.time += dt
        f.step(integrate=False)
        force = f.force.copy()
        f.velocities += ((0.5 * dt) * force) * inv_m[:, None]
    return f


def test_verlet_launch_order_equals_the_reference_step_sequence():
    """The device arrangement (one force evaluation per step, the previous
    step's second half kick folded in) is the reference's step sequence:
    after N steps the positions and the full-step velocities are identical,
    and after one more device step the springs are too (the reference's
    N-th step ends with the force evaluation the device does next)."""
    x, r = column_packing(120, R, seed=4)
    x = x * 0.82                         # overlapping at the start: springs from step 1
    x[:, 1] += r.max() - x[:, 1].min()   # resting on the floor
    cfg = BedConfig(integrator="verlet", right_wall=True,
                    right_wall_x=float((x[:, 0] + r).max()) - 2e-4, right_wall_speed=-0.05,
                    friction=0.4, wall_friction=0.3)
    v = np.random.default_rng(4).normal(scale=0.2, size=x.shape)
    dt = cfg.stable_dt(r.min())
    steps = 300
    ref = _reference_floe_steps(cfg, x, r, v, dt, steps)
    dev = OracleBed(cfg, x, r, dt=dt)
    dev.velocities[:] = v
    for _ in range(steps):
        dev.step()
    assert len(ref.hist_keys) > 25
    assert (ref.xi_right != 0).any() and (ref.xi_floor != 0).any()
    assert np.array_equal(dev.positions, ref.positions)
    assert np.array_equal(dev.full_velocity(), ref.velocities)
    dev.step()
    assert np.array_equal(dev.hist_keys, ref.hist_keys)
    assert np.array_equal(dev.hist_xi, ref.hist_xi)
    for a, b in ((dev.xi_floor, ref.xi_floor), (dev.xi_left, ref.xi_left),
                 (dev.xi_right, ref.xi_right)):
        assert np.array_equal(a, b)


def test_verlet_full_velocity_commits_nothing():
    cfg = BedConfig(integrator="verlet")
    x, r = column_packing(60, R, seed=2)
    x = x * 0.82
    b = OracleBed(cfg, x, r)
    for _ in range(50):
        b.step()
    keys, xi, xf = b.hist_keys.copy(), b.hist_xi.copy(), b.xi_floor.copy()
    v1 = b.full_velocity()
    v2 = b.full_velocity()
    assert np.array_equal(v1, v2)
    assert np.array_equal(keys, b.hist_keys) and np.array_equal(xi, b.hist_xi)
    assert np.array_equal(xf, b.xi_floor)
    assert len(keys) > 20
    assert not np.array_equal(v1, b.velocities)   # the half kick is there

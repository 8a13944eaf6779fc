"""The 3-D oracle (oracle/dem3d.c) pinned against the reference's recorded
outputs and against closed forms.

Recorded outputs (tests/golden/dem3d_fixtures.json, from the reference's
docs/dem3d_ci_regression_metrics.json): the three-point bending fixture
(24 steps; midspan deflection bit for bit, support reaction within 1e-13 N,
which is the cancellation of summing the support forces in another order),
the load-convergence peaks (8 and 16 steps) and the bottom-platen damper
probe.  Closed forms: lattice and bond counts, recommended dt, the
objective shear strain of affine deformations, Hooke's law of a bond,
linear and Hertz contact springs, the restitution of a linear
spring-dashpot collision, Cundall-Strack stick and slip, tool and platen
springs, free fall and drag decay of the semi-implicit Euler update.
Behaviours of the reference's tests/test_dem3d.py: equilibrium, rigid
rotation, irreversible tensile failure, broken bonds become contacts.
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from oracle.dem3d import OracleDEM3D, reference_lattice
from tensordem_amd.dem3dconfig import DEM3DConfig, build_lattice

GOLD = json.loads((Path(__file__).parent / "golden" / "dem3d_fixtures.json").read_text())


def bending_config():
    p = GOLD["bending_case"]
    return DEM3DConfig(
        nx=p["nx"], ny=p["ny"], nz=p["nz"], radius=p["radius_m"], density=p["density_kg_m3"],
        bond_stiffness=p["bond_stiffness_N_m"], contact_stiffness=p["contact_stiffness_N_m"],
        breaking_strain=p["breaking_strain"], shear_breaking_strain=p["shear_breaking_strain"],
        contact_damping=p["contact_damping_Ns_m"], drag=p["drag"], fix_x_edges=False,
        fix_bottom=False, tool_gap=1.0, tool_radius=0.01, tool_speed=0.01)


def bending_run(sim, steps, load_max):
    """docs/ci_regression_evidence.py:45-99 on any solver with the IceDEM3D surface
    (numpy state)."""
    sim.tool_start[:] = 1e6
    sim.tool_velocity[:] = 0.0
    x = sim.initial_positions[:, 0]
    z = sim.initial_positions[:, 2]
    xv = np.unique(x)
    lx, rx = xv[max(1, len(xv) // 4)], xv[min(len(xv) - 2, (3 * len(xv)) // 4)]
    supports = (z == z.min()) & ((x == lx) | (x == rx))
    loaded = (z == z.max()) & (x == xv[len(xv) // 2])
    sim.fixed[:] = supports
    nload = int(loaded.sum())
    rows = []
    for s in range(steps + 1):
        applied = load_max * s / steps
        loads = np.zeros_like(sim.positions)
        loads[loaded, 2] = -applied / nload
        if s:
            sim.step(loads)
        force, _ = sim.forces(loads, update_fracture=False)
        reaction = -force[supports].sum(0)
        rows.append((applied, reaction[2],
                     (sim.positions[loaded, 2] - sim.initial_positions[loaded, 2]).mean()))
    return rows, int(supports.sum()), nload


def test_bending_fixture_rows():
    fx = GOLD["bending_fixture"]
    cfg = bending_config()
    assert cfg.recommended_dt == fx["dt_s"]
    sim = OracleDEM3D(cfg)
    assert (len(sim.positions), len(sim.pairs)) == (fx["particle_count"], fx["bond_count"])
    rows, ns, nl = bending_run(sim, fx["steps"], GOLD["bending_case"]["peak_load_N"])
    assert (ns, nl) == (fx["support_particles"], fx["loaded_particles"])
    for (applied, reac, defl), want in zip(rows, fx["rows"]):
        assert applied == want["applied_load_N"]
        assert defl == want["midspan_deflection_m"]
        assert abs(reac - want["support_reaction_z_N"]) <= 1e-13
    assert sim.broken_bonds == 0


@pytest.mark.parametrize("run", GOLD["load_convergence"]["runs"], ids=lambda r: str(r["steps"]))
def test_load_convergence_peaks(run):
    sim = OracleDEM3D(bending_config())
    rows, _, _ = bending_run(sim, run["steps"], GOLD["load_convergence"]["peak_load_N"])
    assert max(abs(r[2]) for r in rows) == run["peak_midspan_deflection_m"]
    assert max(abs(r[1]) for r in rows) == pytest.approx(run["peak_support_reaction_N"],
                                                         rel=0, abs=1e-13)


def test_bottom_platen_damper_probe():
    g = GOLD["bottom_platen_damper"]
    sim = OracleDEM3D(DEM3DConfig(**g["config"]))
    sim.tool_start[:] = 1e6
    sim.tool_velocity[:] = 0.0
    bottom = sim.positions[:, 2] == sim.positions[:, 2].min()
    sim.positions[bottom, 2] += g["bottom_shift_m"]
    got = {}
    for label, vz in (("static", 0.0), ("approaching", -g["speed_m_s"]),
                      ("separating", g["speed_m_s"])):
        sim.velocities[:] = 0.0
        sim.velocities[bottom, 2] = vz
        force, _ = sim.forces(update_fracture=False)
        got[label] = force[:, 2].sum()
    want = g["result"]
    assert int(bottom.sum()) == want["bottom_particles"]
    for label in ("static", "approaching", "separating"):
        assert got[label] == pytest.approx(want[f"total_fz_{label}_N"], rel=1e-12)


@pytest.mark.parametrize("dims,geometry", [((3, 3, 2), "box"), ((9, 3, 3), "box"),
                                           ((4, 5, 6), "box"), ((7, 3, 5), "cylinder_xz"),
                                           ((5, 2, 5), "cylinder_xz"), ((2, 2, 2), "box")])
def test_lattice_matches_reference_loops(dims, geometry):
    nx, ny, nz = dims
    cfg = DEM3DConfig(nx=nx, ny=ny, nz=nz, geometry=geometry, fix_bottom=True)
    pos, fixed, pairs = reference_lattice(cfg)
    lat = build_lattice(cfg)
    assert np.array_equal(lat.positions, pos)
    assert np.array_equal(lat.fixed, fixed.astype(bool))
    assert np.array_equal(lat.pairs, pairs)
    n = len(pos)
    # neighbour table: every bond once above its lower sphere, once below its upper one
    for p in range(n):
        up = lat.nbr[p, :13]
        dn = lat.nbr[p, 13:]
        want_up = pairs[pairs[:, 0] == p, 1]
        want_dn = pairs[pairs[:, 1] == p, 0]
        assert np.array_equal(up[up >= 0], want_up)
        assert np.array_equal(dn[dn >= 0], want_dn)
        for s in range(26):
            if lat.nbr[p, s] >= 0:
                b = lat.bond_of_slot[p, s]
                assert sorted(pairs[b]) == sorted((p, lat.nbr[p, s]))


def test_box_bond_count_closed_form():
    # 13 half-space offsets: sum over offsets of prod(n_k - |d_k|)
    for nx, ny, nz in ((9, 3, 3), (4, 5, 6), (2, 2, 2)):
        want = 0
        for dz in (0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    if dz > 0 or (dz == 0 and dy > 0) or (dz == 0 and dy == 0 and dx > 0):
                        want += (nx - abs(dx)) * (ny - abs(dy)) * (nz - dz)
        assert len(build_lattice(DEM3DConfig(nx=nx, ny=ny, nz=nz)).pairs) == want
    assert want == 28   # 2x2x2: the complete graph on 8 spheres


def affine_shear(F):
    """sqrt(2/3 dev E : dev E) of E = (F^T F - I)/2."""
    E = 0.5 * (F.T @ F - np.eye(3))
    dev = E - np.trace(E) / 3 * np.eye(3)
    return math.sqrt(2.0 / 3.0 * (dev * dev).sum())


@pytest.mark.parametrize("F", [
    np.array([[1.0, 0.01, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]),
    np.array([[1.02, 0.003, -0.004], [0.001, 0.99, 0.006], [0.002, -0.003, 1.01]]),
    np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]]) @ np.diag([1.01, 1.0, 0.98]),
])
def test_shear_strain_of_affine_deformation(F):
    """Every sphere, boundary ones included (anisotropic sum r r^T), sees the
    exact F of an affine map; a transposed product or a dropped term fails."""
    sim = OracleDEM3D(DEM3DConfig(nx=4, ny=3, nz=3))
    sim.positions = sim.initial_positions @ F.T + np.array([0.3, -0.2, 0.1])
    got = sim.shear_strain()
    assert np.allclose(got, affine_shear(F), rtol=1e-10, atol=1e-12)
    # and the sample is not isotropic: cross @ inv and inv @ cross differ on the boundary
    assert not np.allclose(sim.inv[0].reshape(3, 3), sim.inv[0, 0] * np.eye(3))


def test_shear_strain_closed_form_simple_shear():
    g = 0.02
    F = np.eye(3)
    F[0, 1] = g
    assert affine_shear(F) == pytest.approx(math.sqrt(2 / 3 * (g * g / 2 + g ** 4 / 6)), rel=1e-12)


def small(**kw):
    base = dict(nx=3, ny=3, nz=2, drag=0.0, fix_x_edges=False)
    base.update(kw)
    return OracleDEM3D(DEM3DConfig(**base))


def test_initial_equilibrium_and_force_balance():
    sim = small()
    force, reaction = sim.forces()
    assert np.allclose(force.sum(0), -reaction, atol=1e-10, rtol=0)
    assert np.abs(force).max() < 1e-8


def test_rigid_rotation_does_not_break_bonds():
    sim = small()
    a, b = 0.41, -0.27
    rx = np.array([[1, 0, 0], [0, math.cos(a), -math.sin(a)], [0, math.sin(a), math.cos(a)]])
    ry = np.array([[math.cos(b), 0, math.sin(b)], [0, 1, 0], [-math.sin(b), 0, math.cos(b)]])
    sim.positions = sim.initial_positions @ (rx @ ry).T + np.array([2.0, -3.0, 4.0])
    force, _ = sim.forces()
    assert sim.broken_bonds == 0
    assert np.abs(force).max() < 1e-7
    assert sim.shear_strain().max() < 1e-7


def test_tension_failure_is_irreversible_and_recorded():
    sim = small(breaking_strain=0.01, shear_breaking_strain=0.5)
    sim.positions = sim.initial_positions * 1.03
    sim.time = 0.125
    sim.forces()
    broken = sim.broken_bonds
    assert broken == len(sim.pairs)     # uniform 3 % stretch breaks every bond
    assert (sim.failure_mode == 1).all() and (sim.failure_time_s == 0.125).all()
    rest = np.linalg.norm(sim.initial_positions[sim.pairs[:, 1]]
                          - sim.initial_positions[sim.pairs[:, 0]], axis=1)
    assert np.allclose(sim.failure_extension_m, 0.03 * rest, rtol=1e-12)
    k = sim.config.bond_stiffness
    assert np.allclose(sim.failure_energy_J, 0.5 * k * (0.03 * rest) ** 2, rtol=1e-12)
    sim.positions = sim.initial_positions.copy()
    sim.forces()
    assert sim.broken_bonds == broken


def test_shear_failure_mode():
    sim = small(breaking_strain=0.5, shear_breaking_strain=0.01)
    F = np.eye(3)
    F[0, 1] = 0.05
    sim.positions = sim.initial_positions @ F.T
    sim.forces()
    assert sim.broken_bonds == len(sim.pairs)
    assert (sim.failure_mode == 2).all()


def test_update_fracture_false_keeps_bonds():
    sim = small(breaking_strain=0.01)
    sim.positions = sim.initial_positions * 1.03
    sim.forces(update_fracture=False)
    assert sim.broken_bonds == 0


def test_bond_hookes_law():
    sim = OracleDEM3D(DEM3DConfig(nx=2, ny=2, nz=2, drag=0.0, fix_x_edges=False))
    assert tuple(sim.pairs[0]) == (0, 1)
    sim.alive[1:] = 0
    sim.tool_start[:] = 1e6
    sim.positions[2:] = np.arange(6)[:, None] * np.array([1.0, 0.0, 0.0]) + 10.0
    L0 = 2 * sim.config.radius
    d = np.array([L0 * 1.004, 3e-4, -2e-4])
    sim.positions[1] = sim.positions[0] + d
    force, _ = sim.forces(update_fracture=False)
    L = np.linalg.norm(d)
    want = sim.config.bond_stiffness * (L - L0) * d / L
    assert force[0] == pytest.approx(want, rel=1e-10)
    assert force[1] == pytest.approx(-want, rel=1e-10)


def two_free(cfg_kw, gap):
    """Spheres 0 and 1 of a 2x2x2 lattice with every bond broken, the rest far away."""
    sim = OracleDEM3D(DEM3DConfig(nx=2, ny=2, nz=2, drag=0.0, fix_x_edges=False, **cfg_kw))
    sim.alive[:] = 0
    sim.tool_start[:] = 1e6
    sim.tool_velocity[:] = 0.0
    r = sim.config.radius
    sim.positions[:] = np.arange(8)[:, None] * np.array([1.0, 0.0, 0.0]) + 10.0
    sim.positions[0] = (0.0, 0.0, 0.0)
    sim.positions[1] = (2 * r - gap, 0.0, 0.0)
    return sim


def test_linear_contact_spring_and_energy():
    sim = two_free({}, 1e-4)
    force, _ = sim.forces()
    k = sim.config.contact_stiffness
    ov = 2 * sim.config.radius - sim.positions[1, 0]
    assert force[0] == pytest.approx([-k * ov, 0, 0], rel=1e-12)
    assert force[1] == pytest.approx([k * ov, 0, 0], rel=1e-12)
    assert sim.pair_energy() == pytest.approx(0.5 * k * ov * ov, rel=1e-12)


def test_hertz_contact_force_and_potential():
    sim = two_free(dict(contact_model="hertz", young_modulus=2e8, poisson_ratio=0.25), 2e-4)
    force, _ = sim.forces()
    c = sim.config
    e_star = c.young_modulus / (2 * (1 - c.poisson_ratio ** 2))
    ov = 2 * c.radius - sim.positions[1, 0]
    f = 4 / 3 * e_star * math.sqrt(c.radius / 2) * ov ** 1.5
    assert force[1, 0] == pytest.approx(f, rel=1e-12)
    u = 8 / 15 * e_star * math.sqrt(c.radius / 2) * ov ** 2.5
    assert sim.pair_energy() == pytest.approx(u, rel=1e-12)


def clamped_restitution(e, k, m_eff):
    """Rebound ratio of the non-cohesive linear spring-dashpot: the overlap
    d(t) = (v0/w) exp(-zeta w0 t) sin(w t) until k d + gamma d' = 0, after
    which the spheres separate freely (root found on the closed form)."""
    from scipy.optimize import brentq
    zeta = -math.log(e) / math.sqrt(math.pi ** 2 + math.log(e) ** 2)
    w0 = math.sqrt(k / m_eff)
    w = w0 * math.sqrt(1 - zeta * zeta)
    gam = 2 * zeta * math.sqrt(k * m_eff)

    def d(t):
        return math.exp(-zeta * w0 * t) * math.sin(w * t) / w

    def dd(t):
        return math.exp(-zeta * w0 * t) * (math.cos(w * t) - zeta * w0 / w * math.sin(w * t))

    t_end = brentq(lambda t: k * d(t) + gam * dd(t), 0.5 * math.pi / w, math.pi / w)
    return -dd(t_end)


def test_linear_restitution_of_a_collision():
    """Head-on collision with the restitution-mapped dashpot: the rebound
    ratio is e for the ideal oscillator and a little more for the clamped
    law, which the semi-implicit Euler update reproduces at small dt."""
    e = 0.6
    sim = two_free(dict(restitution_coefficient=e, contact_stiffness=5000.0), -1e-5)
    sim.dt = sim.dt / 20
    v0 = 0.05
    sim.velocities[0] = (v0, 0, 0)
    sim.velocities[1] = (-v0, 0, 0)
    touched = False
    for _ in range(200000):
        sim.step()
        gap = sim.positions[1, 0] - sim.positions[0, 0] - 2 * sim.config.radius
        touched |= gap < 0
        if touched and gap > 1e-6:
            break
    assert touched
    rel = sim.velocities[1, 0] - sim.velocities[0, 0]
    want = clamped_restitution(e, 5000.0, sim.config.mass / 2)
    assert want == pytest.approx(0.632, abs=0.01)
    assert rel / (2 * v0) == pytest.approx(want, rel=2e-3)
    assert sim.velocities[:, 0].sum() == pytest.approx(0.0, abs=1e-12)   # momentum


def test_pinned_pair_uses_full_mass():
    e = 0.5
    sim = two_free(dict(restitution_coefficient=e), 1e-4)
    sim.velocities[1] = (-0.01, 0, 0)
    f_mobile, _ = sim.forces(update_fracture=False)
    sim.fixed[0] = 1
    f_pinned, _ = sim.forces(update_fracture=False)
    c = sim.config
    zeta = -math.log(e) / math.sqrt(math.pi ** 2 + math.log(e) ** 2)
    ov = 2 * c.radius - sim.positions[1, 0]
    vn = -0.01
    for f, m_eff in ((f_mobile, c.mass / 2), (f_pinned, c.mass)):
        g = 2 * zeta * math.sqrt(c.contact_stiffness * m_eff)
        assert f[1, 0] == pytest.approx(c.contact_stiffness * ov - g * vn, rel=1e-12)


def test_cundall_strack_stick_then_slip():
    mu, kt = 0.3, 4000.0
    sim = two_free(dict(friction_coefficient=mu, tangential_stiffness=kt), 1e-4)
    c = sim.config
    ov = 2 * c.radius - sim.positions[1, 0]
    fn = c.contact_stiffness * ov
    u = 1e-3
    sim.velocities[1] = (0.0, u, 0.0)
    dt = sim.dt
    # stick: F_t on sphere 0 = k_t * xi along +y, xi = u dt (positions held)
    f, _ = sim.forces(update_fracture=False)
    assert f[0, 1] == pytest.approx(kt * u * dt, rel=1e-12)
    assert f[1, 1] == pytest.approx(-kt * u * dt, rel=1e-12)
    for _ in range(3):
        sim.forces(update_fracture=False)
    assert sim.hist_xi[0, 0, 1] == pytest.approx(4 * u * dt, rel=1e-12)
    # slip: the spring saturates at mu F_n and the stored xi is F_t / k_t
    sim.velocities[1] = (0.0, 1e3, 0.0)
    f, _ = sim.forces(update_fracture=False)
    assert f[0, 1] == pytest.approx(mu * fn, rel=1e-12)
    assert sim.hist_xi[0, 0, 1] == pytest.approx(mu * fn / kt, rel=1e-12)
    # update_history=False reads the stored spring without advancing it
    before = sim.hist_xi.copy()
    sim.forces(update_fracture=False, update_history=False)
    assert np.array_equal(before, sim.hist_xi)
    # separation drops the history
    sim.positions[1, 0] += 1e-3
    sim.forces(update_fracture=False)
    assert sim.hist_cnt[0] == 0 and sim.hist_cnt[1] == 0


def test_tool_linear_spring_and_reaction():
    sim = small(contact_damping=0.0)
    c = sim.config
    top = np.argmax(sim.positions[:, 2] + np.abs(sim.positions[:, 0]) * -1e-3 - np.abs(sim.positions[:, 1]) * 1e-3)
    p = sim.positions[top]
    sim.tool_start[:] = p + np.array([0.0, 0.0, c.radius + c.tool_radius - 1e-4])
    force0, _ = small(contact_damping=0.0).forces()
    force, reaction = sim.forces(update_fracture=False)
    assert reaction == pytest.approx([0.0, 0.0, c.contact_stiffness * 1e-4], rel=1e-9, abs=1e-15)
    assert force[top, 2] - force0[top, 2] == pytest.approx(-c.contact_stiffness * 1e-4, rel=1e-9)


def test_platen_springs_and_reactions():
    sim = small(top_platen_enabled=True, bottom_platen_enabled=True, platen_stiffness=3000.0,
                platen_damping=0.0, lateral_platens=True, lateral_platen_stiffness=2000.0)
    sim.tool_start[:] = 1e6
    c = sim.config
    sim.positions = sim.initial_positions * 1.0001   # every face sphere pokes out
    _, _ = sim.forces(update_fracture=False)
    z = sim.positions[:, 2]
    top_ov = np.clip(z + c.radius - sim.top_platen_start_z, 0, None)
    bot_ov = np.clip(sim.bottom_platen_start_z + c.radius - z, 0, None)
    assert sim.out[3] == pytest.approx(3000.0 * top_ov.sum(), rel=1e-12)
    assert sim.out[4] == pytest.approx(-3000.0 * bot_ov.sum(), rel=1e-12)
    x = sim.positions[:, 0]
    xmin_ov = np.clip(sim.lateral_platen_pos[0] - x + c.radius, 0, None)
    assert sim.out[5] == pytest.approx(-2000.0 * xmin_ov.sum(), rel=1e-12)
    assert xmin_ov.sum() > 0


def test_free_fall_and_drag_closed_forms():
    g = 9.81
    sim = small(gravity=g)
    sim.alive[:] = 0
    sim.tool_start[:] = 1e6
    sim.positions = sim.initial_positions * 3.0   # no contacts
    z0 = sim.positions[:, 2].copy()
    n = 50
    for _ in range(n):
        sim.step()
    dt = sim.dt
    assert np.allclose(sim.positions[:, 2], z0 - g * dt * dt * n * (n + 1) / 2, rtol=0, atol=1e-13)
    sim2 = small(drag=0.7)
    sim2.alive[:] = 0
    sim2.tool_start[:] = 1e6
    sim2.positions = sim2.initial_positions * 3.0
    sim2.velocities[:] = 0.1
    for _ in range(n):
        sim2.step()
    m = sim2.config.mass
    assert np.allclose(sim2.velocities, 0.1 * (1 - 0.7 * sim2.dt / m) ** n, rtol=1e-12)


def test_broken_bond_is_still_eligible_for_contact():
    sim = small()
    i, j = sim.pairs[0]
    sim.positions[j] = sim.positions[i] + np.array([sim.config.radius, 0.0, 0.0])
    sim.alive[0] = 0
    sim.forces(update_fracture=False)
    assert sim.out[10] == 1     # the pair is a contact
    sim.alive[0] = 1
    sim.forces(update_fracture=False)
    assert sim.out[10] == 0     # with its bond alive it is not


def test_fixed_spheres_stay_put():
    sim = OracleDEM3D(DEM3DConfig(nx=4, ny=3, nz=2))
    fixed0 = sim.positions[sim.fixed.astype(bool)].copy()
    for _ in range(20):
        sim.step()
    assert np.array_equal(sim.positions[sim.fixed.astype(bool)], fixed0)
    assert sim.step_count == 20


def test_cell_list_equals_brute_force():
    """dem3d.py:825-858's guarantee (every overlapping pair found): the
    oracle's sorted cell list and an all-pairs scan give identical forces on a
    random cloud with many overlaps (tests/test_dem3d.py:83-98) and over an
    indentation that fractures the lattice."""
    cfg = DEM3DConfig(nx=3, ny=3, nz=2, drag=0.0, fix_x_edges=False, friction_coefficient=0.3)
    a = OracleDEM3D(cfg)
    b = OracleDEM3D(cfg, brute_force=True)
    x = np.random.default_rng(73).normal(size=a.positions.shape) * 0.035
    a.positions, b.positions = x.copy(), x.copy()
    fa, _ = a.forces()
    fb, _ = b.forces()
    assert a.out[10] == b.out[10] > 0
    assert np.array_equal(fa, fb)
    cfg = DEM3DConfig(nx=11, ny=5, nz=4, radius=0.01, tool_radius=0.03, tool_speed=1.0,
                      tool_gap=0.0002, breaking_strain=0.004, shear_breaking_strain=0.01,
                      contact_damping=1.0, friction_coefficient=0.3)
    a, b = OracleDEM3D(cfg), OracleDEM3D(cfg, brute_force=True)
    for _ in range(600):
        a.step()
        b.step()
    assert a.broken_bonds > 50 and a.out[10] > 0
    assert np.array_equal(a.positions, b.positions)
    assert np.array_equal(a.hist_xi, b.hist_xi)

"""The FP64 oracle against the reference's own golden runs and closed forms.

Setups restate the reference's generators: scripts/verify_dem2d_fracture_ledger.py
(cases a-d) and scripts/verify_dem2d_energy_ledger.py (cases A-C).
"""
import csv
import json
import math

import numpy as np
import pytest

from conftest import GOLDEN
from oracle.plate import OraclePlate, lattice, pairs_within
from tensordem_amd.config import DEMConfig

FRACTURE = json.loads((GOLDEN / "dem2d_fracture_ledger_closure.json").read_text())
ENERGY = json.loads((GOLDEN / "dem2d_energy_ledger_closure.json").read_text())


def indentation(dt_scale=1.0, *, shear=None):
    kw = dict(nx=7, ny=3, tool_speed=0.2)
    if shear is not None:
        kw["shear_breaking_strain"] = shear
    dt0 = DEMConfig(**kw).recommended_dt
    sim = OraclePlate(DEMConfig(**kw, dt=dt0 * dt_scale))
    e0 = sim.mechanical_energy()["mechanical_J"]
    steps = int(round(0.12 / sim.dt))
    events, seen = [], set()
    for _ in range(steps):
        pre = sim.positions.copy()
        sim.step()
        for b in sorted(set(np.nonzero(sim.alive == 0)[0].tolist()) - seen):
            seen.add(b)
            i, j = int(sim.bi[b]), int(sim.bj[b])
            d = float(np.sqrt(((pre[j] - pre[i]) ** 2).sum()))
            events.append((i, j, d, float(sim.failure_extension_m[b]), int(sim.failure_mode[b])))
    return sim, e0, steps, events


def golden_events():
    with open(GOLDEN / "dem2d_fracture_events.csv") as fh:
        return list(csv.DictReader(fh))


def test_indentation_break_events_match_reference():
    sim, _, steps, events = indentation()
    gold = golden_events()
    assert steps == FRACTURE["case_c"]["n_steps"] == 345
    assert sim.dt == FRACTURE["case_c"]["dt"]
    assert len(events) == len(gold) == 21
    for (i, j, d, ext, mode), g in zip(events, gold):
        assert (i, j, mode) == (int(g["pair_i"]), int(g["pair_j"]), int(g["mode"]))
        assert abs(d - float(g["distance_m"])) < 6e-10
        assert abs(ext - float(g["extension_m"])) < 6e-10
        assert abs(0.5 * 2000 * ext**2 - float(g["recorded_energy_J"])) < 1e-12


@pytest.mark.parametrize("dt_scale", [1.0, 0.5, 0.25])
def test_indentation_ledger_matches_reference(dt_scale):
    sim, e0, steps, _ = indentation(dt_scale)
    gold = next(c for c in ENERGY["cases"]
                if c["case"] == "C_tool_indentation" and c["dt_scale"] == dt_scale)
    acc = sim.accounting
    assert steps == gold["steps"]
    assert sim.broken_bonds == gold["broken_bonds"]
    for key in ("fracture_release_J", "pair_contact_acquired_J", "damping_dissipation_J",
                "tool_work_J"):
        assert acc[key] == pytest.approx(gold[key], rel=1e-11), key
    e1 = sim.mechanical_energy()["mechanical_J"]
    assert e1 == pytest.approx(gold["e_end_J"], rel=1e-11)
    residual = (e1 - e0 + acc["damping_dissipation_J"] + acc["fracture_release_J"]
                - acc["tool_work_J"] - acc["external_work_J"] - acc["pair_contact_acquired_J"])
    assert residual == pytest.approx(gold["residual_J"], rel=1e-7, abs=1e-15)


def test_planted_shear_threshold_error_is_caught():
    """Negative control: a 1 % wrong shear threshold changes the event list."""
    _, _, _, events = indentation(shear=0.03 * 1.01)
    gold = [(int(g["pair_i"]), int(g["pair_j"]), int(g["mode"])) for g in golden_events()]
    assert [(i, j, m) for i, j, _, _, m in events] != gold


def single_bond(k, xi0):
    cfg = DEMConfig(nx=2, ny=2, radius=0.025, bond_stiffness=k, contact_stiffness=2000.0,
                    breaking_strain=0.015, shear_breaking_strain=0.03, contact_damping=0.0,
                    drag=0.0, tool_speed=1e-6, tool_gap=1.0, fix_edges=False)
    sim = OraclePlate(cfg)
    sim.alive[:] = 0
    b01 = int(np.nonzero((sim.bi == 0) & (sim.bj == 1))[0][0])
    sim.alive[b01] = 1
    sim.positions[1, 0] += xi0
    return sim, b01


def test_controlled_break_releases_closed_form_energy():
    gold = FRACTURE["case_a"]
    sim, b01 = single_bond(gold["k"], gold["xi0"])
    u0 = sim.mechanical_energy()["mechanical_J"]
    sim.step()
    released = sim.accounting["fracture_release_J"]
    assert sim.failure_mode[b01] > 0
    assert released == pytest.approx(0.5 * gold["k"] * gold["xi0"] ** 2, rel=1e-12)
    assert released == pytest.approx(gold["ledger_released_J"], rel=1e-14)
    assert sim.failure_extension_m[b01] == pytest.approx(gold["recorded_extension_m"], rel=1e-14)
    assert sim.accounting["pair_contact_acquired_J"] == 0.0
    u1 = sim.mechanical_energy()["mechanical_J"]
    assert u0 == pytest.approx(gold["u_before_J"], rel=1e-14)
    assert (u1 - u0) + released == pytest.approx(0.0, abs=1e-18)


@pytest.mark.parametrize("row", FRACTURE["case_b"], ids=lambda r: f"k{r['k']}-xi{r['xi0']}")
def test_break_scaling_grid(row):
    sim, b01 = single_bond(row["k"], row["xi0"])
    sim.step()
    assert bool(sim.failure_mode[b01] > 0) == row["controlled_bond_broken"]
    assert sim.accounting["fracture_release_J"] == pytest.approx(row["ledger_released_J"],
                                                                 rel=1e-12, abs=1e-18)


def test_subthreshold_bond_survives():
    sim, b01 = single_bond(2000.0, 0.99 * 0.05 * 0.015)
    for _ in range(50):
        sim.step()
    assert sim.alive[b01] == 1
    assert sim.accounting["fracture_release_J"] == 0.0


def park(sim, spacing=1.0):
    n = len(sim.positions)
    sim.positions[:] = np.stack((np.arange(n) * spacing, np.zeros(n)), axis=1)
    sim.velocities[:] = 0
    sim.fixed[:] = 0
    sim.alive[:] = 0


@pytest.mark.parametrize("dt_scale", [1.0, 0.5, 0.25])
def test_drag_decay_matches_reference(dt_scale):
    base = dict(nx=2, ny=2, drag=2.0, contact_damping=0.0, tool_speed=1e-9, fix_edges=False)
    dt0 = DEMConfig(**base).recommended_dt
    sim = OraclePlate(DEMConfig(**base, dt=dt0 * dt_scale))
    park(sim)
    sim.velocities[0, 0] = 0.5
    gold = next(c for c in ENERGY["cases"]
                if c["case"] == "A_drag_decay" and c["dt_scale"] == dt_scale)
    steps = int(round(1.5 * (sim.config.mass / 2.0) / sim.dt))
    assert steps == gold["steps"]
    for _ in range(steps):
        sim.step()
    assert sim.velocities[0, 0] == pytest.approx(gold["v_sim"], rel=1e-13)
    assert sim.accounting["damping_dissipation_J"] == pytest.approx(
        gold["damping_dissipation_J"], rel=1e-12)
    # closed form v0 exp(-t/tau) to the discretisation error the reference reports
    tau = sim.config.mass / 2.0
    v_ref = 0.5 * math.exp(-sim.time / tau)
    assert abs(sim.velocities[0, 0] - v_ref) / v_ref == pytest.approx(gold["v_relerr"], rel=1e-9)


def test_bonded_oscillator_matches_reference():
    base = dict(nx=2, ny=2, bond_stiffness=2000.0, contact_stiffness=2000.0, drag=0.0,
                contact_damping=0.0, tool_speed=1e-9, fix_edges=False,
                breaking_strain=10.0, shear_breaking_strain=10.0)
    sim = OraclePlate(DEMConfig(**base))
    rest = 0.05
    amp = 0.005 * rest
    b01 = int(np.nonzero((sim.bi == 0) & (sim.bj == 1))[0][0])
    sim.alive[:] = 0
    sim.alive[b01] = 1
    sim.fixed[:] = 0
    sim.positions[0] = [-rest / 2 - amp / 2, 0.0]
    sim.positions[1] = [rest / 2 + amp / 2, 0.0]
    m = sim.config.mass
    omega = math.sqrt(2 * 2000.0 / m)
    e0 = sim.mechanical_energy()["mechanical_J"]
    steps = int(round(3 * 2 * math.pi / omega / sim.dt))
    gold = next(c for c in ENERGY["cases"]
                if c["case"] == "B_bonded_oscillator" and c["dt_scale"] == 1.0)
    assert steps == gold["steps"]
    xi_err = drift = 0.0
    for _ in range(steps):
        sim.step()
        ext = float(np.linalg.norm(sim.positions[1] - sim.positions[0])) - rest
        xi_err = max(xi_err, abs(ext - amp * math.cos(omega * sim.time)) / amp)
        drift = max(drift, abs(sim.mechanical_energy()["mechanical_J"] - e0) / e0)
    assert xi_err == pytest.approx(gold["xi_relerr_max"], rel=1e-8)
    assert drift == pytest.approx(gold["energy_drift_max"], rel=1e-8)


def isolated(**kw):
    base = dict(nx=6, ny=3, fix_edges=False, tool_gap=10.0, contact_damping=0.0, drag=0.0)
    base.update(kw)
    return OraclePlate(DEMConfig(**base))


def test_contract_impulse_momentum():
    """binding_contract.py:1571-1585: m*dv == n*dt*F to 1e-12."""
    sim = isolated()
    load = np.zeros_like(sim.positions)
    load[:, 0], load[:, 1] = 0.25, -0.1
    for _ in range(16):
        sim.step(load)
    dp = sim.config.mass * sim.velocities.sum(0)
    ref = 16 * sim.dt * load.sum(0)
    assert np.abs(dp - ref).max() / np.abs(ref).max() < 1e-12
    acc = sim.accounting
    assert np.allclose(acc["external_impulse_Ns"], ref, rtol=1e-12)


def test_contract_free_flight_and_clock():
    """Semi-implicit Euler: x_n - x_0 = a dt^2 n(n+1)/2; t = n dt."""
    sim = isolated()
    park(sim, 10.0)
    a = np.array([0.25, -0.1]) / sim.config.mass
    load = np.tile([0.25, -0.1], (len(sim.positions), 1))
    x0 = sim.positions.copy()
    n = 37
    for _ in range(n):
        sim.step(load)
    ref = a * sim.dt ** 2 * n * (n + 1) / 2
    got = sim.positions[0] - x0[0]  # disk 0 sits at the origin: no cancellation
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-12
    assert abs(sim.time - n * sim.dt) <= 1e-12 * n * sim.dt


def test_equilibrium_and_rigid_motion_do_not_break():
    sim = OraclePlate(DEMConfig(nx=3, ny=2, drag=0, fix_edges=False))
    f, r = sim.forces()
    assert np.abs(f).max() < 1e-10 and np.abs(r).max() == 0
    ang = 0.6
    rot = np.array([[math.cos(ang), -math.sin(ang)], [math.sin(ang), math.cos(ang)]])
    sim.positions[:] = sim.initial_positions @ rot.T + np.array([4.0, -4.0])
    f, _ = sim.forces()
    assert sim.broken_bonds == 0 and np.abs(f).max() < 1e-10


def test_tension_breaks_compression_does_not():
    sim = OraclePlate(DEMConfig(nx=3, ny=2, drag=0, fix_edges=False))
    sim.positions[:] = sim.initial_positions * 0.9
    sim.forces()
    assert sim.broken_bonds == 0
    sim.positions[:] = sim.initial_positions * 1.02
    f, _ = sim.forces()
    assert sim.broken_bonds == len(sim.bi) and np.abs(f).max() < 1e-12
    sim.positions[:] = sim.initial_positions * 0.9
    f, _ = sim.forces()
    assert np.abs(f).sum() > 0 and np.abs(f.sum(0)).max() < 1e-12


def test_shear_breaks_but_rotation_does_not():
    cfg = DEMConfig(nx=3, ny=2, drag=0, fix_edges=False, shear_breaking_strain=0.02,
                    breaking_strain=0.5)
    sheared = OraclePlate(cfg)
    sheared.positions[:, 0] += 0.08 * sheared.positions[:, 1]
    sheared.forces()
    assert sheared.broken_bonds > 0
    assert set(sheared.failure_mode[sheared.failure_mode > 0].tolist()) == {2}


def test_contact_repels_and_coincident_disks_stay_finite():
    sim = OraclePlate(DEMConfig(nx=3, ny=2, drag=0, fix_edges=False))
    sim.alive[:] = 0
    sim.positions[1, 0] -= 0.005
    f, _ = sim.forces()
    assert f[0, 0] < 0 < f[1, 0]
    sim.positions[1] = sim.positions[0]
    f, _ = sim.forces()
    assert np.isfinite(f).all() and np.abs(f).sum() > 0


def test_cell_search_equals_brute_force():
    """Unbonded contacts found by the oracle's grid equal the O(n^2) pair set."""
    rng = np.random.default_rng(3)
    sim = OraclePlate(DEMConfig(nx=8, ny=6, drag=0, fix_edges=False))
    sim.positions[:] += rng.normal(scale=0.02, size=sim.positions.shape)
    sim.alive[:] = 0
    f, _ = sim.forces()
    # rebuild the expected force from the brute-force contact set
    pairs = pairs_within(sim.positions, 0.05)
    exp = np.zeros_like(f)
    for i, j in pairs:
        d = sim.positions[j] - sim.positions[i]
        dist = np.sqrt(d[0] * d[0] + d[1] * d[1])
        mag = max(2000.0 * (0.05 - dist), 0.0)
        exp[i] -= mag * d / dist
        exp[j] += mag * d / dist
    assert len(pairs) > 10
    assert np.abs(f - exp).max() < 1e-9


def test_lattice_bonds_equal_distance_definition():
    """Bonds = every pair within 2r*sqrt(2)*(1+1e-10) (dem.py:119-124), triu order."""
    from tensordem_amd.lattice import bonds
    for nx, ny in ((2, 2), (3, 2), (7, 3), (5, 6)):
        pos, _, bi, bj, rest = lattice(nx, ny, 0.025)
        cut = 2 * 0.025 * math.sqrt(2) * (1 + 1e-10)
        ref = [(i, j) for i in range(nx * ny) for j in range(i + 1, nx * ny)
               if math.sqrt(((pos[j] - pos[i]) ** 2).sum()) <= cut]
        assert list(zip(bi.tolist(), bj.tolist())) == ref
        pi, pj, prest, fwd = bonds(nx, ny, 0.025)
        assert list(zip(pi.tolist(), pj.tolist())) == ref
        assert np.array_equal(prest, rest)
        assert fwd[0] == 0 and all(pi[fwd[p]] == p for p in range(nx * ny - 1) if fwd[p] < len(pi)
                                   and p in set(pi.tolist()))

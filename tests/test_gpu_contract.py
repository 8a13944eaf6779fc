"""The behavioural binding contract of `IceDEM` on the HIP path.

Restates, for `tensordem_amd.IceDEM` (FP64 forces), the 2-D probes of the
reference's contract (src/tensordem/binding_contract.py:605-720) and the 2-D
rows of its physics benchmark (binding_contract.py:1467-1651), each against
its analytic or bitwise reference, with the reference's tolerances:

* guards: wrong load shape -> ValueError, non-tensor loads -> TypeError,
  non-finite loads -> ValueError, stale load stamp -> ValueError;
* impulse-momentum identity m*sum(dv) = n*dt*sum(F) to 1e-12 for n = 1, 4,
  16, 64 on an isolated lattice (loads held over a step, not accumulated);
* ledger cross-path: accounting external impulse == n*dt*sum(F) to 1e-12;
* clock: time == n*dt to 1e-12;
* restart round trip bitwise; forces() accepts the step's load;
* free flight: displacement == a dt^2 n(n+1)/2 (the semi-implicit Euler
  closed form) to 1e-12, and first-order convergence to 0.5 a t^2 over dt,
  dt/2, dt/4 (each within the 15 % acceptance of binding_contract.py:1455);
* energy drift of an isolated conservative lattice over 2 000 steps within
  the same 15 % acceptance.
And the frozen state surface (binding_contract.py:93-110) on an instance.
"""
import math

import pytest
import torch

from tensordem_amd import DEMConfig
from tensordem_amd.icedem import IceDEM

pytestmark = pytest.mark.gpu
ACCEPT = 0.15


def isolated(**kw) -> IceDEM:
    args = dict(nx=6, ny=3, fix_edges=False, tool_gap=10.0, contact_damping=0.0, drag=0.0,
                device="cuda")
    args.update(kw)
    return IceDEM(DEMConfig(**args))


def test_state_surface():
    m = isolated()
    n = m.config.nx * m.config.ny
    assert isinstance(m.dt, float) and isinstance(m.time, float)
    assert isinstance(m.step_count, int)
    for name in ("positions", "velocities", "initial_positions"):
        t = getattr(m, name)
        assert isinstance(t, torch.Tensor) and tuple(t.shape) == (n, 2), name
        assert t.dtype == torch.float64
    assert tuple(m.fixed.shape) == (n,) and m.fixed.dtype == torch.bool
    assert isinstance(m.accounting, dict) and isinstance(m.last_step, dict)
    assert tuple(m.tool_position.shape) == (2,)
    assert isinstance(m.broken_bonds, int)
    assert m.snapshot()["schema"] == "tensordem.restart/1"
    assert m.exchange_state()["schema"] == "tensordem.exchange/1"


def test_guards():
    m = isolated()
    with pytest.raises(ValueError):
        m.step(external_forces=torch.zeros((len(m.positions) + 1, 2), dtype=torch.float64,
                                           device="cuda"))
    with pytest.raises(TypeError):
        m.step(external_forces=[[0.0, 0.0]] * len(m.positions))
    bad = torch.zeros_like(m.positions)
    bad[0, 0] = float("nan")
    with pytest.raises(ValueError):
        m.step(external_forces=bad)
    m = isolated()
    with pytest.raises(ValueError):
        m.step(external_forces=torch.zeros_like(m.positions), load_time=m.time + 1e-6)


@pytest.mark.parametrize("n_steps", [1, 4, 16, 64])
def test_impulse_momentum_identity_and_no_accumulation(n_steps):
    m = isolated()
    load = torch.zeros_like(m.positions)
    load[:, 0] = 0.25
    load[:, 1] = -0.1
    v0 = m.velocities.clone()
    for _ in range(n_steps):
        m.step(external_forces=load, load_time=m.time)
    dp = m.config.mass * (m.velocities - v0).sum(dim=0)
    ref = n_steps * m.dt * load.sum(dim=0)
    err = float((dp - ref).abs().max() / ref.abs().max())
    assert err < 1e-12
    assert abs(float((dp / ref).mean()) - 1.0) < 1e-12


def test_ledger_cross_path_and_clock():
    m = isolated()
    load = torch.zeros_like(m.positions)
    load[:, 1] = 0.3
    for _ in range(8):
        m.step(external_forces=load, load_time=m.time)
    ref = (8 * m.dt * load.sum(dim=0)).tolist()
    got = m.accounting["external_impulse_Ns"]
    assert max(abs(a - b) / max(abs(b), 1e-300) for a, b in zip(got, ref)) < 1e-12
    m = isolated()
    for _ in range(37):
        m.step()
    assert abs(m.time - 37 * m.dt) / (37 * m.dt) < 1e-12 and m.step_count == 37


def test_restart_round_trip_and_forces_load_path():
    m = isolated()
    load = torch.zeros_like(m.positions)
    load[:, 1] = 0.3
    for _ in range(3):
        m.step(external_forces=load, load_time=m.time)
    clone = IceDEM.from_snapshot(m.snapshot())
    assert clone.positions.equal(m.positions) and clone.velocities.equal(m.velocities)
    assert clone.time == m.time and clone.step_count == m.step_count
    a = isolated()
    fa, _ = a.forces(external_forces=load)
    assert tuple(fa.shape) == tuple(a.positions.shape)


def _free_flight(dt_factor, total_time=0.01, load=(0.25, -0.1)):
    base = DEMConfig(nx=6, ny=3, fix_edges=False, tool_gap=10.0, contact_damping=0.0,
                     drag=0.0, device="cuda")
    dt = base.recommended_dt * dt_factor
    m = isolated(dt=dt)
    n = max(1, int(round(total_time / dt)))
    f = torch.zeros_like(m.positions)
    f[:, 0], f[:, 1] = load
    x0 = m.positions.clone()
    for _ in range(n):
        m.step(external_forces=f, load_time=m.time)
    dx = (m.positions - x0).mean(dim=0)
    a = (load[0] / m.config.mass, load[1] / m.config.mass)
    sym = [a[i] * dt * dt * n * (n + 1) / 2.0 for i in (0, 1)]
    cont = [0.5 * a[i] * (n * dt) ** 2 for i in (0, 1)]
    got = [float(dx[0]), float(dx[1])]
    return (max(abs(got[i] - sym[i]) / abs(sym[i]) for i in (0, 1)),
            max(abs(got[i] - cont[i]) / abs(cont[i]) for i in (0, 1)))


def test_free_flight_closed_form_and_first_order_convergence():
    e1, c1 = _free_flight(1.0)
    e2, c2 = _free_flight(0.5)
    e4, c4 = _free_flight(0.25)
    assert max(e1, e2, e4) < 1e-12
    assert max(c1, c2, c4) <= ACCEPT
    for coarse, fine in ((c1, c2), (c2, c4)):
        assert abs(math.log(coarse / fine) / math.log(2.0) - 1.0) < 0.1


def test_energy_drift_of_conservative_lattice():
    m = isolated()
    n = len(m.positions)
    m.velocities[:, 0] = 0.05 * torch.sin(torch.arange(n, dtype=torch.float64, device="cuda")
                                          * (2 * math.pi / m.config.nx))
    e0 = float(m.mechanical_energy()["mechanical_J"])
    worst = 0.0
    for _ in range(2000):
        m.step()
        worst = max(worst, abs(float(m.mechanical_energy()["mechanical_J"]) - e0))
    assert worst / e0 <= ACCEPT


def test_restart_layouts_with_broken_bonds():
    """The 7x3 indentation run to step 345 (21 bonds broken): the reference
    layout (all N(N-1)/2 pairs) and the bond layout restore the same state,
    and both continue bit for bit; a reference-default DEMConfig (device
    'cpu') runs on the ROCm device."""
    cfg = DEMConfig(nx=7, ny=3, tool_speed=0.2)
    assert cfg.device == "cpu"
    m = IceDEM(cfg)
    assert m.device.type == "cuda"
    for _ in range(345):
        m.step()
    assert m.broken_bonds == 21
    n = 21
    ref = m.snapshot()
    assert ref["schema"] == "tensordem.restart/1"
    assert len(ref["alive"]) == len(ref["failure_time_s"]) == n * (n - 1) // 2
    assert sum(not a for a in ref["alive"]) == n * (n - 1) // 2 - m.alive.numel() + m.broken_bonds
    bonds = m.snapshot(layout="bonds")
    assert bonds["schema"] == "tensordem_amd.restart_bonds/1"
    assert len(bonds["alive"]) == m.alive.numel()
    clones = [IceDEM.from_snapshot(ref), IceDEM.from_snapshot(bonds)]
    for c in clones:
        assert torch.equal(c.alive, m.alive) and torch.equal(c.failure_mode, m.failure_mode)
        assert torch.equal(c.failure_time_s, m.failure_time_s)
    for _ in range(100):
        m.step()
        for c in clones:
            c.step()
    for c in clones:
        assert torch.equal(c.positions, m.positions) and torch.equal(c.alive, m.alive)

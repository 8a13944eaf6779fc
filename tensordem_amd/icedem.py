"""Drop-in `IceDEM` whose step runs in the HIP kernels.

Same constructor, attributes, methods, exceptions and restart schema as the
reference class (src/tensordem/dem.py:75-624).  State tensors live on the ROCm
device; `step()` and `forces()` are single kernel launches
(tensordem_amd/csrc/plate.hip).  Differences a caller can observe:

* bond-indexed views: `pairs`, `alive`, `bonded`, `rest_lengths` and the
  `failure_*` arrays are indexed by bond (the reference indexes them over all
  N(N-1)/2 pairs, dem.py:116-136; only bonded entries can ever be nonzero
  there).  `pairs[k]` is the k-th bond in the reference's pair order;
  `fracture_events()["pair_index"]` is the bond index.
* `positions` / `velocities` are double-buffered: each step returns a
  different tensor object (assigning to them, or editing them in place, works
  as in the reference).
* `precision="fp32"` evaluates bond, contact and tool forces in FP32 (state,
  sums and ledger stay FP64; decisions within `fp32_margin` of a break
  threshold are re-checked in FP64).
* `dt_policy="degree"` bounds dt by 8 partners per disk instead of all N-1
  (the reference bound shrinks like 1/N, dem.py:61-72).
* `DEMConfig(device="cpu")`, the reference's default, runs on the ROCm
  device as well (`self.device` is the ROCm device; `config` is kept as
  given, so restarts carry the caller's setting).
* `snapshot()` writes the reference's all-pairs `tensordem.restart/1` up to
  `restart.MAX_REFERENCE_PAIRS` pairs and a bond-indexed layout under
  `tensordem_amd.restart_bonds/1` above it; `from_snapshot` reads both.
"""
from __future__ import annotations

import copy
import json
import math
from dataclasses import asdict

import numpy as np
import torch

from . import _lib, restart
from .config import DEMConfig
from .engine import PlateEngine, PlateSpec

ADDITIVE_ACCOUNTING_KEYS = frozenset({"pair_contact_acquired_J"})
_ACC_SCALARS = (("external_work_J", _lib.ACC_EXT_WORK), ("fracture_release_J", _lib.ACC_RELEASE),
                ("pair_contact_acquired_J", _lib.ACC_ACQUIRED),
                ("tool_work_J", _lib.ACC_TOOL_WORK),
                ("damping_dissipation_J", _lib.ACC_DAMPING))
_ACC_VECTORS = (("external_impulse_Ns", _lib.ACC_EXT_IMP),
                ("support_impulse_Ns", _lib.ACC_SUP_IMP),
                ("drag_impulse_Ns", _lib.ACC_DRAG_IMP),
                ("tool_impulse_on_ice_Ns", _lib.ACC_TOOL_IMP))


class IceDEM:
    """A small ice lattice; only tensile axial strain breaks bonds.

    Positions, velocities and externally applied forces have shape (N, 2).
    Tensile failure uses axial bond strain; shear failure uses an objective
    local Green-Lagrange strain invariant (dem.py:75-84).
    """

    def __init__(self, config: DEMConfig = DEMConfig(), *,
                 hull_profile: torch.Tensor | None = None,
                 hull_start: tuple[float, float] = (0.0, 0.0),
                 hull_velocity: tuple[float, float] = (0.0, 0.0),
                 precision: str = "fp64", dt_policy: str = "all_pairs",
                 fp32_margin: float = 1e-3) -> None:
        self.config = config
        self.device = torch.device(config.device)
        if self.device.type not in ("cpu", "cuda"):
            raise ValueError("Only CPU and CUDA devices are supported")
        if not torch.cuda.is_available():
            raise ValueError("tensordem_amd.IceDEM runs on the ROCm device, which is unavailable")
        if self.device.type == "cpu":
            # the reference's default DEMConfig(device="cpu") runs here too:
            # the solver only exists on the ROCm device
            self.device = torch.device("cuda")
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32'")
        if dt_policy not in ("all_pairs", "degree"):
            raise ValueError("dt_policy must be 'all_pairs' or 'degree'")
        self.precision, self.dt_policy = precision, dt_policy
        bound = config.recommended_dt if dt_policy == "all_pairs" else config.degree_bound_dt()
        self.dt = config.dt if config.dt is not None else bound
        if self.dt > bound:
            raise ValueError(f"dt exceeds conservative limit {bound:.6g} s")
        f64 = dict(dtype=torch.float64, device=self.device)
        self.hull_profile = None
        self.hull_start = torch.tensor(hull_start, **f64)
        self.hull_velocity = torch.tensor(hull_velocity, **f64)
        if not bool(torch.isfinite(self.hull_start).all() and torch.isfinite(self.hull_velocity).all()):
            raise ValueError("hull_start and hull_velocity must be finite")
        if hull_profile is not None:
            if not isinstance(hull_profile, torch.Tensor):
                raise TypeError("hull_profile must be a torch.Tensor with shape (M, 2)")
            if hull_profile.ndim != 2 or hull_profile.shape[1] != 2 or hull_profile.shape[0] < 2:
                raise ValueError("hull_profile must have shape (M, 2), M >= 2")
            self.hull_profile = hull_profile.to(**f64).clone()
            if not bool(torch.isfinite(self.hull_profile).all()):
                raise ValueError("hull_profile must contain only finite values")
            seg = torch.linalg.vector_norm(self.hull_profile[1:] - self.hull_profile[:-1], dim=1)
            if bool((seg <= torch.finfo(torch.float64).eps).any()):
                raise ValueError("hull_profile must not contain zero-length segments")
        spec = PlateSpec(config=config, dt=self.dt,
                         hull_profile=None if self.hull_profile is None
                         else self.hull_profile.cpu().numpy(),
                         hull_start=tuple(float(v) for v in hull_start),
                         hull_velocity=tuple(float(v) for v in hull_velocity))
        self._eng = PlateEngine([spec], device=self.device, fp32=precision == "fp32",
                                fp32_margin=fp32_margin)
        e = self._eng
        self.initial_positions = e.initial_positions
        self.pairs = torch.stack((e.bond_i, e.bond_j), dim=1)
        self.bond_indices = self.pairs
        self.bonded = torch.ones(e.nbonds, dtype=torch.bool, device=self.device)
        self.rest_lengths = e.rest_lengths
        d = e.initial_positions[e.bond_j] - e.initial_positions[e.bond_i]
        self.reference_bond_vectors = d
        self.initial_normals = d / e.rest_lengths[:, None]
        self.failure_mode = e.failure_mode
        self.failure_time_s = e.failure_time_s
        self.failure_extension_m = e.failure_extension_m
        self.failure_energy_J = e.failure_energy_J
        self.tool_start = torch.tensor(
            [0.0, (config.ny - 1) * 2 * config.radius + config.radius + config.tool_radius
             + config.tool_gap], **f64)
        self.tool_velocity = torch.tensor([0.0, -config.tool_speed], **f64)
        self._time = 0.0
        self.step_count = 0
        self._last_override = None
        self._note_versions()

    # ---- state as tensors on the device -------------------------------------
    def _note_versions(self) -> None:
        e = self._eng
        self._seen = (id(e.x), e.x._version, id(e.alive), e.alive._version)

    def _sync_inputs(self) -> None:
        """Push caller edits of positions/alive into the kernels' derived state."""
        e = self._eng
        x_id, x_ver, a_id, a_ver = self._seen
        if id(e.alive) != a_id or e.alive._version != a_ver:
            e.sync_links()
        if id(e.x) != x_id or e.x._version != x_ver:
            e.rebin()
        self._note_versions()

    @property
    def positions(self) -> torch.Tensor:
        return self._eng.x

    @positions.setter
    def positions(self, value: torch.Tensor) -> None:
        self._eng.x.copy_(torch.as_tensor(value, dtype=torch.float64).reshape(self._eng.x.shape))

    @property
    def velocities(self) -> torch.Tensor:
        return self._eng.v

    @velocities.setter
    def velocities(self, value: torch.Tensor) -> None:
        self._eng.v.copy_(torch.as_tensor(value, dtype=torch.float64).reshape(self._eng.v.shape))

    @property
    def fixed(self) -> torch.Tensor:
        return self._eng.fixed

    @fixed.setter
    def fixed(self, value: torch.Tensor) -> None:
        self._eng.fixed.copy_(torch.as_tensor(value).to(torch.bool).reshape(self._eng.fixed.shape))

    @property
    def alive(self) -> torch.Tensor:
        return self._eng.alive

    @alive.setter
    def alive(self, value: torch.Tensor) -> None:
        value = torch.as_tensor(value)
        if tuple(value.shape) != tuple(self._eng.alive.shape):
            raise ValueError(f"alive must have shape {tuple(self._eng.alive.shape)} (one per bond)")
        self._eng.alive.copy_(value.to(torch.bool))

    @property
    def time(self) -> float:
        return self._time

    @time.setter
    def time(self, value: float) -> None:
        self._time = float(value)
        self._eng.set_double(0, "time", self._time)

    @property
    def tool_position(self) -> torch.Tensor:
        """Return the circular tool position or prescribed hull translation."""
        if self.hull_profile is not None:
            return self.hull_start + self.time * self.hull_velocity
        return self.tool_start + self.time * self.tool_velocity

    @property
    def broken_bonds(self) -> int:
        return int((~self._eng.alive).sum().item())

    # ---- ledger ----------------------------------------------------------------
    @property
    def accounting(self) -> dict:
        a = self._eng.accounting[0].tolist()
        out = {k: a[i] for k, i in _ACC_SCALARS}
        for k, i in _ACC_VECTORS:
            out[k] = [a[i], a[i + 1]]
        return _LedgerDict(self, out)

    @accounting.setter
    def accounting(self, value: dict) -> None:
        row = torch.zeros(_lib.NACC, dtype=torch.float64)
        for k, i in _ACC_SCALARS:
            row[i] = float(value.get(k, 0.0))
        for k, i in _ACC_VECTORS:
            row[i], row[i + 1] = (float(t) for t in value[k])
        self._eng.accounting[0].copy_(row)

    @property
    def last_step(self) -> dict:
        if self._last_override is not None:
            return self._last_override
        if self.step_count == 0:
            return {}
        l = self._eng.last[0].tolist()
        return dict(time_start_s=l[0], time_end_s=l[1], external_work_J=l[2],
                    support_force_N=[l[3], l[4]], tool_reaction_N=[l[5], l[6]],
                    external_force_N=[l[7], l[8]], external_moment_Nm=l[9])

    @last_step.setter
    def last_step(self, value: dict) -> None:
        self._last_override = copy.deepcopy(value)

    # ---- validation -------------------------------------------------------------
    def _validate_external_forces(self, external_forces):
        if external_forces is None:
            return None
        if not isinstance(external_forces, torch.Tensor):
            raise TypeError("external_forces must be a torch.Tensor with shape (N, 2)")
        if tuple(external_forces.shape) != tuple(self.positions.shape):
            raise ValueError(f"external_forces must have shape {tuple(self.positions.shape)}")
        loads = external_forces.to(device=self.device, dtype=torch.float64)
        if not bool(torch.isfinite(loads).all()):
            raise ValueError("external_forces must contain only finite values")
        return loads

    def _set_loads(self, loads) -> None:
        e = self._eng
        if loads is None:
            if e.has_loads(0):
                e.set_int(0, "has_loads", 0)
        else:
            e.loads.copy_(loads)
            if not e.has_loads(0):
                e.set_int(0, "has_loads", 1)

    def _raise_status(self) -> None:
        st = self._eng.read_status(0)
        if st & (_lib.ST_STALE_LIST | _lib.ST_CAND_OVERFLOW | _lib.ST_CELL_OVERFLOW):
            self._eng.clear_status(0)
            raise RuntimeError(
                "contact search invalid (status %d): a disk moved more than half the Verlet "
                "skin before the list was rebuilt, or a hash cell / candidate list overflowed; "
                "reduce dt" % st)
        if st & _lib.ST_NONFINITE:
            self._eng.clear_status(0)
            raise FloatingPointError("Non-finite simulation state; reduce dt")

    # ---- physics -----------------------------------------------------------------
    @torch.no_grad()
    def forces(self, external_forces: torch.Tensor | None = None):
        """Return total particle force and ice-on-tool reaction (dem.py:226-345)."""
        loads = self._validate_external_forces(external_forces)
        self._sync_inputs()
        self._set_loads(loads)
        self._eng.forces()
        self._raise_status()
        return self._eng.force.clone(), self._eng.reaction[0].clone()

    @torch.no_grad()
    def step(self, external_forces: torch.Tensor | None = None, *,
             load_time: float | None = None) -> None:
        """Semi-implicit Euler; optional nodal loads are supplied in newtons."""
        if load_time is not None and (not math.isfinite(load_time) or
                abs(load_time - self.time) > 1e-12 * max(1.0, abs(self.time))):
            raise ValueError("load_time must equal current DEM time in seconds")
        loads = self._validate_external_forces(external_forces)
        self._sync_inputs()
        self._set_loads(loads)
        self._eng.step(1)
        self._last_override = None
        self._note_versions()
        self._time += self.dt
        self.step_count += 1
        self._raise_status()

    def run(self, nsteps: int) -> None:
        """Advance nsteps without per-step host synchronisation (no loads)."""
        self._sync_inputs()
        self._set_loads(None)
        if nsteps <= 0:
            return
        self._eng.step(int(nsteps))
        self._last_override = None
        self._note_versions()
        for _ in range(int(nsteps)):
            self._time += self.dt
        self.step_count += int(nsteps)
        self._raise_status()

    @torch.no_grad()
    def diagnostics(self, external_forces: torch.Tensor | None = None) -> dict:
        vec = self._diagnostics_device(self._validate_external_forces(external_forces))
        self._raise_status()
        return self._diagnostics_row(vec.tolist(), self.time)

    # Device vector behind diagnostics(): tool position, reaction, boundary
    # reaction, broken-bond count, kinetic energy, member status, ledger row.
    _DIAG_LEN = 9 + _lib.NACC

    @torch.no_grad()
    def _diagnostics_device(self, loads=None) -> torch.Tensor:
        """Enqueue the force evaluation and the reductions of diagnostics() on
        the current stream and return their values as one FP64 device vector,
        without waiting for them (the member status rides along, unchecked)."""
        self._sync_inputs()
        self._set_loads(loads)
        e = self._eng
        e.forces()
        boundary = -e.force[self.fixed].sum(dim=0)
        kinetic = 0.5 * self.config.mass * self.velocities.square().sum()
        broken = (~e.alive).sum().to(torch.float64)
        off = _lib.Member.status.offset
        status = e.members[off:off + 4].view(torch.int32).to(torch.float64)
        return torch.cat([self.tool_position, e.reaction[0], boundary, broken[None],
                          kinetic[None], status, e.accounting[0]])

    @staticmethod
    def _diagnostics_row(v: list, time: float) -> dict:
        acc = v[9:]
        vec = dict(_ACC_VECTORS)
        sc = dict(_ACC_SCALARS)
        return {
            "time": time,
            "tool_x": v[0],
            "tool_y": v[1],
            "reaction_x": v[2],
            "reaction_y": v[3],
            "boundary_reaction_x": v[4],
            "boundary_reaction_y": v[5],
            "broken_bonds": int(v[6]),
            "kinetic_energy": v[7],
            "tool_impulse_on_ice_x_Ns": acc[vec["tool_impulse_on_ice_Ns"]],
            "tool_impulse_on_ice_y_Ns": acc[vec["tool_impulse_on_ice_Ns"] + 1],
            "support_impulse_x_Ns": acc[vec["support_impulse_Ns"]],
            "support_impulse_y_Ns": acc[vec["support_impulse_Ns"] + 1],
            "drag_impulse_x_Ns": acc[vec["drag_impulse_Ns"]],
            "drag_impulse_y_Ns": acc[vec["drag_impulse_Ns"] + 1],
            "external_impulse_x_Ns": acc[vec["external_impulse_Ns"]],
            "external_impulse_y_Ns": acc[vec["external_impulse_Ns"] + 1],
            "fracture_release_J": acc[sc["fracture_release_J"]],
            "pair_contact_acquired_J": acc[sc["pair_contact_acquired_J"]],
            "tool_work_J": acc[sc["tool_work_J"]],
            "damping_dissipation_J": acc[sc["damping_dissipation_J"]],
            "external_work_J": acc[sc["external_work_J"]],
        }

    def mechanical_energy(self) -> dict:
        """Actual current spring/contact potential; does not advance failure."""
        self._sync_inputs()
        self._eng.mechanical_energy()
        self._raise_status()
        k, b, c, t, total = self._eng.energy[0].tolist()
        return dict(kinetic_J=k, bond_J=b, pair_contact_J=c, tool_contact_J=t, mechanical_J=total)

    def exchange_state(self) -> dict:
        """SI global x/y state at t_n (dem.py:418-435, schema tensordem.exchange/1)."""
        return dict(schema="tensordem.exchange/1", time_s=self.time, dt_s=self.dt,
                    positions_m=self.positions.tolist(), velocities_m_s=self.velocities.tolist(),
                    masses_kg=[self.config.mass] * len(self.positions),
                    fixed=self.fixed.tolist(), particle_radius_m=self.config.radius,
                    thickness_m=self.config.thickness)

    def fracture_events(self) -> list[dict]:
        """Deterministic per-bond failure records (dem.py:437-460)."""
        labels = {1: "tensile", 2: "shear", 3: "mixed"}
        mode = self.failure_mode.cpu()
        idx = torch.nonzero(mode > 0, as_tuple=False).flatten().tolist()
        pairs = self.pairs.cpu()
        pos = self.positions.cpu()
        ft = self.failure_time_s.cpu()
        fe = self.failure_extension_m.cpu()
        fw = self.failure_energy_J.cpu()
        events = []
        for p in idx:
            i, j = pairs[p].tolist()
            events.append({
                "pair_index": p, "particle_i": i, "particle_j": j,
                "time_s": float(ft[p]), "mode": labels[int(mode[p])],
                "midpoint_m": ((pos[i] + pos[j]) * 0.5).tolist(),
                "extension_at_failure_m": float(fe[p]),
                "released_spring_energy_J": float(fw[p]),
            })
        return events

    def fragments(self) -> list[dict]:
        """Connected components of surviving bonds, including isolated disks
        (labelled on the device, tdm_plate_fragments)."""
        from .fragments import fragment_records, label_plate
        self._sync_inputs()
        return fragment_records(label_plate(self._eng, 0), self.config.mass)

    # ---- restart (dem.py:486-601) ------------------------------------------------
    def snapshot(self, layout: str | None = None) -> dict:
        """Complete JSON restart; derived lattice topology is rebuilt and checked.

        layout "reference" writes the reference's `tensordem.restart/1`
        (alive and failure records over all N(N-1)/2 pairs, dem.py:486-601),
        which the reference loads; "bonds" writes them per bond under
        `tensordem_amd.restart_bonds/1`.  The default is "reference" while
        the pair arrays stay within restart.MAX_REFERENCE_PAIRS, else "bonds".
        """
        n = self._eng.n
        if layout is None:
            layout = "reference" if restart.pair_count(n) <= restart.MAX_REFERENCE_PAIRS \
                else "bonds"
        if layout not in ("reference", "bonds"):
            raise ValueError("layout must be 'reference' or 'bonds'")
        per_bond = dict(alive=self.alive.cpu().numpy(),
                        failure_mode=self.failure_mode.cpu().numpy(),
                        failure_time_s=self.failure_time_s.cpu().numpy(),
                        failure_extension_m=self.failure_extension_m.cpu().numpy(),
                        failure_energy_J=self.failure_energy_J.cpu().numpy())
        if layout == "reference":
            e = self._eng
            fields = restart.bonds_to_pairs(per_bond, n, e.bond_i.cpu().numpy(),
                                            e.bond_j.cpu().numpy())
            schema = restart.REFERENCE_SCHEMA
        else:
            fields = {k: v.tolist() for k, v in per_bond.items()}
            schema = restart.BOND_SCHEMA
        cfg = asdict(self.config)
        return dict(schema=schema, config=cfg,
                    dt_s=self.dt, time_s=self.time, step_count=self.step_count,
                    positions_m=self.positions.tolist(), velocities_m_s=self.velocities.tolist(),
                    initial_positions_m=self.initial_positions.tolist(), fixed=self.fixed.tolist(),
                    alive=fields["alive"],
                    failure_mode=fields["failure_mode"],
                    failure_time_s=fields["failure_time_s"],
                    failure_extension_m=fields["failure_extension_m"],
                    failure_energy_J=fields["failure_energy_J"],
                    tool_start_m=self.tool_start.tolist(),
                    tool_velocity_m_s=self.tool_velocity.tolist(),
                    hull_profile_m=None if self.hull_profile is None else self.hull_profile.tolist(),
                    hull_start_m=self.hull_start.tolist(), hull_velocity_m_s=self.hull_velocity.tolist(),
                    accounting=copy.deepcopy(dict(self.accounting)),
                    last_step=copy.deepcopy(self.last_step))

    @classmethod
    def from_snapshot(cls, state: dict, *, device: str | None = None, **kwargs) -> "IceDEM":
        """Fail closed on missing fields, incompatible topology or nonfinite data.

        Reads both the reference's all-pairs `tensordem.restart/1` and the
        bond-indexed `tensordem_amd.restart_bonds/1` (see `snapshot`)."""
        if not isinstance(state, dict) or state.get("schema") not in (restart.REFERENCE_SCHEMA,
                                                                      restart.BOND_SCHEMA):
            raise ValueError("unsupported DEM restart schema")
        required = {"schema", "config", "dt_s", "time_s", "step_count", "positions_m",
                    "velocities_m_s", "initial_positions_m", "fixed", "alive", "failure_mode",
                    "failure_time_s", "failure_extension_m", "failure_energy_J", "tool_start_m",
                    "tool_velocity_m_s", "hull_profile_m", "hull_start_m", "hull_velocity_m_s",
                    "accounting", "last_step"}
        if set(state) != required:
            raise ValueError("restart fields missing or unknown")
        cfg = dict(state["config"])
        if device is not None:
            cfg["device"] = device
        config = DEMConfig(**cfg)
        hull = state["hull_profile_m"]
        result = cls(config, hull_profile=None if hull is None else torch.tensor(hull, dtype=torch.float64),
                     hull_start=state["hull_start_m"], hull_velocity=state["hull_velocity_m_s"],
                     **kwargs)
        if state["dt_s"] != result.dt or not math.isfinite(state["time_s"]) or state["time_s"] < 0:
            raise ValueError("invalid restart time or timestep")
        count = state["step_count"]
        if isinstance(count, bool) or not isinstance(count, int) or count < 0:
            raise ValueError("invalid restart step_count")
        if abs(state["time_s"] - count * result.dt) > 1e-10 * max(1.0, state["time_s"]):
            raise ValueError("restart time inconsistent with step count")
        dev = result.device
        if state["schema"] == restart.REFERENCE_SCHEMA:
            e = result._eng
            state = dict(state)
            state.update(restart.pairs_to_bonds(state, e.n, e.bond_i.cpu().numpy(),
                                                e.bond_j.cpu().numpy()))
        for key, attribute in (("positions_m", "positions"), ("velocities_m_s", "velocities"),
                               ("initial_positions_m", "initial_positions"),
                               ("tool_start_m", "tool_start"), ("tool_velocity_m_s", "tool_velocity")):
            original = getattr(result, attribute)
            tensor = torch.tensor(state[key], dtype=torch.float64, device=dev)
            if tensor.shape != original.shape or not bool(torch.isfinite(tensor).all()):
                raise ValueError(f"invalid restart {key}")
            if attribute in ("initial_positions", "tool_start", "tool_velocity"):
                if not torch.equal(tensor, original):
                    raise ValueError("reference lattice mismatch" if attribute == "initial_positions"
                                     else f"restart {key} does not match the configuration")
                continue
            setattr(result, attribute, tensor)
        for key in ("fixed", "alive"):
            values = state[key]
            if not isinstance(values, list) or any(type(v) is not bool for v in values):
                raise ValueError(f"invalid boolean restart {key}")
            tensor = torch.tensor(values, dtype=torch.bool, device=dev)
            if tensor.shape != getattr(result, key).shape:
                raise ValueError(f"invalid restart {key} shape")
            if key == "fixed" and not torch.equal(tensor, result.fixed):
                raise ValueError("fixed boundary mismatch")
            setattr(result, key, tensor)
        for key in ("failure_mode", "failure_time_s", "failure_extension_m", "failure_energy_J"):
            dtype = torch.uint8 if key == "failure_mode" else torch.float64
            tensor = torch.as_tensor(state[key], dtype=dtype, device=dev)
            target = getattr(result, key)
            if tensor.shape != target.shape:
                raise ValueError(f"invalid restart {key} shape")
            if key == "failure_mode":
                if bool((tensor > 3).any()):
                    raise ValueError("invalid restart failure modes")
                if not torch.equal(tensor > 0, ~result.alive):
                    raise ValueError("failure ledger inconsistent with bond state")
            else:
                if not bool(torch.isfinite(tensor).all()):
                    raise ValueError(f"nonfinite restart {key}")
                fm = result.failure_mode
                if key == "failure_time_s" and bool((tensor[fm > 0] < 0).any()):
                    raise ValueError("failed bond has invalid failure time")
                if key == "failure_time_s" and bool((tensor[fm == 0] != -1).any()):
                    raise ValueError("intact bond has failure time")
                if key == "failure_energy_J" and bool((tensor < 0).any()):
                    raise ValueError("negative restart failure energy")
                if key == "failure_energy_J" and bool((tensor[fm == 0] != 0).any()):
                    raise ValueError("intact bond has failure energy")
            target.copy_(tensor)
        if not torch.equal(result.positions[result.fixed], result.initial_positions[result.fixed]) \
                or bool((result.velocities[result.fixed] != 0).any()):
            raise ValueError("fixed particle state mismatch")
        accounting = state["accounting"]
        reference_keys = {k for k, _ in _ACC_SCALARS} | {k for k, _ in _ACC_VECTORS}
        missing = reference_keys - set(accounting)
        if missing - ADDITIVE_ACCOUNTING_KEYS or set(accounting) - reference_keys:
            raise ValueError("restart accounting fields mismatch")
        for key, value in accounting.items():
            tensor = torch.as_tensor(value, dtype=torch.float64)
            expected_shape = (2,) if "impulse" in key else ()
            if tuple(tensor.shape) != expected_shape or not bool(torch.isfinite(tensor).all()):
                raise ValueError("invalid restart accounting")
            if key == "fracture_release_J" and value < 0:
                raise ValueError("negative fracture release")
        json.dumps(state["last_step"], allow_nan=False)
        acc = copy.deepcopy(accounting)
        for key in missing:
            acc[key] = 0.0
        result.accounting = acc
        result.last_step = copy.deepcopy(state["last_step"])
        result.time = state["time_s"]
        result.step_count = count
        result._eng.set_int64(0, "step_count", count)
        return result


class _LedgerDict(dict):
    """`accounting` snapshot; item assignment writes through to the device ledger."""

    def __init__(self, owner: IceDEM, data: dict):
        super().__init__(data)
        self._owner = owner

    def __setitem__(self, key, value):
        super().__setitem__(key, value)
        self._owner.accounting = dict(self)

"""Drop-in `IceDEM3D` whose force evaluation and step run in the HIP kernels.

Same constructor, attributes, methods, exceptions and restart dictionary as
the reference class (src/tensordem/dem3d.py:564-1967) for the options the
device solver evaluates (tensordem_amd/csrc/dem3d.hip): linear or Hertz
normal law, restitution-mapped or absolute pair dashpot, Cundall-Strack
friction, the spherical tool, top/bottom and lateral platens, drag,
gravity, nodal loads, fixed spheres, tensile and shear fracture with
optional Weibull thresholds, box and cylinder_xz lattices.  Options outside
that set (rolling resistance, tanh/dual friction with distinct static and
dynamic coefficients, Mindlin partial slip, platen friction, buoyancy,
Maxwell-Wiechert bonds, creep damage, the calibrated Hertz dashpot) raise
NotImplementedError at construction instead of being ignored.

Differences a caller can observe:

* `DEM3DConfig(device="cpu")`, the reference's default, runs on the ROCm
  device too; every state tensor is a device tensor.
* a pair's friction spring is dropped when the pair separates or its bond
  is alive (the reference keeps stale springs of separated pairs in a dict,
  dem3d.py:1152-1185);
* a sphere's contacts are summed by ascending partner (the reference
  visits them in its cell-dictionary order), so forces agree to rounding;
* `run(steps)` advances many steps per call, replaying a captured graph.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import asdict
from typing import Any

import numpy as np
import torch

from . import _lib
from .dem3dconfig import (DEM3DConfig, build_lattice, hertz_effective_radius,
                          hertz_reduced_modulus, hertz_reduced_shear_modulus)
from .strength import weibull_strengths

CELL_CAP = 32
HISTORY_SLOTS = 16
SKIN_RADII = 0.2


def _unsupported(cfg: DEM3DConfig) -> list[str]:
    out = []
    if cfg.rolling_resistance_coefficient > 0.0:
        out.append("rolling_resistance_coefficient > 0")
    mu_s, mu_d = cfg.static_friction_coefficient, cfg.dynamic_friction_coefficient
    if cfg.friction_regularization != "none":
        mu = cfg.friction_coefficient
        if (mu if mu_s is None else mu_s) != (mu if mu_d is None else mu_d):
            out.append(f"friction_regularization={cfg.friction_regularization!r} with "
                       "distinct static and dynamic coefficients")
    if cfg.tangential_model != "cundall_strack":
        out.append(f"tangential_model={cfg.tangential_model!r}")
    if cfg.platen_friction_coefficient > 0.0:
        out.append("platen_friction_coefficient > 0")
    if cfg.water_level is not None:
        out.append("water_level (buoyancy)")
    if cfg.bond_model != "linear":
        out.append(f"bond_model={cfg.bond_model!r}")
    if cfg.bond_creep_exponent > 0:
        out.append("bond_creep_exponent > 0")
    if cfg.contact_damping_model == "restitution_calibrated":
        out.append("contact_damping_model='restitution_calibrated'")
    return out


def _friction_mu(cfg: DEM3DConfig) -> float:
    # "tanh" and "dual" with equal static and dynamic coefficients reduce to
    # mu_s N (dem3d.py:474-478)
    mu = float(cfg.friction_coefficient)
    if cfg.friction_regularization == "none":
        return mu
    return mu if cfg.static_friction_coefficient is None else float(cfg.static_friction_coefficient)


class IceDEM3D:
    """3-D bonded-sphere DEM with a spherical prescribed indenter.

    State arrays have shape (N, 3); external forces are instantaneous nodal
    loads in newtons.  The indenter translates at constant velocity.  Bond
    failure is irreversible (dem3d.py:564-577)."""

    def __init__(self, config: DEM3DConfig = DEM3DConfig()) -> None:
        self.config = config
        bad = _unsupported(config)
        if bad:
            raise NotImplementedError("IceDEM3D on the ROCm device does not evaluate: "
                                      + ", ".join(bad))
        if not torch.cuda.is_available():
            raise ValueError("tensordem_amd.IceDEM3D runs on the ROCm device, which is unavailable")
        self.lib = _lib.load()
        self.device = torch.device("cuda")
        self.dtype = torch.float64
        self.dt = config.dt if config.dt is not None else config.recommended_dt
        if self.dt > config.recommended_dt:
            raise ValueError(f"dt exceeds conservative limit {config.recommended_dt:.6g} s")
        lat = build_lattice(config)
        dev, f64 = self.device, torch.float64
        n = len(lat.positions)
        self._x = torch.tensor(lat.positions, dtype=f64, device=dev)
        self._v = torch.zeros_like(self._x)
        self._x_alt = torch.empty_like(self._x)
        self._v_alt = torch.empty_like(self._x)
        self.initial_positions = self._x.clone()
        self._fixed = torch.tensor(lat.fixed, dtype=torch.bool, device=dev)
        self.pairs = torch.tensor(lat.pairs, dtype=torch.long, device=dev)
        nb = len(lat.pairs)
        i, j = self.pairs.T
        ref = self.initial_positions[j] - self.initial_positions[i]
        self.rest_lengths = torch.linalg.vector_norm(ref, dim=1)
        self.bonded = torch.ones(nb, dtype=torch.bool, device=dev)
        self._alive = self.bonded.clone()
        self.bond_indices = self.pairs
        self.reference_bond_vectors = ref.clone()
        self.initial_normals = ref / self.rest_lengths.clamp_min(torch.finfo(f64).eps)[:, None]
        self.failure_mode = torch.zeros(nb, dtype=torch.uint8, device=dev)
        self.failure_time_s = torch.full((nb,), -1.0, dtype=f64, device=dev)
        self.failure_extension_m = torch.zeros(nb, dtype=f64, device=dev)
        self.failure_energy_J = torch.zeros(nb, dtype=f64, device=dev)
        self.bond_law = None
        self.branch_offsets = torch.zeros((nb, 0), dtype=f64, device=dev)
        self._bond_elongation_prev = torch.zeros(nb, dtype=f64, device=dev)
        self.bond_damage = torch.zeros(nb, dtype=f64, device=dev)
        self.bond_breaking_strain: torch.Tensor | None = None
        self.bond_strength_factors: torch.Tensor | None = None
        if config.bond_strength_weibull_modulus > 0.0 and nb:
            # dem3d.py:686-697: median-1 factors from a private CPU generator
            factors = torch.from_numpy(weibull_strengths(
                nb, float(config.bond_strength_weibull_modulus),
                seed=int(config.bond_strength_seed), median=1.0))
            self.bond_strength_factors = factors.to(device=dev)
            self.bond_breaking_strain = (config.breaking_strain * factors).to(device=dev)
        self.cumulative_bond_dissipation_J = 0.0
        self._bond_keys = self.pairs[:, 0] * n + self.pairs[:, 1]
        self.energy_audit = None
        self._time = 0.0
        self.step_count = 0
        # device scalars the kernels read between steps
        self._tool = torch.zeros(6, dtype=f64, device=dev)
        self._tool[2] = ((config.nz - 1) * 2 * config.radius + config.radius
                         + config.tool_radius + config.tool_gap)
        self._tool[5] = -config.tool_speed
        self._clock = torch.zeros(10, dtype=f64, device=dev)
        self.last_reaction = torch.zeros(3, dtype=f64, device=dev)
        self.top_platen_start_z = (float(lat.positions[:, 2].max()) + config.radius
                                   + config.top_platen_gap)
        self.bottom_platen_start_z = (float(lat.positions[:, 2].min()) - config.radius
                                      - config.bottom_platen_gap)
        if config.lateral_platens:
            g = config.lateral_platen_gap
            p = lat.positions
            self._clock[2:6] = torch.tensor(
                [float(p[:, 0].min()) - config.radius - g, float(p[:, 0].max()) + config.radius + g,
                 float(p[:, 1].min()) - config.radius - g, float(p[:, 1].max()) + config.radius + g],
                dtype=f64)
        self._rolling_torque = torch.zeros((n, 3), dtype=f64, device=dev)
        self._last_friction_power = 0.0
        self._last_rolling_power = 0.0
        self.rolling_resistance_torque_physical = False
        self._top_platen_tang_disp = torch.zeros((n, 3), dtype=f64, device=dev)
        self._bottom_platen_tang_disp = torch.zeros((n, 3), dtype=f64, device=dev)
        self.last_top_platen_friction = torch.zeros((n, 3), dtype=f64, device=dev)
        self.last_bottom_platen_friction = torch.zeros((n, 3), dtype=f64, device=dev)
        # kernel buffers
        self._nbr = torch.tensor(lat.nbr, dtype=torch.int32, device=dev)
        self._bos = torch.tensor(lat.bond_of_slot, dtype=torch.int32, device=dev)
        self._pairs32 = torch.tensor(lat.pairs, dtype=torch.int32, device=dev)
        self._force = torch.zeros((n, 3), dtype=f64, device=dev)
        self._pair_energy = torch.zeros(n, dtype=f64, device=dev)
        self._out = torch.zeros(_lib.D3_NOUT, dtype=f64, device=dev)
        self._loads_t: torch.Tensor | None = None
        self._graphs: dict = {}
        # about 8 buckets per sphere: a cell of width 2r + skin holds up to 8
        # lattice spheres, so collisions must stay rare
        ncells = 1 << max(10, (8 * n - 1).bit_length())
        kh = HISTORY_SLOTS if _friction_mu(config) > 0.0 else 1
        self._params = self._make_params(n, nb, ncells, kh)
        ws = int(self.lib.tdm_d3_workspace_bytes(n, nb, ncells, CELL_CAP, kh))
        self._ws = torch.empty(ws + 256, dtype=torch.uint8, device=dev)
        self._arrays = _lib.D3Arrays()
        self._bind_arrays()
        _lib.check(self.lib.tdm_d3_setup(ctypes.byref(self._arrays), ctypes.byref(self._params),
                                         self._stream_ptr))
        self._seen_version = self._x._version

    # ------------------------------------------------------------------ set-up
    @property
    def _stream_ptr(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _make_params(self, n, nb, ncells, kh) -> _lib.D3Params:
        c = self.config
        p = _lib.D3Params()
        e_p = hertz_reduced_modulus(c.young_modulus, c.poisson_ratio)
        r_p = hertz_effective_radius(c.radius)
        e_r = hertz_reduced_modulus(c.young_modulus, c.poisson_ratio, rigid_counterpart=True)
        r_r = hertz_effective_radius(c.radius, rigid_counterpart=True)
        g_p = hertz_reduced_shear_modulus(c.young_modulus, c.poisson_ratio)
        zeta = 0.0
        if c.restitution_coefficient < 1.0:
            a = -math.log(c.restitution_coefficient)
            zeta = a / math.sqrt(math.pi ** 2 + a ** 2)
        m = c.mass
        p.radius, p.mass, p.dt, p.drag = c.radius, m, self.dt, c.drag
        p.kb, p.kc, p.kt, p.mu = (c.bond_stiffness, c.contact_stiffness,
                                  c.effective_tangential_stiffness, _friction_mu(c))
        p.eps_t, p.eps_s = c.breaking_strain, c.shear_breaking_strain
        p.hz_pair = (4.0 / 3.0) * e_p * math.sqrt(r_p)
        p.hz_pair_u = (8.0 / 15.0) * e_p * math.sqrt(r_p)
        p.hz_rigid = (4.0 / 3.0) * e_r * math.sqrt(r_r)
        p.hz_kt = 8.0 * g_p * math.sqrt(r_p)
        p.zeta2 = 2.0 * zeta
        p.gamma_mobile = 2.0 * zeta * math.sqrt(c.contact_stiffness * (m / 2.0))
        p.gamma_pinned = 2.0 * zeta * math.sqrt(c.contact_stiffness * m)
        p.cdamp, p.tool_r = c.contact_damping, c.tool_radius
        p.platen_k, p.platen_c = c.platen_stiffness, c.platen_damping
        p.top_v, p.bot_v = c.top_platen_velocity, c.bottom_platen_velocity
        p.lat_k = (c.lateral_platen_stiffness if c.lateral_platen_stiffness is not None
                   else c.platen_stiffness)
        p.lat_c = (c.lateral_platen_damping if c.lateral_platen_damping is not None
                   else c.platen_damping)
        p.mg = c.mass * c.gravity
        # candidate lists of partners within 2r + skin, rebuilt on the device
        # when a sphere has moved 0.45 skin
        p.skin = SKIN_RADII * c.radius
        p.cell_w = 2.0 * c.radius + p.skin
        p.hertz = int(c.contact_model == "hertz")
        p.damp = 2 if c.contact_damping_model == "absolute" else (
            1 if c.restitution_coefficient < 1.0 else 0)
        p.friction = int(_friction_mu(c) > 0.0)
        p.mindlin_kt = int(c.contact_model == "hertz" and c.tangential_stiffness is None)
        p.top, p.bot = int(c.top_platen_enabled), int(c.bottom_platen_enabled)
        p.lateral, p.gravity = int(c.lateral_platens), int(c.gravity > 0.0)
        p.n, p.nbonds, p.ncells, p.cell_cap, p.kh = n, nb, ncells, CELL_CAP, kh
        p.nblocks = int(self.lib.tdm_d3_blocks(n))
        return p

    def _bind_arrays(self) -> None:
        a = self._arrays
        a.x, a.v = self._x.data_ptr(), self._v.data_ptr()
        a.x_alt, a.v_alt = self._x_alt.data_ptr(), self._v_alt.data_ptr()
        a.x0 = self.initial_positions.data_ptr()
        a.fixed = self._fixed.data_ptr()
        a.nbr, a.bond_of_slot = self._nbr.data_ptr(), self._bos.data_ptr()
        a.pairs = self._pairs32.data_ptr() if len(self.pairs) else None
        nb = len(self.pairs)
        a.alive = self._alive.data_ptr() if nb else None
        a.thr = self.bond_breaking_strain.data_ptr() if self.bond_breaking_strain is not None \
            else None
        a.failure_mode = self.failure_mode.data_ptr() if nb else None
        a.failure_time_s = self.failure_time_s.data_ptr() if nb else None
        a.failure_extension_m = self.failure_extension_m.data_ptr() if nb else None
        a.failure_energy_J = self.failure_energy_J.data_ptr() if nb else None
        a.force, a.pair_energy = self._force.data_ptr(), self._pair_energy.data_ptr()
        a.tool, a.clock, a.out = self._tool.data_ptr(), self._clock.data_ptr(), self._out.data_ptr()
        base = self._ws.data_ptr()
        a.workspace = (base + 255) & ~255
        a.workspace_bytes = self._ws.numel() - (a.workspace - base)
        a.loads = None

    def _sync_params(self) -> None:
        # platen starts are plain attributes a caller may move between steps
        self._params.top_z0 = float(self.top_platen_start_z)
        self._params.bot_z0 = float(self.bottom_platen_start_z)
        # positions written by the caller (assigned or edited in place, which
        # torch counts in the tensor's version) invalidate the candidate lists
        if self._x._version != self._seen_version:
            _lib.check(self.lib.tdm_d3_rebin(ctypes.byref(self._arrays),
                                             ctypes.byref(self._params), self._stream_ptr))
            self._seen_version = self._x._version

    @property
    def rebuilds(self) -> int:
        """Candidate-list rebuilds so far."""
        return int(self._out[_lib.D3_REBUILDS].item())

    # ------------------------------------------------------------------ state
    def _assign(self, dst: torch.Tensor, value, name: str) -> None:
        value = torch.as_tensor(value)
        if tuple(value.shape) != tuple(dst.shape):
            raise ValueError(f"{name} must have shape {tuple(dst.shape)}")
        dst.copy_(value.to(device=dst.device, dtype=dst.dtype))

    positions = property(lambda self: self._x,
                         lambda self, v: self._assign(self._x, v, "positions"))
    velocities = property(lambda self: self._v,
                          lambda self, v: self._assign(self._v, v, "velocities"))
    fixed = property(lambda self: self._fixed,
                     lambda self, v: self._assign(self._fixed, v, "fixed"))
    alive = property(lambda self: self._alive,
                     lambda self, v: self._assign(self._alive, v, "alive"))

    @property
    def time(self) -> float:
        return self._time

    @time.setter
    def time(self, value: float) -> None:
        self._time = float(value)
        self._clock[0] = self._time

    @property
    def tool_start(self) -> torch.Tensor:
        return self._tool[0:3]

    @tool_start.setter
    def tool_start(self, value) -> None:
        self._assign(self._tool[0:3], value, "tool_start")

    @property
    def tool_velocity(self) -> torch.Tensor:
        return self._tool[3:6]

    @tool_velocity.setter
    def tool_velocity(self, value) -> None:
        self._assign(self._tool[3:6], value, "tool_velocity")

    @property
    def lateral_platen_pos(self) -> torch.Tensor:
        return self._clock[2:6]

    @lateral_platen_pos.setter
    def lateral_platen_pos(self, value) -> None:
        self._assign(self._clock[2:6], value, "lateral_platen_pos")

    @property
    def lateral_platen_velocity(self) -> torch.Tensor:
        return self._clock[6:10]

    @lateral_platen_velocity.setter
    def lateral_platen_velocity(self, value) -> None:
        self._assign(self._clock[6:10], value, "lateral_platen_velocity")

    @property
    def last_top_platen_reaction_z(self) -> torch.Tensor:
        return self._out[_lib.D3_TOP_REACTION]

    @property
    def last_bottom_platen_reaction_z(self) -> torch.Tensor:
        return self._out[_lib.D3_BOTTOM_REACTION]

    @property
    def last_lateral_platen_reaction(self) -> torch.Tensor:
        return self._out[_lib.D3_LATERAL:_lib.D3_LATERAL + 4]

    @property
    def tool_position(self) -> torch.Tensor:
        return self.tool_start + self.time * self.tool_velocity

    @property
    def broken_bonds(self) -> int:
        return int((self.bonded & ~self.alive).sum().item())

    @property
    def unrelaxed_bond_stiffness(self) -> float:
        return float(self.config.bond_stiffness)

    @property
    def platen_friction_k(self) -> float:
        k_t = self.config.platen_friction_stiffness
        return float(self.config.platen_stiffness if k_t is None else k_t)

    @property
    def buoyancy_active(self) -> bool:
        return False

    # ------------------------------------------------------------------ evaluation
    def _validate_loads(self, external_forces: torch.Tensor | None) -> torch.Tensor | None:
        if external_forces is None:
            return None
        if not isinstance(external_forces, torch.Tensor) or \
                tuple(external_forces.shape) != tuple(self._x.shape):
            raise ValueError(f"external_forces must be a tensor of shape {tuple(self._x.shape)}")
        loads = external_forces.to(device=self.device, dtype=self.dtype).contiguous()
        if not bool(torch.isfinite(loads).all()):
            raise ValueError("external_forces must contain only finite values")
        return loads

    def _set_loads(self, loads: torch.Tensor | None) -> None:
        self._loads_t = loads
        self._arrays.loads = loads.data_ptr() if loads is not None else None

    def _check_status(self) -> int:
        st = int(self._out[_lib.D3_STATUS].item())
        if st & (_lib.ST_CELL_OVERFLOW | _lib.ST_CAND_OVERFLOW | _lib.ST_HIST_OVERFLOW):
            _lib.check(self.lib.tdm_d3_clear_status(ctypes.byref(self._arrays),
                                                    ctypes.byref(self._params), self._stream_ptr))
            raise RuntimeError(f"IceDEM3D contact capacity exceeded (status {st}): more spheres "
                               "in one cell, contacts or friction springs per sphere than the "
                               "kernel holds")
        return st

    @torch.no_grad()
    def forces(self, external_forces: torch.Tensor | None = None, *,
               update_fracture: bool = True,
               update_history: bool = True) -> tuple[torch.Tensor, torch.Tensor]:
        """Forces at the current state (dem3d.py:1300-1642); commits fracture
        unless update_fracture=False and advances friction springs unless
        update_history=False."""
        self._set_loads(self._validate_loads(external_forces))
        self._sync_params()
        flags = (_lib.D3_UPDATE_FRACTURE if update_fracture else 0) | \
            (_lib.D3_UPDATE_HISTORY if update_history else 0)
        _lib.check(self.lib.tdm_d3_forces(ctypes.byref(self._arrays), ctypes.byref(self._params),
                                          flags, self._stream_ptr))
        self._set_loads(None)
        reaction = self._out[_lib.D3_REACTION:_lib.D3_REACTION + 3].clone()
        if update_fracture:
            self.last_reaction.copy_(reaction)
        if update_history:
            self._last_friction_power = float(self._out[_lib.D3_FRICTION_POWER])
        self._check_status()
        return self._force.clone(), reaction

    @torch.no_grad()
    def step(self, external_forces: torch.Tensor | None = None) -> None:
        """One explicit (semi-implicit Euler) step (dem3d.py:1644-1662)."""
        self._steps(1, external_forces)
        if self.energy_audit is not None:
            self.energy_audit.observe(self)

    @torch.no_grad()
    def run(self, steps: int, external_forces: torch.Tensor | None = None, *,
            graph: bool = True, pairs: int = 32) -> None:
        """`steps` calls of step(external_forces), replaying a captured graph
        of 2*pairs steps where it can; the state is checked once at the end."""
        if steps < 0:
            raise ValueError("steps must be nonnegative")
        if self.energy_audit is not None:
            for _ in range(steps):
                self.step(external_forces)
            return
        loads = self._validate_loads(external_forces)
        self._set_loads(loads)
        self._sync_params()
        done = 0
        if graph and steps >= 2 * pairs:
            g = self._graph(pairs, loads)
            for _ in range(steps // (2 * pairs)):
                _lib.check(self.lib.tdm_graph_launch(g, self._stream_ptr))
                self._arrays.evals += 2 * pairs
                self._arrays.steps += 2 * pairs
                if self._params.friction:
                    self._arrays.hist_flips += 2 * pairs
                done += 2 * pairs
        rest = steps - done
        if rest:
            _lib.check(self.lib.tdm_d3_step(ctypes.byref(self._arrays),
                                            ctypes.byref(self._params), rest, self._stream_ptr))
        self._advance_clock(steps)
        self._set_loads(None)
        self._finish_steps()

    def _graph(self, pairs: int, loads: torch.Tensor | None):
        key = (pairs, bytes(self._params), None if loads is None else loads.data_ptr(),
               self._arrays.evals & 1, self._arrays.hist_flips & 1)
        g = self._graphs.get(key)
        if g is None:
            for old in self._graphs.values():
                self.lib.tdm_graph_destroy(old)
            self._graphs.clear()
            out = ctypes.c_void_p()
            side = torch.cuda.Stream(self.device)
            side.wait_stream(torch.cuda.current_stream(self.device))
            with torch.cuda.stream(side):
                _lib.check(self.lib.tdm_d3_graph_create(ctypes.byref(self._arrays),
                                                        ctypes.byref(self._params), pairs,
                                                        ctypes.c_void_p(side.cuda_stream),
                                                        ctypes.byref(out)))
            torch.cuda.current_stream(self.device).wait_stream(side)
            g = self._graphs[key] = out
            # the graph's loads buffer must outlive it
            self._graph_loads = loads
        return g

    def _steps(self, k: int, external_forces) -> None:
        self._set_loads(self._validate_loads(external_forces))
        self._sync_params()
        _lib.check(self.lib.tdm_d3_step(ctypes.byref(self._arrays), ctypes.byref(self._params),
                                        k, self._stream_ptr))
        self._set_loads(None)
        self._advance_clock(k)
        self._finish_steps()

    def _advance_clock(self, k: int) -> None:
        for _ in range(k):
            self._time += self.dt
        self.step_count += k

    def _finish_steps(self) -> None:
        self.last_reaction.copy_(self._out[_lib.D3_REACTION:_lib.D3_REACTION + 3])
        st = self._check_status()
        self._last_friction_power = float(self._out[_lib.D3_FRICTION_POWER])
        if st & _lib.ST_NONFINITE:
            raise FloatingPointError("non-finite DEM3D state; reduce dt or contact stiffness")

    def step_timed(self, steps: int) -> tuple[float, float, float]:
        """`steps` (even) steps with HIP events on each step, strain and bond
        kernel launch; mean durations in microseconds (measurement helper)."""
        self._sync_params()
        step_us, strain_us, bond_us = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _lib.check(self.lib.tdm_d3_step_timed(ctypes.byref(self._arrays),
                                              ctypes.byref(self._params), int(steps),
                                              self._stream_ptr, ctypes.byref(step_us),
                                              ctypes.byref(strain_us), ctypes.byref(bond_us)))
        self._advance_clock(int(steps))
        self._finish_steps()
        return step_us.value, strain_us.value, bond_us.value

    def shear_strain(self) -> torch.Tensor:
        """Objective shear strain of every sphere at the current state."""
        out = torch.empty(len(self._x), dtype=self.dtype, device=self.device)
        _lib.check(self.lib.tdm_d3_shear(ctypes.byref(self._arrays), ctypes.byref(self._params),
                                         ctypes.c_void_p(out.data_ptr()), self._stream_ptr))
        return out

    # ------------------------------------------------------------------ diagnostics
    @torch.no_grad()
    def bond_stored_energy(self, extension: torch.Tensor | None = None) -> torch.Tensor:
        if extension is None:
            i, j = self.pairs.T
            extension = torch.linalg.vector_norm(self._x[j] - self._x[i], dim=1) - self.rest_lengths
        return 0.5 * self.config.bond_stiffness * extension.square()

    def _pair_contact_energy(self) -> float:
        self._sync_params()
        _lib.check(self.lib.tdm_d3_energy(ctypes.byref(self._arrays), ctypes.byref(self._params),
                                          self._stream_ptr))
        return float(self._pair_energy.sum())

    def _rigid_potential(self, overlap: torch.Tensor, k: float) -> torch.Tensor:
        c = self.config
        if c.contact_model == "hertz":
            e_r = hertz_reduced_modulus(c.young_modulus, c.poisson_ratio, rigid_counterpart=True)
            r_r = hertz_effective_radius(c.radius, rigid_counterpart=True)
            return (8.0 / 15.0) * e_r * math.sqrt(r_r) * overlap.clamp_min(0.0).pow(2.5)
        return 0.5 * k * overlap.square()

    def _boundary_energies(self) -> tuple[float, float]:
        c = self.config
        tool_overlap = (c.radius + c.tool_radius - torch.linalg.vector_norm(
            self._x - self.tool_position, dim=1)).clamp_min(0)
        tool = float(self._rigid_potential(tool_overlap, c.contact_stiffness).sum())
        platen = 0.0
        if c.top_platen_enabled:
            z = self.top_platen_start_z + self.time * c.top_platen_velocity
            ov = (self._x[:, 2] + c.radius - z).clamp_min(0.0)
            platen += float(self._rigid_potential(ov, c.platen_stiffness).sum())
        if c.bottom_platen_enabled:
            z = self.bottom_platen_start_z + self.time * c.bottom_platen_velocity
            ov = (z + c.radius - self._x[:, 2]).clamp_min(0.0)
            platen += float(self._rigid_potential(ov, c.platen_stiffness).sum())
        return tool, platen

    @torch.no_grad()
    def lateral_platen_energy(self) -> float:
        c = self.config
        if not c.lateral_platens:
            return 0.0
        k = c.lateral_platen_stiffness if c.lateral_platen_stiffness is not None \
            else c.platen_stiffness
        px, py = self._x[:, 0], self._x[:, 1]
        lp = self.lateral_platen_pos
        total = ((lp[0] - px + c.radius).clamp_min(0.0).square().sum()
                 + (px + c.radius - lp[1]).clamp_min(0.0).square().sum()
                 + (lp[2] - py + c.radius).clamp_min(0.0).square().sum()
                 + (py + c.radius - lp[3]).clamp_min(0.0).square().sum())
        return float(0.5 * k * total)

    @torch.no_grad()
    def submerged_fraction(self) -> torch.Tensor:
        return torch.zeros(len(self._x), dtype=self.dtype, device=self.device)

    @torch.no_grad()
    def body_forces(self) -> torch.Tensor:
        body = torch.zeros_like(self._x)
        if self.config.gravity > 0.0:
            body[:, 2] -= self.config.mass * self.config.gravity
        return body

    @torch.no_grad()
    def body_force_ledger(self) -> dict[str, float]:
        gravity = -self.config.mass * self.config.gravity * len(self._x)
        return {"gravity_force_z_N": float(gravity), "buoyancy_force_z_N": 0.0,
                "wet_drag_force_z_N": 0.0, "net_body_force_z_N": float(gravity),
                "submerged_volume_m3": 0.0}

    @torch.no_grad()
    def diagnostics(self, external_forces: torch.Tensor | None = None) -> dict[str, Any]:
        """Forces, reactions and energies without committing anything
        (dem3d.py:1665-1738)."""
        force, reaction = self.forces(external_forces, update_fracture=False,
                                      update_history=False)
        support = -force[self._fixed].sum(dim=0)
        kinetic = 0.5 * self.config.mass * self._v.square().sum()
        bond_energy = (self.bond_stored_energy() * self._alive).sum()
        contact = self._pair_contact_energy()
        tool, platen = self._boundary_energies()
        lateral = self.lateral_platen_energy()
        tp = self.tool_position
        return {
            "time": self.time,
            "tool_x": float(tp[0]), "tool_y": float(tp[1]), "tool_z": float(tp[2]),
            "reaction_x": float(reaction[0]), "reaction_y": float(reaction[1]),
            "reaction_z": float(reaction[2]),
            "top_platen_reaction_z_N": float(self.last_top_platen_reaction_z),
            "bottom_platen_reaction_z_N": float(self.last_bottom_platen_reaction_z),
            "boundary_reaction_x": float(support[0]), "boundary_reaction_y": float(support[1]),
            "boundary_reaction_z": float(support[2]),
            "broken_bonds": self.broken_bonds,
            "particle_count": len(self._x),
            "kinetic_energy_J": float(kinetic),
            "bond_elastic_energy_J": float(bond_energy),
            "particle_contact_energy_J": contact,
            "tool_contact_energy_J": tool,
            "platen_contact_energy_J": platen,
            "released_bond_energy_J": float(self.failure_energy_J.sum()),
            "lateral_platen_reaction": [float(v) for v in self.last_lateral_platen_reaction],
            "lateral_platen_energy_J": lateral,
            "mechanical_energy_J": float(kinetic + bond_energy) + contact + tool + platen + lateral,
            **self.body_force_ledger(),
        }

    @torch.no_grad()
    def energy(self) -> dict[str, float]:
        """Mechanical energy components (dem3d.py:1740-1805)."""
        c = self.config
        kinetic = 0.5 * c.mass * float(self._v.square().sum())
        i, j = self.pairs.T
        extension = torch.linalg.vector_norm(self._x[j] - self._x[i], dim=1) - self.rest_lengths
        bond = float(self.bond_stored_energy(extension)[self._alive].sum())
        contact = self._pair_contact_energy()
        tool, platen = self._boundary_energies()
        platen += self.lateral_platen_energy()
        released = 0.5 * c.bond_stiffness * float(extension[~self._alive].square().sum())
        return {
            "time_s": self.time, "step": self.step_count, "broken_bonds": self.broken_bonds,
            "particle_count": len(self._x), "kinetic_energy_J": kinetic,
            "bond_elastic_energy_J": bond, "particle_contact_energy_J": contact,
            "tool_contact_energy_J": tool, "platen_contact_energy_J": platen,
            "released_bond_energy_J": released,
            "mechanical_energy_J": kinetic + bond + contact + tool + platen,
        }

    def bond_law_report(self) -> dict[str, Any]:
        cfg = self.config
        out: dict[str, Any] = {
            "bond_model": cfg.bond_model,
            "unrelaxed_stiffness_N_m": float(cfg.bond_stiffness),
            "relaxed_stiffness_N_m": float(cfg.bond_stiffness),
            "creep_exponent": float(cfg.bond_creep_exponent),
            "creep_time_s": float(cfg.bond_creep_time_s),
            "cumulative_bond_dissipation_J": float(self.cumulative_bond_dissipation_J),
            "creep_failed_bonds": int((self.failure_mode == 4).sum()),
            "strength_weibull_modulus": float(cfg.bond_strength_weibull_modulus),
            "strength_weibull_seed": int(cfg.bond_strength_seed),
            "max_bond_damage": float(self.bond_damage.max()) if len(self.bond_damage) else 0.0,
        }
        if self.bond_breaking_strain is not None:
            out["bond_breaking_strain_median"] = float(self.bond_breaking_strain.median())
            out["bond_breaking_strain_min"] = float(self.bond_breaking_strain.min())
            out["bond_breaking_strain_max"] = float(self.bond_breaking_strain.max())
        return out

    # ------------------------------------------------------------------ restart
    def state_dict(self) -> dict[str, Any]:
        """Restart dictionary of dem3d.py:1828-1852 (tensors on the device)."""
        return {
            "dimension": 3, "config": asdict(self.config), "dt": self.dt,
            "time": self.time, "step_count": self.step_count,
            "positions": self._x.clone(), "velocities": self._v.clone(),
            "initial_positions": self.initial_positions.clone(), "fixed": self._fixed.clone(),
            "pairs": self.pairs.clone(), "bonded": self.bonded.clone(),
            "alive": self._alive.clone(),
            "failure_mode": self.failure_mode.clone(),
            "failure_time_s": self.failure_time_s.clone(),
            "failure_extension_m": self.failure_extension_m.clone(),
            "failure_energy_J": self.failure_energy_J.clone(),
            "bond_law": {
                "branch_offsets": self.branch_offsets.clone(),
                "elongation_prev": self._bond_elongation_prev.clone(),
                "damage": self.bond_damage.clone(),
                "cumulative_bond_dissipation_J": self.cumulative_bond_dissipation_J,
            },
        }

    @torch.no_grad()
    def load_state_dict(self, state: dict[str, Any]) -> None:
        """Validated restore (dem3d.py:1854-1967); friction springs restart
        from zero, as the reference's dictionary does not carry them."""
        if not isinstance(state, dict):
            raise TypeError("state must be a dictionary returned by state_dict()")
        required = {"dimension", "config", "dt", "time", "step_count", "positions", "velocities",
                    "initial_positions", "fixed", "pairs", "bonded", "alive", "failure_mode",
                    "failure_time_s", "failure_extension_m", "failure_energy_J"}
        missing = required.difference(state)
        if missing:
            raise ValueError(f"checkpoint is missing keys: {sorted(missing)}")
        if state["dimension"] != 3:
            raise ValueError("checkpoint dimension must be 3")
        saved = state["config"]
        if not isinstance(saved, dict):
            raise ValueError("checkpoint config must be a dictionary")
        if {**asdict(DEM3DConfig()), **saved} != asdict(self.config):
            raise ValueError("checkpoint config does not match this solver configuration")
        if not math.isfinite(float(state["dt"])) or float(state["dt"]) != self.dt:
            raise ValueError("checkpoint dt does not match this solver")
        time_value = float(state["time"])
        step_value = state["step_count"]
        if not math.isfinite(time_value) or time_value < 0:
            raise ValueError("checkpoint time must be finite and nonnegative")
        if isinstance(step_value, bool) or not isinstance(step_value, int) or step_value < 0:
            raise ValueError("checkpoint step_count must be a nonnegative integer")
        targets = {
            "positions": self._x, "velocities": self._v,
            "initial_positions": self.initial_positions, "fixed": self._fixed,
            "pairs": self.pairs, "bonded": self.bonded, "alive": self._alive,
            "failure_mode": self.failure_mode, "failure_time_s": self.failure_time_s,
            "failure_extension_m": self.failure_extension_m,
            "failure_energy_J": self.failure_energy_J,
        }
        staged: dict[str, torch.Tensor] = {}
        for name, target in targets.items():
            value = state[name]
            if not isinstance(value, torch.Tensor) or tuple(value.shape) != tuple(target.shape):
                raise ValueError(f"checkpoint {name} must be a tensor with shape "
                                 f"{tuple(target.shape)}")
            if value.dtype != target.dtype:
                raise ValueError(f"checkpoint {name} dtype must be {target.dtype}, got {value.dtype}")
            value = value.to(device=self.device)
            if value.is_floating_point() and not bool(torch.isfinite(value).all()):
                raise ValueError(f"checkpoint {name} contains non-finite values")
            staged[name] = value
        if not torch.equal(staged["initial_positions"], self.initial_positions):
            raise ValueError("checkpoint reference geometry does not match this solver")
        if not torch.equal(staged["pairs"], self.pairs):
            raise ValueError("checkpoint pair topology does not match this solver")
        if not torch.equal(staged["bonded"], self.bonded):
            raise ValueError("checkpoint bond topology does not match this solver")
        if bool((staged["alive"] & ~staged["bonded"]).any()):
            raise ValueError("checkpoint marks a non-bond pair as alive")
        if not torch.equal(staged["fixed"], self._fixed):
            raise ValueError("checkpoint fixed-boundary mask does not match this solver")
        for name, target in targets.items():
            target.copy_(staged[name])
        self.time = time_value
        self.step_count = step_value
        self.last_reaction.zero_()
        _lib.check(self.lib.tdm_d3_clear_history(ctypes.byref(self._arrays),
                                                 ctypes.byref(self._params), self._stream_ptr))

    def __del__(self):
        try:
            for g in self._graphs.values():
                self.lib.tdm_graph_destroy(g)
        except Exception:
            pass

"""FP64 CPU oracle of IceDEM3D (test infrastructure only).

`OracleDEM3D` keeps numpy state and drives `orc3_eval` of oracle/dem3d.c,
which documents the law and the summation order.  The lattice is built
here with the reference's own loops (src/tensordem/dem3d.py:589-656):
meshgrid order, cylinder mask, 13 half-space offsets, compaction, bonds
sorted by i * n + j.  Only tests, the smoke check and the benchmark's CPU
baseline use this module.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from .plate import _p, library

(D3_RADIUS, D3_MASS, D3_DT, D3_DRAG, D3_KB, D3_KC, D3_KT, D3_MU, D3_EPS_T, D3_EPS_S,
 D3_HZ_PAIR, D3_HZ_PAIR_U, D3_HZ_RIGID, D3_HZ_KT, D3_ZETA2, D3_GAMMA_MOBILE, D3_GAMMA_PINNED,
 D3_CDAMP, D3_TOOL_R) = range(19)
D3_TOOL_X0 = 19
D3_TOOL_V = 22
(D3_PLATEN_K, D3_PLATEN_C, D3_TOP_Z0, D3_BOT_Z0, D3_TOP_V, D3_BOT_V, D3_LAT_K, D3_LAT_C,
 D3_MG, D3_TIME) = range(25, 35)
D3_LAT_POS = 35
D3_LAT_VEL = 39
D3_NPARAM = 43
(I3_HERTZ, I3_DAMP, I3_FRICTION, I3_MINDLIN_KT, I3_TOP, I3_BOT, I3_LATERAL, I3_GRAVITY,
 I3_UPDATE_FRACTURE, I3_UPDATE_HISTORY, I3_OP, I3_KH, I3_BRUTE, I3_NIPARAM) = range(14)
DAMP_NONE, DAMP_RESTITUTION, DAMP_ABSOLUTE = 0, 1, 2
OP_STEP, OP_FORCES, OP_ENERGY = 0, 1, 2
O3_REACTION, O3_TOP_Z, O3_BOT_Z, O3_LAT = 0, 3, 4, 5
O3_FRICTION_POWER, O3_CONTACTS, O3_BREAKS, O3_STATUS, O3_NOUT = 9, 10, 11, 12, 13
KH = 16


def _bind():
    lib = library()
    if not hasattr(lib, "_d3_bound"):
        d = ctypes.POINTER(ctypes.c_double)
        u8 = ctypes.POINTER(ctypes.c_uint8)
        i32 = ctypes.POINTER(ctypes.c_int32)
        ci = ctypes.POINTER(ctypes.c_int)
        c_int = ctypes.c_int
        lib.orc3_covinv.argtypes = [c_int, c_int, i32, d, d, u8]
        lib.orc3_covinv.restype = None
        lib.orc3_shear.argtypes = [c_int, c_int, i32, d, d, d, u8, d]
        lib.orc3_shear.restype = None
        lib.orc3_eval.argtypes = [c_int, c_int, d, ci, d, d, d, u8, i32, u8, d, d, u8,
                                  u8, d, d, d, d, d, i32, i32, d, i32, i32, d, d, d]
        lib.orc3_eval.restype = c_int
        lib._d3_bound = True
    return lib


def reference_lattice(cfg):
    """dem3d.py:589-656 restated with plain loops: (positions, fixed, pairs)."""
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    full = []
    keep = []
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                full.append((float(x), float(y), float(z)))
                if cfg.geometry == "cylinder_xz":
                    xc, zc = (nx - 1) / 2.0, (nz - 1) / 2.0
                    rad = min(nx - 1, nz - 1) / 2.0
                    keep.append((x - xc) * (x - xc) + (z - zc) * (z - zc) <= rad**2 + 1e-12)
                else:
                    keep.append(True)
    compact = {}
    grid = []
    for f, k in enumerate(keep):
        if k:
            compact[f] = len(grid)
            grid.append(full[f])
    z_origin = (nz - 1) / 2.0 if cfg.geometry == "cylinder_xz" else 0.0
    centre = ((nx - 1) / 2, (ny - 1) / 2, z_origin)
    pos = np.array([[(g[k] - centre[k]) * (2 * cfg.radius) for k in range(3)] for g in grid])
    fixed = np.zeros(len(grid), dtype=np.uint8)
    for p, g in enumerate(grid):
        if cfg.fix_x_edges and cfg.geometry == "box" and (g[0] == 0 or g[0] == nx - 1):
            fixed[p] = 1
        if cfg.fix_bottom and g[2] == 0:
            fixed[p] = 1
    pairs = []
    for dz in range(0, 2):
        for dy in range(-1, 2):
            for dx in range(-1, 2):
                if dx == dy == dz == 0:
                    continue
                if not (dz > 0 or (dz == 0 and dy > 0) or (dz == 0 and dy == 0 and dx > 0)):
                    continue
                for z in range(nz):
                    if not 0 <= z + dz < nz:
                        continue
                    for y in range(ny):
                        if not 0 <= y + dy < ny:
                            continue
                        for x in range(nx):
                            if not 0 <= x + dx < nx:
                                continue
                            a = compact.get((z * ny + y) * nx + x, -1)
                            b = compact.get(((z + dz) * ny + y + dy) * nx + x + dx, -1)
                            if a >= 0 and b >= 0:
                                pairs.append((min(a, b), max(a, b)))
    n = len(grid)
    pairs.sort(key=lambda ab: ab[0] * n + ab[1])
    return pos, fixed, np.array(pairs, dtype=np.int32).reshape(-1, 2)


def pair_constants(cfg):
    """Scalars of the contact law computed as the reference does (dem3d.py:1352-1419)."""
    from tensordem_amd.dem3dconfig import (hertz_effective_radius, hertz_reduced_modulus,
                                           hertz_reduced_shear_modulus)
    e_p = hertz_reduced_modulus(cfg.young_modulus, cfg.poisson_ratio)
    r_p = hertz_effective_radius(cfg.radius)
    e_r = hertz_reduced_modulus(cfg.young_modulus, cfg.poisson_ratio, rigid_counterpart=True)
    r_r = hertz_effective_radius(cfg.radius, rigid_counterpart=True)
    g_p = hertz_reduced_shear_modulus(cfg.young_modulus, cfg.poisson_ratio)
    zeta = 0.0
    if cfg.restitution_coefficient < 1.0:
        a = -math.log(cfg.restitution_coefficient)
        zeta = a / math.sqrt(math.pi ** 2 + a ** 2)
    m = cfg.mass
    return dict(
        hz_pair=(4.0 / 3.0) * e_p * math.sqrt(r_p),
        hz_pair_u=(8.0 / 15.0) * e_p * math.sqrt(r_p),
        hz_rigid=(4.0 / 3.0) * e_r * math.sqrt(r_r),
        hz_kt=8.0 * g_p * math.sqrt(r_p),
        zeta2=2.0 * zeta,
        gamma_mobile=2.0 * zeta * math.sqrt(cfg.contact_stiffness * (m / 2.0)),
        gamma_pinned=2.0 * zeta * math.sqrt(cfg.contact_stiffness * m),
    )


class OracleDEM3D:
    """numpy mirror of IceDEM3D driven by the C oracle."""

    def __init__(self, config, *, dt=None, bond_thresholds=None, brute_force=False):
        c = self.config = config
        self.brute_force = bool(brute_force)
        self.lib = _bind()
        self.dt = float(dt if dt is not None else (c.dt if c.dt is not None else c.recommended_dt))
        pos, fixed, pairs = reference_lattice(c)
        self.positions = pos.copy()
        self.initial_positions = pos.copy()
        self.velocities = np.zeros_like(pos)
        self.fixed = fixed
        self.pairs = pairs
        nb = len(pairs)
        n = len(pos)
        self.alive = np.ones(nb, dtype=np.uint8)
        self.failure_mode = np.zeros(nb, dtype=np.uint8)
        self.failure_time_s = np.full(nb, -1.0)
        self.failure_extension_m = np.zeros(nb)
        self.failure_energy_J = np.zeros(nb)
        self.thresholds = None if bond_thresholds is None else \
            np.ascontiguousarray(bond_thresholds, dtype=np.float64)
        self.inv = np.zeros((n, 9))
        self.degenerate = np.zeros(n, dtype=np.uint8)
        self.lib.orc3_covinv(n, nb, _p(self.pairs, ctypes.c_int32),
                             _p(self.initial_positions, ctypes.c_double),
                             _p(self.inv, ctypes.c_double), _p(self.degenerate, ctypes.c_uint8))
        self.hist_xi = np.zeros((n, KH, 3))
        self.hist_q = np.full((n, KH), -1, dtype=np.int32)
        self.hist_cnt = np.zeros(n, dtype=np.int32)
        self.time = 0.0
        self.step_count = 0
        self.tool_start = np.array([0.0, 0.0, (c.nz - 1) * 2 * c.radius + c.radius
                                    + c.tool_radius + c.tool_gap])
        self.tool_velocity = np.array([0.0, 0.0, -c.tool_speed])
        self.top_platen_start_z = float(pos[:, 2].max()) + c.radius + c.top_platen_gap
        self.bottom_platen_start_z = float(pos[:, 2].min()) - c.radius - c.bottom_platen_gap
        self.lateral_platen_pos = np.zeros(4)
        self.lateral_platen_velocity = np.zeros(4)
        if c.lateral_platens:
            g = c.lateral_platen_gap
            self.lateral_platen_pos[:] = (pos[:, 0].min() - c.radius - g, pos[:, 0].max() + c.radius + g,
                                          pos[:, 1].min() - c.radius - g, pos[:, 1].max() + c.radius + g)
        self.out = np.zeros(O3_NOUT)
        self.last_reaction = np.zeros(3)

    def _params(self, op, update_fracture, update_history):
        c = self.config
        k = pair_constants(c)
        dp = np.zeros(D3_NPARAM)
        dp[D3_RADIUS], dp[D3_MASS], dp[D3_DT], dp[D3_DRAG] = c.radius, c.mass, self.dt, c.drag
        dp[D3_KB], dp[D3_KC], dp[D3_KT] = c.bond_stiffness, c.contact_stiffness, \
            c.effective_tangential_stiffness
        dp[D3_MU], dp[D3_EPS_T], dp[D3_EPS_S] = c.friction_coefficient, c.breaking_strain, \
            c.shear_breaking_strain
        dp[D3_HZ_PAIR], dp[D3_HZ_PAIR_U], dp[D3_HZ_RIGID], dp[D3_HZ_KT] = \
            k["hz_pair"], k["hz_pair_u"], k["hz_rigid"], k["hz_kt"]
        dp[D3_ZETA2], dp[D3_GAMMA_MOBILE], dp[D3_GAMMA_PINNED] = \
            k["zeta2"], k["gamma_mobile"], k["gamma_pinned"]
        dp[D3_CDAMP], dp[D3_TOOL_R] = c.contact_damping, c.tool_radius
        dp[D3_TOOL_X0:D3_TOOL_X0 + 3] = self.tool_start
        dp[D3_TOOL_V:D3_TOOL_V + 3] = self.tool_velocity
        dp[D3_PLATEN_K], dp[D3_PLATEN_C] = c.platen_stiffness, c.platen_damping
        dp[D3_TOP_Z0], dp[D3_BOT_Z0] = self.top_platen_start_z, self.bottom_platen_start_z
        dp[D3_TOP_V], dp[D3_BOT_V] = c.top_platen_velocity, c.bottom_platen_velocity
        dp[D3_LAT_K] = c.lateral_platen_stiffness if c.lateral_platen_stiffness is not None \
            else c.platen_stiffness
        dp[D3_LAT_C] = c.lateral_platen_damping if c.lateral_platen_damping is not None \
            else c.platen_damping
        dp[D3_MG] = c.mass * c.gravity
        dp[D3_TIME] = self.time
        dp[D3_LAT_POS:D3_LAT_POS + 4] = self.lateral_platen_pos
        dp[D3_LAT_VEL:D3_LAT_VEL + 4] = self.lateral_platen_velocity
        ip = np.zeros(I3_NIPARAM, dtype=np.intc)
        ip[I3_HERTZ] = c.contact_model == "hertz"
        if c.contact_damping_model == "absolute":
            ip[I3_DAMP] = DAMP_ABSOLUTE
        elif c.restitution_coefficient < 1.0:
            ip[I3_DAMP] = DAMP_RESTITUTION
        ip[I3_FRICTION] = c.friction_coefficient > 0.0
        ip[I3_MINDLIN_KT] = c.contact_model == "hertz" and c.tangential_stiffness is None
        ip[I3_TOP], ip[I3_BOT] = c.top_platen_enabled, c.bottom_platen_enabled
        ip[I3_LATERAL] = c.lateral_platens
        ip[I3_GRAVITY] = c.gravity > 0.0
        ip[I3_UPDATE_FRACTURE], ip[I3_UPDATE_HISTORY] = update_fracture, update_history
        ip[I3_OP], ip[I3_KH] = op, KH
        ip[I3_BRUTE] = self.brute_force
        return dp, ip

    def _eval(self, op, external_forces=None, update_fracture=True, update_history=True):
        n, nb = len(self.positions), len(self.pairs)
        dp, ip = self._params(op, update_fracture, update_history)
        loads = None if external_forces is None else \
            np.ascontiguousarray(external_forces, dtype=np.float64)
        force = np.zeros((n, 3))
        energy = np.zeros(n)
        hx, hq, hc = np.zeros_like(self.hist_xi), np.full_like(self.hist_q, -1), \
            np.zeros_like(self.hist_cnt)
        D, U8, I32 = ctypes.c_double, ctypes.c_uint8, ctypes.c_int32
        st = self.lib.orc3_eval(
            n, nb, _p(dp, D), _p(ip, ctypes.c_int), _p(self.positions, D), _p(self.velocities, D),
            _p(self.initial_positions, D), _p(self.fixed, U8), _p(self.pairs, I32),
            _p(self.alive, U8), _p(self.thresholds, D), _p(self.inv, D), _p(self.degenerate, U8),
            _p(self.failure_mode, U8), _p(self.failure_time_s, D), _p(self.failure_extension_m, D),
            _p(self.failure_energy_J, D), _p(loads, D), _p(self.hist_xi, D), _p(self.hist_q, I32),
            _p(self.hist_cnt, I32), _p(hx, D), _p(hq, I32), _p(hc, I32), _p(force, D),
            _p(energy, D), _p(self.out, D))
        if ip[I3_FRICTION] and update_history and op != OP_ENERGY:
            self.hist_xi, self.hist_q, self.hist_cnt = hx, hq, hc
        if st & 16:
            raise RuntimeError("oracle friction history overflow")
        return force, energy, st

    def forces(self, external_forces=None, *, update_fracture=True, update_history=True):
        force, _, _ = self._eval(OP_FORCES, external_forces, update_fracture, update_history)
        reaction = self.out[O3_REACTION:O3_REACTION + 3].copy()
        if update_fracture:
            self.last_reaction = reaction
        return force, reaction

    def step(self, external_forces=None):
        force, _, st = self._eval(OP_STEP, external_forces)
        self.last_reaction = self.out[O3_REACTION:O3_REACTION + 3].copy()
        if self.config.lateral_platens:
            self.lateral_platen_pos = self.lateral_platen_pos + self.lateral_platen_velocity * self.dt
        self.time += self.dt
        self.step_count += 1
        if st & 1:
            raise FloatingPointError("non-finite DEM3D state; reduce dt or contact stiffness")
        return force

    def shear_strain(self):
        n = len(self.positions)
        out = np.zeros(n)
        self.lib.orc3_shear(n, len(self.pairs), _p(self.pairs, ctypes.c_int32),
                            _p(self.positions, ctypes.c_double),
                            _p(self.initial_positions, ctypes.c_double),
                            _p(self.inv, ctypes.c_double), _p(self.degenerate, ctypes.c_uint8),
                            _p(out, ctypes.c_double))
        return out

    def pair_energy(self):
        _, energy, _ = self._eval(OP_ENERGY)
        return float(energy.sum())

    @property
    def broken_bonds(self) -> int:
        return int((self.alive == 0).sum())

"""FP64 CPU oracle for the bonded ice plate (test infrastructure only).

`OraclePlate` drives the C restatement in `oracle/dem2d.c` (built into
`oracle/liboracle.so` by `oracle/Makefile`) with numpy state.  It mirrors the
reference solver `IceDEM` (src/tensordem/dem.py:75-624): the same lattice, the
same fracture commit, the same force law, integrator and ledger channels.

Only tests, the smoke check and the CPU-baseline leg of the benchmark use this
module; the product path never imports it.
"""
from __future__ import annotations

import ctypes
import math
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = None

P_RADIUS, P_KB, P_KC, P_EPS_T, P_EPS_S, P_CDAMP, P_DRAG, P_TOOL_R, P_DT, P_MASS, \
    P_TOOL_X, P_TOOL_Y, P_TOOL_VX, P_TOOL_VY, P_TIME, P_NPARAM = range(16)

ACC_KEYS = ("external_work_J", "fracture_release_J", "pair_contact_acquired_J",
            "tool_work_J", "damping_dissipation_J")
IMP_KEYS = ("external_impulse_Ns", "support_impulse_Ns", "drag_impulse_Ns",
            "tool_impulse_on_ice_Ns")


def library() -> ctypes.CDLL:
    """Load (building on first use in a source tree) the oracle shared library."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = _HERE / "liboracle.so"
    if not path.exists():
        import subprocess
        subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
    lib = ctypes.CDLL(str(path))
    d = ctypes.POINTER(ctypes.c_double)
    i32 = ctypes.POINTER(ctypes.c_int32)
    u8 = ctypes.POINTER(ctypes.c_uint8)
    c_int = ctypes.c_int
    lib.orc_shear_strain.argtypes = [c_int, d, d, c_int, i32, i32, d]
    lib.orc_shear_strain.restype = None
    lib.orc_plate_forces.argtypes = [d, c_int, d, d, d, c_int, i32, i32, d, u8, u8, d, d, d,
                                     d, d, c_int, d, d, d]
    lib.orc_plate_forces.restype = None
    lib.orc_plate_step.argtypes = [d, c_int, d, d, d, u8, c_int, i32, i32, d, u8, u8, d, d, d,
                                   d, d, c_int, d, d, d]
    lib.orc_plate_step.restype = c_int
    lib.orc_plate_step_thr.argtypes = lib.orc_plate_step.argtypes + [d, d]
    lib.orc_plate_step_thr.restype = c_int
    lib.orc_plate_forces_thr.argtypes = lib.orc_plate_forces.argtypes + [d, d]
    lib.orc_plate_forces_thr.restype = None
    lib.orc_mech_energy.argtypes = [d, c_int, d, d, c_int, i32, i32, d, u8, d, c_int, d]
    lib.orc_mech_energy.restype = None
    lib.orc_pairs_within.argtypes = [c_int, d, ctypes.c_double, i32, c_int]
    lib.orc_pairs_within.restype = c_int
    c_dbl = ctypes.c_double
    lib.orc_hull_forces.argtypes = [c_int, d, d, d, c_int, d, c_dbl, c_dbl, c_dbl, c_dbl, d, d]
    lib.orc_hull_forces.restype = None
    _LIB = lib
    return lib


def _p(a: np.ndarray | None, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def lattice(nx: int, ny: int, radius: float, fix_edges: bool = True):
    """Reference lattice (dem.py:105-129): positions, fixed mask, bonds, rest lengths.

    Bonds are every pair closer than 2r*sqrt(2)*(1+1e-10), listed in the
    `torch.triu_indices` order of the reference (ascending i, then j).  On a
    square grid of spacing 2r only the eight grid neighbours can qualify, so
    they are the only candidates tested.
    """
    x = np.arange(nx, dtype=np.float64)
    y = np.arange(ny, dtype=np.float64)
    yy, xx = np.meshgrid(y, x, indexing="ij")
    pos = np.stack((xx.ravel() - (nx - 1) / 2, yy.ravel()), axis=1) * (2 * radius)
    fixed = np.zeros(nx * ny, dtype=np.uint8)
    if fix_edges:
        fixed = ((xx.ravel() == 0) | (xx.ravel() == nx - 1)).astype(np.uint8)
    cutoff = 2 * radius * math.sqrt(2) * (1 + 1e-10)
    ii, jj = [], []
    idx = np.arange(nx * ny).reshape(ny, nx)
    # candidate partners with larger flat index: (0,+1), (+1,-1), (+1,0), (+1,+1)
    for oy, ox in ((0, 1), (1, -1), (1, 0), (1, 1)):
        ys, xs = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
        ok = (ys + oy < ny) & (xs + ox >= 0) & (xs + ox < nx)
        a = idx[ys[ok], xs[ok]]
        b = idx[ys[ok] + oy, xs[ok] + ox]
        ii.append(a)
        jj.append(b)
    bi = np.concatenate(ii).astype(np.int64)
    bj = np.concatenate(jj).astype(np.int64)
    delta = pos[bj] - pos[bi]
    dist = np.sqrt(delta[:, 0] * delta[:, 0] + delta[:, 1] * delta[:, 1])
    keep = dist <= cutoff
    bi, bj, dist = bi[keep], bj[keep], dist[keep]
    order = np.lexsort((bj, bi))
    return (pos, fixed, bi[order].astype(np.int32), bj[order].astype(np.int32),
            dist[order].copy())


class OraclePlate:
    """numpy-state mirror of `IceDEM` driven by the C oracle.

    `config` is any object with the `DEMConfig` fields (dem.py:17-72).
    """

    def __init__(self, config, *, hull_profile=None, hull_start=(0.0, 0.0),
                 hull_velocity=(0.0, 0.0), dt=None, eps_t_disk=None, eps_s_disk=None):
        self.config = config
        self.lib = library()
        c = config
        self.dt = float(dt if dt is not None else (c.dt if c.dt is not None else c.recommended_dt))
        pos, fixed, bi, bj, rest = lattice(c.nx, c.ny, c.radius, c.fix_edges)
        self.positions = pos.copy()
        self.initial_positions = pos.copy()
        self.velocities = np.zeros_like(pos)
        self.fixed = fixed
        self.bi, self.bj, self.rest = bi, bj, rest
        nb = len(bi)
        self.alive = np.ones(nb, dtype=np.uint8)
        self.failure_mode = np.zeros(nb, dtype=np.uint8)
        self.failure_time_s = np.full(nb, -1.0)
        self.failure_extension_m = np.zeros(nb)
        self.failure_energy_J = np.zeros(nb)
        self.time = 0.0
        self.step_count = 0
        self.acc = np.zeros(13)
        self.last = np.zeros(10)
        self.tool_start = np.array([0.0, (c.ny - 1) * 2 * c.radius + c.radius + c.tool_radius
                                    + c.tool_gap])
        self.tool_velocity = np.array([0.0, -c.tool_speed])
        self.hull_profile = None if hull_profile is None else np.asarray(hull_profile, np.float64)
        self.hull_start = np.asarray(hull_start, np.float64)
        self.hull_velocity = np.asarray(hull_velocity, np.float64)
        self.force = np.zeros_like(pos)
        # per-disk strengths: a bond breaks at the weaker of its two disks
        self.eps_t_bond = self.eps_s_bond = None
        if eps_t_disk is not None or eps_s_disk is not None:
            et = np.full(len(pos), c.breaking_strain) if eps_t_disk is None else \
                np.asarray(eps_t_disk, np.float64)
            es = np.full(len(pos), c.shear_breaking_strain) if eps_s_disk is None else \
                np.asarray(eps_s_disk, np.float64)
            self.eps_t_bond = np.ascontiguousarray(np.minimum(et[bi], et[bj]))
            self.eps_s_bond = np.ascontiguousarray(np.minimum(es[bi], es[bj]))

    # -- helpers -------------------------------------------------------------
    @property
    def tool_position(self) -> np.ndarray:
        if self.hull_profile is not None:
            return self.hull_start + self.time * self.hull_velocity
        return self.tool_start + self.time * self.tool_velocity

    def _params(self) -> np.ndarray:
        c = self.config
        p = np.zeros(P_NPARAM)
        p[P_RADIUS], p[P_KB], p[P_KC] = c.radius, c.bond_stiffness, c.contact_stiffness
        p[P_EPS_T], p[P_EPS_S] = c.breaking_strain, c.shear_breaking_strain
        p[P_CDAMP], p[P_DRAG], p[P_TOOL_R] = c.contact_damping, c.drag, c.tool_radius
        p[P_DT], p[P_MASS] = self.dt, c.density * math.pi * c.radius ** 2 * c.thickness
        tp = self.tool_position
        p[P_TOOL_X], p[P_TOOL_Y] = tp
        tv = self.hull_velocity if self.hull_profile is not None else self.tool_velocity
        p[P_TOOL_VX], p[P_TOOL_VY] = tv
        p[P_TIME] = self.time
        return p

    def _hull(self):
        if self.hull_profile is None:
            return None, 0
        verts = np.ascontiguousarray(self.hull_profile + self.tool_position)
        return verts, len(verts)

    def _loads(self, loads):
        if loads is None:
            return None
        a = np.ascontiguousarray(np.asarray(loads, dtype=np.float64))
        assert a.shape == self.positions.shape
        return a

    # -- reference surface -----------------------------------------------------
    def shear_strain(self) -> np.ndarray:
        s = np.zeros(len(self.positions))
        L = self.lib
        L.orc_shear_strain(len(s), _p(self.positions, ctypes.c_double),
                           _p(self.initial_positions, ctypes.c_double), len(self.bi),
                           _p(self.bi, ctypes.c_int32), _p(self.bj, ctypes.c_int32),
                           _p(s, ctypes.c_double))
        return s

    def forces(self, external_forces=None):
        loads = self._loads(external_forces)
        hull, m = self._hull()
        force = np.zeros_like(self.positions)
        reaction = np.zeros(2)
        scal = np.zeros(5)
        d, i32, u8 = ctypes.c_double, ctypes.c_int32, ctypes.c_uint8
        self.lib.orc_plate_forces_thr(
            _p(self._params(), d), len(self.positions), _p(self.positions, d),
            _p(self.velocities, d), _p(self.initial_positions, d), len(self.bi),
            _p(self.bi, i32), _p(self.bj, i32), _p(self.rest, d), _p(self.alive, u8),
            _p(self.failure_mode, u8), _p(self.failure_time_s, d),
            _p(self.failure_extension_m, d), _p(self.failure_energy_J, d), _p(loads, d),
            _p(hull, d), m, _p(force, d), _p(reaction, d), _p(scal, d),
            _p(self.eps_t_bond, d), _p(self.eps_s_bond, d))
        self.acc[1] += scal[0]
        self.acc[2] += scal[1]
        self._damping_power, self._tool_power = scal[2], scal[3]
        return force, reaction

    def step(self, external_forces=None):
        loads = self._loads(external_forces)
        hull, m = self._hull()
        d, i32, u8 = ctypes.c_double, ctypes.c_int32, ctypes.c_uint8
        bad = self.lib.orc_plate_step_thr(
            _p(self._params(), d), len(self.positions), _p(self.positions, d),
            _p(self.velocities, d), _p(self.initial_positions, d), _p(self.fixed, u8),
            len(self.bi), _p(self.bi, i32), _p(self.bj, i32), _p(self.rest, d),
            _p(self.alive, u8), _p(self.failure_mode, u8), _p(self.failure_time_s, d),
            _p(self.failure_extension_m, d), _p(self.failure_energy_J, d), _p(loads, d),
            _p(hull, d), m, _p(self.acc, d), _p(self.last, d), _p(self.force, d),
            _p(self.eps_t_bond, d), _p(self.eps_s_bond, d))
        self.time += self.dt
        self.step_count += 1
        if bad:
            raise FloatingPointError("Non-finite simulation state; reduce dt")

    @property
    def accounting(self) -> dict:
        a = self.acc
        out = {k: float(a[i]) for i, k in enumerate(ACC_KEYS)}
        for t, k in enumerate(IMP_KEYS):
            out[k] = [float(a[5 + 2 * t]), float(a[6 + 2 * t])]
        return out

    @property
    def broken_bonds(self) -> int:
        return int((self.alive == 0).sum())

    def mechanical_energy(self) -> dict:
        hull, m = self._hull()
        out = np.zeros(5)
        d, i32, u8 = ctypes.c_double, ctypes.c_int32, ctypes.c_uint8
        self.lib.orc_mech_energy(_p(self._params(), d), len(self.positions),
                                 _p(self.positions, d), _p(self.velocities, d), len(self.bi),
                                 _p(self.bi, i32), _p(self.bj, i32), _p(self.rest, d),
                                 _p(self.alive, u8), _p(hull, d), m, _p(out, d))
        return dict(kinetic_J=out[0], bond_J=out[1], pair_contact_J=out[2],
                    tool_contact_J=out[3], mechanical_J=out[4])


def pairs_within(x: np.ndarray, cutoff: float) -> np.ndarray:
    """All pairs (i<j) with |x_j - x_i| < cutoff, by brute force (small n only)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = len(x)
    lib = library()
    cap = 16 * n + 16
    while True:
        out = np.zeros((cap, 2), dtype=np.int32)
        cnt = lib.orc_pairs_within(n, _p(x, ctypes.c_double), float(cutoff),
                                   _p(out, ctypes.c_int32), cap)
        if cnt <= cap:
            return out[:cnt].copy()
        cap = cnt


def hull_forces(positions, velocities, vertices, hull_velocity, *, particle_radius,
                hull_radius, stiffness, damping=0.0):
    """The oracle's polyline hull law on its own (hull_contact.py:7-75):
    per-particle forces (N,2) and the hull reaction (2,)."""
    x = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 2)
    v = np.ascontiguousarray(velocities, dtype=np.float64).reshape(-1, 2)
    vert = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 2)
    hv = np.ascontiguousarray(hull_velocity, dtype=np.float64).reshape(2)
    force = np.zeros_like(x)
    reaction = np.zeros(2)
    d = ctypes.c_double
    library().orc_hull_forces(len(x), _p(x, d), _p(v, d), _p(vert, d), len(vert), _p(hv, d),
                              float(particle_radius), float(hull_radius), float(stiffness),
                              float(damping), _p(force, d), _p(reaction, d))
    return force, reaction

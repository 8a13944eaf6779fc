"""FP64 CPU oracle for the Hertz-Mindlin disk bed (test infrastructure only).

`OracleBed` drives `orb_step` of oracle/bed2d.c with numpy state; the law is
documented there.  Only tests, the smoke check and the benchmark's CPU
baseline use this module.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .plate import _p, library

B_ESTAR, B_GSTAR, B_CDAMP, B_MU, B_MU_WALL, B_RHO_PI_T, B_GRAV_Y, B_DT, B_Y_FLOOR, B_X_LEFT, \
    B_USE_FLOOR, B_USE_LEFT, B_X_RIGHT, B_V_RIGHT, B_USE_RIGHT, B_TIME, B_NPARAM = range(17)
BS_DASH, BS_FRIC, BS_WALL_DASH, BS_WALL_FRIC, BS_WALL_FX, BS_WALL_FY, BS_NPAIRS, BS_NWALL, \
    BS_WALL_POWER, BS_NSCAL = range(10)


def _bind():
    lib = library()
    if not hasattr(lib, "_bed_bound"):
        d = ctypes.POINTER(ctypes.c_double)
        i64 = ctypes.POINTER(ctypes.c_int64)
        i32 = ctypes.POINTER(ctypes.c_int32)
        lib.orb_step.argtypes = [d, ctypes.c_int, d, d, d, ctypes.c_int, i64, d, ctypes.c_int,
                                 i64, d, d, d, d, d, d, ctypes.c_int]
        lib.orb_step.restype = ctypes.c_int
        lib.orb_contacts.argtypes = [ctypes.c_int, d, d, ctypes.c_int, i32]
        lib.orb_contacts.restype = ctypes.c_int
        lib._bed_bound = True
    return lib


class OracleBed:
    """config.integrator "euler": semi-implicit Euler.  "verlet": the floe bed's
    kick-drift-kick velocity-Verlet (scale_ceiling/kernel.py:410-421) in the
    device's arrangement, one force evaluation per step: a step evaluates
    the force of the current (x, v) and commits its springs, kicks v by
    (dt/2 F)/m twice -- the second half kick of the previous step, skipped
    while `verlet_start` marks v as a full-step velocity, then the first half
    kick of this step -- and drifts x by dt v.  `velocities` is then the
    half-step velocity; `full_velocity()` gives the reference's `vel`."""

    def __init__(self, config, x, r, dt=None):
        self.config = config
        self.integrator = getattr(config, "integrator", "euler")
        self.verlet_start = True
        self.lib = _bind()
        self.positions = np.ascontiguousarray(x, dtype=np.float64).copy()
        self.radii = np.ascontiguousarray(r, dtype=np.float64).copy()
        self.velocities = np.zeros_like(self.positions)
        n = len(self.radii)
        self.dt = float(dt if dt is not None else (config.dt if config.dt is not None
                                                   else config.stable_dt(self.radii.min())))
        self.hist_keys = np.zeros(0, dtype=np.int64)
        self.hist_xi = np.zeros((0, 2))
        self.xi_floor = np.zeros((n, 2))
        self.xi_left = np.zeros((n, 2))
        self.xi_right = np.zeros((n, 2))
        # right wall: x = right_x + right_v * t (the config's, changeable)
        self.right_wall = bool(config.right_wall)
        self.right_x = float(config.right_wall_x)
        self.right_v = float(config.right_wall_speed)
        self.force = np.zeros((n, 2))
        self.time = 0.0
        self.step_count = 0
        # dash, fric, wall dash, wall fric (J); wall impulse x, y (N s); wall work (J)
        self.acc = np.zeros(7)
        self.last = np.zeros(BS_NSCAL)

    def _params(self) -> np.ndarray:
        c = self.config
        p = np.zeros(B_NPARAM)
        p[B_ESTAR], p[B_GSTAR], p[B_CDAMP] = c.e_star, c.g_star, c.damping_coefficient
        p[B_MU], p[B_MU_WALL], p[B_RHO_PI_T] = c.friction, c.wall_friction, c.rho_pi_t
        p[B_GRAV_Y], p[B_DT] = -c.gravity, self.dt
        p[B_Y_FLOOR], p[B_X_LEFT] = c.floor_y, c.back_wall_x
        p[B_USE_FLOOR], p[B_USE_LEFT] = float(c.floor), float(c.back_wall)
        p[B_X_RIGHT], p[B_V_RIGHT] = self.right_x, self.right_v
        p[B_USE_RIGHT], p[B_TIME] = float(self.right_wall), self.time
        return p

    def set_right_wall(self, *, x=None, speed=None, enabled=None) -> None:
        """Move, re-time or switch the right wall now: it is at x from the
        current time on, moving at `speed`."""
        if speed is not None:
            if x is None:
                x = self.right_x + self.right_v * self.time
            self.right_v = float(speed)
        if x is not None:
            self.right_x = float(x) - self.right_v * self.time
        if enabled is not None:
            self.right_wall = bool(enabled)

    @property
    def right_wall_position(self) -> float:
        return self.right_x + self.right_v * self.time

    def _evaluate(self, integrate: bool, walls=None):
        """orb_step on the current state: the force into self.force, the new
        pair springs and the step scalars returned; wall springs updated in
        the arrays of `walls` (default: the oracle's own)."""
        n = len(self.radii)
        cap = max(16, 4 * n)
        ok = np.zeros(cap, dtype=np.int64)
        ox = np.zeros((cap, 2))
        scal = np.zeros(BS_NSCAL)
        xf, xl, xr = walls if walls is not None else (self.xi_floor, self.xi_left, self.xi_right)
        d, i64 = ctypes.c_double, ctypes.c_int64
        cnt = self.lib.orb_step(_p(self._params(), d), n, _p(self.positions, d),
                                _p(self.velocities, d), _p(self.radii, d), len(self.hist_keys),
                                _p(self.hist_keys, i64), _p(self.hist_xi, d), cap, _p(ok, i64),
                                _p(ox, d), _p(xf, d), _p(xl, d), _p(xr, d), _p(self.force, d),
                                _p(scal, d), int(integrate))
        if cnt == -1:
            raise RuntimeError("oracle contact capacity exceeded")
        if cnt == -2:
            raise FloatingPointError("Non-finite simulation state; reduce dt")
        return ok[:cnt].copy(), ox[:cnt].copy(), scal

    def _verlet_kicks(self):
        """(dt/2 F)/m per disk, in the reference's order of operations."""
        inv_m = 1.0 / self.masses
        return ((0.5 * self.dt) * self.force) * inv_m[:, None]

    def full_velocity(self) -> np.ndarray:
        """The full-step velocity of the current state: for a velocity-Verlet
        state v + (dt/2 F)/m with the force of the current (x, v) (springs
        not committed), else v."""
        if self.integrator != "verlet" or self.verlet_start:
            return self.velocities.copy()
        keep = self.force.copy()
        walls = (self.xi_floor.copy(), self.xi_left.copy(), self.xi_right.copy())
        self._evaluate(False, walls)
        out = self.velocities + self._verlet_kicks()
        self.force[:] = keep
        return out

    def step(self, integrate: bool = True) -> None:
        verlet = self.integrator == "verlet" and integrate
        ok, ox, scal = self._evaluate(integrate and not verlet)
        self.hist_keys = ok
        self.hist_xi = ox
        self.last = scal.copy()
        if verlet:
            kick = self._verlet_kicks()
            u = self.velocities if self.verlet_start else self.velocities + kick
            v = u + kick
            x = self.positions + self.dt * v
            if not (np.isfinite(v).all() and np.isfinite(x).all()):
                raise FloatingPointError("Non-finite simulation state; reduce dt")
            self.velocities[:] = v
            self.positions[:] = x
            self.verlet_start = False
        if integrate:
            dt = self.dt
            self.acc[0] += dt * scal[BS_DASH]
            self.acc[1] += dt * scal[BS_FRIC]
            self.acc[2] += dt * scal[BS_WALL_DASH]
            self.acc[3] += dt * scal[BS_WALL_FRIC]
            self.acc[4] += dt * scal[BS_WALL_FX]
            self.acc[5] += dt * scal[BS_WALL_FY]
            self.acc[6] += dt * scal[BS_WALL_POWER]
            self.time += dt
            self.step_count += 1

    def contacts(self) -> np.ndarray:
        n = len(self.radii)
        cap = 8 * n + 16
        out = np.zeros((cap, 2), dtype=np.int32)
        cnt = self.lib.orb_contacts(n, _p(self.positions, ctypes.c_double),
                                    _p(self.radii, ctypes.c_double), cap,
                                    _p(out, ctypes.c_int32))
        return out[:min(cnt, cap)].copy()

    @property
    def masses(self) -> np.ndarray:
        return self.config.rho_pi_t * (self.radii * self.radii)

    def energies(self) -> dict:
        """Kinetic (of the full-step velocity), gravitational (relative to
        y = 0) and Hertz normal potential (2/5) f_el d of the current contacts."""
        c = self.config
        m = self.masses
        kin = 0.5 * float((m[:, None] * self.full_velocity() ** 2).sum())
        pot = float((m * c.gravity * self.positions[:, 1]).sum())
        el = 0.0
        for i, j in self.contacts():
            dvec = self.positions[j] - self.positions[i]
            d = (self.radii[i] + self.radii[j]) - np.sqrt((dvec ** 2).sum())
            reff = self.radii[i] * self.radii[j] / (self.radii[i] + self.radii[j])
            el += 0.4 * (2.0 / 3.0) * 2 * c.e_star * np.sqrt(reff * d) * d * d
        for gap, use in ((self.positions[:, 1] - c.floor_y, c.floor),
                         (self.positions[:, 0] - c.back_wall_x, c.back_wall),
                         (self.right_wall_position - self.positions[:, 0], self.right_wall)):
            if use:
                d = np.maximum(self.radii - gap, 0.0)
                el += float((0.4 * (2.0 / 3.0) * 2 * c.e_star * np.sqrt(self.radii * d) * d * d).sum())
        return dict(kinetic_J=kin, potential_J=pot, elastic_normal_J=el)

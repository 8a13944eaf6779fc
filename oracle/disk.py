"""FP64 CPU oracle for the 2-D rigid frictional disk assembly (test
infrastructure only).

`OracleDisk` drives `ord_eval` of oracle/disk2d.c with numpy state and
mirrors `DiskSystem` of granular2d/engine.py:144-684: masses and inertias
(engine.py:170-171), the stability bound (engine.py:231-235), the contact
state carried by key between steps, the energy ledger (engine.py:578-588, 630-673)
and the state digest (engine.py:289-295).  Only tests, the smoke check and
the benchmark's CPU baseline use this module.
"""
from __future__ import annotations

import ctypes
import hashlib
import math

import numpy as np

from .plate import _p, library

(OD_KN, OD_ZETA2, OD_KT, OD_MU, OD_MUR, OD_KR_COEF, OD_DT, OD_GX, OD_GY, OD_DRIFT_X,
 OD_DRIFT_Y, OD_DRAG_TIME, OD_DECAY, OD_LX, OD_LY, OD_USE_GRAVITY, OD_USE_DECAY,
 OD_NPARAM) = range(18)
(ODS_DASH, ODS_SLIDE, ODS_ROLL, ODS_EL_N, ODS_EL_T, ODS_EL_R, ODS_WALL_FX, ODS_WALL_FY,
 ODS_HULL_FX, ODS_HULL_FY, ODS_NPAIR, ODS_NWALL, ODS_NHULL, ODS_NSCAL) = range(14)
(ODR_I, ODR_J, ODR_KIND, ODR_NX, ODR_NY, ODR_LX, ODR_LY, ODR_FX, ODR_FY, ODR_FN, ODR_FT,
 ODR_OV, ODR_PX, ODR_PY, ODR_NREC) = range(15)


def _bind():
    lib = library()
    if not hasattr(lib, "_disk_bound"):
        d = ctypes.POINTER(ctypes.c_double)
        i64 = ctypes.POINTER(ctypes.c_int64)
        c_int, c_i64 = ctypes.c_int, ctypes.c_int64
        lib.ord_eval.argtypes = [d, c_i64, d, d, d, d, d, d, d, d, c_int, d, d, c_int, c_i64,
                                 i64, d, d, c_i64, i64, d, d, d, d, c_int, c_int]
        lib.ord_eval.restype = c_i64
        lib._disk_bound = True
    return lib


def _segments(value) -> np.ndarray:
    if value is None:
        return np.zeros((0, 2, 2))
    return np.ascontiguousarray(np.asarray(value, dtype=np.float64).reshape(-1, 2, 2))


class OracleDisk:
    def __init__(self, positions, radii, config, *, walls=None, hull=None, periodic_x=None,
                 periodic_y=None, wall_velocity=None):
        self.config = cfg = config
        self.lib = _bind()
        self.pos = np.array(positions, dtype=np.float64).reshape(-1, 2)
        self.radii = np.array(radii, dtype=np.float64).reshape(-1)
        n = len(self.radii)
        self.vel = np.zeros_like(self.pos)
        self.theta = np.zeros(n)
        self.omega = np.zeros(n)
        self.mass = cfg.density * math.pi * self.radii**2 * cfg.thickness
        self.inertia = 0.5 * self.mass * self.radii**2
        self.walls = _segments(walls)
        self.hull = _segments(hull)
        self.wall_velocity = (None if wall_velocity is None else
                              np.ascontiguousarray(np.asarray(wall_velocity, np.float64)
                                                   .reshape(-1, 2)))
        self.periodic_x = None if periodic_x is None else float(periodic_x)
        self.periodic_y = None if periodic_y is None else float(periodic_y)
        m_min = float(self.mass.min())
        self.recommended_dt = 0.1 / math.sqrt(2.0 * cfg.kn / (0.5 * m_min))
        self.dt = cfg.dt if cfg.dt is not None else self.recommended_dt
        self.time = 0.0
        self.steps = 0
        self.st_key = np.zeros(0, np.int64)
        self.st_xi = np.zeros((0, 2))
        self.st_roll = np.zeros(0)
        self.energy_ledger = None
        self.last = np.zeros(ODS_NSCAL)
        self.wall_force = np.zeros(2)
        self.hull_force = np.zeros(2)

    def _params(self) -> np.ndarray:
        c = self.config
        p = np.zeros(OD_NPARAM)
        k_t = c.kt_ratio * c.kn
        p[OD_KN], p[OD_ZETA2], p[OD_KT] = c.kn, 2.0 * c.zeta, k_t
        p[OD_MU], p[OD_MUR], p[OD_KR_COEF], p[OD_DT] = c.mu, c.mu_rolling, 0.25 * k_t, self.dt
        g = np.asarray(c.gravity, np.float64)
        p[OD_GX], p[OD_GY] = g
        p[OD_DRIFT_X], p[OD_DRIFT_Y] = c.drift_velocity
        p[OD_DRAG_TIME] = c.drag_time
        if c.damping_time > 0.0:
            p[OD_DECAY], p[OD_USE_DECAY] = math.exp(-self.dt / c.damping_time), 1.0
        p[OD_LX] = self.periodic_x or 0.0
        p[OD_LY] = self.periodic_y or 0.0
        p[OD_USE_GRAVITY] = float(bool(np.any(g != 0.0)))
        return p

    def _eval(self, advance: bool, records: bool):
        n = len(self.radii)
        nseg = len(self.walls) + len(self.hull)
        cap = 64 + 8 * n + n * nseg
        nk = np.zeros(cap, np.int64)
        nx = np.zeros((cap, 2))
        nr = np.zeros(cap)
        rec = np.zeros((cap, ODR_NREC)) if records else None
        scal = np.zeros(ODS_NSCAL)
        d, i64 = ctypes.c_double, ctypes.c_int64
        wv = _p(self.wall_velocity, d) if self.wall_velocity is not None else None
        ledger = advance and self.energy_ledger is not None
        m = self.lib.ord_eval(_p(self._params(), d), n, _p(self.pos, d), _p(self.vel, d),
                              _p(self.theta, d), _p(self.omega, d), _p(self.radii, d),
                              _p(self.mass, d), _p(self.inertia, d), _p(self.walls, d),
                              len(self.walls), wv, _p(self.hull, d), len(self.hull),
                              len(self.st_key), _p(self.st_key, i64), _p(self.st_xi, d),
                              _p(self.st_roll, d), cap, _p(nk, i64), _p(nx, d), _p(nr, d),
                              _p(rec, d) if records else None, _p(scal, d), int(advance),
                              int(ledger))
        if m < 0:
            raise RuntimeError("oracle contact capacity exceeded")
        self.last = scal
        self.wall_force = scal[[ODS_WALL_FX, ODS_WALL_FY]].copy()
        self.hull_force = scal[[ODS_HULL_FX, ODS_HULL_FY]].copy()
        if advance:
            self.st_key, self.st_xi, self.st_roll = nk[:m].copy(), nx[:m].copy(), nr[:m].copy()
            if ledger:
                led = self.energy_ledger
                led["normal_dashpot"] += float(scal[ODS_DASH])
                led["friction"] -= float(scal[ODS_SLIDE])
                led["rolling"] -= float(scal[ODS_ROLL])
        return rec[:m] if records else None

    def step(self, n: int = 1) -> None:
        """n steps; `contacts` keeps the records of the last step's evaluation."""
        for k in range(n):
            r = self._eval(True, k == n - 1)
            self.time += self.dt
            self.steps += 1
        if n > 0:
            self.contacts = self._records(r)

    def contact_forces(self) -> dict:
        self.contacts = self._records(self._eval(False, True))
        return self.contacts

    @staticmethod
    def _records(r: np.ndarray) -> dict:
        f = r[:, [ODR_FX, ODR_FY]]
        return {
            "i": r[:, ODR_I].astype(np.int64), "j": r[:, ODR_J].astype(np.int64),
            "kind": r[:, ODR_KIND].astype(np.int8), "n": r[:, [ODR_NX, ODR_NY]],
            "lever": r[:, [ODR_LX, ODR_LY]], "f": f, "f_on_i": f, "f_on_j": -f,
            "fn": r[:, ODR_FN], "ft": r[:, ODR_FT], "overlap": r[:, ODR_OV],
            "point": r[:, [ODR_PX, ODR_PY]],
        }

    def kinetic_energy(self) -> float:
        return float(np.sum(0.5 * self.mass * np.sum(self.vel**2, axis=1))
                     + np.sum(0.5 * self.inertia * self.omega**2))

    def momentum(self) -> np.ndarray:
        return (self.mass[:, None] * self.vel).sum(axis=0)

    def enable_energy_ledger(self) -> dict:
        self.energy_ledger = {"normal_dashpot": 0.0, "friction": 0.0, "rolling": 0.0,
                              "kinetic_initial": self.kinetic_energy()}
        return self.energy_ledger

    def energies(self) -> dict:
        g = np.asarray(self.config.gravity, dtype=np.float64)
        pe = -float(np.sum(self.mass * (self.pos @ g))) if np.any(g != 0.0) else 0.0
        return {"kinetic": self.kinetic_energy(), "potential": pe,
                "elastic_normal": float(self.last[ODS_EL_N]),
                "elastic_tangential": float(self.last[ODS_EL_T]),
                "elastic_rolling": float(self.last[ODS_EL_R])}

    def dissipated_work(self) -> float:
        return -float(sum(self.energy_ledger[k] for k in ("normal_dashpot", "friction",
                                                          "rolling")))

    def digest(self) -> str:
        h = hashlib.sha256()
        for arr in (self.pos, self.vel, self.theta, self.omega):
            h.update(np.ascontiguousarray(arr, dtype=np.float64).tobytes())
        h.update(np.float64(self.time).tobytes())
        return h.hexdigest()

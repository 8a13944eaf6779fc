"""Rigid frictional disks with rotation on the device: the reference's
`tensordem.granular2d` engine (src/tensordem/granular2d/engine.py:1-684).

`DiskConfig`, `DiskSystem`, `box_walls`, `polyline_segments` and
`ship_segments` keep the reference's names, arguments, defaults, validation
and exceptions.  `DiskSystem.step` advances the disks with the HIP kernels of
tensordem_amd/csrc/disk.hip (two launches per step, no host work in between);
trajectories match the reference bit for bit in FP64.

Differences a caller can see:

* `pos`, `vel`, `theta` and `omega` are float64 tensors on the ROCm device
  (`device="cuda"`), not numpy arrays; in-place edits and assignment work as
  in the reference.  `pos`, `vel` and `omega` are double-buffered: the tensor
  objects alternate between steps, so read them from the system after a step
  instead of keeping an old reference.
* `contact_forces()` and `contacts` return numpy arrays, as the reference
  does.  `contacts` after `step()` holds the records of that call's last step
  once per-contact records are enabled (by a `contact_forces()` call or the
  first read of `contacts`); before that, the first read evaluates the
  current state.
* A disk has at most `cand_slots` (24) candidate partners within
  2*max(r) + skin and a grid cell holds at most `cell_cap` disks (by default
  the packing bound of the smallest disks in one cell, at least 16); beyond
  that `step` raises RuntimeError instead of computing a wrong answer.
"""
from __future__ import annotations

import ctypes
import hashlib
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib

__all__ = ["DiskConfig", "DiskSystem", "box_walls", "polyline_segments", "ship_segments"]


@dataclass(frozen=True)
class DiskConfig:
    """Material and contact parameters (SI); engine.py:51-82."""

    kn: float = 3.0e4
    kt_ratio: float = 0.5
    zeta: float = 0.2
    mu: float = 0.5
    mu_rolling: float = 0.05
    density: float = 917.0
    thickness: float = 0.03
    gravity: tuple[float, float] = (0.0, 0.0)
    drift_velocity: tuple[float, float] = (0.0, 0.0)
    drag_time: float = 0.0
    damping_time: float = 0.0
    dt: float | None = None
    skin: float = 0.0035

    def __post_init__(self) -> None:
        for name in ("kn", "kt_ratio", "density", "thickness"):
            value = float(getattr(self, name))
            if not math.isfinite(value) or value <= 0.0:
                raise ValueError(f"{name} must be finite and positive")
        for name in ("zeta", "mu", "mu_rolling", "drag_time", "damping_time", "skin"):
            value = float(getattr(self, name))
            if not math.isfinite(value) or value < 0.0:
                raise ValueError(f"{name} must be finite and nonnegative")
        if self.dt is not None and (not math.isfinite(self.dt) or self.dt <= 0.0):
            raise ValueError("dt must be finite and positive")
        if len(tuple(self.gravity)) != 2 or len(tuple(self.drift_velocity)) != 2:
            raise ValueError("gravity and drift_velocity must have two components")
        if not all(math.isfinite(v) for v in tuple(self.gravity) + tuple(self.drift_velocity)):
            raise ValueError("gravity and drift_velocity must be finite")


def polyline_segments(points) -> np.ndarray:
    """(S, 2, 2) segments of the closed polyline through `points` (M >= 3)."""
    pts = np.asarray(points, dtype=np.float64)
    if pts.ndim != 2 or pts.shape[1] != 2 or pts.shape[0] < 3:
        raise ValueError("points must have shape (M, 2) with M >= 3")
    if not np.isfinite(pts).all():
        raise ValueError("points must be finite")
    nxt = np.concatenate([pts[1:], pts[:1]])
    return np.stack([pts, nxt], axis=1)


def box_walls(width: float, height: float, *, left: bool = True, right: bool = True,
              bottom: bool = True, top: bool = True, margin: float = 0.0) -> np.ndarray:
    """Axis-aligned walls of a width x height box as (S, 2, 2) segments, in the
    order bottom, top, left, right; `margin` extends each past the corners."""
    for name, value in (("width", width), ("height", height)):
        if not math.isfinite(value) or value <= 0.0:
            raise ValueError(f"{name} must be finite and positive")
    w, h, e = float(width), float(height), float(margin)
    table = ((bottom, ((-e, 0.0), (w + e, 0.0))), (top, ((-e, h), (w + e, h))),
             (left, ((0.0, -e), (0.0, h + e))), (right, ((w, -e), (w, h + e))))
    segs = [s for keep, s in table if keep]
    return np.asarray(segs, dtype=np.float64).reshape(-1, 2, 2)


def ship_segments(length: float, beam: float, origin=(0.0, 0.0), *, bow: float = 0.0,
                  stern: float = 0.0, n_bow: int = 6) -> np.ndarray:
    """Closed hull polyline in the x-y plane, bow towards +x (engine.py:119-141):
    flat sides of half-beam, a transom whose corners are set in by `stern`, and
    a bow of length `bow` tapering linearly to the stem in n_bow pieces."""
    for name, value in (("length", length), ("beam", beam)):
        if not math.isfinite(value) or value <= 0.0:
            raise ValueError(f"{name} must be finite and positive")
    if not (0.0 <= bow < length) or not (0.0 <= stern < length):
        raise ValueError("bow/stern must be smaller than the hull length")
    x0, y0 = float(origin[0]), float(origin[1])
    half = 0.5 * beam
    xs = x0 + max(stern, 0.0)
    xb = x0 + length - bow
    starboard = [(x0, y0), (xs, y0 - half), (xb, y0 - half)]
    if bow > 0.0:
        starboard += [(xb + (k / n_bow) * bow, y0 - half * (1.0 - k / n_bow))
                      for k in range(1, n_bow)]
    stem = [(x0 + length, y0)]
    port = []
    if bow > 0.0:
        port += [(xb + (k / n_bow) * bow, y0 + half * (1.0 - k / n_bow))
                 for k in range(n_bow - 1, 0, -1)]
    port += [(xb, y0 + half), (xs, y0 + half)]
    return polyline_segments(np.asarray(starboard + stem + port, dtype=np.float64))


def jittered_box(side: int, radius: float = 0.01, seed: int = 0, speed: float = 0.05):
    """side x side disks of radius U[0.9, 1.1] x radius on a square grid of
    pitch 2.1 radius, each jittered inside its cell (the largest neighbours
    overlap slightly), with seeded Gaussian velocities of `speed` rms per
    axis, and the four walls of the side*pitch square box.  Returns
    (positions, radii, velocities, walls)."""
    n = side * side
    rng = np.random.default_rng(seed)
    r = radius * rng.uniform(0.9, 1.1, size=n)
    pitch = 2.1 * radius
    idx = np.arange(n)
    slack = np.maximum(0.5 * pitch - r, 0.0)[:, None]
    pos = (np.stack([idx % side, idx // side], 1) + 0.5) * pitch \
        + rng.uniform(-1, 1, size=(n, 2)) * slack
    vel = rng.normal(scale=speed, size=(n, 2))
    return pos, r, vel, box_walls(side * pitch, side * pitch)


def _as_segments(value, name: str) -> np.ndarray:
    if value is None:
        return np.zeros((0, 2, 2))
    arr = np.asarray(value, dtype=np.float64)
    if arr.ndim != 3 or arr.shape[1:] != (2, 2):
        raise ValueError(f"{name} must have shape (S, 2, 2)")
    if not np.isfinite(arr).all():
        raise ValueError(f"{name} must be finite")
    return arr


class DiskParams(ctypes.Structure):
    """Mirror of `tdm_disk_params`."""
    _fields_ = [
        ("kn", ctypes.c_double), ("zeta2", ctypes.c_double), ("kt", ctypes.c_double),
        ("mu", ctypes.c_double), ("mu_rolling", ctypes.c_double), ("kr_coef", ctypes.c_double),
        ("dt", ctypes.c_double), ("gravity", ctypes.c_double * 2),
        ("drift", ctypes.c_double * 2), ("drag_time", ctypes.c_double),
        ("decay", ctypes.c_double), ("period", ctypes.c_double * 2), ("reach", ctypes.c_double),
        ("skin", ctypes.c_double), ("cell_h", ctypes.c_double * 2),
        ("grid_origin", ctypes.c_double * 2), ("ncell", ctypes.c_int32 * 2),
        ("use_gravity", ctypes.c_int32), ("use_decay", ctypes.c_int32), ("n", ctypes.c_int32),
        ("n_walls", ctypes.c_int32), ("n_hull", ctypes.c_int32),
        ("cand_slots", ctypes.c_int32), ("cell_cap", ctypes.c_int32),
        ("nblocks", ctypes.c_int32), ("seg_box", (ctypes.c_double * 4) * 8),
    ]


class DiskArrays(ctypes.Structure):
    """Mirror of `tdm_disk_arrays`."""
    _fields_ = [(name, ctypes.c_void_p) for name in (
        "x", "v", "omega", "x_alt", "v_alt", "omega_alt", "theta", "radius", "mass", "inertia",
        "walls", "hull", "wall_velocity", "out", "records", "workspace")] + [
        ("workspace_bytes", ctypes.c_size_t), ("launches", ctypes.c_int64),
        ("steps", ctypes.c_int64)]


NOUT, NREC = 16, 14
OUT_DASH, OUT_FRIC, OUT_ROLL, OUT_ELASTIC, OUT_WALL, OUT_HULL, OUT_COUNTS = 0, 1, 2, 3, 6, 8, 10
OUT_STATUS, OUT_REBUILDS, OUT_CANDIDATES = 13, 14, 15
FLAG_LEDGER, FLAG_RECORD = 1, 2
LEDGER_KEYS = ("normal_dashpot", "friction", "rolling")


class _Ledger(dict):
    """The energy ledger dict (engine.py:630-653); the three work entries are
    read from the device whenever they are looked up."""

    def __init__(self, system: "DiskSystem", kinetic_initial: float):
        super().__init__(normal_dashpot=0.0, friction=0.0, rolling=0.0,
                         kinetic_initial=kinetic_initial)
        self._system = system

    def _sync(self) -> None:
        if self._system.energy_ledger is self:
            vals = self._system.out[:3].tolist()
            for k, v in zip(LEDGER_KEYS, vals):
                dict.__setitem__(self, k, v)

    def __getitem__(self, key):
        self._sync()
        return dict.__getitem__(self, key)

    def get(self, key, default=None):
        self._sync()
        return dict.get(self, key, default)

    def items(self):
        self._sync()
        return dict.items(self)

    def values(self):
        self._sync()
        return dict.values(self)

    def __repr__(self) -> str:
        self._sync()
        return dict.__repr__(self)


class DiskSystem:
    """A deterministic 2D rigid-disk assembly with optional static boundaries
    (engine.py:144-684), stepped on the device."""

    def __init__(self, positions, radii, config: DiskConfig | None = None, *,
                 walls=None, hull=None, periodic_x: float | None = None,
                 periodic_y: float | None = None, wall_velocity=None, device="cuda",
                 cand_slots: int = 24, cell_cap: int | None = None) -> None:
        self.config = cfg = config or DiskConfig()
        pos = np.array(positions, dtype=np.float64)
        rad = np.array(radii, dtype=np.float64).reshape(-1)
        if pos.ndim != 2 or pos.shape[1] != 2:
            raise ValueError("positions must have shape (N, 2)")
        if pos.shape[0] != rad.shape[0]:
            raise ValueError("positions and radii must have the same length")
        if pos.shape[0] < 1:
            raise ValueError("at least one disk is required")
        if not np.isfinite(pos).all() or not np.isfinite(rad).all():
            raise ValueError("positions and radii must be finite")
        if (rad <= 0.0).any():
            raise ValueError("radii must be positive")
        mass = cfg.density * math.pi * rad**2 * cfg.thickness
        inertia = 0.5 * mass * rad**2
        self.walls = _as_segments(walls, "walls")
        self.hull = _as_segments(hull, "hull")
        if wall_velocity is None:
            wv = None
        else:
            wv = np.array(wall_velocity, dtype=np.float64).reshape(-1, 2)
            if wv.shape != (len(self.walls), 2):
                raise ValueError("wall_velocity must have shape (n_walls, 2)")
            if not np.isfinite(wv).all():
                raise ValueError("wall_velocity must be finite")
        self.periodic_x = None if periodic_x is None else float(periodic_x)
        self.periodic_y = None if periodic_y is None else float(periodic_y)
        for name, value in (("periodic_x", self.periodic_x), ("periodic_y", self.periodic_y)):
            if value is not None and (not math.isfinite(value) or value <= 0.0):
                raise ValueError(f"{name} must be a positive period or None")
        if len(self.walls) > 255 or len(self.hull) > 255:
            raise ValueError("at most 255 wall and 255 hull segments")
        self._mass_np = mass
        self.time = 0.0
        self.steps = 0
        self.dt = cfg.dt if cfg.dt is not None else self.recommended_dt
        if self.dt > self.recommended_dt * (1.0 + 1e-12):
            raise ValueError(f"dt exceeds the stability limit {self.recommended_dt:.6g} s")

        self.lib = _lib.load()
        self.device = torch.device(device)
        if self.device.type != "cuda" or not torch.cuda.is_available():
            raise ValueError("DiskSystem runs on the ROCm device (device='cuda')")
        f64 = dict(dtype=torch.float64, device=self.device)
        n = len(rad)
        self._x = torch.tensor(pos, **f64)
        self._v = torch.zeros((n, 2), **f64)
        self._om = torch.zeros(n, **f64)
        self._x_alt = torch.empty_like(self._x)
        self._v_alt = torch.empty_like(self._v)
        self._om_alt = torch.empty_like(self._om)
        self._theta = torch.zeros(n, **f64)
        self.radii = torch.tensor(rad, **f64)
        self.mass = torch.tensor(mass, **f64)
        self.inertia = torch.tensor(inertia, **f64)
        self._walls_d = torch.tensor(self.walls.reshape(-1, 4), **f64) if len(self.walls) else None
        self._hull_d = torch.tensor(self.hull.reshape(-1, 4), **f64) if len(self.hull) else None
        self._wall_velocity = wv
        self._wv_d = torch.tensor(wv, **f64) if wv is not None else None
        # what the device holds, to notice in-place edits of walls, hull and
        # wall velocities between calls (the reference reads them afresh at
        # every evaluation, engine.py:352-375, 411-412)
        self._uploaded = self._segment_state()
        self.out = torch.zeros(NOUT, **f64)
        self.records = None
        self._records_on = False
        self._records_valid = False  # records describe the last evaluation
        self._contacts = None
        self.energy_ledger: dict | None = None
        self.hull_force_history: list = []
        self.contact_history: list = []

        p = self._params = DiskParams()
        k_t = cfg.kt_ratio * cfg.kn
        p.kn, p.zeta2, p.kt = cfg.kn, 2.0 * cfg.zeta, k_t
        p.mu, p.mu_rolling, p.kr_coef, p.dt = cfg.mu, cfg.mu_rolling, 0.25 * k_t, self.dt
        g = np.asarray(cfg.gravity, dtype=np.float64)
        p.gravity[0], p.gravity[1] = g
        p.drift[0], p.drift[1] = (float(c) for c in cfg.drift_velocity)
        p.drag_time = cfg.drag_time
        p.use_gravity = int(bool(np.any(g != 0.0)))
        if cfg.damping_time > 0.0:
            p.decay, p.use_decay = math.exp(-self.dt / cfg.damping_time), 1
        p.skin = cfg.skin
        p.reach = 2.0 * float(rad.max()) + cfg.skin
        self._grid(pos)
        p.n, p.n_walls, p.n_hull = n, len(self.walls), len(self.hull)
        if cell_cap is None:
            # the packing bound of the smallest disks in one cell (at least 16)
            from .bed import cell_capacity
            cell_cap = cell_capacity(p.cell_h[0], float(rad.min()), floor=16,
                                     cell_size_y=p.cell_h[1])
        p.cand_slots, p.cell_cap = int(cand_slots), int(cell_cap)
        p.nblocks = self.lib.tdm_disk_blocks(n)
        wsb = self.lib.tdm_disk_workspace_bytes(n, p.ncell[0] * p.ncell[1], p.cell_cap,
                                                p.cand_slots, p.n_walls, p.n_hull)
        self.workspace = torch.zeros(int(wsb) + 256, dtype=torch.uint8, device=self.device)
        self._arrays = DiskArrays()
        self._refresh()
        _lib.check(self.lib.tdm_disk_setup(ctypes.byref(self._arrays), ctypes.byref(p),
                                           self._stream))
        self._xver = self._x._version
        self._graphs = {}

    # ---- set-up helpers --------------------------------------------------------
    def _grid(self, pos: np.ndarray) -> None:
        """Cell grid of edge >= reach: the period on a periodic axis, otherwise
        the box around the disks with room to move (segments are not binned;
        disks that leave the box are clamped into its edge cells, which stays
        correct).  At most a few cells per disk in total."""
        p = self._params
        lo, hi = pos.min(0), pos.max(0)
        span = float(max(hi - lo)) + p.reach
        n = len(pos)
        h0 = p.reach * (1.0 + 1e-9)
        ext, org = [0.0, 0.0], [0.0, 0.0]
        for a, period in enumerate((self.periodic_x, self.periodic_y)):
            p.period[a] = period or 0.0
            if period is not None:
                ext[a], org[a] = period, 0.0
            else:
                org[a] = float(lo[a]) - 0.25 * span - 2.0 * h0
                ext[a] = float(hi[a]) + 0.25 * span + 2.0 * h0 - org[a]
        cap = max(1 << 12, 4 * n)
        scale = 1.0
        while True:
            nc = [max(1, int(math.floor(ext[a] / (h0 * scale)))) if p.period[a] > 0 else
                  int(math.ceil(ext[a] / (h0 * scale))) + 1 for a in range(2)]
            if nc[0] * nc[1] <= cap:
                break
            scale *= 1.25
        for a in range(2):
            if p.period[a] > 0:
                p.ncell[a], p.cell_h[a], p.grid_origin[a] = nc[a], ext[a] / nc[a], 0.0
            else:
                p.ncell[a], p.cell_h[a], p.grid_origin[a] = nc[a], h0 * scale, org[a]

    @property
    def _stream(self) -> int:
        return torch.cuda.current_stream(self.device).cuda_stream

    def _refresh(self) -> None:
        a = self._arrays
        a.x, a.v, a.omega = self._x.data_ptr(), self._v.data_ptr(), self._om.data_ptr()
        a.x_alt, a.v_alt = self._x_alt.data_ptr(), self._v_alt.data_ptr()
        a.omega_alt, a.theta = self._om_alt.data_ptr(), self._theta.data_ptr()
        a.radius, a.mass = self.radii.data_ptr(), self.mass.data_ptr()
        a.inertia = self.inertia.data_ptr()
        a.walls = self._walls_d.data_ptr() if self._walls_d is not None else None
        a.hull = self._hull_d.data_ptr() if self._hull_d is not None else None
        a.wall_velocity = self._wv_d.data_ptr() if self._wv_d is not None else None
        a.out = self.out.data_ptr()
        a.records = self.records.data_ptr() if self.records is not None else None
        base = self.workspace.data_ptr()
        a.workspace = (base + 255) & ~255
        a.workspace_bytes = self.workspace.numel() - (a.workspace - base)

    def _swap(self) -> None:
        self._x, self._x_alt = self._x_alt, self._x
        self._v, self._v_alt = self._v_alt, self._v
        self._om, self._om_alt = self._om_alt, self._om
        self._refresh()
        self._xver = self._x._version

    def _segment_state(self):
        wv = self._wall_velocity
        return (self.walls.copy(), self.hull.copy(), None if wv is None else wv.copy())

    @property
    def wall_velocity(self):
        """(n_walls, 2) velocities of the wall segments, or None; assign it or
        edit it in place between steps."""
        return self._wall_velocity

    @wall_velocity.setter
    def wall_velocity(self, value) -> None:
        if value is None:
            self._wall_velocity = None
            return
        wv = np.array(value, dtype=np.float64).reshape(-1, 2)
        if wv.shape != (len(self.walls), 2):
            raise ValueError("wall_velocity must have shape (n_walls, 2)")
        if not np.isfinite(wv).all():
            raise ValueError("wall_velocity must be finite")
        self._wall_velocity = wv

    def _sync_segments(self) -> None:
        """Upload walls, hull and wall velocities the caller changed since the
        last call, refresh the segment boxes and drop captured graphs (their
        launch parameters hold the old boxes)."""
        walls, hull, wv = self._uploaded
        cur_wv = self._wall_velocity
        same_wv = (wv is None and cur_wv is None) or (
            wv is not None and cur_wv is not None and np.array_equal(wv, cur_wv))
        if np.array_equal(walls, self.walls) and np.array_equal(hull, self.hull) and same_wv:
            return
        new_walls = _as_segments(self.walls, "walls")
        new_hull = _as_segments(self.hull, "hull")
        if len(new_walls) != len(walls) or len(new_hull) != len(hull):
            raise ValueError("the number of wall and hull segments is fixed at construction; "
                             "move segments in place instead")
        if not np.isfinite(new_walls).all() or not np.isfinite(new_hull).all():
            raise ValueError("walls and hull must be finite")
        if cur_wv is not None:
            cur_wv = np.asarray(cur_wv, dtype=np.float64).reshape(-1, 2)
            if cur_wv.shape != (len(new_walls), 2) or not np.isfinite(cur_wv).all():
                raise ValueError("wall_velocity must have shape (n_walls, 2) and be finite")
        f64 = dict(dtype=torch.float64, device=self.device)
        if self._walls_d is not None:
            self._walls_d.copy_(torch.as_tensor(new_walls.reshape(-1, 4)))
        if self._hull_d is not None:
            self._hull_d.copy_(torch.as_tensor(new_hull.reshape(-1, 4)))
        if cur_wv is None:
            self._wv_d = None
        elif self._wv_d is None:
            self._wv_d = torch.tensor(cur_wv, **f64)
        else:
            self._wv_d.copy_(torch.as_tensor(cur_wv))
        self._refresh()
        _lib.check(self.lib.tdm_disk_segments(ctypes.byref(self._arrays),
                                              ctypes.byref(self._params), self._stream))
        self.close()
        self._uploaded = self._segment_state()

    def _sync_edits(self) -> None:
        """A caller that moved disks gets a fresh candidate search; moved or
        re-timed segments are uploaded."""
        self._sync_segments()
        if self._x._version != self._xver:
            _lib.check(self.lib.tdm_disk_rebin(ctypes.byref(self._arrays),
                                               ctypes.byref(self._params), self._stream))
            self._xver = self._x._version

    def _check_status(self) -> None:
        st = int(self.out[OUT_STATUS].item())
        if st & _lib.ST_CAND_OVERFLOW:
            raise RuntimeError("a disk has more candidate partners than cand_slots; "
                               "construct the system with a larger cand_slots")
        if st & _lib.ST_CELL_OVERFLOW:
            raise RuntimeError("a grid cell holds more disks than cell_cap; construct the "
                               "system with a larger cell_cap")

    def _ensure_records(self) -> None:
        if self.records is None:
            p = self._params
            slots = p.cand_slots + p.n_walls + p.n_hull
            self.records = torch.zeros((slots, p.n, NREC), dtype=torch.float64,
                                       device=self.device)
            self._refresh()
        self._records_on = True

    # ---- state -----------------------------------------------------------------
    @property
    def pos(self) -> torch.Tensor:
        return self._x

    @pos.setter
    def pos(self, value) -> None:
        self._x.copy_(torch.as_tensor(np.asarray(value, np.float64) if not
                                      isinstance(value, torch.Tensor) else value)
                      .reshape(self._x.shape))

    @property
    def vel(self) -> torch.Tensor:
        return self._v

    @vel.setter
    def vel(self, value) -> None:
        self._v.copy_(torch.as_tensor(np.asarray(value, np.float64) if not
                                      isinstance(value, torch.Tensor) else value)
                      .reshape(self._v.shape))

    @property
    def omega(self) -> torch.Tensor:
        return self._om

    @omega.setter
    def omega(self, value) -> None:
        self._om.copy_(torch.as_tensor(np.asarray(value, np.float64) if not
                                       isinstance(value, torch.Tensor) else value)
                       .reshape(self._om.shape))

    @property
    def theta(self) -> torch.Tensor:
        return self._theta

    @theta.setter
    def theta(self, value) -> None:
        self._theta.copy_(torch.as_tensor(np.asarray(value, np.float64) if not
                                          isinstance(value, torch.Tensor) else value)
                          .reshape(self._theta.shape))

    @property
    def n(self) -> int:
        return int(self._params.n)

    @property
    def mean_radius(self) -> float:
        return float(self.radii.mean().item())

    @property
    def recommended_dt(self) -> float:
        m_min = float(self._mass_np.min())
        omega = math.sqrt(2.0 * self.config.kn / (0.5 * m_min))
        return 0.1 / omega

    @property
    def periodic(self) -> tuple[float | None, float | None]:
        return (self.periodic_x, self.periodic_y)

    @property
    def n_periodic_axes(self) -> int:
        return int(self.periodic_x is not None) + int(self.periodic_y is not None)

    def domain_area(self) -> float | None:
        if self.periodic_x is not None and self.periodic_y is not None:
            return float(self.periodic_x * self.periodic_y)
        if len(self.walls) == 0:
            return None
        pts = self.walls.reshape(-1, 2)
        return float(np.prod(pts.max(axis=0) - pts.min(axis=0)))

    def area_fraction(self, area: float | None = None) -> float:
        disk_area = float(torch.sum(math.pi * self.radii**2).item())
        if area is None:
            area = self.domain_area()
        if area is None or area <= 0.0:
            raise ValueError("an explicit positive area is required")
        return disk_area / area

    def set_uniform_velocity(self, velocity) -> None:
        v = torch.as_tensor(np.asarray(velocity, dtype=np.float64).reshape(1, 2),
                            device=self.device)
        self._v[:] = v

    def zero_velocities(self) -> None:
        self._v.zero_()
        self._om.zero_()

    def momentum(self) -> np.ndarray:
        return (self.mass[:, None] * self._v).sum(dim=0).cpu().numpy()

    def kinetic_energy(self) -> float:
        ke = torch.sum(0.5 * self.mass * torch.sum(self._v**2, dim=1)) \
            + torch.sum(0.5 * self.inertia * self._om**2)
        return float(ke.item())

    def mean_speed(self) -> float:
        return float(torch.mean(torch.linalg.vector_norm(self._v, dim=1)).item())

    def digest(self) -> str:
        """SHA-256 of the state bytes and the clock (engine.py:289-295)."""
        h = hashlib.sha256()
        for t in (self._x, self._v, self._theta, self._om):
            h.update(np.ascontiguousarray(t.cpu().numpy(), dtype=np.float64).tobytes())
        h.update(np.float64(self.time).tobytes())
        return h.hexdigest()

    @property
    def wall_force(self) -> np.ndarray:
        return self.out[OUT_WALL:OUT_WALL + 2].cpu().numpy()

    @property
    def hull_force(self) -> np.ndarray:
        return self.out[OUT_HULL:OUT_HULL + 2].cpu().numpy()

    @property
    def rebuilds(self) -> int:
        return int(self.out[OUT_REBUILDS].item())

    def list_stats(self) -> dict:
        """Candidates, pair, wall and hull contacts per disk at the last
        evaluation (pair contacts counted once)."""
        o = self.out.tolist()
        n = self.n
        return dict(cand_mean=o[OUT_CANDIDATES] / n, pairs_per_disk=o[OUT_COUNTS] / n,
                    walls_per_disk=o[OUT_COUNTS + 1] / n, hull_per_disk=o[OUT_COUNTS + 2] / n)

    # ---- stepping ----------------------------------------------------------------
    def step(self, n: int = 1, *, record: bool = False) -> None:
        """Advance n steps with semi-implicit Euler (engine.py:599-627)."""
        if n < 0:
            raise ValueError("n must be nonnegative")
        if n == 0:
            return
        if record:
            for _ in range(n):
                self._advance(1)
                wf, hf = self.wall_force, self.hull_force
                self.contact_history.append({
                    "time": self.time, "hull_fx": float(hf[0]), "hull_fy": float(hf[1]),
                    "wall_fx": float(wf[0]), "wall_fy": float(wf[1]),
                    "kinetic_energy": self.kinetic_energy()})
            return
        self._advance(n)

    def _advance(self, n: int) -> None:
        self._sync_edits()
        flags = (FLAG_LEDGER if self.energy_ledger is not None else 0) | \
            (FLAG_RECORD if self._records_on else 0)
        _lib.check(self.lib.tdm_disk_step(ctypes.byref(self._arrays), ctypes.byref(self._params),
                                          int(n), flags, self._stream))
        for _ in range(n):
            self.time += self.dt
        self.steps += n
        if n % 2:
            self._swap()
        self._contacts = None
        self._records_valid = self._records_on
        self._check_status()

    def graph(self, pairs: int):
        """2*pairs steps captured as a graph (state ends in the same buffers)."""
        a = self._arrays
        key = (pairs, self.energy_ledger is not None, self._x.data_ptr(), a.launches & 1,
               a.steps & 1)
        if key not in self._graphs:
            self._sync_edits()
            g = ctypes.c_void_p()
            flags = FLAG_LEDGER if self.energy_ledger is not None else 0
            stream = torch.cuda.Stream(self.device)
            stream.wait_stream(torch.cuda.current_stream(self.device))
            with torch.cuda.stream(stream):
                _lib.check(self.lib.tdm_disk_graph_create(
                    ctypes.byref(self._arrays), ctypes.byref(self._params), int(pairs), flags,
                    stream.cuda_stream, ctypes.byref(g)))
            torch.cuda.current_stream(self.device).wait_stream(stream)
            self._graphs[key] = g
        return self._graphs[key]

    def run_graph(self, pairs: int, repeats: int = 1) -> None:
        self._sync_edits()
        g = self.graph(pairs)
        for _ in range(repeats):
            _lib.check(self.lib.tdm_graph_launch(g, self._stream))
            # the library's launch and step counts, which the graph baked in
            self._arrays.launches += 2 * pairs
            self._arrays.steps += 2 * pairs
        for _ in range(2 * pairs * repeats):
            self.time += self.dt
        self.steps += 2 * pairs * repeats
        self._contacts = None
        self._records_valid = False  # captured steps write no records
        self._check_status()  # the replays' flush wrote the status word

    def step_timed(self, nsteps: int) -> tuple[float, float]:
        """nsteps (even) steps with each kernel timed by HIP events; returns the
        mean step-kernel and prep-kernel durations in microseconds."""
        self._sync_edits()
        a, b = ctypes.c_double(), ctypes.c_double()
        _lib.check(self.lib.tdm_disk_step_timed(ctypes.byref(self._arrays),
                                                ctypes.byref(self._params), int(nsteps),
                                                self._stream, ctypes.byref(a), ctypes.byref(b)))
        for _ in range(nsteps):
            self.time += self.dt
        self.steps += nsteps
        self._contacts = None
        self._records_valid = False
        return a.value, b.value

    # ---- contacts, energies --------------------------------------------------------
    def _compact_records(self) -> dict:
        p = self._params
        S = p.cand_slots
        rec = self.records
        parts = [rec[:S], rec[S:S + p.n_walls], rec[S + p.n_walls:]]
        rows = []
        for r in parts:
            r = r.permute(1, 0, 2).reshape(-1, NREC)   # disk-major
            rows.append(r[r[:, 2] >= 0.0])
        r = torch.cat(rows).cpu().numpy()
        f = r[:, 7:9].copy()
        return {
            "i": r[:, 0].astype(np.int64), "j": r[:, 1].astype(np.int64),
            "kind": r[:, 2].astype(np.int8), "n": r[:, 3:5].copy(), "lever": r[:, 5:7].copy(),
            "f": f, "f_on_i": f, "f_on_j": -f, "fn": r[:, 9].copy(), "ft": r[:, 10].copy(),
            "overlap": r[:, 11].copy(), "point": r[:, 12:14].copy(),
        }

    def contact_forces(self) -> dict:
        """Evaluate the contacts at the current state without integrating or
        advancing any spring (engine.py:675-678); per-contact records."""
        self._ensure_records()
        self._sync_edits()
        _lib.check(self.lib.tdm_disk_inspect(ctypes.byref(self._arrays),
                                             ctypes.byref(self._params), self._stream))
        self._check_status()
        self._records_valid = True
        self._contacts = self._compact_records()
        return self._contacts

    @property
    def contacts(self) -> dict:
        """Records of the last evaluation (the last step of the last `step`
        call).  Where none were written (before records were first requested,
        after graph replays) they describe the current state, evaluated without
        touching the energies, wall and hull forces or counts of the last step;
        from then on every step writes records, as the reference's does."""
        if self._contacts is None:
            if not self._records_valid:
                saved = self.out.clone()
                rec = self.contact_forces()
                self.out.copy_(saved)
                return rec
            self._contacts = self._compact_records()
        return self._contacts

    def enable_energy_ledger(self) -> dict:
        """Start (or reset) the non-conservative work ledger (engine.py:630-653)."""
        self.out[:3] = 0.0
        self.energy_ledger = _Ledger(self, self.kinetic_energy())
        return self.energy_ledger

    def energies(self) -> dict:
        """Kinetic, gravitational and stored spring energies (engine.py:655-666)."""
        g = np.asarray(self.config.gravity, dtype=np.float64)
        if np.any(g != 0.0):
            gt = torch.as_tensor(g, device=self.device)
            pe = -float(torch.sum(self.mass * (self._x @ gt)).item())
        else:
            pe = 0.0
        el = self.out[OUT_ELASTIC:OUT_ELASTIC + 3].tolist()
        return {"kinetic": self.kinetic_energy(), "potential": pe, "elastic_normal": el[0],
                "elastic_tangential": el[1], "elastic_rolling": el[2]}

    def dissipated_work(self) -> float:
        if self.energy_ledger is None:
            raise ValueError("call enable_energy_ledger() first")
        return -float(sum(self.energy_ledger[k] for k in LEDGER_KEYS))

    def close(self) -> None:
        for g in self._graphs.values():
            self.lib.tdm_graph_destroy(g)
        self._graphs = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

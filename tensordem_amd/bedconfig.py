"""Unbonded Hertz-Mindlin disk bed: configuration and column packing.

The law restates, in two dimensions, the reference's floe-bed contact model
(src/tensordem/scale_ceiling/kernel.py:38-129, 216-227, 277-401): Hertz
normal force with a restitution-derived viscous term, Cundall-Strack
tangential spring capped by Coulomb friction, floor, back-wall and moving
right-wall (indenter or gate) contacts, gravity.  The column geometry follows the 2-D column-collapse protocol
(src/tensordem/granular2d/column_collapse.py:196-266): a column of aspect
ratio H/L resting on a rough floor against a back wall, released by the
absence of the gate wall.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class BedConfig:
    young: float = 1.0e9            # E [Pa]
    poisson: float = 0.3            # nu
    friction: float = 0.5           # mu, disk-disk
    wall_friction: float = 0.5      # mu, disk-wall
    restitution: float = 0.5        # e, realised through the viscous normal term
    density: float = 917.0          # [kg/m^3]
    thickness: float = 0.02         # out-of-plane disk thickness [m]
    gravity: float = 9.81           # [m/s^2], acting along -y
    floor_y: float = 0.0            # floor position [m]
    back_wall_x: float = 0.0        # back wall position [m]
    floor: bool = True
    back_wall: bool = True
    dt: float | None = None
    # right wall: the floe bed's indenter (scale_ceiling/kernel.py:156-197,
    # 338-401, 514-519), or the column's gate when it stands still; at
    # x = right_wall_x + right_wall_speed * t, pushing towards -x
    right_wall: bool = False
    right_wall_x: float = 0.0       # position at t = 0 [m]
    right_wall_speed: float = 0.0   # velocity along +x [m/s]
    # "euler": semi-implicit Euler (the plate's integrator); "verlet": the
    # floe bed's kick-drift-kick velocity-Verlet (scale_ceiling/kernel.py:
    # 410-421), whose state between steps holds the half-step velocity
    integrator: str = "euler"

    def __post_init__(self) -> None:
        for name in ("young", "density", "thickness"):
            v = getattr(self, name)
            if not math.isfinite(v) or v <= 0:
                raise ValueError(f"{name} must be finite and positive")
        if not 0.0 <= self.poisson < 0.5:
            raise ValueError("poisson must lie in [0, 0.5)")
        if not 0.0 < self.restitution <= 1.0:
            raise ValueError("restitution must lie in (0, 1]")
        for name in ("friction", "wall_friction", "gravity"):
            v = getattr(self, name)
            if not math.isfinite(v) or v < 0:
                raise ValueError(f"{name} must be finite and nonnegative")
        if self.dt is not None and (not math.isfinite(self.dt) or self.dt <= 0):
            raise ValueError("dt must be finite and positive")
        for name in ("right_wall_x", "right_wall_speed"):
            if not math.isfinite(getattr(self, name)):
                raise ValueError(f"{name} must be finite")
        if self.integrator not in ("euler", "verlet"):
            raise ValueError("integrator must be 'euler' or 'verlet'")

    # -- contact constants (scale_ceiling/kernel.py:119-129, 216-227) --------
    @property
    def e_star(self) -> float:
        """Pair reduced modulus E/(2(1-nu^2)) for equal materials."""
        return self.young / (2.0 * (1.0 - self.poisson ** 2))

    @property
    def g_star(self) -> float:
        """Mindlin pair shear modulus G/(2(2-nu)), G = E/(2(1+nu))."""
        return (self.young / (2.0 * (1.0 + self.poisson))) / (2.0 * (2.0 - self.poisson))

    @property
    def beta(self) -> float:
        """-ln e / sqrt(ln^2 e + pi^2); 0 for a perfectly elastic contact."""
        if self.restitution >= 1.0:
            return 0.0
        ln_e = math.log(self.restitution)
        return -ln_e / math.sqrt(ln_e * ln_e + math.pi ** 2)

    @property
    def damping_coefficient(self) -> float:
        """2 sqrt(5/6) beta: the viscous normal term is this x sqrt(s_n m*) x v_n."""
        return 2.0 * math.sqrt(5.0 / 6.0) * self.beta

    @property
    def rho_pi_t(self) -> float:
        """Disk mass per squared radius: m = rho_pi_t * r^2."""
        return (self.density * math.pi) * self.thickness

    def stable_dt(self, r_min: float, overlap_fraction: float = 0.01,
                  safety: float = 0.03) -> float:
        """Hertz contact period at a given overlap, times a safety factor
        (the floe bed's rule, scale_ceiling/kernel.py:131-141, for 2-D disks)."""
        m = self.rho_pi_t * r_min ** 2
        m_eff = 0.5 * m
        delta = overlap_fraction * r_min
        k = 2.0 * self.e_star * math.sqrt(0.5 * r_min * delta)
        return safety * 2.0 * math.pi / math.sqrt(k / m_eff)


def column_packing(n: int, r0: float, *, polydispersity: float = 0.2, aspect: float = 2.0,
                   seed: int = 0, gap: float = 0.02):
    """Loose random column of n disks: radii uniform in [(1-p) r0, (1+p) r0],
    centres on a hexagonal grid of spacing 2 r_max (1 + gap) shaken by a
    seeded jitter that keeps every pair apart, filled row by row from the
    floor against the back wall with H / L = aspect.  Returns (x, r) with x of
    shape (n, 2), both float64, in row-major (spatially coherent) order."""
    rng = np.random.default_rng(seed)
    r = r0 * rng.uniform(1.0 - polydispersity, 1.0 + polydispersity, size=n)
    rmax = r0 * (1.0 + polydispersity)
    a = 2.0 * rmax * (1.0 + gap)
    dy = a * math.sqrt(3.0) / 2.0
    # columns per row from the aspect ratio: rows * dy ~ aspect * cols * a
    cols = max(2, int(round(math.sqrt(n * dy / (aspect * a)))))
    rows = -(-n // cols)
    k = np.arange(n)
    row, col = k // cols, k % cols
    x = rmax + col * a + np.where(row % 2 == 1, 0.5 * a, 0.0)
    y = rmax + row * dy
    slack = max(0.5 * (a - 2.0 * rmax), 0.0)  # gap < 0 packs disks into contact
    x = x + rng.uniform(-slack, slack, size=n)
    y = y + rng.uniform(-slack, slack, size=n) * 0.5
    del rows
    return np.stack((x, y), axis=1), r


def column_deposit(n: int, r0: float, *, polydispersity: float = 0.2, aspect: float = 2.0,
                   seed: int = 0, gap: float = 0.15, packing: float = 0.755):
    """Random loose column of n disks built by sequential deposition: radii
    uniform in [(1-p) r0, (1+p) r0]; row by row from the back wall (x = 0),
    each disk placed right of its row neighbour with a random gap in
    [0, gap r0) and lowered until it rests on a disk of the rows below or on
    the floor (y = 0).  Every disk touches at least one support, no two
    overlap, the area fraction comes out near 0.755 and the width
    L = sqrt(n pi <r^2> / (aspect packing)) makes H / L close to `aspect`.
    Returns (x (n,2), r (n,), L), in row-major (spatially coherent) order;
    the column is not yet in equilibrium (settle it behind a gate at x = L,
    `bed.settled_column`)."""
    rng = np.random.default_rng(seed)
    r = r0 * rng.uniform(1.0 - polydispersity, 1.0 + polydispersity, size=n)
    width = math.sqrt(n * math.pi * float(np.mean(r * r)) / (aspect * packing))
    rmax = r0 * (1.0 + polydispersity)
    # a disk's supports lie within r + rmax of its centre; row neighbours are
    # at least 2 r_min apart, so five consecutive disks of a row cover that reach
    reach_w = 2 + int(math.ceil((r0 + rmax) / (r0 * (1.0 - polydispersity))))
    x = np.empty(n)
    y = np.empty(n)
    rows: list[tuple[np.ndarray, np.ndarray, np.ndarray]] = []
    per_row = int(width / (2.0 * r0 * (1.0 - polydispersity))) + 2
    k = 0
    while k < n:
        m = min(n - k, per_row)
        rr = r[k:k + m]
        g = gap * r0 * rng.uniform(0.0, 1.0, size=m)
        xs = rr[0] + g[0] + np.concatenate(([0.0], np.cumsum(rr[:-1] + rr[1:] + g[1:])))
        cnt = max(1, int(np.count_nonzero(xs + rr <= width)))
        rx, rr = xs[:cnt], rr[:cnt]
        ry = rr.copy()
        # the rows below that can reach this one: disks of a row are at most
        # 2 rmax above the lowest support of the row before it
        for px, py, pr in rows[-4:]:
            lo = np.searchsorted(px, rx - (rr + rmax))
            idx = lo[:, None] + np.arange(reach_w)[None, :]
            ok = idx < len(px)
            idc = np.minimum(idx, len(px) - 1)
            dx = px[idc] - rx[:, None]
            s = pr[idc] + rr[:, None]
            h2 = s * s - dx * dx
            ok &= h2 > 0.0
            need = np.where(ok, py[idc] + np.sqrt(np.maximum(h2, 0.0)), -np.inf)
            ry = np.maximum(ry, need.max(axis=1))
        rows.append((rx, ry, rr))
        x[k:k + cnt] = rx
        y[k:k + cnt] = ry
        k += cnt
    return np.stack((x, y), axis=1), r, width

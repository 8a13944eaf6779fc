"""DEM3DConfig and the bonded-sphere lattice of IceDEM3D (host logic).

`DEM3DConfig` keeps the reference's fields, defaults and validation
(src/tensordem/dem3d.py:115-432) so configurations, restart files and
callers carry over unchanged; `recommended_dt` follows dem3d.py:530-561.
`build_lattice` builds the reference's sphere lattice (dem3d.py:589-656):
box or cylinder_xz geometry compacted to contiguous indices, the 13
half-space neighbour offsets, bonds sorted by (i, j), fixed edges.  It
also lays out the per-sphere neighbour table the device kernels read:
26 slots, the partners above the sphere (ascending) then the partners
below it (ascending), each with its bond index, -1 where absent.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


def hertz_reduced_modulus(young_modulus: float, poisson_ratio: float, *,
                          rigid_counterpart: bool = False) -> float:
    """E* = E / (2 (1 - nu^2)) for equal spheres, E / (1 - nu^2) against a
    rigid surface (dem3d.py:71-79)."""
    if not math.isfinite(young_modulus) or young_modulus <= 0.0:
        raise ValueError("young_modulus must be finite and positive")
    if not (math.isfinite(poisson_ratio) and 0.0 <= poisson_ratio < 0.5):
        raise ValueError("poisson_ratio must be in [0, 0.5)")
    compliance = 1.0 - poisson_ratio ** 2
    return float(young_modulus / (compliance if rigid_counterpart else 2.0 * compliance))


def hertz_effective_radius(radius: float, *, rigid_counterpart: bool = False) -> float:
    """R* = R/2 for equal spheres, R against a plane (dem3d.py:82-86)."""
    if not math.isfinite(radius) or radius <= 0.0:
        raise ValueError("radius must be finite and positive")
    return float(radius) if rigid_counterpart else 0.5 * float(radius)


def hertz_reduced_shear_modulus(young_modulus: float, poisson_ratio: float, *,
                                rigid_counterpart: bool = False) -> float:
    """G* = E / (4 (2 - nu)(1 + nu)) for equal spheres, E / (2 (2 - nu)(1 + nu))
    against a rigid surface (hertz_mindlin.py:51-92)."""
    denom = (2.0 - poisson_ratio) * (1.0 + poisson_ratio)
    return float(young_modulus / ((2.0 if rigid_counterpart else 4.0) * denom))


def hertz_secant_coefficient(e_star: float, r_eff: float) -> float:
    """c of the secant stiffness k = c sqrt(d), c = (4/3) E* sqrt(R*)."""
    return (4.0 / 3.0) * float(e_star) * math.sqrt(float(r_eff))


@dataclass(frozen=True)
class DEM3DConfig:
    """Numerical and physical parameters of the 3-D bonded-sphere model."""
    nx: int = 9
    ny: int = 5
    nz: int = 3
    radius: float = 0.025
    density: float = 917.0
    bond_stiffness: float = 2_000.0
    contact_stiffness: float = 5_000.0
    tangential_stiffness: float | None = None
    breaking_strain: float = 0.015
    shear_breaking_strain: float = 0.03
    contact_damping: float = 5.0
    drag: float = 2.0
    tool_radius: float = 0.1
    tool_speed: float = 0.5
    tool_gap: float = 0.01
    dt: float | None = None
    device: str = "cpu"
    geometry: str = "box"
    fix_x_edges: bool = True
    fix_bottom: bool = False
    top_platen_enabled: bool = False
    bottom_platen_enabled: bool = False
    platen_stiffness: float = 5_000.0
    platen_damping: float = 0.0
    platen_friction_coefficient: float = 0.0
    platen_friction_stiffness: float | None = None
    top_platen_gap: float = 0.0
    bottom_platen_gap: float = 0.0
    top_platen_velocity: float = 0.0
    bottom_platen_velocity: float = 0.0
    lateral_platens: bool = False
    lateral_platen_stiffness: float | None = None
    lateral_platen_damping: float | None = None
    lateral_platen_gap: float = 0.0
    friction_coefficient: float = 0.0
    rolling_resistance_coefficient: float = 0.0
    restitution_coefficient: float = 1.0
    friction_regularization: str = "none"
    static_friction_coefficient: float | None = None
    dynamic_friction_coefficient: float | None = None
    friction_transition_speed: float = 1.0e-3
    contact_model: str = "linear"
    young_modulus: float = 1.0e9
    poisson_ratio: float = 0.3
    hertz_stability_overlap_fraction: float = 0.1
    contact_damping_model: str = "restitution"
    tangential_model: str = "cundall_strack"
    gravity: float = 0.0
    water_level: float | None = None
    water_density: float = 1025.0
    wet_drag_coefficient: float = 0.0
    bond_model: str = "linear"
    bond_equilibrium_fraction: float = 0.35
    bond_delayed_elastic_fraction: tuple[float, ...] = (0.45, 0.20)
    bond_relaxation_time_s: tuple[float, ...] = (0.5, 30.0)
    bond_creep_exponent: int = 0
    bond_creep_time_s: float = 1.0
    bond_creep_force_fraction: float = 1.0
    bond_strength_weibull_modulus: float = 0.0
    bond_strength_seed: int = 0

    def __post_init__(self) -> None:   # dem3d.py:273-432
        for name in ("nx", "ny", "nz"):
            value = getattr(self, name)
            if isinstance(value, bool) or not isinstance(value, int) or value < 2:
                raise ValueError(f"{name} must be an integer >= 2")
        for name in ("radius", "density", "bond_stiffness", "contact_stiffness",
                     "breaking_strain", "shear_breaking_strain", "tool_radius", "tool_speed"):
            if not math.isfinite(getattr(self, name)) or getattr(self, name) <= 0:
                raise ValueError(f"{name} must be finite and positive")
        for name in ("contact_damping", "drag", "tool_gap", "platen_damping",
                     "top_platen_gap", "bottom_platen_gap"):
            if not math.isfinite(getattr(self, name)) or getattr(self, name) < 0:
                raise ValueError(f"{name} must be finite and nonnegative")
        if not math.isfinite(self.platen_stiffness) or self.platen_stiffness <= 0:
            raise ValueError("platen_stiffness must be finite and positive")
        if not math.isfinite(self.platen_friction_coefficient) or self.platen_friction_coefficient < 0:
            raise ValueError("platen_friction_coefficient must be finite and nonnegative")
        if self.platen_friction_stiffness is not None and (
                not math.isfinite(self.platen_friction_stiffness) or self.platen_friction_stiffness <= 0):
            raise ValueError("platen_friction_stiffness must be finite and positive when given")
        for name in ("top_platen_velocity", "bottom_platen_velocity"):
            if not math.isfinite(getattr(self, name)):
                raise ValueError(f"{name} must be finite")
        for name in ("top_platen_enabled", "bottom_platen_enabled", "lateral_platens"):
            if not isinstance(getattr(self, name), bool):
                raise ValueError(f"{name} must be a boolean")
        if self.dt is not None and (not math.isfinite(self.dt) or self.dt <= 0):
            raise ValueError("dt must be finite and positive")
        if self.lateral_platen_stiffness is not None and (
                not math.isfinite(self.lateral_platen_stiffness) or self.lateral_platen_stiffness <= 0):
            raise ValueError("lateral_platen_stiffness must be positive when given")
        if self.lateral_platen_damping is not None and (
                not math.isfinite(self.lateral_platen_damping) or self.lateral_platen_damping < 0):
            raise ValueError("lateral_platen_damping must be nonnegative when given")
        if not math.isfinite(self.lateral_platen_gap) or self.lateral_platen_gap < 0:
            raise ValueError("lateral_platen_gap must be nonnegative")
        if self.device not in ("cpu", "cuda"):
            raise ValueError("device must be 'cpu' or 'cuda'")
        if self.geometry not in ("box", "cylinder_xz"):
            raise ValueError("geometry must be 'box' or 'cylinder_xz'")
        if self.geometry == "cylinder_xz" and (self.nx < 3 or self.nz < 3):
            raise ValueError("cylinder_xz geometry requires nx and nz >= 3")
        if not math.isfinite(self.friction_coefficient) or self.friction_coefficient < 0:
            raise ValueError("friction_coefficient must be finite and nonnegative")
        if not math.isfinite(self.rolling_resistance_coefficient) or self.rolling_resistance_coefficient < 0:
            raise ValueError("rolling_resistance_coefficient must be finite and nonnegative")
        if self.friction_regularization not in ("none", "tanh", "dual"):
            raise ValueError("friction_regularization must be 'none', 'tanh' or 'dual'")
        mu_s, mu_d = self.static_friction_coefficient, self.dynamic_friction_coefficient
        for name, value in (("static_friction_coefficient", mu_s),
                            ("dynamic_friction_coefficient", mu_d)):
            if value is not None and (not math.isfinite(value) or value < 0.0):
                raise ValueError(f"{name} must be finite and nonnegative, or None")
        if mu_s is not None and mu_d is not None and mu_d > mu_s:
            raise ValueError("dynamic_friction_coefficient must not exceed static_friction_coefficient")
        if not math.isfinite(self.friction_transition_speed) or self.friction_transition_speed <= 0.0:
            raise ValueError("friction_transition_speed must be finite and positive")
        if not (0.0 < self.restitution_coefficient <= 1.0):
            raise ValueError("restitution_coefficient must be in (0, 1]")
        if self.tangential_stiffness is not None and (
                not math.isfinite(self.tangential_stiffness) or self.tangential_stiffness <= 0.0):
            raise ValueError("tangential_stiffness must be finite and positive, or None")
        if self.contact_damping_model not in ("restitution", "absolute", "restitution_calibrated"):
            raise ValueError("contact_damping_model must be 'restitution', 'absolute' or "
                             "'restitution_calibrated'")
        if self.contact_damping_model == "restitution_calibrated" and self.contact_model != "hertz":
            raise ValueError("contact_damping_model='restitution_calibrated' requires "
                             "contact_model='hertz'")
        if self.tangential_model not in ("cundall_strack", "mindlin_partial_slip"):
            raise ValueError("tangential_model must be 'cundall_strack' or 'mindlin_partial_slip'")
        if self.tangential_model == "mindlin_partial_slip":
            if self.contact_model != "hertz":
                raise ValueError("tangential_model='mindlin_partial_slip' requires contact_model='hertz'")
            if self.friction_regularization != "none":
                raise ValueError("tangential_model='mindlin_partial_slip' requires "
                                 "friction_regularization='none'")
        if self.contact_damping_model == "absolute" and self.restitution_coefficient < 1.0:
            raise ValueError("contact_damping_model='absolute' and restitution_coefficient < 1 "
                             "specify the same dashpot twice; pick one convention")
        if self.contact_model not in ("linear", "hertz"):
            raise ValueError("contact_model must be 'linear' or 'hertz'")
        if not math.isfinite(self.young_modulus) or self.young_modulus <= 0.0:
            raise ValueError("young_modulus must be finite and positive")
        if not (math.isfinite(self.poisson_ratio) and 0.0 <= self.poisson_ratio < 0.5):
            raise ValueError("poisson_ratio must be in [0, 0.5)")
        if not (math.isfinite(self.hertz_stability_overlap_fraction)
                and 0.0 < self.hertz_stability_overlap_fraction < 1.0):
            raise ValueError("hertz_stability_overlap_fraction must be in (0, 1)")
        if not math.isfinite(self.gravity) or self.gravity < 0.0:
            raise ValueError("gravity must be finite and nonnegative")
        if self.water_level is not None and not math.isfinite(self.water_level):
            raise ValueError("water_level must be finite or None")
        if not math.isfinite(self.water_density) or self.water_density <= 0.0:
            raise ValueError("water_density must be finite and positive")
        if not math.isfinite(self.wet_drag_coefficient) or self.wet_drag_coefficient < 0.0:
            raise ValueError("wet_drag_coefficient must be finite and nonnegative")
        if self.bond_model not in ("linear", "maxwell_wiechert"):
            raise ValueError("bond_model must be 'linear' or 'maxwell_wiechert'")
        if not (math.isfinite(self.bond_equilibrium_fraction) and 0.0 < self.bond_equilibrium_fraction < 1.0):
            raise ValueError("bond_equilibrium_fraction must be in (0, 1)")
        if len(self.bond_delayed_elastic_fraction) != len(self.bond_relaxation_time_s):
            raise ValueError("bond_delayed_elastic_fraction and bond_relaxation_time_s "
                             "must have equal length")
        if self.bond_model == "maxwell_wiechert" and not self.bond_relaxation_time_s:
            raise ValueError("bond_model='maxwell_wiechert' needs at least one relaxation time")
        for value in self.bond_relaxation_time_s:
            if not (math.isfinite(value) and value > 0.0):
                raise ValueError("bond_relaxation_time_s entries must be finite and positive")
        for value in self.bond_delayed_elastic_fraction:
            if not (math.isfinite(value) and value > 0.0):
                raise ValueError("bond_delayed_elastic_fraction entries must be finite and positive")
        if isinstance(self.bond_creep_exponent, bool) or not isinstance(self.bond_creep_exponent, int) \
                or self.bond_creep_exponent < 0:
            raise ValueError("bond_creep_exponent must be a nonnegative integer")
        if self.bond_creep_exponent > 0:
            if not (math.isfinite(self.bond_creep_time_s) and self.bond_creep_time_s > 0.0):
                raise ValueError("bond_creep_time_s must be finite and positive")
            if not (math.isfinite(self.bond_creep_force_fraction) and self.bond_creep_force_fraction > 0.0):
                raise ValueError("bond_creep_force_fraction must be finite and positive")
        if isinstance(self.bond_strength_weibull_modulus, bool) or not isinstance(
                self.bond_strength_weibull_modulus, (int, float)):
            raise TypeError("bond_strength_weibull_modulus must be a real number")
        if not math.isfinite(self.bond_strength_weibull_modulus) or self.bond_strength_weibull_modulus < 0.0:
            raise ValueError("bond_strength_weibull_modulus must be finite and nonnegative")
        if isinstance(self.bond_strength_seed, bool) or not isinstance(self.bond_strength_seed, int):
            raise TypeError("bond_strength_seed must be an integer")

    @property
    def mass(self) -> float:
        return self.density * (4.0 / 3.0) * math.pi * self.radius**3

    @property
    def effective_tangential_stiffness(self) -> float:
        return float(self.contact_stiffness if self.tangential_stiffness is None
                     else self.tangential_stiffness)

    def _bound_stiffness(self, linear_stiffness: float, *, rigid_counterpart: bool) -> float:
        # dem3d.py:506-527: the Hertz law is bounded by its secant stiffness at
        # d_ref = hertz_stability_overlap_fraction * radius
        if self.contact_model == "linear":
            return float(linear_stiffness)
        delta_ref = self.hertz_stability_overlap_fraction * self.radius
        e_star = hertz_reduced_modulus(self.young_modulus, self.poisson_ratio,
                                       rigid_counterpart=rigid_counterpart)
        r_eff = hertz_effective_radius(self.radius, rigid_counterpart=rigid_counterpart)
        return float(hertz_secant_coefficient(e_star, r_eff) * math.sqrt(max(delta_ref, 0.0)))

    @property
    def recommended_dt(self) -> float:   # dem3d.py:530-561
        n = self.nx * self.ny * self.nz
        degree = min(n - 1, 26)
        platen_k = (self.platen_stiffness if self.platen_friction_stiffness is None
                    else max(self.platen_stiffness, self.platen_friction_stiffness))
        lat_k = (self.lateral_platen_stiffness if self.lateral_platen_stiffness is not None
                 else self.platen_stiffness)
        lat_c = (self.lateral_platen_damping if self.lateral_platen_damping is not None
                 else self.platen_damping)
        n_platens = (int(self.top_platen_enabled) + int(self.bottom_platen_enabled)
                     + (4 if self.lateral_platens else 0))
        stiffness = max(
            self.bond_stiffness,
            self._bound_stiffness(self.contact_stiffness, rigid_counterpart=False),
            self._bound_stiffness(self.effective_tangential_stiffness, rigid_counterpart=False),
            self._bound_stiffness(max(platen_k, lat_k if self.lateral_platens else 0.0),
                                  rigid_counterpart=True) if n_platens else 0.0,
        )
        k = (2 * degree + 1) * stiffness
        c = ((2 * degree + 1) * self.contact_damping + self.drag
             + n_platens * max(self.platen_damping, lat_c if self.lateral_platens else 0.0))
        elastic = 0.15 * math.sqrt(self.mass / k)
        damping = 0.15 * self.mass / c if c > 0 else math.inf
        travel = 0.1 * self.radius / self.tool_speed
        return min(elastic, damping, travel)


# forward half-space offsets (dx, dy, dz) of dem3d.py:621-629, in the
# reference's loop order (dz, dy, dx)
FORWARD_OFFSETS = tuple((dx, dy, dz) for dz in range(0, 2) for dy in range(-1, 2)
                        for dx in range(-1, 2)
                        if (dz > 0 or (dz == 0 and dy > 0) or (dz == 0 and dy == 0 and dx > 0)))
NSLOTS = 26


@dataclass
class Lattice:
    positions: np.ndarray      # (n, 3) initial positions
    fixed: np.ndarray          # (n,) bool
    pairs: np.ndarray          # (B, 2) int64, sorted by (i, j)
    grid: np.ndarray           # (n, 3) int64 lattice coordinates
    nbr: np.ndarray            # (n, 26) int32 partners: above ascending, then below ascending
    bond_of_slot: np.ndarray   # (n, 26) int32 bond index of each slot


def build_lattice(cfg: DEM3DConfig) -> Lattice:
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    zz, yy, xx = np.meshgrid(np.arange(nz, dtype=np.float64), np.arange(ny, dtype=np.float64),
                             np.arange(nx, dtype=np.float64), indexing="ij")
    full = np.stack((xx, yy, zz), axis=-1).reshape(-1, 3)
    if cfg.geometry == "cylinder_xz":
        xc, zc = (nx - 1) / 2.0, (nz - 1) / 2.0
        disk = min(nx - 1, nz - 1) / 2.0
        keep = (((xx - xc) * (xx - xc) + (zz - zc) * (zz - zc)) <= disk**2 + 1e-12).reshape(-1)
    else:
        keep = np.ones(len(full), dtype=bool)
    compact_to_full = np.nonzero(keep)[0]
    full_to_compact = np.full(len(full), -1, dtype=np.int64)
    full_to_compact[compact_to_full] = np.arange(len(compact_to_full))
    grid = full[compact_to_full]
    z_origin = (nz - 1) / 2.0 if cfg.geometry == "cylinder_xz" else 0.0
    centre = np.array([(nx - 1) / 2, (ny - 1) / 2, z_origin])
    positions = (grid - centre) * (2 * cfg.radius)
    gi = grid.astype(np.int64)
    fixed = np.zeros(len(grid), dtype=bool)
    if cfg.fix_x_edges and cfg.geometry == "box":
        fixed |= (gi[:, 0] == 0) | (gi[:, 0] == nx - 1)
    if cfg.fix_bottom:
        fixed |= gi[:, 2] == 0
    n = len(grid)
    chunks = []
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    z, y, x = z.reshape(-1), y.reshape(-1), x.reshape(-1)
    for dx, dy, dz in FORWARD_OFFSETS:
        ok = (x + dx >= 0) & (x + dx < nx) & (y + dy >= 0) & (y + dy < ny) & (z + dz < nz)
        fa = (z * ny + y) * nx + x
        fb = ((z + dz) * ny + (y + dy)) * nx + (x + dx)
        a, b = full_to_compact[fa[ok]], full_to_compact[fb[ok]]
        both = (a >= 0) & (b >= 0)
        chunks.append(np.stack((np.minimum(a, b)[both], np.maximum(a, b)[both]), axis=1))
    pairs = np.concatenate(chunks) if chunks else np.zeros((0, 2), dtype=np.int64)
    pairs = pairs[np.argsort(pairs[:, 0] * n + pairs[:, 1], kind="stable")]
    nbr = np.full((n, NSLOTS), -1, dtype=np.int32)
    bos = np.full((n, NSLOTS), -1, dtype=np.int32)
    bidx = np.arange(len(pairs))
    # partners above: bonds are sorted by (i, j), so each i's bonds are ascending in j
    i, j = pairs[:, 0], pairs[:, 1]
    start = np.searchsorted(i, np.arange(n))
    rank = bidx - start[i]
    nbr[i, rank] = j
    bos[i, rank] = bidx
    # partners below: bonds of one j in ascending i
    order = np.lexsort((i, j))
    js, is_ = j[order], i[order]
    startj = np.searchsorted(js, np.arange(n))
    rankj = np.arange(len(js)) - startj[js]
    nbr[js, 13 + rankj] = is_
    bos[js, 13 + rankj] = order
    return Lattice(positions, fixed, pairs, gi, nbr, bos)

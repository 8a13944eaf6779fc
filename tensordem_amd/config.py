"""Solver configuration for the bonded ice plate.

`DEMConfig` keeps the field names, defaults, validation and derived
quantities of the reference dataclass (src/tensordem/dem.py:17-72) so that
existing callers construct it unchanged.  `dt_policy` is an addition: the
reference's automatic step bounds every particle by all N-1 possible
partners (dem.py:61-72), which shrinks dt like 1/N; "degree" bounds it by the
`max_degree` partners a disk in a square lattice can touch instead.
"""
from __future__ import annotations

import math
from dataclasses import dataclass


@dataclass(frozen=True)
class DEMConfig:
    nx: int = 21
    ny: int = 5
    radius: float = 0.025
    thickness: float = 0.2
    density: float = 917.0
    bond_stiffness: float = 2_000.0
    contact_stiffness: float = 2_000.0
    breaking_strain: float = 0.015
    shear_breaking_strain: float = 0.03
    contact_damping: float = 5.0
    drag: float = 2.0
    tool_radius: float = 0.1
    tool_speed: float = 1.0
    tool_gap: float = 0.01
    dt: float | None = None
    device: str = "cpu"
    fix_edges: bool = True

    def __post_init__(self) -> None:
        for name in ("nx", "ny"):
            value = getattr(self, name)
            if isinstance(value, bool) or not isinstance(value, int) or value < 2:
                raise ValueError(f"{name} must be an integer >= 2")
        for name in (
            "radius", "thickness", "density", "bond_stiffness",
            "contact_stiffness", "breaking_strain", "shear_breaking_strain",
            "tool_radius", "tool_speed",
        ):
            value = getattr(self, name)
            if not math.isfinite(value) or value <= 0:
                raise ValueError(f"{name} must be finite and positive")
        for name in ("contact_damping", "drag", "tool_gap"):
            value = getattr(self, name)
            if not math.isfinite(value) or value < 0:
                raise ValueError(f"{name} must be finite and nonnegative")
        if self.dt is not None and (not math.isfinite(self.dt) or self.dt <= 0):
            raise ValueError("dt must be finite and positive")

    @property
    def mass(self) -> float:
        return self.density * math.pi * self.radius**2 * self.thickness

    def _dt_bound(self, neighbors: int) -> float:
        stiffness = (2 * neighbors + 1) * max(self.bond_stiffness, self.contact_stiffness)
        damping = (2 * neighbors + 1) * self.contact_damping + self.drag
        elastic_dt = 0.2 * math.sqrt(self.mass / stiffness)
        damping_dt = 0.2 * self.mass / damping if damping else math.inf
        travel_dt = 0.1 * self.radius / self.tool_speed
        return min(elastic_dt, damping_dt, travel_dt)

    @property
    def recommended_dt(self) -> float:
        """Conservative bound over all N-1 partners (dem.py:61-72)."""
        return self._dt_bound(self.nx * self.ny - 1)

    def degree_bound_dt(self, max_degree: int = 8) -> float:
        """The same bound with at most `max_degree` partners per disk.

        A disk in the square lattice has 8 bonds; a disk of radius r can touch
        at most 6 equal disks without overlap beyond the bond range, so 8
        bounds both.  Used for plates too large for the all-partner bound.
        """
        return self._dt_bound(min(max_degree, self.nx * self.ny - 1))

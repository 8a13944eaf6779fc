"""Device-resident batch of bonded ice plates driven by the HIP kernels.

`PlateEngine` owns every device array of one or more independent plates
(members) as torch tensors on the ROCm device, fills the `tdm_plate_arrays`
descriptor with their addresses and calls the C ABI.  One kernel launch
advances all members by one step of `IceDEM.step` (src/tensordem/dem.py:
347-382); members never interact, and each member's ledger is reduced in an
order that depends only on that member, so a member gives the same result
alone and inside any batch.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .config import DEMConfig
from .lattice import bonds, fixed_mask, positions

THREADS = 256


@dataclass
class PlateSpec:
    """One plate: reference config, time step, tool or hull, optional per-disk strengths."""
    config: DEMConfig
    dt: float | None = None
    hull_profile: np.ndarray | None = None
    hull_start: tuple[float, float] = (0.0, 0.0)
    hull_velocity: tuple[float, float] = (0.0, 0.0)
    eps_t_disk: np.ndarray | None = None
    eps_s_disk: np.ndarray | None = None
    tile_x: int | None = None
    extra: dict = field(default_factory=dict)


def default_tile(nx: int) -> int:
    """Tile width for a lattice row of nx disks (256 disks per tile)."""
    if nx >= 48:
        return 32
    return 16


@dataclass
class PlateParts:
    """One member's static data on the device and its member-record template
    (batch-relative offsets left at zero).  Built once per plate; a batch is
    assembled from parts by concatenation on the device."""
    member: object
    n: int
    nbonds: int
    ntiles: int
    ncells: int
    x0: torch.Tensor
    fixed: torch.Tensor
    fwd: torch.Tensor
    bi: torch.Tensor
    bj: torch.Tensor
    rest: torch.Tensor
    eps_t: torch.Tensor | None
    eps_s: torch.Tensor | None
    hull: torch.Tensor | None


def plate_parts(s: PlateSpec, *, device="cuda", skin_factor: float = 1.0,
                grid_margin: float = 0.25, fp32_margin: float = 1e-3) -> PlateParts:
    lib = _lib.load()
    dev = torch.device(device)
    c = s.config
    pos = positions(c.nx, c.ny, c.radius)
    bi, bj, rl, fw = bonds(c.nx, c.ny, c.radius)
    m = _lib.Member()
    m.radius, m.bond_stiffness, m.contact_stiffness = c.radius, c.bond_stiffness, c.contact_stiffness
    m.breaking_strain, m.shear_breaking_strain = c.breaking_strain, c.shear_breaking_strain
    m.contact_damping, m.drag, m.tool_radius = c.contact_damping, c.drag, c.tool_radius
    m.dt = float(s.dt if s.dt is not None else (c.dt if c.dt is not None else c.recommended_dt))
    m.mass = c.mass
    hull = None
    if s.hull_profile is None:
        m.tool_start[0] = 0.0
        m.tool_start[1] = (c.ny - 1) * 2 * c.radius + c.radius + c.tool_radius + c.tool_gap
        m.tool_velocity[0], m.tool_velocity[1] = 0.0, -c.tool_speed
        m.hull_m = 0
    else:
        hp = np.asarray(s.hull_profile, dtype=np.float64)
        m.tool_start[0], m.tool_start[1] = s.hull_start
        m.tool_velocity[0], m.tool_velocity[1] = s.hull_velocity
        m.hull_m = len(hp)
        hull = torch.tensor(hp, dtype=torch.float64, device=dev)
    m.time = 0.0
    m.skin = skin_factor * c.radius
    m.cell_size = 2 * c.radius + m.skin
    w = (c.nx - 1) * 2 * c.radius
    h = (c.ny - 1) * 2 * c.radius
    mx = grid_margin * max(w, h) + 2 * m.cell_size
    m.grid_origin[0] = pos[:, 0].min() - mx
    m.grid_origin[1] = -mx
    m.ncx = int(math.ceil((w + 2 * mx) / m.cell_size)) + 1
    m.ncy = int(math.ceil((h + 2 * mx) / m.cell_size)) + 1
    m.fp32_margin = fp32_margin
    m.step_count = 0
    m.nx, m.ny = c.nx, c.ny
    m.nbonds = len(bi)
    m.tile_x = s.tile_x or default_tile(c.nx)
    nt = lib.tdm_plate_tiles(c.nx, c.ny, m.tile_x)
    _lib.check(nt if nt < 0 else 0)
    m.nblocks = nt
    m.mode, m.tab, m.status, m.has_loads = 1, 0, 0, 0
    m.per_disk_strength = int(s.eps_t_disk is not None)
    f64 = dict(dtype=torch.float64, device=dev)
    n = c.nx * c.ny
    et = es = None
    if s.eps_t_disk is not None or s.eps_s_disk is not None:
        et = torch.tensor(np.full(n, c.breaking_strain) if s.eps_t_disk is None
                          else np.asarray(s.eps_t_disk, np.float64), **f64)
        es = torch.tensor(np.full(n, c.shear_breaking_strain) if s.eps_s_disk is None
                          else np.asarray(s.eps_s_disk, np.float64), **f64)
        m.per_disk_strength = 1
    return PlateParts(
        member=m, n=n, nbonds=len(bi), ntiles=nt, ncells=m.ncx * m.ncy,
        x0=torch.tensor(pos, **f64),
        fixed=torch.tensor(fixed_mask(c.nx, c.ny, c.fix_edges), dtype=torch.bool, device=dev),
        fwd=torch.tensor(fw, dtype=torch.int32, device=dev),
        bi=torch.tensor(bi, dtype=torch.int64, device=dev),
        bj=torch.tensor(bj, dtype=torch.int64, device=dev),
        rest=torch.tensor(rl, **f64), eps_t=et, eps_s=es, hull=hull)


class PlateEngine:
    def __init__(self, specs: list[PlateSpec], *, device: str | torch.device = "cuda",
                 fp32: bool = False, skin_factor: float = 1.0, cand_slots: int = 16,
                 cell_cap: int = 8, grid_margin: float = 0.25, fp32_margin: float = 1e-3,
                 parts: list[PlateParts] | None = None):
        self.lib = _lib.load()
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise ValueError("PlateEngine runs on the ROCm device (device='cuda')")
        if not torch.cuda.is_available():
            raise RuntimeError("no ROCm device available")
        self.specs = list(specs)
        self.fp32 = bool(fp32)
        if parts is None:
            parts = [plate_parts(s, device=self.device, skin_factor=skin_factor,
                                 grid_margin=grid_margin, fp32_margin=fp32_margin)
                     for s in self.specs]
        if len(parts) != len(self.specs):
            raise ValueError("one PlateParts per spec")
        nm = len(parts)
        members = (_lib.Member * nm)()
        n = nb = blocks = cells = hull_n = 0
        per_disk = any(pt.eps_t is not None for pt in parts)
        fwd, bond_i, bond_j, eps_t, eps_s, hull = [], [], [], [], [], []
        for k, pt in enumerate(parts):
            ctypes.memmove(ctypes.byref(members[k]), ctypes.byref(pt.member),
                           ctypes.sizeof(_lib.Member))
            m = members[k]
            m.first, m.bond_first, m.block_first, m.cell_first = n, nb, blocks, cells
            m.hull_first = hull_n
            if pt.hull is not None:
                hull.append(pt.hull)
                hull_n += len(pt.hull)
            fwd.append(pt.fwd + nb)
            bond_i.append(pt.bi + n)
            bond_j.append(pt.bj + n)
            if per_disk:
                c = self.specs[k].config
                eps_t.append(pt.eps_t if pt.eps_t is not None else
                             torch.full((pt.n,), c.breaking_strain, dtype=torch.float64,
                                        device=self.device))
                eps_s.append(pt.eps_s if pt.eps_s is not None else
                             torch.full((pt.n,), c.shear_breaking_strain, dtype=torch.float64,
                                        device=self.device))
            n += pt.n
            nb += pt.nbonds
            blocks += pt.ntiles
            cells += pt.ncells
        self.n, self.nbonds, self.nmembers, self.nblocks, self.ncells = n, nb, nm, blocks, cells
        self.max_tile_x = max(int(pt.member.tile_x) for pt in parts)
        self.cand_slots, self.cell_cap = cand_slots, cell_cap
        dev = self.device
        f64 = dict(dtype=torch.float64, device=dev)
        self.initial_positions = torch.cat([pt.x0 for pt in parts])
        self.x = self.initial_positions.clone()
        self.v = torch.zeros_like(self.x)
        self.x_alt = torch.empty_like(self.x)
        self.v_alt = torch.empty_like(self.x)
        self.fixed = torch.cat([pt.fixed for pt in parts])
        self.links = torch.zeros(n, dtype=torch.uint8, device=dev)
        self.fwd_bond = torch.cat(fwd)
        self.bond_i = torch.cat(bond_i)
        self.bond_j = torch.cat(bond_j)
        self.rest_lengths = torch.cat([pt.rest for pt in parts])
        self.alive = torch.ones(nb, dtype=torch.bool, device=dev)
        self.failure_mode = torch.zeros(nb, dtype=torch.uint8, device=dev)
        self.failure_time_s = torch.full((nb,), -1.0, **f64)
        self.failure_extension_m = torch.zeros(nb, **f64)
        self.failure_energy_J = torch.zeros(nb, **f64)
        self.loads = torch.zeros((n, 2), **f64)
        self.hull = torch.cat(hull) if hull else torch.zeros((1, 2), **f64)
        self.eps_t_disk = torch.cat(eps_t) if per_disk else None
        self.eps_s_disk = torch.cat(eps_s) if per_disk else None
        self.accounting = torch.zeros((nm, _lib.NACC), **f64)
        self.last = torch.zeros((nm, _lib.NLAST), **f64)
        self.force = torch.zeros((n, 2), **f64)
        self.reaction = torch.zeros((nm, 2), **f64)
        self.energy = torch.zeros((nm, 5), **f64)
        self._members_host = members
        self._has_loads = [0] * nm
        self.members = torch.frombuffer(bytearray(bytes(members)), dtype=torch.uint8).to(dev)
        wsb = self.lib.tdm_plate_workspace_bytes(n, blocks, nm, cells, cand_slots, cell_cap)
        self.workspace = torch.zeros(int(wsb) + 256, dtype=torch.uint8, device=dev)
        self._arrays = _lib.Arrays()
        self._refresh()
        _lib.check(self.lib.tdm_plate_setup(ctypes.byref(self._arrays), self.stream))
        self._graphs: dict = {}

    # ---- descriptor ----------------------------------------------------------
    @property
    def stream(self) -> int:
        return torch.cuda.current_stream(self.device).cuda_stream

    def _refresh(self, loads: bool = False) -> None:
        a = self._arrays
        p = lambda t: None if t is None else t.data_ptr()  # noqa: E731
        a.x, a.v, a.x_alt, a.v_alt = p(self.x), p(self.v), p(self.x_alt), p(self.v_alt)
        a.fixed, a.links, a.fwd_bond, a.alive = p(self.fixed), p(self.links), p(self.fwd_bond), p(self.alive)
        a.failure_mode, a.failure_time_s = p(self.failure_mode), p(self.failure_time_s)
        a.failure_extension_m, a.failure_energy_J = p(self.failure_extension_m), p(self.failure_energy_J)
        a.loads, a.hull = p(self.loads), p(self.hull)
        a.eps_t_disk, a.eps_s_disk = p(self.eps_t_disk), p(self.eps_s_disk)
        a.accounting, a.last, a.force, a.reaction = p(self.accounting), p(self.last), p(self.force), p(self.reaction)
        a.energy = p(self.energy)
        a.members = p(self.members)
        base = self.workspace.data_ptr()
        aligned = (base + 255) & ~255
        a.workspace = aligned
        a.workspace_bytes = self.workspace.numel() - (aligned - base)
        a.n, a.nbonds, a.nmembers, a.nblocks, a.ncells = self.n, self.nbonds, self.nmembers, self.nblocks, self.ncells
        a.cand_slots, a.cell_cap = self.cand_slots, self.cell_cap
        a.max_tile_x = self.max_tile_x

    def _swap(self) -> None:
        self.x, self.x_alt = self.x_alt, self.x
        self.v, self.v_alt = self.v_alt, self.v
        self._refresh()

    # ---- member table ----------------------------------------------------------
    def read_members(self):
        raw = self.members.cpu().numpy().tobytes()
        return (_lib.Member * self.nmembers).from_buffer_copy(raw)

    def write_members(self, members) -> None:
        self.members.copy_(torch.frombuffer(bytearray(bytes(members)), dtype=torch.uint8))

    def _field_bytes(self, member: int, name: str, ctype, value) -> None:
        off = member * ctypes.sizeof(_lib.Member) + getattr(_lib.Member, name).offset
        raw = bytes(ctype(value))
        self.members[off:off + len(raw)].copy_(torch.frombuffer(bytearray(raw), dtype=torch.uint8))

    def set_double(self, member: int, name: str, value: float) -> None:
        self._field_bytes(member, name, ctypes.c_double, float(value))

    def set_int(self, member: int, name: str, value: int) -> None:
        self._field_bytes(member, name, ctypes.c_int32, int(value))
        if name == "has_loads":
            self._has_loads[member] = int(value)

    def set_int64(self, member: int, name: str, value: int) -> None:
        self._field_bytes(member, name, ctypes.c_int64, int(value))

    def has_loads(self, member: int) -> bool:
        return bool(self._has_loads[member])

    def read_status(self, member: int) -> int:
        off = member * ctypes.sizeof(_lib.Member) + _lib.Member.status.offset
        return int(self.members[off:off + 4].cpu().view(torch.int32).item())

    def clear_status(self, member: int) -> None:
        self.set_int(member, "status", 0)

    def set_field(self, name: str, values) -> None:
        """Set one member-table field for all members (host round trip)."""
        ms = self.read_members()
        for k, m in enumerate(ms):
            val = values[k] if isinstance(values, (list, tuple, np.ndarray)) else values
            if isinstance(getattr(m, name), ctypes.Array):
                for t, e in enumerate(val):
                    getattr(m, name)[t] = e
            else:
                setattr(m, name, val)
        self.write_members(ms)

    # ---- operations ----------------------------------------------------------
    def step(self, nsteps: int = 1) -> None:
        _lib.check(self.lib.tdm_plate_step(ctypes.byref(self._arrays), int(nsteps), 0,
                                           int(self.fp32), self.stream))
        if nsteps % 2:
            self._swap()

    def step_timed(self, nsteps: int) -> float:
        """Advance nsteps (even) and return the mean step-kernel time in us,
        from HIP events recorded around each step-kernel launch."""
        if nsteps % 2:
            raise ValueError("step_timed needs an even step count")
        out = ctypes.c_double()
        _lib.check(self.lib.tdm_plate_step_timed(ctypes.byref(self._arrays), int(nsteps),
                                                 int(self.fp32), self.stream, ctypes.byref(out)))
        return out.value

    def forces(self) -> None:
        _lib.check(self.lib.tdm_plate_forces(ctypes.byref(self._arrays), 0, int(self.fp32),
                                             self.stream))

    def mechanical_energy(self) -> None:
        _lib.check(self.lib.tdm_plate_energy(ctypes.byref(self._arrays), 0, self.stream))

    def sync_links(self) -> None:
        _lib.check(self.lib.tdm_plate_sync_links(ctypes.byref(self._arrays), self.stream))

    def rebin(self) -> None:
        _lib.check(self.lib.tdm_plate_rebin(ctypes.byref(self._arrays), self.stream))

    def graph(self, pairs: int):
        """Capture 2*pairs steps (state returns to x/v) as a replayable graph."""
        key = (pairs, self.fp32, self.x.data_ptr())
        if key not in self._graphs:
            g = ctypes.c_void_p()
            stream = torch.cuda.Stream(self.device)
            stream.wait_stream(torch.cuda.current_stream(self.device))
            with torch.cuda.stream(stream):
                _lib.check(self.lib.tdm_plate_graph_create(
                    ctypes.byref(self._arrays), int(pairs), int(self.fp32), stream.cuda_stream,
                    ctypes.byref(g)))
            torch.cuda.current_stream(self.device).wait_stream(stream)
            self._graphs[key] = g
        return self._graphs[key]

    def run_graph(self, pairs: int, repeats: int = 1) -> None:
        g = self.graph(pairs)
        for _ in range(repeats):
            _lib.check(self.lib.tdm_graph_launch(g, self.stream))

    def status(self) -> np.ndarray:
        return np.array([m.status for m in self.read_members()], dtype=np.int32)

    def close(self) -> None:
        for g in self._graphs.values():
            self.lib.tdm_graph_destroy(g)
        self._graphs.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

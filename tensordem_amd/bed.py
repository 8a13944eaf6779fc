"""Device-resident batch of Hertz-Mindlin disk beds driven by the HIP kernels.

`BedEngine` owns the device arrays of one or more independent beds (members)
and calls the bed entry points of include/tensordem_amd.h.  One launch
advances every member by one step of the bed law (tensordem_amd/csrc/bed.hip);
a member's ledger is reduced in an order that depends only on that member.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .bedconfig import BedConfig

THREADS = 256
INT32_MAX = 2 ** 31 - 1


def morton_keys(cx: torch.Tensor, cy: torch.Tensor) -> torch.Tensor:
    """Morton (Z-order) code of integer cell coordinates below 2**16."""
    def spread(v):
        v = v.to(torch.int64) & 0xFFFF
        v = (v | (v << 8)) & 0x00FF00FF
        v = (v | (v << 4)) & 0x0F0F0F0F
        v = (v | (v << 2)) & 0x33333333
        v = (v | (v << 1)) & 0x55555555
        return v
    return spread(cx) | (spread(cy) << 1)


def spatial_order(x: torch.Tensor, origin, h: float, ncx: int, ncy: int) -> torch.Tensor:
    """Permutation putting disks in Morton order of their cells (side h from
    `origin`, clamped into the ncx x ncy grid), ties by current index."""
    cx = torch.clamp(torch.floor((x[:, 0] - origin[0]) / h), 0, ncx - 1).to(torch.int64)
    cy = torch.clamp(torch.floor((x[:, 1] - origin[1]) / h), 0, ncy - 1).to(torch.int64)
    n = len(x)
    key = morton_keys(cx, cy) * n + torch.arange(n, device=x.device)
    return torch.argsort(key)


def remap_lists(cand, xi, mask, count, perm, inv, first):
    """Candidate lists of the disks of one member after renumbering.

    cand (S, m) int32 global indices, xi (S, m, 2), mask (m,) int32 and count
    (m,) uint8 are the member's current lists in the old numbering; perm (m,)
    gives the old global index of each new local position, inv (N,) the new
    global index of every old one.  Returns the lists of the new positions:
    partners renumbered and re-sorted ascending with their spring histories
    and contact bits, each history negated where the pair's lo -> hi
    orientation flips, wall bits (16, 17, 18) kept."""
    S, m = cand.shape
    dev = cand.device
    loc = perm - first
    c = cand[:, loc].to(torch.int64)
    h = xi[:, loc].clone()
    mk = mask[loc].to(torch.int64)
    k = count[loc].to(torch.int64)
    slot = torch.arange(S, device=dev)[:, None]
    valid = slot < k[None, :]
    c_old = torch.where(valid, c, perm[None, :])
    c_new = inv[c_old]
    self_new = first + torch.arange(m, device=dev)
    flip = valid & ((perm[None, :] < c_old) != (self_new[None, :] < c_new))
    h = torch.where(flip[..., None], -h, h)
    key = torch.where(valid, c_new, torch.full_like(c_new, INT32_MAX))
    order = torch.argsort(key, dim=0, stable=True)
    c_sorted = torch.gather(key, 0, order)
    h_sorted = torch.gather(h, 0, order[..., None].expand(-1, -1, 2))
    bits = (mk[None, :] >> order) & 1
    bits = torch.where(valid, bits, torch.zeros_like(bits))
    new_mask = (bits << slot).sum(dim=0) | (mk & (7 << 16))
    out_cand = torch.where(valid, c_sorted, torch.zeros_like(c_sorted)).to(torch.int32)
    return out_cand, h_sorted, new_mask.to(torch.int32), k.to(torch.uint8)


@dataclass
class BedSpec:
    """One bed: law, initial positions (n,2) and radii (n,), time step."""
    config: BedConfig
    positions: np.ndarray
    radii: np.ndarray
    dt: float | None = None
    velocities: np.ndarray | None = None


@dataclass
class BedParts:
    """One bed's static data on the device and its member-record template
    (batch-relative offsets left at zero); a batch is assembled from parts by
    concatenation on the device."""
    member: object
    n: int
    nblocks: int
    ncells: int
    x0: torch.Tensor
    v0: torch.Tensor
    radius: torch.Tensor


def bed_parts(s: BedSpec, *, device="cuda", skin_factor: float = 0.4,
              runout: float = 1.0) -> BedParts:
    lib = _lib.load()
    dev = torch.device(device)
    c = s.config
    x = np.asarray(s.positions, np.float64)
    r = np.asarray(s.radii, np.float64)
    if x.ndim != 2 or x.shape[1] != 2 or r.shape != (len(x),) or len(x) == 0:
        raise ValueError("positions must be (n,2) and radii (n,)")
    m = _lib.BedMember()
    m.e_star, m.g_star, m.damping = c.e_star, c.g_star, c.damping_coefficient
    m.friction, m.wall_friction, m.rho_pi_t = c.friction, c.wall_friction, c.rho_pi_t
    m.gravity_y = -c.gravity
    m.dt = float(s.dt if s.dt is not None else
                 (c.dt if c.dt is not None else c.stable_dt(float(r.min()))))
    m.floor_y, m.back_wall_x = c.floor_y, c.back_wall_x
    m.right_wall_x, m.right_wall_v = c.right_wall_x, c.right_wall_speed
    m.use_right_wall = int(c.right_wall)
    m.integrator = int(c.integrator == "verlet")
    m.verlet_start = 1          # the given velocities are full-step ones
    m.time, m.step_count = 0.0, 0
    r0 = float(r.mean())
    m.skin = skin_factor * r0
    m.cell_size = 2.0 * float(r.max()) + m.skin
    lo = x.min(0) - r.max()
    hi = x.max(0) + r.max()
    w, h = hi - lo
    span = max(w, h)
    x0 = min(lo[0], c.back_wall_x) - 2 * m.cell_size
    y0 = min(lo[1], c.floor_y) - 2 * m.cell_size
    m.grid_origin[0], m.grid_origin[1] = x0, y0
    m.ncx = int(math.ceil((hi[0] + runout * span - x0) / m.cell_size)) + 1
    m.ncy = int(math.ceil((hi[1] + 0.25 * span - y0) / m.cell_size)) + 1
    m.n = len(x)
    m.use_floor, m.use_back_wall = int(c.floor), int(c.back_wall)
    m.nblocks = lib.tdm_bed_blocks(len(x))
    m.mode, m.tab, m.list_parity, m.status = 1, 0, 0, 0
    f64 = dict(dtype=torch.float64, device=dev)
    v = np.zeros_like(x) if s.velocities is None else np.asarray(s.velocities, np.float64)
    return BedParts(member=m, n=len(x), nblocks=m.nblocks, ncells=m.ncx * m.ncy,
                    x0=torch.tensor(x, **f64), v0=torch.tensor(v, **f64),
                    radius=torch.tensor(r, **f64))


def cell_capacity(cell_size: float, r_min: float, floor: int = 8,
                  cell_size_y: float | None = None) -> int:
    """Entries per hash cell that no packing can overflow: disks of radius
    >= r_min whose pairwise overlaps stay below 2 % of r_min hold disjoint
    concentric disks of radius r = 0.99 r_min; those centred in an h x h_y
    cell lie inside the rectangle (h + 2 r) x (h_y + 2 r), at a packing
    density of at most pi / sqrt(12), so at most that area x pi/sqrt(12) /
    (pi r^2) of them fit; never below `floor`.  (Contact overlaps of a
    settled bed are of order 1e-4 of a radius; a cell that overflows anyway
    is flagged, never dropped silently.)"""
    hy = cell_size if cell_size_y is None else cell_size_y
    r = 0.99 * r_min
    area = (cell_size + 2.0 * r) * (hy + 2.0 * r)
    bound = int(math.floor(area / (math.sqrt(12.0) * r * r) + 1e-9))
    return max(int(floor), bound)


class BedEngine:
    def __init__(self, specs: list[BedSpec], *, device="cuda", fp32: bool = False,
                 skin_factor: float = 0.4, cand_slots: int = 16, cell_cap: int | None = None,
                 runout: float = 1.0, parts: list[BedParts] | None = None):
        """cell_cap None: the largest `cell_capacity` of the members (8 for
        radii in [0.8, 1.2] r0; 22 for the floe bed's widest spread, radii in
        [0.5, 1.5] r0)."""
        self.lib = _lib.load()
        self.device = torch.device(device)
        if self.device.type != "cuda" or not torch.cuda.is_available():
            raise ValueError("BedEngine runs on the ROCm device (device='cuda')")
        self.specs = list(specs)
        self.fp32 = bool(fp32)
        if parts is None:
            parts = [bed_parts(s, device=self.device, skin_factor=skin_factor, runout=runout)
                     for s in self.specs]
        if len(parts) != len(self.specs):
            raise ValueError("one BedParts per spec")
        nm = len(parts)
        members = (_lib.BedMember * nm)()
        n = blocks = cells = 0
        for k, pt in enumerate(parts):
            ctypes.memmove(ctypes.byref(members[k]), ctypes.byref(pt.member),
                           ctypes.sizeof(_lib.BedMember))
            m = members[k]
            m.first, m.block_first, m.cell_first = n, blocks, cells
            n += pt.n
            blocks += pt.nblocks
            cells += pt.ncells
        self.n, self.nmembers, self.nblocks, self.ncells = n, nm, blocks, cells
        if cell_cap is None:
            cell_cap = max(cell_capacity(pt.member.cell_size, float(pt.radius.min()))
                           for pt in parts)
        self.cand_slots, self.cell_cap = cand_slots, int(cell_cap)
        dev = self.device
        f64 = dict(dtype=torch.float64, device=dev)
        self.x = torch.cat([pt.x0 for pt in parts])
        self.v = torch.cat([pt.v0 for pt in parts])
        self.x_alt = torch.empty_like(self.x)
        self.v_alt = torch.empty_like(self.v)
        self.radius = torch.cat([pt.radius for pt in parts])
        # original index of the disk at each storage position (renumber)
        self.ids = torch.arange(n, dtype=torch.int64, device=dev)
        self.xi_wall = torch.zeros((n, 6), **f64)   # floor, back wall, right wall
        self.accounting = torch.zeros((nm, _lib.BED_NACC), **f64)
        self.energy = torch.zeros((nm, 4), **f64)
        self._members_host = members
        self._verlet = any(members[k].integrator for k in range(nm))
        self.members = torch.frombuffer(bytearray(bytes(members)), dtype=torch.uint8).to(dev)
        args = (n, blocks, nm, cells, cand_slots, cell_cap)
        wsb = self.lib.tdm_bed_workspace_bytes(*args)
        self.workspace = torch.zeros(int(wsb) + 256, dtype=torch.uint8, device=dev)
        offs = [ctypes.c_size_t() for _ in range(5)]
        _lib.check(self.lib.tdm_bed_workspace_layout(*args, *[ctypes.byref(o) for o in offs]))
        self._layout = [o.value for o in offs]
        self._arrays = _lib.BedArrays()
        self._refresh()
        _lib.check(self.lib.tdm_bed_setup(ctypes.byref(self._arrays), self.stream))
        self._graphs = {}

    @property
    def stream(self) -> int:
        return torch.cuda.current_stream(self.device).cuda_stream

    def _ws_base(self) -> int:
        base = self.workspace.data_ptr()
        return (base + 255) & ~255

    def _refresh(self) -> None:
        a = self._arrays
        a.x, a.v, a.x_alt, a.v_alt = (t.data_ptr() for t in (self.x, self.v, self.x_alt, self.v_alt))
        a.radius, a.xi_wall = self.radius.data_ptr(), self.xi_wall.data_ptr()
        a.accounting, a.energy = self.accounting.data_ptr(), self.energy.data_ptr()
        a.members = self.members.data_ptr()
        a.workspace = self._ws_base()
        a.workspace_bytes = self.workspace.numel() - (a.workspace - self.workspace.data_ptr())
        a.n, a.nmembers, a.nblocks, a.ncells = self.n, self.nmembers, self.nblocks, self.ncells
        a.cand_slots, a.cell_cap = self.cand_slots, self.cell_cap
        a.fp32_histories = int(self.fp32)
        a.verlet = int(self._verlet)

    def _swap(self) -> None:
        self.x, self.x_alt = self.x_alt, self.x
        self.v, self.v_alt = self.v_alt, self.v
        self._refresh()

    # ---- operations ------------------------------------------------------------
    def step(self, nsteps: int = 1) -> None:
        _lib.check(self.lib.tdm_bed_step(ctypes.byref(self._arrays), int(nsteps), int(self.fp32),
                                         self.stream))
        if nsteps % 2:
            self._swap()

    def step_timed(self, nsteps: int) -> float:
        out = ctypes.c_double()
        _lib.check(self.lib.tdm_bed_step_timed(ctypes.byref(self._arrays), int(nsteps),
                                               int(self.fp32), self.stream, ctypes.byref(out)))
        return out.value

    def rebin(self) -> None:
        _lib.check(self.lib.tdm_bed_rebin(ctypes.byref(self._arrays), self.stream))

    def full_velocities(self) -> torch.Tensor:
        """(n,2) full-step velocities of the current state: `v` for Euler
        members; for velocity-Verlet members v + (dt/2 F)/m with the force of
        the current state (the floe bed's `vel` after a step,
        scale_ceiling/kernel.py:417-418), springs and ledger untouched."""
        self._refresh()
        _lib.check(self.lib.tdm_bed_velocity(ctypes.byref(self._arrays), int(self.fp32),
                                             self.stream))
        return self.v_alt.clone()

    def energies(self) -> torch.Tensor:
        """(nmembers, 4): kinetic, gravitational, Hertz normal and tangential
        spring energy of the current state; the kinetic term of a
        velocity-Verlet member is that of its full-step velocity (the floe
        bed's `kinetic_energy`, scale_ceiling/kernel.py:455-457)."""
        arr = self._arrays
        if self._verlet:
            self._refresh()
            _lib.check(self.lib.tdm_bed_velocity(ctypes.byref(self._arrays), int(self.fp32),
                                                 self.stream))
            arr = _lib.BedArrays.from_buffer_copy(self._arrays)
            arr.v = self.v_alt.data_ptr()
        _lib.check(self.lib.tdm_bed_energy(ctypes.byref(arr), self.stream))
        return self.energy.clone()

    def graph(self, pairs: int):
        key = (pairs, self.fp32, self.x.data_ptr())
        if key not in self._graphs:
            g = ctypes.c_void_p()
            stream = torch.cuda.Stream(self.device)
            stream.wait_stream(torch.cuda.current_stream(self.device))
            with torch.cuda.stream(stream):
                _lib.check(self.lib.tdm_bed_graph_create(ctypes.byref(self._arrays), int(pairs),
                                                         int(self.fp32), stream.cuda_stream,
                                                         ctypes.byref(g)))
            torch.cuda.current_stream(self.device).wait_stream(stream)
            self._graphs[key] = g
        return self._graphs[key]

    def run_graph(self, pairs: int, repeats: int = 1) -> None:
        g = self.graph(pairs)
        for _ in range(repeats):
            _lib.check(self.lib.tdm_graph_launch(g, self.stream))

    def renumber(self) -> torch.Tensor:
        """Store every member's disks in Morton order of their current cells.

        As a collapse flows, disks that touch drift apart in memory and the
        step's partner gathers lose cache locality; renumbering restores it.
        All per-disk state is permuted in place (so captured graphs stay
        valid): positions, velocities, radii, wall springs, list origins and
        the current candidate lists with their spring histories (partners
        renumbered and re-sorted, histories negated where a pair's lo -> hi
        orientation flips); then the hash is rebuilt, so the next step
        searches new lists and carries the springs over.  `ids` maps storage
        positions to the caller's original disk indices.  Returns the
        permutation (new position -> old position)."""
        ms = self.read_members()
        dev = self.device
        perm = torch.empty(self.n, dtype=torch.int64, device=dev)
        for m in ms:
            sl = slice(m.first, m.first + m.n)
            perm[sl] = m.first + spatial_order(self.x[sl], m.grid_origin, m.cell_size,
                                               m.ncx, m.ncy)
        inv = torch.empty_like(perm)
        inv[perm] = torch.arange(self.n, dtype=torch.int64, device=dev)
        S = self.cand_slots
        cand = self._ws_view(0, torch.int32, (2, S, self.n))
        xi = self.hist_view()
        mask = self._ws_view(2, torch.int32, (2, self.n))
        count = self._ws_view(3, torch.uint8, (2, self.n))
        xb = self._ws_view(4, torch.float64, (self.n, 2))
        for m in ms:
            sl = slice(m.first, m.first + m.n)
            lp = m.list_parity
            c, h, mk, k = remap_lists(cand[lp][:, sl], xi[lp][:, sl], mask[lp][sl],
                                      count[lp][sl], perm[sl], inv, m.first)
            cand[lp][:, sl] = c
            xi[lp][:, sl] = h
            mask[lp][sl] = mk
            count[lp][sl] = k
        for t in (self.x, self.v, self.radius, self.xi_wall, xb, self.ids):
            t.copy_(t[perm])
        self.rebin()
        return perm

    def set_right_wall(self, member: int = 0, *, x=None, speed=None, enabled=None) -> None:
        """Move, re-time or switch member's right wall (the indenter, or the
        column's gate) between calls: from the member's current time on it
        stands at x and moves at `speed` along +x (scale_ceiling/kernel.py:
        514-519 sets the indenter velocity and advances it by v dt per step)."""
        ms = self.read_members()
        m = ms[member]
        t = m.time
        if speed is not None:
            if x is None:
                x = m.right_wall_x + m.right_wall_v * t
            m.right_wall_v = float(speed)
        if x is not None:
            m.right_wall_x = float(x) - m.right_wall_v * t
        if enabled is not None:
            m.use_right_wall = int(bool(enabled))
        if not (math.isfinite(m.right_wall_x) and math.isfinite(m.right_wall_v)):
            raise ValueError("right wall position and speed must be finite")
        self.write_members(ms)

    def right_wall_position(self, member: int = 0) -> float:
        m = self.read_members()[member]
        return m.right_wall_x + m.right_wall_v * m.time

    def read_members(self):
        raw = self.members.cpu().numpy().tobytes()
        return (_lib.BedMember * self.nmembers).from_buffer_copy(raw)

    def write_members(self, members) -> None:
        self.members.copy_(torch.frombuffer(bytearray(bytes(members)), dtype=torch.uint8))

    def status(self) -> np.ndarray:
        return np.array([m.status for m in self.read_members()], dtype=np.int32)

    def hist_view(self) -> torch.Tensor:
        """The spring histories of both list buffers, (2, slots, n, 2), in the
        precision the bed steps in (float32 with FP32 forces)."""
        dt = torch.float32 if self.fp32 else torch.float64
        return self._ws_view(1, dt, (2, self.cand_slots, self.n, 2))

    def _ws_view(self, which: int, dtype, shape):
        off = self._ws_base() - self.workspace.data_ptr() + self._layout[which]
        nbytes = int(np.prod(shape)) * torch.empty((), dtype=dtype).element_size()
        return self.workspace[off: off + nbytes].view(dtype).view(*shape)

    def list_stats(self) -> dict:
        """Mean / max candidate-list length and mean pair contacts per disk
        (both copies of a pair counted) of the current lists, over the batch."""
        ms = self.read_members()
        lp = torch.tensor([m.list_parity for m in ms], device=self.device)
        sel = torch.repeat_interleave(lp, torch.tensor([m.n for m in ms], device=self.device))
        idx = torch.arange(self.n, device=self.device)
        nc = self._ws_view(3, torch.uint8, (2, self.n))[sel, idx].float()
        mk = self._ws_view(2, torch.int32, (2, self.n))[sel, idx]
        pc = sum(((mk >> b) & 1).sum() for b in range(self.cand_slots)).item()
        return dict(cand_mean=nc.mean().item(), cand_max=int(nc.max().item()),
                    contacts_per_disk=pc / self.n)

    # ---- restart ----------------------------------------------------------------
    RESTART_SCHEMA = "tensordem_amd.bed_restart/1"

    def snapshot(self, member: int = 0) -> dict:
        """Complete state of one member, on the host: positions, velocities,
        radii, law, clock, ledger, every active pair spring history (pairs
        i < j with their xi, in the member's local indices) and the wall
        springs.  Restoring it continues the run bit for bit."""
        from dataclasses import asdict, replace
        ms = self.read_members()
        m = ms[member]
        sl = slice(m.first, m.first + m.n)
        S = self.cand_slots
        lp = m.list_parity
        cand = self._ws_view(0, torch.int32, (2, S, self.n))[lp][:, sl].cpu().numpy()
        xi = self.hist_view()[lp][:, sl].double().cpu().numpy()
        mask = self._ws_view(2, torch.int32, (2, self.n))[lp][sl].cpu().numpy().astype(np.int64)
        nc = self._ws_view(3, torch.uint8, (2, self.n))[lp][sl].cpu().numpy()
        p_loc, q_loc, h = [], [], []
        for e in range(S):
            act = (((mask >> e) & 1) == 1) & (e < nc)
            p = np.nonzero(act)[0]
            q = cand[e][act] - m.first
            keep = q > p
            p_loc.append(p[keep]); q_loc.append(q[keep]); h.append(xi[e][act][keep])
        pi, qi, hx = np.concatenate(p_loc), np.concatenate(q_loc), np.concatenate(h)
        order = np.lexsort((qi, pi))
        spec = self.specs[member]
        cfg = replace(spec.config, right_wall=bool(m.use_right_wall),
                      right_wall_x=float(m.right_wall_x), right_wall_speed=float(m.right_wall_v))
        return {
            "schema": self.RESTART_SCHEMA,
            "config": asdict(cfg), "dt": m.dt, "time": m.time,
            "step_count": int(m.step_count),
            "positions": self.x[sl].cpu().numpy().copy(),
            "velocities": self.v[sl].cpu().numpy().copy(),
            "radii": self.radius[sl].cpu().numpy().copy(),
            "accounting": self.accounting[member].cpu().numpy().copy(),
            "pairs": np.stack((pi[order], qi[order]), 1).astype(np.int64),
            "pair_xi": hx[order].copy(),
            "wall_xi": self.xi_wall[sl].cpu().numpy().copy(),
            "wall_active": ((mask >> 16) & 7).astype(np.uint8),
            # velocity-Verlet: 0 while `velocities` holds the half-step velocity
            "verlet_start": int(m.verlet_start),
        }

    @classmethod
    def from_snapshot(cls, state: dict, **kw) -> "BedEngine":
        """A one-member engine continuing `state` (from `snapshot`).  The
        spring histories are written into the current candidate lists, so the
        first step's list rebuild carries them over like any other."""
        validate_bed_restart(state)
        cfg = BedConfig(**state["config"])
        eng = cls([BedSpec(cfg, state["positions"], state["radii"], dt=state["dt"],
                           velocities=state["velocities"])], **kw)
        n, S, dev = eng.n, eng.cand_slots, eng.device
        pairs = np.asarray(state["pairs"], np.int64).reshape(-1, 2)
        hxi = np.asarray(state["pair_xi"], np.float64).reshape(-1, 2)
        disk = np.concatenate([pairs[:, 0], pairs[:, 1]])
        part = np.concatenate([pairs[:, 1], pairs[:, 0]])
        hh = np.concatenate([hxi, hxi])
        order = np.lexsort((part, disk))
        disk, part, hh = disk[order], part[order], hh[order]
        counts = np.bincount(disk, minlength=n)
        if counts.max(initial=0) > S:
            raise ValueError("restart holds more contacts per disk than candidate slots")
        slot = np.arange(len(disk)) - np.repeat(np.cumsum(counts) - counts, counts)
        cand = np.zeros((S, n), np.int32)
        xi = np.zeros((S, n, 2))
        mask = np.zeros(n, np.int64)
        cand[slot, disk] = part
        xi[slot, disk] = hh
        np.add.at(mask, disk, (1 << slot).astype(np.int64))
        mask |= np.asarray(state["wall_active"], np.int64) << 16
        eng._ws_view(0, torch.int32, (2, S, n))[0].copy_(torch.from_numpy(cand).to(dev))
        eng.hist_view()[0].copy_(torch.from_numpy(xi).to(dev))
        eng._ws_view(2, torch.int32, (2, n))[0].copy_(
            torch.from_numpy(mask.astype(np.uint32).view(np.int32)).to(dev))
        eng._ws_view(3, torch.uint8, (2, n))[0].copy_(
            torch.from_numpy(counts.astype(np.uint8)).to(dev))
        eng._ws_view(4, torch.float64, (n, 2)).copy_(eng.x)
        wxi = np.zeros((n, 6))
        old = np.asarray(state["wall_xi"], np.float64)
        wxi[:, :old.shape[1]] = old      # files of two-wall beds hold (n, 4)
        eng.xi_wall.copy_(torch.from_numpy(wxi).to(dev))
        acc = np.zeros(_lib.BED_NACC)
        old_acc = np.asarray(state["accounting"], np.float64)
        acc[:len(old_acc)] = old_acc     # files written before the wall-work channel
        eng.accounting[0].copy_(torch.from_numpy(acc).to(dev))
        ms = eng.read_members()
        ms[0].time, ms[0].step_count = float(state["time"]), int(state["step_count"])
        ms[0].verlet_start = int(state.get("verlet_start", 1))
        eng.write_members(ms)
        return eng

    def contact_pairs(self) -> np.ndarray:
        """Pairs (i<j) in contact during the last step, from the kernels' own
        per-slot contact masks (sorted)."""
        ms = self.read_members()
        S, n = self.cand_slots, self.n
        u8 = self.workspace
        base = self._ws_base() - self.workspace.data_ptr()
        out = []
        for m in ms:
            lp = m.list_parity
            co, xo, mo, no, _ = self._layout
            cand = u8[base + co: base + co + 2 * S * n * 4].view(torch.int32).view(2, S, n)[lp]
            mask = u8[base + mo: base + mo + 2 * n * 4].view(torch.int32).view(2, n)[lp]
            cnt = u8[base + no: base + no + 2 * n].view(2, n)[lp]
            sl = slice(m.first, m.first + m.n)
            c = cand[:, sl].cpu().numpy()
            mk = mask[sl].cpu().numpy().astype(np.int64)
            nc = cnt[sl].cpu().numpy()
            for e in range(S):
                sel = ((mk >> e) & 1).astype(bool) & (e < nc)
                p = np.nonzero(sel)[0] + m.first
                q = c[e][sel]
                keep = q > p
                out.append(np.stack((p[keep], q[keep]), axis=1))
        if not out:
            return np.zeros((0, 2), dtype=np.int64)
        pairs = np.concatenate(out)
        return pairs[np.lexsort((pairs[:, 1], pairs[:, 0]))]

    def close(self) -> None:
        for g in self._graphs.values():
            self.lib.tdm_graph_destroy(g)
        self._graphs.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def validate_bed_restart(state: dict) -> None:
    """Raise ValueError unless `state` is a complete, self-consistent bed
    restart: shapes agree, values are finite, pairs are i < j in range with
    no repeats, the clock is non-negative."""
    if not isinstance(state, dict) or state.get("schema") != BedEngine.RESTART_SCHEMA:
        raise ValueError("not a tensordem_amd bed restart state")
    need = {"schema", "config", "dt", "time", "step_count", "positions", "velocities", "radii",
            "accounting", "pairs", "pair_xi", "wall_xi", "wall_active"}
    if not need <= set(state) <= need | {"verlet_start"}:
        raise ValueError("bed restart fields missing or unknown")
    if "verlet_start" in state and state["verlet_start"] not in (0, 1):
        raise ValueError("invalid bed restart verlet_start")
    x = np.asarray(state["positions"], np.float64)
    n = len(x)
    shapes = {"positions": [(n, 2)], "velocities": [(n, 2)], "radii": [(n,)],
              "accounting": [(_lib.BED_NACC,), (_lib.BED_NACC - 1,)],
              "wall_xi": [(n, 6), (n, 4)], "wall_active": [(n,)]}
    for k, shp in shapes.items():
        a = np.asarray(state[k])
        if a.shape not in shp:
            raise ValueError(f"invalid bed restart {k} shape")
        if not np.isfinite(a.astype(np.float64)).all():
            raise ValueError(f"nonfinite bed restart {k}")
    if n == 0 or (np.asarray(state["radii"]) <= 0).any():
        raise ValueError("bed restart needs disks with positive radii")
    pairs = np.asarray(state["pairs"])
    xi = np.asarray(state["pair_xi"], np.float64)
    if pairs.ndim != 2 or pairs.shape[1] != 2 or xi.shape != (len(pairs), 2):
        raise ValueError("invalid bed restart pairs / pair_xi shape")
    if not np.isfinite(xi).all():
        raise ValueError("nonfinite bed restart pair_xi")
    if len(pairs):
        if pairs.min() < 0 or pairs.max() >= n or (pairs[:, 0] >= pairs[:, 1]).any():
            raise ValueError("bed restart pairs must be i < j within the bed")
        if len(np.unique(pairs[:, 0] * n + pairs[:, 1])) != len(pairs):
            raise ValueError("repeated bed restart pair")
    wa = np.asarray(state["wall_active"])
    if wa.dtype.kind not in "iu" or (wa.astype(np.int64) & ~7).any():
        raise ValueError("invalid bed restart wall_active")
    if np.asarray(state["wall_xi"]).shape[1] == 4 and (wa.astype(np.int64) & 4).any():
        raise ValueError("bed restart marks right wall contacts but stores no right wall springs")
    dt, t, sc = state["dt"], state["time"], state["step_count"]
    if not (math.isfinite(dt) and dt > 0 and math.isfinite(t) and t >= 0):
        raise ValueError("invalid bed restart time or timestep")
    if isinstance(sc, bool) or int(sc) != sc or sc < 0:
        raise ValueError("invalid bed restart step_count")


def save_bed_restart(state: dict, path) -> None:
    """Write a `BedEngine.snapshot` to an .npz file (arrays) with the scalars
    and the law as JSON inside it; read back by `load_bed_restart` without
    unpickling anything."""
    import json
    meta = {k: state[k] for k in ("schema", "config", "dt", "time", "step_count", "verlet_start")
            if k in state}
    arrays = {k: np.asarray(v) for k, v in state.items() if k not in meta}
    np.savez(path, meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8), **arrays)


def load_bed_restart(path) -> dict:
    import json
    with np.load(path, allow_pickle=False) as z:
        state = json.loads(bytes(z["meta"]).decode())
        for k in z.files:
            if k != "meta":
                state[k] = z[k]
    if state.get("schema") != BedEngine.RESTART_SCHEMA:
        raise ValueError(f"{path}: not a tensordem_amd bed restart file")
    validate_bed_restart(state)
    return state


def settled_column(n: int, r0: float = 0.01, *, seed: int = 0, config: BedConfig | None = None,
                   aspect: float = 2.0, fp32: bool = True, chunk: int = 2000,
                   max_steps: int = 200_000, rest: float = 1e-6, device="cuda",
                   log=None) -> tuple[dict, dict]:
    """The release state of a column collapse: a deposited random loose
    column (`column_deposit`) settled under gravity between the back wall,
    the floor and a gate at its width (a right wall at rest), on the device,
    until its kinetic energy is below `rest` times m g H summed over the
    disks (or `max_steps`).  Returns the bed restart state with the gate
    already removed, the clock at zero and an empty ledger: the column then
    stands as the reference's protocol releases it (granular2d/
    column_collapse.py:236-266, the gate is the absence of a right wall),
    with its contact springs; and a dict describing the settling."""
    from dataclasses import asdict, replace
    from .bedconfig import column_deposit
    cfg = config or BedConfig()
    x, r, width = column_deposit(n, r0, seed=seed, aspect=aspect)
    gated = replace(cfg, right_wall=True, right_wall_x=width, right_wall_speed=0.0)
    eng = BedEngine([BedSpec(gated, x, r)], fp32=fp32, device=device)
    height = float((x[:, 1] + r).max())
    m = gated.rho_pi_t * (r * r)
    scale = float((m * gated.gravity * height).sum())
    done = 0
    ke = math.inf
    while done < max_steps:
        eng.run_graph(chunk // 2)
        done += chunk
        if done % (5 * chunk) == 0:
            ke = float(eng.energies()[0, 0].item())
            if log:
                log(f"settle: {done} steps, kinetic {ke:.3e} J ({ke / scale:.2e} of m g H)")
            if ke < rest * scale:
                break
    if eng.status()[0]:
        raise RuntimeError(f"settling the column raised status bits {eng.status()[0]}")
    state = eng.snapshot()
    eng.close()
    # remove the gate: no right wall, no right-wall springs
    state["config"] = asdict(replace(BedConfig(**state["config"]), right_wall=False))
    state["wall_xi"] = np.asarray(state["wall_xi"]).copy()
    state["wall_xi"][:, 4:6] = 0.0
    state["wall_active"] = (np.asarray(state["wall_active"]) & 3).astype(np.uint8)
    state["time"], state["step_count"] = 0.0, 0
    state["accounting"] = np.zeros_like(np.asarray(state["accounting"]))
    info = dict(settle_steps=done, kinetic_J=ke, kinetic_rel=ke / scale, width_m=width,
                contact_pairs=len(state["pairs"]), height_m=height)
    return state, info

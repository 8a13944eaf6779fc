"""Continuous-batching request loop for plate and bed simulations.

Requests (one simulation each: a bonded plate under an indenter, or an
unbonded Hertz-Mindlin bed, for a given number of steps) arrive over time.
The server keeps two device-resident batches, one per kind, each a single
engine whose one launch per step advances all of its members.  At every
chunk boundary:

  * finished members leave and their results are taken;
  * waiting requests are admitted in arrival order while the batch has a
    free slot and the HBM budget covers the request's estimated footprint;
  * if membership changed, the batch is repacked: a new engine is built and
    every continuing member's state is copied into it on the device --
    positions, velocities, bond and fracture records, ledgers, clocks and,
    for beds, the candidate lists with their spring histories (partner
    indices shifted to the member's new place).  Results therefore do not
    depend on who else is in the batch: a request run in the mix is bit for
    bit the request run alone.
  * one chunk of steps is launched for each batch, the two batches on their
    own streams so that their kernels overlap.

Latency of a request = time of the chunk boundary at which it completed
minus its arrival time.
"""
from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass, replace

import numpy as np
import torch

from . import _lib
from .bed import BedEngine, BedSpec, bed_parts
from .bedconfig import BedConfig, column_packing
from .config import DEMConfig
from .engine import PlateEngine, PlateSpec, plate_parts


@dataclass
class Request:
    rid: int
    kind: str                 # "plate" or "bed"
    spec: object              # PlateSpec or BedSpec
    steps: int
    arrival: float = 0.0      # seconds after the window opens
    admitted: float | None = None
    finished: float | None = None
    result: dict | None = None

    @property
    def n(self) -> int:
        if self.kind == "plate":
            return self.spec.config.nx * self.spec.config.ny
        return len(self.spec.radii)

    @property
    def latency(self) -> float | None:
        return None if self.finished is None else self.finished - self.arrival


# ---------------------------------------------------------------------------
# request construction
# ---------------------------------------------------------------------------
def plate_request(rid: int, n: int, rng: np.random.Generator, steps: int, arrival: float = 0.0):
    """A plate of about n disks (ny ~ sqrt(n/8), a strip) with a swept tool
    speed and breaking strains; degree-bounded dt."""
    ny = int(min(max(round(math.sqrt(n / 8.0)), 4), 128))
    nx = max(2, int(round(n / ny)))
    cfg = DEMConfig(nx=nx, ny=ny, tool_speed=float(rng.uniform(0.2, 1.0)),
                    breaking_strain=float(rng.uniform(0.008, 0.015)),
                    shear_breaking_strain=float(rng.uniform(0.015, 0.03)))
    return Request(rid, "plate", PlateSpec(cfg, dt=cfg.degree_bound_dt()), steps, arrival)


def bed_request(rid: int, n: int, rng: np.random.Generator, steps: int, arrival: float = 0.0):
    """A loose column of n disks (seeded by the request id) with swept
    friction and restitution."""
    cfg = BedConfig(friction=float(rng.uniform(0.3, 0.7)),
                    restitution=float(rng.uniform(0.3, 0.8)))
    x, r = column_packing(n, 0.01, seed=rid)
    return Request(rid, "bed", BedSpec(cfg, x, r), steps, arrival)


def sweep_requests(*, rate: float = 50.0, window: float = 60.0, steps: int = 5000,
                   n_range=(1000, 200_000), seed: int = 0, max_requests: int | None = None):
    """Poisson arrivals at `rate` per second over `window` seconds; each
    request is a plate or a bed with equal probability, with a particle count
    log-uniform on n_range."""
    rng = np.random.default_rng(seed)
    out, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > window or (max_requests is not None and len(out) >= max_requests):
            break
        n = int(round(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
        make = plate_request if rng.random() < 0.5 else bed_request
        out.append(make(len(out), n, rng, steps, t))
    return out


def device_bytes(req: Request, cand_slots: int = 16) -> int:
    """Upper estimate of a request's device footprint (state, lists, grid)."""
    n = req.n
    if req.kind == "plate":
        c = req.spec.config
        per_disk = 5 * 16 + 2 + 16 + 16 + 16 + 4 + 8 + 4 * cand_slots + 1
        per_bond = 16 + 8 + 1 + 1 + 8 + 8 + 8
        w, h = (c.nx - 1) * 2 * c.radius, (c.ny - 1) * 2 * c.radius
        cell = 3 * c.radius
        m = 0.25 * max(w, h) + 2 * cell
        cells = (w / cell + 2 * m / cell + 2) * (h / cell + 2 * m / cell + 2)
        return int(n * per_disk + 4 * n * per_bond + cells * 2 * (4 + 8 * 4) + (1 << 20))
    r = req.spec.radii
    x = req.spec.positions
    per_disk = 4 * 16 + 8 + 32 + 2 * cand_slots * (4 + 16) + 8 + 2 + 16
    span = float(np.ptp(x, axis=0).max()) + 4 * float(r.max())
    cell = 2.0 * float(r.max())
    cells = (2.0 * span / cell + 8) * (1.25 * span / cell + 8)
    return int(n * per_disk + cells * 2 * (4 + 8 * 4) + (1 << 20))


# ---------------------------------------------------------------------------
# batches and repacking
# ---------------------------------------------------------------------------
def _ranges(firsts, counts, device):
    if not firsts:
        return torch.zeros(0, dtype=torch.int64, device=device)
    return torch.cat([torch.arange(f, f + c, device=device) for f, c in zip(firsts, counts)])


def _field(struct, name):
    f = getattr(struct, name)
    return f.offset, f.size


def _copy_fields(dst_members, src_members, dst_idx, src_idx, struct, names):
    """Device-to-device copy of member-record fields (no host round trip)."""
    sz = ctypes.sizeof(struct)
    d = dst_members.view(-1, sz)
    s_ = src_members.view(-1, sz)
    di = torch.tensor(dst_idx, device=d.device)
    si = torch.tensor(src_idx, device=d.device)
    for nm in names:
        off, size = _field(struct, nm)
        d[di, off:off + size] = s_[si, off:off + size]


def _zero_fields(members, idx, struct, names):
    sz = ctypes.sizeof(struct)
    d = members.view(-1, sz)
    di = torch.tensor(idx, device=d.device)
    for nm in names:
        off, size = _field(struct, nm)
        d[di, off:off + size] = 0


def _carry_plate(old: PlateEngine, old_members, eng: PlateEngine, members, keep) -> None:
    """Queue the continuing members' state copies (device only)."""
    om, nm = old._members_host, eng._members_host
    dev = eng.x.device
    si = [old_members.index(r) for r in keep]
    di = [members.index(r) for r in keep]
    cnt_d = [om[i].nx * om[i].ny for i in si]
    sd = _ranges([om[i].first for i in si], cnt_d, dev)
    dd = _ranges([nm[j].first for j in di], cnt_d, dev)
    cnt_b = [om[i].nbonds for i in si]
    sb = _ranges([om[i].bond_first for i in si], cnt_b, dev)
    db = _ranges([nm[j].bond_first for j in di], cnt_b, dev)
    eng.x[dd] = old.x[sd]
    eng.v[dd] = old.v[sd]
    for name in ("alive", "failure_mode", "failure_time_s", "failure_extension_m",
                 "failure_energy_J"):
        getattr(eng, name)[db] = getattr(old, name)[sb]
    ti, tj = torch.tensor(si, device=dev), torch.tensor(di, device=dev)
    eng.accounting[tj] = old.accounting[ti]
    eng.last[tj] = old.last[ti]
    _copy_fields(eng.members, old.members, di, si, _lib.Member, ("time", "step_count", "status"))
    eng.sync_links()
    eng.rebin()


def _carry_bed(old: BedEngine, old_members, eng: BedEngine, members, keep) -> None:
    """Queue the continuing members' state copies (device only): state, wall
    springs, ledger, clocks, and the current candidate lists with their spring
    histories and list origins (partner indices shifted to the new place)."""
    om, nm = old._members_host, eng._members_host
    dev = eng.x.device
    S = old.cand_slots
    si = [old_members.index(r) for r in keep]
    di = [members.index(r) for r in keep]
    cnt = [om[i].n for i in si]
    sd = _ranges([om[i].first for i in si], cnt, dev)
    dd = _ranges([nm[j].first for j in di], cnt, dev)
    cnt_t = torch.tensor(cnt, device=dev)
    shift = torch.tensor([nm[j].first - om[i].first for i, j in zip(si, di)], device=dev)
    sh = torch.repeat_interleave(shift, cnt_t).to(torch.int32)
    off, _ = _field(_lib.BedMember, "list_parity")
    sz = ctypes.sizeof(_lib.BedMember)
    lp_m = old.members.view(-1, sz)[:, off:off + 4].contiguous().view(torch.int32).view(-1)
    lp = torch.repeat_interleave(lp_m[torch.tensor(si, device=dev)].long(), cnt_t)
    eng.x[dd] = old.x[sd]
    eng.v[dd] = old.v[sd]
    eng.xi_wall[dd] = old.xi_wall[sd]
    n_old, n_new = old.n, eng.n
    oc = old._ws_view(0, torch.int32, (2, S, n_old))
    eng._ws_view(0, torch.int32, (2, S, n_new))[0][:, dd] = oc.permute(1, 0, 2)[:, lp, sd] + sh
    ox = old.hist_view()
    eng.hist_view()[0][:, dd] = ox.permute(1, 0, 2, 3)[:, lp, sd]
    eng._ws_view(2, torch.int32, (2, n_new))[0][dd] = old._ws_view(2, torch.int32, (2, n_old))[lp, sd]
    eng._ws_view(3, torch.uint8, (2, n_new))[0][dd] = old._ws_view(3, torch.uint8, (2, n_old))[lp, sd]
    eng._ws_view(4, torch.float64, (n_new, 2))[dd] = old._ws_view(4, torch.float64, (n_old, 2))[sd]
    ti, tj = torch.tensor(si, device=dev), torch.tensor(di, device=dev)
    eng.accounting[tj] = old.accounting[ti]
    _copy_fields(eng.members, old.members, di, si, _lib.BedMember,
                 ("time", "step_count", "status", "rebuilds", "verlet_start"))
    # the carried lists are current: no rebuild is pending for these members
    _zero_fields(eng.members, di, _lib.BedMember, ("mode", "tab", "list_parity"))


class _Batch:
    """One kind's device batch: the engine, its members (requests) and the
    steps each has done.  All device work of a round is queued on the batch
    stream; a new engine is built on a side stream meanwhile."""

    def __init__(self, kind: str, fp32: bool, device, keep_state):
        self.kind, self.fp32, self.device = kind, fp32, torch.device(device)
        self.keep_state = keep_state
        self.engine = None
        self.members: list[Request] = []
        self.done: dict[int, int] = {}
        self.parts: dict[int, object] = {}
        self.retired: list = []     # (engine, event): closed once the event has passed
        self.pending: list = []     # (requests, buffers, event): results in flight
        self._stream = self._prep = None

    @property
    def stream(self):
        if self._stream is None:
            self._stream = torch.cuda.Stream(self.device)
        return self._stream

    @property
    def prep(self):
        if self._prep is None:
            self._prep = torch.cuda.Stream(self.device)
        return self._prep

    def _parts_of(self, r: Request):
        if r.rid not in self.parts:
            with torch.cuda.stream(self.prep):
                self.parts[r.rid] = (plate_parts(r.spec, device=self.device)
                                     if self.kind == "plate"
                                     else bed_parts(r.spec, device=self.device))
        return self.parts[r.rid]

    # -- one scheduling round ----------------------------------------------------
    def round(self, join: list[Request], chunk: int, now: float):
        fin = [r for r in self.members if self.done[r.rid] >= r.steps]
        keep = [r for r in self.members if self.done[r.rid] < r.steps]
        if fin:
            self._extract(fin, now)
        if fin or join:
            self._repack(keep, join, chunk)
        if self.engine is not None:
            with torch.cuda.stream(self.stream):
                self.engine.run_graph(chunk // 2)
            for r in self.members:
                self.done[r.rid] = self.done.get(r.rid, 0) + chunk
        ev = torch.cuda.Event(enable_timing=True)
        ev.record(self.stream)
        self._close_retired()
        return ev

    def _extract(self, fin: list[Request], now: float) -> None:
        eng = self.engine
        hm = eng._members_host
        sz = ctypes.sizeof(type(hm[0]))
        bufs = []
        with torch.cuda.stream(self.stream):
            for r in fin:
                k = self.members.index(r)
                m = hm[k]
                n = m.nx * m.ny if self.kind == "plate" else m.n
                x = eng.x[m.first:m.first + n]
                vals = [x.sum().reshape(1), eng.accounting[k]]
                extra = {}
                keep = self.keep_state is True or (
                    not isinstance(self.keep_state, bool) and r.rid in self.keep_state)
                if self.kind == "plate":
                    alive = eng.alive[m.bond_first:m.bond_first + m.nbonds]
                    vals.insert(1, (~alive).sum().to(torch.float64).reshape(1))
                    if keep:
                        extra["alive"] = alive
                if keep:
                    extra["x"], extra["v"] = x, eng.v[m.first:m.first + n]
                dev_parts = {"scalars": torch.cat(vals),
                             "member": eng.members.view(-1, sz)[k].clone(), **extra}
                host = {key: torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
                        for key, t in dev_parts.items()}
                for key, t in dev_parts.items():
                    host[key].copy_(t, non_blocking=True)
                bufs.append((r, host, dev_parts))
        ev = torch.cuda.Event(enable_timing=True)
        ev.record(self.stream)
        for r in fin:
            del self.done[r.rid]
            self.parts.pop(r.rid, None)
        self.pending.append((bufs, ev, now))

    def collect(self, to_time) -> list[Request]:
        """Results whose copies have landed (to_time maps the event to the
        request's completion time)."""
        out, keep = [], []
        for bufs, ev, now in self.pending:
            if not ev.query():
                keep.append((bufs, ev, now))
                continue
            t = to_time(ev, now)
            for r, host, _ in bufs:
                st = _lib.Member if self.kind == "plate" else _lib.BedMember
                m = st.from_buffer_copy(host["member"].numpy().tobytes())
                sc = host["scalars"].numpy()
                res = {"x_sum": float(sc[0]), "time": m.time, "step_count": int(m.step_count),
                       "status": int(m.status)}
                if self.kind == "plate":
                    res["broken_bonds"] = int(sc[1])
                    res["accounting"] = sc[2:].copy()
                else:
                    res["rebuilds"] = int(m.rebuilds)
                    res["accounting"] = sc[1:].copy()
                for key in ("x", "v", "alive"):
                    if key in host:
                        res[key] = host[key].numpy().copy()
                r.result, r.finished = res, t
                out.append(r)
        self.pending = keep
        return out

    def _repack(self, keep: list[Request], join: list[Request], chunk: int) -> None:
        old, old_members = self.engine, self.members
        members = keep + join
        eng = None
        if members:
            parts = [self._parts_of(r) for r in members]
            with torch.cuda.stream(self.prep):
                if self.kind == "plate":
                    eng = PlateEngine([r.spec for r in members], device=self.device,
                                      fp32=self.fp32, parts=parts)
                else:
                    eng = BedEngine([r.spec for r in members], device=self.device,
                                    fp32=self.fp32, parts=parts)
                eng.graph(chunk // 2)
            self.stream.wait_stream(self.prep)
            if old is not None and keep:
                with torch.cuda.stream(self.stream):
                    if self.kind == "plate":
                        _carry_plate(old, old_members, eng, members, keep)
                    else:
                        _carry_bed(old, old_members, eng, members, keep)
        if old is not None:
            ev = torch.cuda.Event()
            ev.record(self.stream)
            self.retired.append((old, ev))
        self.engine, self.members = eng, members

    def _close_retired(self, wait: bool = False) -> None:
        keep = []
        for eng, ev in self.retired:
            if wait:
                ev.synchronize()
            if ev.query():
                eng.close()
            else:
                keep.append((eng, ev))
        self.retired = keep


# ---------------------------------------------------------------------------
# the loop
# ---------------------------------------------------------------------------
class Server:
    """Continuous batching over two device batches (plates, beds).

    hbm_budget: bytes the admitted requests' estimated footprints may use;
    max_active: members per batch; chunk: steps between scheduling points
    (every request's step count must be a multiple of it); keep_state: True
    copies every finished request's positions, velocities (and alive mask)
    into its result, a set of request ids only those requests'.  The host
    runs at most one round ahead of the device."""

    def __init__(self, *, hbm_budget: float = 288e9, max_active: int = 64, chunk: int = 250,
                 fp32: bool = True, device="cuda", keep_state=False):
        if chunk <= 0 or chunk % 2:
            raise ValueError("chunk must be a positive even number of steps")
        self.hbm_budget, self.max_active, self.chunk = float(hbm_budget), int(max_active), chunk
        self.batches = {k: _Batch(k, fp32, device, keep_state) for k in ("plate", "bed")}
        self.used = 0.0
        self.footprint: dict[int, int] = {}
        self.log: list[tuple] = []   # (time, active plates, active beds, waiting)
        self.host_seconds = 0.0      # host time spent issuing rounds (builds, copies, launches)
        self.rounds = 0

    def _admit(self, waiting: list[Request], now: float) -> dict:
        join = {"plate": [], "bed": []}
        while waiting:
            r = waiting[0]
            b = self.batches[r.kind]
            active = len([m for m in b.members if b.done.get(m.rid, 0) < m.steps]) + \
                len(join[r.kind])
            need = device_bytes(r)
            if need > self.hbm_budget:
                raise ValueError(f"request {r.rid} needs {need} B, more than the HBM budget")
            if active >= self.max_active or self.used + need > self.hbm_budget:
                break
            waiting.pop(0)
            self.used += need
            self.footprint[r.rid] = need
            r.admitted = now
            join[r.kind].append(r)
        return join

    def run(self, requests: list[Request], *, realtime: bool = True) -> list[Request]:
        """Serve `requests` until all are done.  realtime: arrivals are in
        seconds of wall time and completion times come from device events;
        otherwise arrivals and completions count scheduling rounds."""
        for r in requests:
            if r.steps <= 0 or r.steps % self.chunk:
                raise ValueError(f"request {r.rid}: steps must be a positive multiple of "
                                 f"{self.chunk}")
        pending = sorted(requests, key=lambda r: r.arrival)
        waiting: list[Request] = []
        done: list[Request] = []
        batches = list(self.batches.values())
        t0 = time.perf_counter()
        refs = {}
        for b in batches:
            ev = torch.cuda.Event(enable_timing=True)
            ev.record(b.stream)
            refs[id(b)] = ev
        torch.cuda.synchronize()
        t_ref = time.perf_counter() - t0
        rounds = 0

        def clock():
            return time.perf_counter() - t0 if realtime else float(rounds)

        def to_time_of(b):
            if not realtime:
                return lambda ev, now: now
            ref = refs[id(b)]
            return lambda ev, now: t_ref + ref.elapsed_time(ev) * 1e-3

        def busy():
            return any(b.members or b.pending for b in batches)

        prev = []
        while pending or waiting or busy():
            now = clock()
            while pending and pending[0].arrival <= now:
                waiting.append(pending.pop(0))
            if not waiting and not any(b.members for b in batches):
                if any(b.pending for b in batches):
                    for ev in prev:
                        ev.synchronize()
                elif pending:
                    if realtime:
                        time.sleep(max(0.0, pending[0].arrival - clock()))
                    else:
                        rounds = int(math.ceil(pending[0].arrival))
                    continue
            join = self._admit(waiting, now)
            t_r = time.perf_counter()
            evs = [b.round(join[b.kind], self.chunk, now) for b in batches]
            self.host_seconds += time.perf_counter() - t_r
            for ev in prev:          # the host stays at most one round ahead
                ev.synchronize()
            prev = evs
            rounds += 1
            for b in batches:
                for r in b.collect(to_time_of(b)):
                    self.used -= self.footprint.pop(r.rid)
                    done.append(r)
            self.log.append((clock(), len(self.batches["plate"].done),
                             len(self.batches["bed"].done), len(waiting)))
        torch.cuda.synchronize()
        for b in batches:
            for r in b.collect(to_time_of(b)):
                self.used -= self.footprint.pop(r.rid)
                done.append(r)
            b._close_retired(wait=True)
        self.rounds = rounds
        return done


def run_alone(req: Request, *, fp32: bool = True, device="cuda", chunk: int = 250) -> dict:
    """The same request in a batch of its own (reference for mix tests)."""
    r = replace(req, admitted=None, finished=None, result=None)
    srv = Server(fp32=fp32, device=device, chunk=chunk, keep_state=True)
    return srv.run([r], realtime=False)[0].result


def latency_summary(done: list[Request]) -> dict:
    lat = np.array([r.latency for r in done if r.latency is not None])
    if len(lat) == 0:
        return {}
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max()), "completed": int(len(lat))}


__all__ = ["Request", "Server", "sweep_requests", "plate_request", "bed_request", "device_bytes",
           "run_alone", "latency_summary", "_lib"]

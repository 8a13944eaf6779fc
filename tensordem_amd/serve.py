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

import math
import time
from dataclasses import dataclass, replace

import numpy as np
import torch

from . import _lib
from .bed import BedEngine, BedSpec
from .bedconfig import BedConfig, column_packing
from .config import DEMConfig
from .engine import PlateEngine, PlateSpec


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


class _Batch:
    """One kind's device batch: the engine, its members (requests) and the
    steps each has done."""

    def __init__(self, kind: str, fp32: bool, device):
        self.kind, self.fp32, self.device = kind, fp32, torch.device(device)
        self.engine = None
        self.members: list[Request] = []
        self.done: dict[int, int] = {}
        self._stream = None

    @property
    def stream(self):
        if self._stream is None:
            self._stream = torch.cuda.Stream(self.device)
        return self._stream

    # -- repack ----------------------------------------------------------------
    def repack(self, keep: list[Request], join: list[Request]) -> None:
        old, old_members = self.engine, self.members
        members = keep + join
        if not members:
            if old is not None:
                old.close()
            self.engine, self.members = None, []
            return
        with torch.cuda.stream(self.stream):
            if self.kind == "plate":
                eng = PlateEngine([r.spec for r in members], device=self.device, fp32=self.fp32)
                if old is not None and keep:
                    _carry_plate(old, old_members, eng, members, keep)
            else:
                eng = BedEngine([r.spec for r in members], device=self.device, fp32=self.fp32)
                if old is not None and keep:
                    _carry_bed(old, old_members, eng, members, keep)
        if old is not None:
            self.stream.synchronize()
            old.close()
        self.engine, self.members = eng, members

    def launch(self, steps: int) -> None:
        if self.engine is None:
            return
        with torch.cuda.stream(self.stream):
            self.engine.run_graph(steps // 2)
        for r in self.members:
            self.done[r.rid] = self.done.get(r.rid, 0) + steps

    def harvest(self, now: float, keep_state: bool) -> list[Request]:
        """Requests that reached their step count: record results."""
        fin = [r for r in self.members if self.done[r.rid] >= r.steps]
        if not fin:
            return []
        eng = self.engine
        ms = eng.read_members()
        for r in fin:
            k = self.members.index(r)
            m = ms[k]
            if self.kind == "plate":
                n = m.nx * m.ny
                x = eng.x[m.first:m.first + n]
                alive = eng.alive[m.bond_first:m.bond_first + m.nbonds]
                res = {"broken_bonds": int((~alive).sum()), "time": m.time,
                       "step_count": int(m.step_count), "status": int(m.status),
                       "accounting": eng.accounting[k].cpu().numpy().copy()}
                if keep_state:
                    res["alive"] = alive.cpu().numpy().copy()
            else:
                x = eng.x[m.first:m.first + m.n]
                res = {"time": m.time, "step_count": int(m.step_count), "status": int(m.status),
                       "rebuilds": int(m.rebuilds),
                       "accounting": eng.accounting[k].cpu().numpy().copy()}
            xs = x.cpu().numpy()
            res["x_sum"] = float(xs.sum())
            if keep_state:
                res["x"] = xs.copy()
                res["v"] = (eng.v[m.first:m.first + len(xs)]).cpu().numpy().copy()
            r.result, r.finished = res, now
            del self.done[r.rid]
        return fin


def _carry_plate(old: PlateEngine, old_members, eng: PlateEngine, members, keep) -> None:
    om, nm = old.read_members(), eng.read_members()
    dev = eng.x.device
    src_d, dst_d, cnt_d, src_b, dst_b, cnt_b = [], [], [], [], [], []
    for r in keep:
        i, j = old_members.index(r), members.index(r)
        n = om[i].nx * om[i].ny
        src_d.append(om[i].first); dst_d.append(nm[j].first); cnt_d.append(n)
        src_b.append(om[i].bond_first); dst_b.append(nm[j].bond_first); cnt_b.append(om[i].nbonds)
        nm[j].time, nm[j].step_count, nm[j].status = om[i].time, om[i].step_count, om[i].status
        eng.accounting[j].copy_(old.accounting[i])
        eng.last[j].copy_(old.last[i])
    sd, dd = _ranges(src_d, cnt_d, dev), _ranges(dst_d, cnt_d, dev)
    sb, db = _ranges(src_b, cnt_b, dev), _ranges(dst_b, cnt_b, dev)
    eng.x[dd] = old.x[sd]
    eng.v[dd] = old.v[sd]
    for name in ("alive", "failure_mode", "failure_time_s", "failure_extension_m",
                 "failure_energy_J"):
        getattr(eng, name)[db] = getattr(old, name)[sb]
    eng.write_members(nm)
    eng.sync_links()
    eng.rebin()


def _carry_bed(old: BedEngine, old_members, eng: BedEngine, members, keep) -> None:
    om, nm = old.read_members(), eng.read_members()
    dev = eng.x.device
    S = old.cand_slots
    src, dst, cnt, shift, lps = [], [], [], [], []
    for r in keep:
        i, j = old_members.index(r), members.index(r)
        src.append(om[i].first); dst.append(nm[j].first); cnt.append(om[i].n)
        shift.append(nm[j].first - om[i].first)
        lps.append(om[i].list_parity)
        m = nm[j]
        m.time, m.step_count, m.status, m.rebuilds = (om[i].time, om[i].step_count, om[i].status,
                                                      om[i].rebuilds)
        # continuing lists are valid: no rebuild is pending for this member
        m.mode, m.tab, m.list_parity = 0, 0, 0
        eng.accounting[j].copy_(old.accounting[i])
    sd, dd = _ranges(src, cnt, dev), _ranges(dst, cnt, dev)
    sh = torch.repeat_interleave(torch.tensor(shift, device=dev), torch.tensor(cnt, device=dev))
    lp = torch.repeat_interleave(torch.tensor(lps, device=dev), torch.tensor(cnt, device=dev))
    eng.x[dd] = old.x[sd]
    eng.v[dd] = old.v[sd]
    eng.xi_wall[dd] = old.xi_wall[sd]
    n_old, n_new = old.n, eng.n
    oc = old._ws_view(0, torch.int32, (2, S, n_old))
    nc_ = eng._ws_view(0, torch.int32, (2, S, n_new))
    nc_[0][:, dd] = oc.permute(1, 0, 2)[:, lp, sd] + sh.to(torch.int32)
    ox = old._ws_view(1, torch.float64, (2, S, n_old, 2))
    nx_ = eng._ws_view(1, torch.float64, (2, S, n_new, 2))
    nx_[0][:, dd] = ox.permute(1, 0, 2, 3)[:, lp, sd]
    om_ = old._ws_view(2, torch.int32, (2, n_old))
    eng._ws_view(2, torch.int32, (2, n_new))[0][dd] = om_[lp, sd]
    on_ = old._ws_view(3, torch.uint8, (2, n_old))
    eng._ws_view(3, torch.uint8, (2, n_new))[0][dd] = on_[lp, sd]
    ob = old._ws_view(4, torch.float64, (n_old, 2))
    eng._ws_view(4, torch.float64, (n_new, 2))[dd] = ob[sd]
    eng.write_members(nm)


# ---------------------------------------------------------------------------
# the loop
# ---------------------------------------------------------------------------
class Server:
    """Continuous batching over two device batches (plates, beds).

    hbm_budget: bytes the admitted requests' estimated footprints may use;
    max_active: members per batch; chunk: steps between scheduling points
    (every request's step count must be a multiple of it)."""

    def __init__(self, *, hbm_budget: float = 288e9, max_active: int = 64, chunk: int = 250,
                 fp32: bool = True, device="cuda", keep_state: bool = False):
        if chunk <= 0 or chunk % 2:
            raise ValueError("chunk must be a positive even number of steps")
        self.hbm_budget, self.max_active, self.chunk = float(hbm_budget), int(max_active), chunk
        self.keep_state = keep_state
        self.batches = {k: _Batch(k, fp32, device) for k in ("plate", "bed")}
        self.used = 0.0
        self.footprint: dict[int, int] = {}
        self.log: list[tuple] = []   # (time, active plates, active beds, waiting)

    def _admit(self, waiting: list[Request], now: float) -> dict:
        join = {"plate": [], "bed": []}
        while waiting:
            r = waiting[0]
            b = self.batches[r.kind]
            active = len(b.members) + len(join[r.kind])
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
        seconds of wall time; otherwise they count scheduling rounds (one
        chunk each) and latencies are in rounds."""
        for r in requests:
            if r.steps <= 0 or r.steps % self.chunk:
                raise ValueError(f"request {r.rid}: steps must be a positive multiple of "
                                 f"{self.chunk}")
        pending = sorted(requests, key=lambda r: r.arrival)
        waiting: list[Request] = []
        done: list[Request] = []
        t0 = time.perf_counter()
        rounds = 0
        clock = (lambda: time.perf_counter() - t0) if realtime else (lambda: float(rounds))
        while pending or waiting or any(b.members for b in self.batches.values()):
            now = clock()
            while pending and pending[0].arrival <= now:
                waiting.append(pending.pop(0))
            idle = not any(b.members for b in self.batches.values())
            if idle and not waiting and pending:
                if realtime:
                    time.sleep(max(0.0, pending[0].arrival - clock()))
                else:
                    rounds = int(math.ceil(pending[0].arrival))
                continue
            join = self._admit(waiting, now)
            for kind, b in self.batches.items():
                if join[kind] or len(b.members) != len(b.done):
                    keep = [r for r in b.members if r.rid in b.done]
                    b.repack(keep, join[kind])
            for b in self.batches.values():
                b.launch(self.chunk)
            for b in self.batches.values():
                if b.engine is not None:
                    b.stream.synchronize()
            rounds += 1
            now = clock()
            for b in self.batches.values():
                for r in b.harvest(now, self.keep_state):
                    self.used -= self.footprint.pop(r.rid)
                    done.append(r)
            self.log.append((now, len(self.batches["plate"].members),
                             len(self.batches["bed"].members), len(waiting)))
        for b in self.batches.values():
            b.repack([], [])
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

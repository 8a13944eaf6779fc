"""Benchmark: particle-steps/s of the FP32-force / FP64-state HIP steps.

Headline workload (`value`): the unbonded granular column collapse of
1 048 576 disks (radii U[0.8, 1.2] x 0.01 m, Hertz-Mindlin law with
E = 1 GPa, nu = 0.3, mu = 0.5, e = 0.5, ice density, aspect-2 loose column
from seed 0, floor and back wall, gravity), timed over --steps steps from
release (default 50 000) after --warmup untimed steps on a throw-away engine;
the FP32-force / FP64-integrate target is stated at this 1M-disk size.
Steps are replayed from captured graphs; inputs are resident in HBM before
the timed region.

Reported beside the kernel rate:
  * plain_fp64_gpu: the same step written as eager PyTorch FP64 on the same
    GPU (sort-based cell list, the reference's ops), timed on a short run;
  * fp64_kernel: the HIP step with FP64 forces;
  * roofline: algorithmic bytes per launch / mean step-kernel time (HIP
    events attached to each dispatch) against the 8 TB/s HBM peak, with the
    counter-measured traffic from the committed profile;
  * cpu_baseline: the C oracle on one host core over a bounded sample;
  * c2_plate: the large single ice plate, a 1024 x 128 reference lattice
    (131 072 disks, 520 834 bonds), both ends clamped, circular indenter of
    radius 0.5 m moving down at 0.5 m/s, strains 0.01 / 0.02, degree-bounded
    dt (the reference's all-pairs bound would give 5.5e-8 s), 20 000 steps;
  * c4_ensemble: 512 independent plates in one launch per step;
  * c5_serve: the continuous-batching request loop over a Poisson window;
  * disk_system: rigid frictional disks with rotation (the granular2d engine),
    1 048 576 disks in a walled box under gravity, FP64.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "particle-steps/s FP32-force vs FP64 plain on 1 GPU; p50/p99 batched solve latency"
HBM_PEAK_GBS = 8000.0
_T0 = time.perf_counter()


def progress(msg: str) -> None:
    """One line on stderr as each phase of the run ends."""
    print(f"bench.py [{time.perf_counter() - _T0:7.1f} s]: {msg}", file=sys.stderr, flush=True)
# compulsory bytes per disk per step of plate_step_kernel: x, v read (32) and
# written (32), link bits read+written (2), fixed flag (1), candidate count (1),
# list-build offset xb (8)
PLATE_BYTES_PER_DISK = 76


def c2_config(device="cuda"):
    from tensordem_amd import DEMConfig
    return DEMConfig(nx=1024, ny=128, radius=0.025, tool_radius=0.5, tool_speed=0.5,
                     breaking_strain=0.01, shear_breaking_strain=0.02, device=device)


# ---------------------------------------------------------------------------
# plain FP64 eager PyTorch step (the speedup denominator)
# ---------------------------------------------------------------------------
class PlainFP64:
    """IceDEM.step (src/tensordem/dem.py:226-382) as eager FP64 PyTorch ops on
    the GPU, with a per-step sort-based cell list instead of all N^2 pairs.
    Ledger sums stay on the device (no host syncs)."""

    def __init__(self, cfg, dt, device="cuda", pinv=True, eps_t_disk=None, eps_s_disk=None):
        import torch
        from tensordem_amd.lattice import bonds, fixed_mask, positions
        dev = torch.device(device)
        self.cfg, self.dt, self.pinv = cfg, dt, pinv
        f64 = dict(dtype=torch.float64, device=dev)
        self.x0 = torch.tensor(positions(cfg.nx, cfg.ny, cfg.radius), **f64)
        self.x = self.x0.clone()
        self.v = torch.zeros_like(self.x)
        self.fixed = torch.tensor(fixed_mask(cfg.nx, cfg.ny, cfg.fix_edges), device=dev)
        bi, bj, rest, _ = bonds(cfg.nx, cfg.ny, cfg.radius)
        self.i = torch.tensor(bi, device=dev)
        self.j = torch.tensor(bj, device=dev)
        self.rest = torch.tensor(rest, **f64)
        self.alive = torch.ones(len(bi), dtype=torch.bool, device=dev)
        # per-disk strengths: a bond breaks at its weaker disk
        self.eps_t, self.eps_s = cfg.breaking_strain, cfg.shear_breaking_strain
        if eps_t_disk is not None:
            et = torch.tensor(eps_t_disk, **f64)
            self.eps_t = torch.minimum(et[self.i], et[self.j])
        if eps_s_disk is not None:
            es = torch.tensor(eps_s_disk, **f64)
            self.eps_s = torch.minimum(es[self.i], es[self.j])
        self.refv = self.x0[self.j] - self.x0[self.i]
        n = len(self.x)
        self.n = n
        self.bond_keys = torch.sort(self.i * n + self.j).values
        self.time = torch.zeros((), **f64)
        self.acc = torch.zeros(5, **f64)
        self.tool_start = torch.tensor([0.0, (cfg.ny - 1) * 2 * cfg.radius + cfg.radius
                                        + cfg.tool_radius + cfg.tool_gap], **f64)
        self.tool_v = torch.tensor([0.0, -cfg.tool_speed], **f64)

    def _contacts(self):
        import torch
        c, x, n = self.cfg, self.x, self.n
        h = 2 * c.radius
        cell = torch.floor(x / h).long()
        cell -= cell.min(0).values
        ncx = int(cell[:, 0].max()) + 3  # host sync, as a plain implementation would
        key = (cell[:, 1] + 1) * ncx + cell[:, 0] + 1
        order = torch.argsort(key)
        skey = key[order]
        pi, pj = [], []
        for oy in (-1, 0, 1):
            for ox in (-1, 0, 1):
                nk = key + oy * ncx + ox
                lo = torch.searchsorted(skey, nk)
                hi = torch.searchsorted(skey, nk, right=True)
                cnt = hi - lo
                src = torch.repeat_interleave(torch.arange(n, device=x.device), cnt)
                start = torch.repeat_interleave(lo, cnt)
                off = torch.arange(len(src), device=x.device) - torch.repeat_interleave(
                    torch.cumsum(cnt, 0) - cnt, cnt)
                dst = order[start + off]
                pi.append(src)
                pj.append(dst)
        a, b = torch.cat(pi), torch.cat(pj)
        keep = a < b
        a, b = a[keep], b[keep]
        d = torch.linalg.vector_norm(x[b] - x[a], dim=1)
        keep = d < 2 * c.radius
        a, b = a[keep], b[keep]
        bonded = torch.isin(a * n + b, self.bond_keys)
        return a[~bonded], b[~bonded]

    def step(self):
        import torch
        c, x, v, i, j = self.cfg, self.x, self.v, self.i, self.j
        n = self.n
        # shear strain (dem.py:182-208)
        cur = x[j] - x[i]
        ro = self.refv[:, :, None] * self.refv[:, None, :]
        cr = cur[:, :, None] * self.refv[:, None, :]
        cov = torch.zeros((n, 2, 2), dtype=x.dtype, device=x.device)
        crs = torch.zeros_like(cov)
        cov.index_add_(0, i, ro)
        cov.index_add_(0, j, ro)
        crs.index_add_(0, i, cr)
        crs.index_add_(0, j, cr)
        if self.pinv:
            inv = torch.linalg.pinv(cov)  # the reference's op (dem.py:202)
        else:
            det = cov[:, 0, 0] * cov[:, 1, 1] - cov[:, 0, 1] * cov[:, 1, 0]
            inv = torch.stack((torch.stack((cov[:, 1, 1], -cov[:, 0, 1]), 1),
                               torch.stack((-cov[:, 1, 0], cov[:, 0, 0]), 1)), 1) / det[:, None, None]
        F = crs @ inv
        E = 0.5 * (F.transpose(1, 2) @ F - torch.eye(2, dtype=x.dtype, device=x.device))
        s = torch.sqrt((E[:, 0, 0] - E[:, 1, 1]).square() + 4.0 * E[:, 0, 1].square())
        # fracture (dem.py:238-277)
        dist = torch.linalg.vector_norm(cur, dim=1)
        nrm = cur / dist.clamp_min(torch.finfo(x.dtype).eps)[:, None]
        ext = dist - self.rest
        fail = ((ext / self.rest) > self.eps_t) | (0.5 * (s[i] + s[j]) > self.eps_s)
        newly = self.alive & fail
        self.acc[0] += (0.5 * c.bond_stiffness * ext[newly].square()).sum()
        self.alive &= ~fail
        # bond + broken-bond contact forces
        ov = (2 * c.radius - dist).clamp_min(0)
        vn = ((v[j] - v[i]) * nrm).sum(1)
        rep = (c.contact_stiffness * ov - c.contact_damping * vn).clamp_min(0)
        rep = torch.where((ov > 0) & ~self.alive, rep, torch.zeros_like(rep))
        mag = torch.where(self.alive, c.bond_stiffness * ext, -rep)
        pf = mag[:, None] * nrm
        force = -c.drag * v
        force.index_add_(0, i, pf)
        force.index_add_(0, j, -pf)
        # unbonded contacts
        a, b = self._contacts()
        d = x[b] - x[a]
        dd = torch.linalg.vector_norm(d, dim=1)
        nn = d / dd.clamp_min(torch.finfo(x.dtype).eps)[:, None]
        ovc = 2 * c.radius - dd
        vnc = ((v[b] - v[a]) * nn).sum(1)
        repc = (c.contact_stiffness * ovc - c.contact_damping * vnc).clamp_min(0)
        cf = -repc[:, None] * nn
        force.index_add_(0, a, cf)
        force.index_add_(0, b, -cf)
        # tool (dem.py:294-312)
        tp = self.tool_start + self.time * self.tool_v
        td = x - tp
        tdist = torch.linalg.vector_norm(td, dim=1)
        tn = td / tdist.clamp_min(torch.finfo(x.dtype).eps)[:, None]
        tov = (c.radius + c.tool_radius - tdist).clamp_min(0)
        tsp = ((v - self.tool_v) * tn).sum(1)
        tmag = torch.where(tov > 0, (c.contact_stiffness * tov - c.contact_damping * tsp).clamp_min(0),
                           torch.zeros_like(tov))
        tf = tmag[:, None] * tn
        force += tf
        self.acc[1] += self.dt * (tf.sum(0) @ self.tool_v)
        # integrate (dem.py:363-371)
        self.v += self.dt * force / c.mass
        self.v[self.fixed] = 0
        self.x += self.dt * self.v
        self.x[self.fixed] = self.x0[self.fixed]
        self.time += self.dt


def time_plain(cfg, dt, steps, warmup, pinv=True):
    import torch
    p = PlainFP64(cfg, dt, pinv=pinv)
    for _ in range(warmup):
        p.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        p.step()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


# ---------------------------------------------------------------------------
# plain FP64 eager PyTorch bed step (the C3 speedup denominator)
# ---------------------------------------------------------------------------
class PlainBedFP64:
    """The Hertz-Mindlin bed step as eager FP64 PyTorch ops on the GPU, the way
    the reference's floe bed runs it (src/tensordem/scale_ceiling/kernel.py:
    250-260, 277-401; search.py:60-162): a Verlet list of pairs within
    r_i + r_j + skin from a sort-based cell list (stable argsort, bincount
    starts, five half-stencil offsets), rebuilt when some disk has moved more
    than skin/4 (infinity norm) since the last build; spring history carried
    over by searchsorted on sorted pair keys i*n+j (granular2d/engine.py:
    388-396); forces scattered with index_add_; semi-implicit Euler.  Same
    law, constants and wall handling as the HIP kernel.  Ledger sums stay on
    the device; the rebuild test is the one host sync per step."""

    def __init__(self, cfg, x, r, dt, skin, v=None, device="cuda"):
        import torch
        dev = torch.device(device)
        f64 = dict(dtype=torch.float64, device=dev)
        self.cfg, self.dt, self.skin = cfg, float(dt), float(skin)
        self.x = torch.as_tensor(x, **f64).clone()
        self.v = torch.zeros_like(self.x) if v is None else torch.as_tensor(v, **f64).clone()
        self.r = torch.as_tensor(r, **f64)
        self.n = len(self.r)
        self.m = cfg.rho_pi_t * (self.r * self.r)
        self.cell = 2.0 * float(self.r.max()) + self.skin
        self.keys = torch.zeros(0, dtype=torch.int64, device=dev)
        self.xi = torch.zeros((0, 2), **f64)
        self.xi_wall = torch.zeros((2, self.n, 2), **f64)
        self.acc = torch.zeros(6, **f64)
        self.rebuilds = 0
        self._build()

    def _build(self):
        import torch
        x, r, n, h = self.x, self.r, self.n, self.cell
        origin = x.min(0).values - h
        ij = torch.floor((x - origin) / h).long()
        ncx = int(ij[:, 0].max()) + 2
        cid = ij[:, 1] * ncx + ij[:, 0]
        ncell = (int(ij[:, 1].max()) + 2) * ncx
        order = torch.argsort(cid, stable=True)
        counts = torch.bincount(cid, minlength=ncell)
        starts = torch.cumsum(counts, 0) - counts
        kmax = int(counts.max())
        ar = torch.arange(n, device=x.device)
        pi, pj = [], []
        for ox, oy in ((0, 0), (1, 0), (-1, 1), (0, 1), (1, 1)):
            nc = cid + ox + oy * ncx
            cnt, st = counts[nc], starts[nc]
            for k in range(kmax):
                ok = k < cnt
                j = order[(st + k).clamp(max=n - 1)]
                if ox == 0 and oy == 0:
                    ok &= j > ar
                pi.append(ar[ok])
                pj.append(j[ok])
        i, j = torch.cat(pi), torch.cat(pj)
        lo, hi = torch.minimum(i, j), torch.maximum(i, j)
        d = torch.linalg.vector_norm(x[hi] - x[lo], dim=1)
        keep = d < (r[lo] + r[hi]) + self.skin
        key = torch.sort(lo[keep] * n + hi[keep]).values
        if len(self.keys):
            pos = torch.searchsorted(self.keys, key).clamp(max=len(self.keys) - 1)
            hit = self.keys[pos] == key
            self.xi = torch.where(hit[:, None], self.xi[pos], torch.zeros_like(self.xi[pos]))
        else:
            self.xi = torch.zeros((len(key), 2), dtype=x.dtype, device=x.device)
        self.keys, self.i, self.j = key, key // n, key % n
        self.xb = x.clone()
        self.rebuilds += 1

    def _wall(self, w, gap, nrm):
        import torch
        c, r, v, dt = self.cfg, self.r, self.v, self.dt
        d = r - gap
        act = d > 0
        dp = d.clamp(min=0.0)
        sn = (2 * c.e_star) * torch.sqrt(r * dp)
        fel = ((2.0 / 3.0) * sn) * dp
        vn = v @ nrm
        fn = (fel - (c.damping_coefficient * torch.sqrt(sn * self.m)) * vn).clamp(min=0.0)
        st = (8 * c.g_star) * torch.sqrt(r * dp)
        e = self.xi_wall[w] + (v - vn[:, None] * nrm) * dt
        mag = torch.linalg.vector_norm(e, dim=1)
        sc = ((c.wall_friction * fn) / (st * mag).clamp(min=1e-300)).clamp(max=1.0)
        e = torch.where(act[:, None], e * torch.where(mag > 0, sc, 1.0)[:, None], 0.0)
        self.xi_wall[w] = e
        f = torch.where(act[:, None], fn[:, None] * nrm - st[:, None] * e, 0.0)
        self.acc[2] += dt * torch.where(act, (fn - fel) * (-vn), 0.0).sum()
        self.acc[3] += dt * torch.where(act, st * (e * v).sum(1), 0.0).sum()
        self.acc[4:6] += dt * f.sum(0)
        return f

    def step(self):
        import torch
        c, dt = self.cfg, self.dt
        if float((self.x - self.xb).abs().max()) > 0.25 * self.skin:
            self._build()
        x, v, r, i, j = self.x, self.v, self.r, self.i, self.j
        dv = x[j] - x[i]
        dist = torch.linalg.vector_norm(dv, dim=1)
        d = (r[i] + r[j]) - dist
        act = d > 0
        dp = d.clamp(min=0.0)
        nrm = dv / dist.clamp(min=1e-300)[:, None]
        reff = (r[i] * r[j]) / (r[i] + r[j])
        meff = (self.m[i] * self.m[j]) / (self.m[i] + self.m[j])
        sn = (2 * c.e_star) * torch.sqrt(reff * dp)
        fel = ((2.0 / 3.0) * sn) * dp
        vr = v[i] - v[j]
        vn = (vr * nrm).sum(1)
        fn = (fel + (c.damping_coefficient * torch.sqrt(sn * meff)) * vn).clamp(min=0.0)
        st = (8 * c.g_star) * torch.sqrt(reff * dp)
        e = self.xi + (vr - vn[:, None] * nrm) * dt
        mag = torch.linalg.vector_norm(e, dim=1)
        sc = ((c.friction * fn) / (st * mag).clamp(min=1e-300)).clamp(max=1.0)
        e = torch.where(act[:, None], e * torch.where(mag > 0, sc, 1.0)[:, None], 0.0)
        self.xi = e
        f = torch.where(act[:, None], fn[:, None] * nrm + st[:, None] * e, 0.0)
        force = torch.zeros_like(x)
        force[:, 1] = self.m * (-c.gravity)
        force.index_add_(0, j, f)
        force.index_add_(0, i, -f)
        self.acc[0] += dt * torch.where(act, (fn - fel) * vn, 0.0).sum()
        self.acc[1] += dt * torch.where(act, st * (e * vr).sum(1), 0.0).sum()
        if c.floor:
            force += self._wall(0, x[:, 1] - c.floor_y, torch.tensor([0.0, 1.0], dtype=x.dtype,
                                                                      device=x.device))
        if c.back_wall:
            force += self._wall(1, x[:, 0] - c.back_wall_x, torch.tensor([1.0, 0.0], dtype=x.dtype,
                                                                         device=x.device))
        self.v = v + (dt * force) / self.m[:, None]
        self.x = x + dt * self.v

    def contact_pairs(self):
        import torch
        x, r, i, j = self.x, self.r, self.i, self.j
        d = (r[i] + r[j]) - torch.linalg.vector_norm(x[j] - x[i], dim=1)
        k = self.keys[d > 0]
        return torch.stack((k // self.n, k % self.n), 1)


# ---------------------------------------------------------------------------
def kernel_event_time(eng, launches):
    """Mean step-kernel duration (s): HIP events recorded by the library on the
    launch stream around each step-kernel launch, continuing the timed run."""
    return eng.step_timed(launches - launches % 2) * 1e-6


def traffic_from_profile(name="plate_step_c2_counters.md"):
    """HBM bytes per step-kernel launch measured with rocprofv3 counters
    (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel), if recorded."""
    path = ROOT / "profiles" / name
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def host_workers():
    """Host cores the parallel CPU baseline uses: the GPU's share of the
    host (16 cores), or every core where there are fewer."""
    return max(1, min(16, os.cpu_count() or 1))


class CpuPool:
    """Single-threaded FP64 oracle processes (oracle/cpu_worker.py), one per
    host core used, started before this process touches the GPU and set up
    while it works; `run` releases them together over a common window and
    returns their aggregate particle-steps/s."""

    def __init__(self, kind, dt, workers):
        import subprocess
        env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1",
                   MKL_NUM_THREADS="1")
        self.kind, self.workers = kind, workers
        self.procs = [subprocess.Popen(
            [sys.executable, "-m", "oracle.cpu_worker", kind, "--dt", repr(dt), "--wait"],
            cwd=str(ROOT), env=env, stdin=subprocess.PIPE, stdout=subprocess.PIPE,
            stderr=subprocess.PIPE, text=True) for _ in range(workers)]

    def run(self, seconds):
        for p in self.procs:
            if p.stdout.readline().strip() != "ready":
                raise RuntimeError(f"cpu worker failed: {p.stderr.read()[-800:]}")
        start_at = time.time() + 0.5
        for p in self.procs:
            p.stdin.write(f"{start_at!r} {seconds!r}\n")
            p.stdin.flush()
        outs = []
        for p in self.procs:
            out, err = p.communicate()
            if p.returncode != 0:
                raise RuntimeError(f"cpu worker failed: {err[-800:]}")
            outs.append(json.loads(out.strip().splitlines()[-1]))
        rate = sum(o["particle_steps"] / o["seconds"] for o in outs)
        steps = sum(o["steps"] for o in outs)
        return rate, steps, max(o["seconds"] for o in outs)

    def close(self):
        for p in self.procs:
            if p.poll() is None:
                p.kill()
                p.wait()


def copy_calibration(mib=512):
    """A plain device copy of known size, run under rocprofv3 --pmc to scale
    FETCH_SIZE / WRITE_SIZE (each byte is read once and written once)."""
    import torch
    x = torch.ones(mib * 2**20 // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": x.numel() * 8, "launches": 3}))


# ---------------------------------------------------------------------------
# C3: unbonded granular column collapse, 1 048 576 disks
# ---------------------------------------------------------------------------
BED_N = 1 << 20
BED_R0 = 0.01
# compulsory bytes per disk per step of bed_step_kernel, before the list- and
# contact-dependent terms: x, v read (32) and written (32), radius (8), list
# origin xb (16), contact mask read (4; written only when a contact forms or
# breaks), candidate count (1)
BED_BYTES_FIXED = 93


def bed_bytes_per_disk(cand_mean, contacts_per_disk, wall_per_disk, hist_bytes):
    """+ 4 B per candidate index, + 2 hist_bytes per pair-history copy in
    contact (read and written: float2, 8 B, with FP32 forces; double2, 16 B,
    with FP64 forces), + 32 B per wall history (double2, read and written)."""
    return (BED_BYTES_FIXED + 4.0 * cand_mean + 2.0 * hist_bytes * contacts_per_disk
            + 32.0 * wall_per_disk)


def graph_steps(k, cap):
    """(steps, chunk): k rounded up to even, and the largest even divisor of
    it not above cap as the steps per captured graph (graphs hold step pairs)."""
    steps = max(2, k + (k & 1))
    chunk = max(c for c in range(2, min(cap, steps) + 1, 2) if steps % c == 0)
    return steps, chunk


BED_FULL_STEPS = 50000
BED_WINDOWS = 10
_C3_RELEASE = None


def c3_inputs():
    """Law, deposited positions and radii of the column (the host baseline's
    sample; the GPU runs start from the settled state of the same column)."""
    from tensordem_amd.bedconfig import BedConfig, column_deposit
    x, r, _ = column_deposit(BED_N, BED_R0, seed=0)
    return BedConfig(), x, r


def c3_release():
    """The release state of the column collapse: the deposited column
    settled to rest behind a gate on the device, gate removed (computed
    once per process, about 250 000 FP32 steps)."""
    global _C3_RELEASE
    if _C3_RELEASE is None:
        from tensordem_amd.bed import settled_column
        t0 = time.perf_counter()
        state, info = settled_column(BED_N, BED_R0, seed=0, max_steps=400000)
        info["seconds"] = time.perf_counter() - t0
        print(f"bench.py: column settled in {info['settle_steps']} steps "
              f"({info['seconds']:.1f} s), {info['contact_pairs']} contact pairs", file=sys.stderr)
        _C3_RELEASE = (state, info)
    return _C3_RELEASE


def bed_windows(steps, full=BED_FULL_STEPS, windows=BED_WINDOWS):
    """[(start, n)]: the timed steps split into windows spread over the
    collapse when fewer than `full` steps are timed (start steps even), else
    one window from release."""
    if steps >= full:
        return [(0, steps)]
    # windows of at least 4 steps: each short window carries one gap between
    # graph launches (a few microseconds)
    w = max(1, min(windows, steps // 4))
    per = [steps // w + (1 if k < steps % w else 0) for k in range(w)]
    span = full - 2 * max(per)   # a short window runs k untimed lead steps first
    starts = [0 if w == 1 else 2 * round(k * span / (w - 1) / 2) for k in range(w)]
    return list(zip(starts, per))


def _advance(eng, k):
    if k > 0:
        eng.step(k)


def bench_bed(args):
    """The 1M-disk column collapse from its settled release state: FP32 and
    FP64 forces over the timed steps (windows spread over the 50 000-step
    collapse when fewer are timed, advanced untimed between them), kernel
    events and list statistics at the same windows, the plain FP64 path
    from the state at the middle window, the oracle on a CPU sample.  With
    --bed-from S (or --bed-start FILE) every engine starts instead from the
    state S steps into the collapse, carried over with its spring histories
    by the bed restart, and the steps run in one window."""
    import torch
    from tensordem_amd import _lib
    from tensordem_amd.bed import BedEngine, load_bed_restart, save_bed_restart

    cfg = None
    if args.bed_start:
        state, info = load_bed_restart(args.bed_start), {}
    else:
        state, info = c3_release()
        if args.bed_from:
            adv = BedEngine.from_snapshot(state, fp32=True)
            _advance(adv, args.bed_from)
            state = adv.snapshot()
            adv.close()
            del adv
            if args.bed_save_start:
                save_bed_restart(state, args.bed_save_start)
    n = len(state["radii"])
    start = int(state["step_count"])
    steps = max(2, int(args.steps))
    wins = bed_windows(steps, windows=args.bed_windows) if start == 0 else [(0, steps)]

    def engine(fp32):
        return BedEngine.from_snapshot(state, fp32=fp32)

    warm = engine(True)
    wsteps = max(0, int(args.warmup))
    _advance(warm, wsteps)
    torch.cuda.synchronize()
    m0 = warm.read_members()[0]
    dt, skin = m0.dt, m0.skin
    warm.close()
    del warm

    def run(eng, timed, label):
        """Walk the windows: advance untimed to each start, then `timed(eng,
        k)`, which returns (result, steps it advanced)."""
        at, out = 0, []
        for w, (s0, k) in enumerate(wins):
            _advance(eng, s0 - at)
            res, adv = timed(eng, k)
            out.append(res)
            at = s0 + adv
            progress(f"bed {label}: window {w + 1}/{len(wins)} done")
        return out

    # a short window replays its graph once untimed first (the first replay
    # of a graph uploads it), so it times steps k .. 2k after its start
    lead = [k - (k & 1) if k < 1000 else 0 for _, k in wins]

    def graph_window(eng, k):
        """k timed steps from graph replays.  A short window first replays its
        graph once untimed and queues the timed replay right behind it, with
        no host wait in between: the GPU goes from the lead replay straight
        into the timed one, so the start event does not count the host's
        launch latency."""
        stream = torch.cuda.current_stream()
        even = k - (k & 1)
        _, chunk = graph_steps(even, 1000) if even >= 2 else (0, 0)
        short = k < 1000
        if chunk:
            eng.graph(chunk // 2)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        if short and chunk:
            eng.run_graph(chunk // 2, repeats=even // chunk)   # untimed lead
        e0.record(stream)
        if chunk:
            eng.run_graph(chunk // 2, repeats=even // chunk)
        if k % 2:
            eng.step(1)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) * 1e-3, k + (even if short else 0)

    out = {}
    for fp32 in (True, False):
        eng = engine(fp32)
        secs = run(eng, graph_window, "FP32 graph" if fp32 else "FP64 graph")
        gpu_s = sum(secs)
        acc = eng.accounting[0].cpu().numpy()
        out["fp32" if fp32 else "fp64"] = dict(
            seconds=gpu_s, ms_per_step=gpu_s / steps * 1e3, rate=n * steps / gpu_s,
            window_ms_per_step=[s_ / k * 1e3 for s_, (_, k) in zip(secs, wins)],
            status=int(eng.status()[0]), contacts_end=int(acc[_lib.BED_CONTACTS]),
            stats_end=eng.list_stats(), rebuilds=int(eng.read_members()[0].rebuilds))
        eng.close()
        del eng
    # kernel-event pass over the same windows (HIP events on every step-kernel
    # dispatch; event timing needs an even step count per call), sampling the
    # lists for the byte count; the plain path starts from the middle window
    eng = engine(True)
    half_at = wins[len(wins) // 2][0]
    half = None

    def event_window(eng, k):
        """Kernel events over the window.  A short window is timed by one call
        queued right behind its untimed lead steps (the GPU stays busy, so its
        clocks do not drop in a host gap); a long one in ten calls, the lists
        sampled after each."""
        nonlocal half
        lk = k - (k & 1) if k < 1000 else 0
        _advance(eng, lk)
        kk = k + (k & 1)
        calls = 1 if k < 1000 else 10
        stride = max(2, (kk // calls) + ((kk // calls) & 1))
        us, byts, done = [], [], 0
        while done < kk:
            c = min(stride, kk - done)
            us.append((eng.step_timed(c), c))
            done += c
            st = eng.list_stats()
            acc = eng.accounting[0].cpu().numpy()
            byts.append(bed_bytes_per_disk(st["cand_mean"], st["contacts_per_disk"],
                                           acc[_lib.BED_WALL_CONTACTS] / n, hist_bytes=8))
            if half is None and eng.read_members()[0].step_count - start >= half_at:
                half = (eng.x.cpu().numpy(), eng.v.cpu().numpy())
        return (us, byts), lk + kk

    ev = run(eng, event_window, "kernel events")
    eng.close()
    del eng
    us_all = [u for w_us, _ in ev for u in w_us]
    kernel_s = sum(u * c for u, c in us_all) / sum(c for _, c in us_all) * 1e-6
    byts = [b for _, w_b in ev for b in w_b]
    bpd = sum(byts) / len(byts)
    achieved = bpd * n / kernel_s / 1e9
    r32, r64 = out["fp32"], out["fp64"]
    windowed = len(wins) > 1
    res = {
        "workload": f"granular column collapse, {n} disks, radii U[0.8,1.2]x{BED_R0} m, "
                    "E=1 GPa, nu=0.3, mu=0.5, e=0.5 (viscous), rho=917, t=0.02 m, "
                    "aspect 2 random loose column (sequential deposition, seed 0) settled to "
                    "rest behind a gate, released, floor + back wall, gravity, "
                    f"dt={dt:.6g} s, "
                    + (f"{steps} steps in {len(wins)} windows spread over the "
                       f"{BED_FULL_STEPS}-step collapse" if windowed else
                       f"{steps} steps from "
                       + ("release" if not start else f"step {start} of the collapse")),
        "value": r32["rate"], "unit": "particle-steps/s", "steps": steps, "warmup": wsteps,
        "ms_per_step": r32["ms_per_step"], "dtype": "f32-force/f64-state",
        "windows": [{"start_step": s0 + ld + start, "steps": k}
                    for (s0, k), ld in zip(wins, lead)],
        "window_ms_per_step": r32["window_ms_per_step"],
        "skin_m": skin, "start_step": start,
        "release_state": {k: info[k] for k in ("settle_steps", "contact_pairs", "kinetic_rel",
                                                "width_m", "height_m") if k in info},
        "status_bits": r32["status"], "contact_pairs_end": r32["contacts_end"],
        "lists_end": r32["stats_end"], "list_rebuilds": r32["rebuilds"],
        "fp64_kernel": {"value": r64["rate"], "ms_per_step": r64["ms_per_step"],
                        "status_bits": r64["status"], "contact_pairs_end": r64["contacts_end"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_from_profile("bed_step_c3_counters.md"),
                     "traffic_sample": "counted over the collapse windows of "
                                       "profiles/bed_step_c3_counters.md",
                     "kernel": "bed_step_kernel<float, true>", "kernel_us": kernel_s * 1e6,
                     "kernel_timing": "HIP events on every step-kernel dispatch of "
                                      "tdm_bed_step_timed calls at the timed windows",
                     "bytes_per_disk": bpd, "bytes_per_launch": bpd * n},
    }
    for name, rr in out.items():
        if rr["status"]:
            print(f"bench.py: bed {name} run ended with status bits {rr['status']}; its rate "
                  "is not a valid measurement", file=sys.stderr)
    if not args.no_plain and half is not None:
        r = np.asarray(state["radii"])
        cfg = BedConfig_from(state)
        p = PlainBedFP64(cfg, half[0], r, dt, skin, v=half[1])
        for _ in range(3):
            p.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.bed_plain_steps):
            p.step()
        torch.cuda.synchronize()
        ps = (time.perf_counter() - t0) / args.bed_plain_steps
        res["plain_fp64_gpu"] = dict(
            value=n / ps, unit="particle-steps/s", ms_per_step=ps * 1e3,
            steps=args.bed_plain_steps, list_rebuilds=p.rebuilds - 1,
            what="eager PyTorch FP64 bed step on the same GPU (sort-based cell list, Verlet "
                 "list rebuilt at |dx|inf > skin/4, searchsorted history remap, index_add_), "
                 f"from the state at step {start + half_at}")
        res["speedup_vs_plain_fp64"] = r32["rate"] * ps / n
        del p
    if not args.no_cpu:
        ccfg, cx, cr = c3_inputs()
        one = cpu_baseline_bed(ccfg, cx, cr, dt, args.cpu_seconds)
        res["cpu_baseline"] = with_pool(one, getattr(args, "cpu_pools", {}).get("bed"),
                                        args.cpu_seconds,
                                        f"the {len(cr)}-disk deposited column (before settling)")
    return res


def save_bed_phases(out: Path):
    """Release state of the column and the states 25 000 and 48 000 steps
    into its collapse (FP32 forces), as bed restart files for profiling runs
    that should not replay the settling or the collapse under a profiler."""
    from tensordem_amd.bed import BedEngine, save_bed_restart
    out.mkdir(parents=True, exist_ok=True)
    state, _ = c3_release()
    save_bed_restart(state, out / "c3_release.npz")
    eng = BedEngine.from_snapshot(state, fp32=True)
    at = 0
    for name, s_ in (("c3_mid.npz", 25000), ("c3_late.npz", 48000)):
        _advance(eng, s_ - at)
        at = s_
        save_bed_restart(eng.snapshot(), out / name)
    eng.close()
    print(f"bench.py: wrote {out}/c3_release.npz, c3_mid.npz, c3_late.npz", file=sys.stderr)


def bench_horizon(args):
    """How long FP32-force runs of the 1024 x 128 plate keep the FP64
    broken-bond set, against the re-check margin and the time step: the FP64
    run (bit-identical to the FP64 oracle over its 20 000 steps) and FP32 runs
    with fp32_margin 1e-3, 1e-2 and 1e-1 step in lockstep; every 100 steps
    their alive masks are compared.  The same at dt / 2 over the same
    simulated time.  Reports each run's first differing checkpoint, its
    broken bonds at the end and its mean time per step."""
    import torch
    from tensordem_amd.engine import PlateEngine, PlateSpec

    cfg = c2_config()
    dt0 = cfg.degree_bound_dt()
    rows = []
    for dt, total, margins in ((dt0, args.plate_steps, (1e-3, 1e-2, 1e-1)),
                               (0.5 * dt0, 2 * args.plate_steps, (1e-3, 1e-1))):
        ref = PlateEngine([PlateSpec(cfg, dt=dt)], fp32=False)
        runs = {m: PlateEngine([PlateSpec(cfg, dt=dt)], fp32=True, fp32_margin=m)
                for m in margins}
        first = {m: None for m in margins}
        secs = {m: 0.0 for m in margins}
        every = 100
        for k in range(every, total + 1, every):
            ref.step(every)
            for m, eng in runs.items():
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                eng.step(every)
                torch.cuda.synchronize()
                secs[m] += time.perf_counter() - t0
                if first[m] is None and not torch.equal(eng.alive, ref.alive):
                    first[m] = k
        nb_ref = int((~ref.alive).sum())
        for m, eng in runs.items():
            rows.append({"dt_s": dt, "steps": total, "fp32_margin": m,
                         "first_difference_step": first[m],
                         "equal_through_step": (first[m] - every) if first[m] else total,
                         "equal_through_time_s": ((first[m] - every) if first[m] else total) * dt,
                         "broken_end": int((~eng.alive).sum()), "broken_end_fp64": nb_ref,
                         "ms_per_step": secs[m] / total * 1e3,
                         "status_bits": int(eng.status()[0])})
            eng.close()
        ref.close()
    return {"workload": "1024x128 plate indentation, FP32 forces vs FP64, broken-bond sets "
                        "compared every 100 steps", "value": 0.0, "steps": args.plate_steps,
            "ms_per_step": 0.0, "dtype": "f32-force/f64-state", "horizon": rows}


def BedConfig_from(state):
    from tensordem_amd.bedconfig import BedConfig
    return BedConfig(**state["config"])


def with_pool(one, pool, seconds, what):
    """The host baseline: the aggregate of the parallel pool when there is
    one (the single-core figure kept beside it), else the single core."""
    if pool is None:
        return one
    try:
        rate, steps, secs = pool.run(seconds)
    except (RuntimeError, OSError, ValueError) as exc:
        # the single-core figure stands if the worker processes fail
        print(f"bench.py: parallel host baseline unavailable ({exc}); reporting one core",
              file=sys.stderr)
        return one
    return dict(value=rate, unit="particle-steps/s", cores=pool.workers, kind="port",
                sample=f"{pool.workers} concurrent single-threaded FP64 C oracle processes, each "
                       f"on its own copy of {what} (cell grid rebuilt every step), {steps} steps "
                       f"in total over a common {secs:.1f} s window",
                single_core=one)


def cpu_baseline_bed(cfg, x, r, dt, budget_s):
    """C oracle (FP64, one core) on the 1M-disk deposited column."""
    from oracle.bed import OracleBed
    sim = OracleBed(cfg, x, r, dt=dt)
    sim.step()
    t0 = time.perf_counter()
    steps = 0
    while time.perf_counter() - t0 < budget_s:
        sim.step()
        steps += 1
    el = time.perf_counter() - t0
    return dict(value=len(r) * steps / el, unit="particle-steps/s", cores=1, kind="port",
                sample=f"{steps} steps of the {len(r)}-disk deposited column (FP64 C oracle, "
                       f"cell grid rebuilt every step), {el:.1f} s")


# ---------------------------------------------------------------------------
# rigid frictional disks with rotation (granular2d engine), FP64
# ---------------------------------------------------------------------------
DISK_SIDE = 1024
DISK_R = 0.01


def disk_inputs(side=DISK_SIDE):
    """side x side disks of radius U[0.9, 1.1] x 0.01 m on a jittered square
    grid of pitch 2.1 r (the largest neighbours overlap slightly) in a walled
    box, seeded velocities of 0.05 m/s rms; gravity, mu = 0.5, mu_r = 0.05,
    the engine's default stiffness, damping, ice density and skin."""
    from tensordem_amd.granular2d import DiskConfig, jittered_box
    pos, r, vel, walls = jittered_box(side, DISK_R, seed=0)
    cfg = DiskConfig(mu=0.5, mu_rolling=0.05, gravity=(0.0, -9.81))
    return cfg, pos, r, vel, walls


def disk_dt():
    """The 1M-disk box's time step (the engine's stability bound, from its
    lightest disk); also valid for the smaller boxes of the host baseline."""
    cfg, _, r, _, _ = disk_inputs()
    m_min = cfg.density * math.pi * float(r.min()) ** 2 * cfg.thickness
    return 0.1 / math.sqrt(2.0 * cfg.kn / (0.5 * m_min))


def disk_bytes_per_disk(cand_mean, pairs_per_disk, boundary_per_disk):
    """Compulsory bytes per disk per launch of disk_step_kernel: own x, v,
    omega read and written (80), radius, mass, inertia (24), angle read and
    written (16), candidate count (4), contact mask read and written (8),
    position at the last search (16); per candidate slot its index (4); per
    pair in contact each disk's copy of its spring read and written
    (2 x 2 x 24); per boundary contact its spring state read and written
    (2 x 32)."""
    return 148.0 + 4.0 * cand_mean + 96.0 * pairs_per_disk + 64.0 * boundary_per_disk


def bench_disk(args):
    """The granular2d engine on 1M disks: graph-replayed FP64 steps, kernel
    events with list statistics, the C oracle on a CPU sample."""
    import torch
    from tensordem_amd.granular2d import DiskSystem

    cfg, pos, r, vel, walls = disk_inputs()
    n = len(r)
    steps, chunk = graph_steps(args.disk_steps, 100)
    sysd = DiskSystem(pos, r, cfg, walls=walls)
    sysd.vel = vel
    sysd.step(2)
    wsteps, wchunk = graph_steps(max(2, args.disk_warmup), 100)
    sysd.run_graph(chunk // 2, repeats=wsteps // chunk)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    sysd.run_graph(chunk // 2, repeats=steps // chunk)
    e1.record(stream)
    torch.cuda.synchronize()
    gpu_s = e0.elapsed_time(e1) * 1e-3
    st = sysd.list_stats()
    status = int(sysd.out[13].item())
    rebuilds = sysd.rebuilds
    step_us, prep_us = sysd.step_timed(200)
    st2 = sysd.list_stats()
    mean = {k: 0.5 * (st[k] + st2[k]) for k in st}
    bpd = disk_bytes_per_disk(mean["cand_mean"], mean["pairs_per_disk"],
                              mean["walls_per_disk"] + mean["hull_per_disk"])
    achieved = bpd * n / (step_us * 1e-6) / 1e9
    res = {
        "workload": f"rigid frictional disks with rotation (granular2d engine), {n} disks, "
                    f"radii U[0.9,1.1]x{DISK_R} m on a jittered grid in a walled box, "
                    f"gravity, mu=0.5, mu_rolling=0.05, kn=3e4, zeta=0.2, dt={sysd.dt:.6g} s, "
                    f"{steps} steps after {wsteps} warm-up steps",
        "value": n * steps / gpu_s, "unit": "particle-steps/s", "steps": steps,
        "warmup": wsteps, "ms_per_step": gpu_s / steps * 1e3, "dtype": "f64",
        "status_bits": status, "list_rebuilds": rebuilds, "lists": st,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_from_profile("disk_step_counters.md"),
                     "kernel": "disk_step_kernel<true, false>", "kernel_us": step_us,
                     "prep_kernel_us": prep_us, "bytes_per_disk": bpd,
                     "bytes_per_launch": bpd * n},
    }
    if status:
        print(f"bench.py: disk run ended with status bits {status}; its rate is not a valid "
              "measurement", file=sys.stderr)
    sysd.close()
    del sysd
    if not args.no_cpu:
        one = cpu_baseline_disk(args.cpu_seconds)
        res["cpu_baseline"] = with_pool(one, getattr(args, "cpu_pools", {}).get("disk"),
                                        args.cpu_seconds, "a 65 536-disk box of the same "
                                        "construction")
    return res


def cpu_baseline_disk(budget_s, side=256):
    """C oracle (FP64, one core) on a 256 x 256 box of the same construction
    (65 536 disks; per-disk work does not depend on the box size)."""
    from oracle.disk import OracleDisk
    from tensordem_amd.granular2d import DiskConfig
    cfg, pos, r, vel, walls = disk_inputs(side)
    sim = OracleDisk(pos, r, DiskConfig(mu=cfg.mu, mu_rolling=cfg.mu_rolling,
                                        gravity=cfg.gravity, dt=disk_dt()), walls=walls)
    sim.vel[:] = vel
    sim.step()
    t0 = time.perf_counter()
    steps = 0
    while time.perf_counter() - t0 < budget_s:
        sim.step()
        steps += 1
    el = time.perf_counter() - t0
    return dict(value=len(r) * steps / el, unit="particle-steps/s", cores=1, kind="port",
                sample=f"{steps} steps of a {len(r)}-disk box of the same construction (FP64 C "
                       f"oracle, cell grid rebuilt every step), {el:.1f} s")


# ---------------------------------------------------------------------------
# IceDEM3D: the reference's 10k-100k scale study, 47^3 bonded spheres
# ---------------------------------------------------------------------------
D3_SIDE = 47
D3_REFERENCE_MS = 15872.0   # docs/dem3d_scale_benchmark.md:62-66, 103 823 spheres, loong64


def d3_config(side=D3_SIDE):
    """scripts/benchmark_dem3d_scale.py:49-70: r = 0.025 m, k_b = 2000, k_c = 5000,
    c = 5, drag = 2, tool r = 0.1 m at 0.01 m/s, 1 cm gap, free edges,
    dt = 0.1 x recommended_dt."""
    import dataclasses
    from tensordem_amd.dem3dconfig import DEM3DConfig
    base = DEM3DConfig(nx=side, ny=side, nz=side, radius=0.025, density=917.0,
                       bond_stiffness=2000.0, contact_stiffness=5000.0, contact_damping=5.0,
                       drag=2.0, tool_radius=0.1, tool_speed=0.01, tool_gap=0.01,
                       breaking_strain=0.015, shear_breaking_strain=0.03, fix_x_edges=False,
                       fix_bottom=False)
    return dataclasses.replace(base, dt=0.10 * base.recommended_dt)


def d3_bytes_per_sphere(bonds_per_sphere, cells_per_sphere, list_per_sphere):
    """Compulsory (unique) bytes per sphere per launch of d3_step_kernel: own x
    and v read (48), the new x and v and the force written (72), initial
    position (24), fixed flag (1), the 26-slot partner and bond-index rows
    (208), each of its bonds' alive byte and bond force once (25 B/N), its
    candidate count and list (4 + 4 per entry), its position at the last
    list build (24), and its share of the cell counts cleared for the next
    evaluation (4 x ncells/N)."""
    return (48.0 + 72.0 + 24.0 + 1.0 + 208.0 + 25.0 * bonds_per_sphere + 4.0
            + 4.0 * list_per_sphere + 24.0 + 4.0 * cells_per_sphere)


def bench_dem3d(args):
    """IceDEM3D on 47^3 = 103 823 spheres: graph-replayed FP64 steps, kernel
    events, the C oracle on a CPU sample."""
    import torch
    from tensordem_amd.dem3d import IceDEM3D

    cfg = d3_config()
    sim = IceDEM3D(cfg)
    n, nb = len(sim.positions), len(sim.pairs)
    steps, chunk = graph_steps(args.d3_steps, 64)
    sim.run(max(64, args.d3_warmup), pairs=chunk // 2)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    sim.run(steps, pairs=chunk // 2)
    e1.record(stream)
    torch.cuda.synchronize()
    gpu_s = e0.elapsed_time(e1) * 1e-3
    step_us, strain_us, bond_us = sim.step_timed(200)
    # the lattice barely moves in this workload: each sphere lists its face
    # neighbours (rest length 2r < 2r + skin), no others
    assert sim.broken_bonds == 0
    listed = 2.0 * float((sim.rest_lengths < 2.0 * cfg.radius * 1.05).sum()) / n
    bps = d3_bytes_per_sphere(nb / n, sim._params.ncells / n, listed)
    achieved = bps * n / (step_us * 1e-6) / 1e9
    ms = gpu_s / steps * 1e3
    res = {
        "workload": f"IceDEM3D bonded-sphere lattice {D3_SIDE}^3 = {n} spheres, {nb} bonds "
                    "(the reference's scale_100k case: r=0.025 m, k_b=2000, k_c=5000, c=5, "
                    f"drag=2, tool r=0.1 m at 0.01 m/s, dt=0.1 x recommended = {sim.dt:.6g} s), "
                    f"{steps} steps",
        "value": n * steps / gpu_s, "unit": "particle-steps/s", "steps": steps,
        "ms_per_step": ms, "dtype": "f64", "broken_bonds": sim.broken_bonds,
        "reference_ms_per_step": D3_REFERENCE_MS,
        "reference_over_this": D3_REFERENCE_MS / ms,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_from_profile("dem3d_step_scale100k_counters.md"),
                     "kernel": "d3_step_kernel", "kernel_us": step_us,
                     "strain_kernel_us": strain_us, "bond_kernel_us": bond_us,
                     "bytes_per_sphere": bps,
                     "bytes_per_launch": bps * n, "list_entries_per_sphere": listed,
                     "list_rebuilds": sim.rebuilds},
    }
    del sim
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_dem3d(args.cpu_seconds)
    return res


def cpu_baseline_dem3d(budget_s):
    """C oracle (FP64, one core, sorted cell list) on the same 47^3 lattice."""
    from oracle.dem3d import OracleDEM3D
    sim = OracleDEM3D(d3_config())
    sim.step()
    t0 = time.perf_counter()
    steps = 0
    while time.perf_counter() - t0 < budget_s:
        sim.step()
        steps += 1
    el = time.perf_counter() - t0
    n = len(sim.positions)
    return dict(value=n * steps / el, unit="particle-steps/s", cores=1, kind="port",
                sample=f"{steps} steps of the same {n}-sphere lattice (FP64 C oracle, sorted "
                       f"cell list rebuilt every step), {el:.1f} s")


# ---------------------------------------------------------------------------
# C4: 512 independent plates with Weibull strengths, one launch per step
# ---------------------------------------------------------------------------
def bench_ensemble(args):
    """512 plates (nx log-uniform 16..256, ny 4..32, tool 0.1..1 m/s, per-disk
    Weibull strains) in one batch: autotuned launch shape, FP32 and FP64 forces
    over graph-captured steps, kernel events, plain FP64 per member, CPU oracle."""
    import torch
    from tensordem_amd.autotune import autotune_plates
    from tensordem_amd.engine import PlateEngine
    from tensordem_amd.ensemble import random_ensemble, total_disks

    specs = random_ensemble(512, seed=0)
    n = total_disks(specs)
    tune = autotune_plates(specs, fp32=True, steps=200, warmup=40)
    tuned = tune.apply(specs)
    chunk = 100
    steps = max(chunk, args.ens_steps - args.ens_steps % chunk)
    out = {}
    for fp32 in (True, False):
        eng = PlateEngine(tuned, fp32=fp32, **tune.engine_kwargs())
        eng.graph(chunk // 2)
        stream = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        eng.run_graph(chunk // 2, repeats=steps // chunk)
        e1.record(stream)
        torch.cuda.synchronize()
        gpu_s = e0.elapsed_time(e1) * 1e-3
        out["fp32" if fp32 else "fp64"] = dict(
            rate=n * steps / gpu_s, ms_per_step=gpu_s / steps * 1e3,
            broken=int((~eng.alive).sum()), status=int(np.bitwise_or.reduce(eng.status())),
            alive=eng.alive.cpu().numpy())
        eng.close()
        del eng
    eng = PlateEngine(tuned, fp32=True, **tune.engine_kwargs())
    kernel_s = eng.step_timed(steps) * 1e-6
    # tiles at most 128 wide take the kernel that stages candidate forces in LDS
    staged = 0 < eng.max_tile_x <= 128
    eng.close()
    del eng
    r32, r64 = out["fp32"], out["fp64"]
    # + the two per-disk strength thresholds read per step
    bytes_per_disk = PLATE_BYTES_PER_DISK + 16
    achieved = bytes_per_disk * n / kernel_s / 1e9
    res = {
        "workload": f"512 independent plates ({n} disks): nx log-uniform [16,256], ny uniform "
                    "[4,32], tool speed U[0.1,1] m/s, per-disk Weibull(5) strains mean "
                    "0.01 axial / 0.02 shear (bond = weaker disk), degree-bounded dt per member, "
                    f"{steps} steps, one launch per step for all members (seed 0)",
        "value": r32["rate"], "unit": "particle-steps/s", "steps": steps, "disks": n,
        "dtype": "f32-force/f64-state",
        "ms_per_step": r32["ms_per_step"], "graph_chunk_steps": chunk,
        "broken_bonds_end": r32["broken"], "status_bits": r32["status"],
        "broken_set_equal_fp64": bool(np.array_equal(r32["alive"], r64["alive"])),
        "autotune": {"tile_policy": tune.tile_policy, "skin_factor": tune.skin_factor,
                     "seconds": tune.seconds, "table": tune.table},
        "fp64_kernel": {"value": r64["rate"], "ms_per_step": r64["ms_per_step"],
                        "broken_bonds_end": r64["broken"], "status_bits": r64["status"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_from_profile("plate_step_c4_counters.md"),
                     "kernel": f"plate_step_kernel<float, true, {str(staged).lower()}>",
                     "kernel_us": kernel_s * 1e6,
                     "bytes_per_disk": bytes_per_disk, "bytes_per_launch": bytes_per_disk * n},
    }
    if not args.no_plain:
        # every 4th member, each from rest: the plain path is launch-bound, so
        # its time per step hardly depends on the member's size
        tot_n = tot_s = 0.0
        for s in specs[::4]:
            p = PlainFP64(s.config, s.dt, pinv=False, eps_t_disk=s.eps_t_disk,
                          eps_s_disk=s.eps_s_disk)
            p.step()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(3):
                p.step()
            torch.cuda.synchronize()
            tot_s += (time.perf_counter() - t0) / 3
            tot_n += s.config.nx * s.config.ny
            del p
        res["plain_fp64_gpu"] = dict(
            value=tot_n / tot_s, unit="particle-steps/s", members_sampled=len(specs[::4]),
            what="eager PyTorch FP64 plate step run member after member (every 4th member, "
                 "3 steps each, closed-form 2x2 inverse), particle-steps over summed step time")
        res["speedup_vs_plain_fp64"] = r32["rate"] / (tot_n / tot_s)
    if not args.no_cpu:
        from oracle.plate import OraclePlate
        t0 = time.perf_counter()
        done_ps, k = 0, 0
        while time.perf_counter() - t0 < args.cpu_seconds and k < len(specs):
            s = specs[k]
            sim = OraclePlate(s.config, dt=s.dt, eps_t_disk=s.eps_t_disk, eps_s_disk=s.eps_s_disk)
            for _ in range(10):
                sim.step()
            done_ps += 10 * s.config.nx * s.config.ny
            k += 1
        el = time.perf_counter() - t0
        res["cpu_baseline"] = dict(value=done_ps / el, unit="particle-steps/s", cores=1,
                                   kind="port",
                                   sample=f"first {k} members, 10 steps each from rest (FP64 C "
                                          f"oracle), {el:.1f} s")
    return res


# ---------------------------------------------------------------------------
# C5: serving, Poisson arrivals, continuous batching
# ---------------------------------------------------------------------------
def _plain_step_seconds(req, steps=3):
    """Seconds per step of the plain FP64 eager path for one request, alone."""
    import torch
    if req.kind == "plate":
        p = PlainFP64(req.spec.config, req.spec.dt, pinv=False)
    else:
        c = req.spec.config
        r = req.spec.radii
        dt = c.dt if c.dt is not None else c.stable_dt(float(r.min()))
        p = PlainBedFP64(c, req.spec.positions, r, dt, 0.4 * float(r.mean()))
    p.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        p.step()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def bench_serve(args):
    """Poisson arrivals (50/s over the window), half plates and half beds,
    1k-200k disks log-uniform, 5000 steps each, served by continuous batching
    in real time; p50/p99 completion latency and aggregate particle-steps/s,
    beside the same requests run one after another on the plain FP64 path
    (estimated from sampled requests) and on the CPU oracle."""
    from tensordem_amd.serve import Server, latency_summary, sweep_requests

    reqs = sweep_requests(rate=50.0, window=args.serve_window, steps=5000, seed=0)
    # warm the kernels, graphs and allocator on a few small requests
    warm = sweep_requests(rate=50.0, window=0.2, steps=500, n_range=(1000, 3000), seed=99)
    Server(chunk=args.serve_chunk).run(warm, realtime=False)
    srv = Server(chunk=args.serve_chunk, max_active=args.serve_slots)
    t0 = time.perf_counter()
    done = srv.run(reqs, realtime=True)
    wall = time.perf_counter() - t0
    total_ps = float(sum(r.n * r.steps for r in done))
    last = max(r.finished for r in done)
    lat = latency_summary(done)
    kinds = {k: [r for r in done if r.kind == k] for k in ("plate", "bed")}
    res = {
        "workload": f"Poisson arrivals 50/s over {args.serve_window:g} s (seed 0), half bonded "
                    "plates (indenter, swept speed and strains) and half Hertz-Mindlin beds "
                    "(swept friction and restitution), 1k-200k disks log-uniform, 5000 steps "
                    f"each; continuous batching with {args.serve_chunk}-step rounds, at most "
                    f"{args.serve_slots} members per batch, 288 GB HBM admission budget",
        "value": total_ps / last, "unit": "particle-steps/s", "requests": len(done),
        "dtype": "f32-force/f64-state", "steps": 5000, "ms_per_step": None,
        "latency_s": lat,
        "latency_s_by_kind": {k: latency_summary(v) for k, v in kinds.items()},
        "offered_particle_steps_per_s": total_ps / args.serve_window,
        "served_particle_steps": total_ps, "last_completion_s": last, "wall_s": wall,
        "rounds": srv.rounds, "host_issue_seconds": srv.host_seconds,
        "max_active": [max(p for _, p, _, _ in srv.log), max(b for _, _, b, _ in srv.log)],
        "max_waiting": max(w for *_, w in srv.log),
        "status_nonzero": sum(1 for r in done if r.result["status"]),
    }
    if not args.no_plain:
        # sequential plain FP64: time per step = a + b n per kind, fitted on a
        # sample spanning the size range
        fits = {}
        for k, rs in kinds.items():
            rs = sorted(rs, key=lambda r: r.n)
            pick = [rs[int(i)] for i in np.linspace(0, len(rs) - 1, 6)]
            ns = np.array([r.n for r in pick], dtype=np.float64)
            ts = np.array([_plain_step_seconds(r) for r in pick])
            b, a = np.polyfit(ns, ts, 1)
            fits[k] = (max(a, 0.0), max(b, 0.0), ns.tolist(), ts.tolist())
        seq_s = sum(r.steps * (fits[r.kind][0] + fits[r.kind][1] * r.n) for r in done)
        res["plain_fp64_gpu"] = dict(
            value=total_ps / seq_s, unit="particle-steps/s", sequential_seconds=seq_s,
            fit_seconds_per_step={k: {"a": f[0], "b_per_disk": f[1], "n": f[2], "t": f[3]}
                                  for k, f in fits.items()},
            what="the same requests one after another on the eager PyTorch FP64 paths "
                 "(time per step fitted as a + b n per kind on 6 sampled requests each)")
        res["speedup_vs_plain_fp64"] = res["value"] / res["plain_fp64_gpu"]["value"]
    if not args.no_cpu:
        from oracle.bed import OracleBed
        from oracle.plate import OraclePlate
        t_c = time.perf_counter()
        ps, k = 0, 0
        small = sorted(done, key=lambda r: r.n)
        while time.perf_counter() - t_c < args.cpu_seconds and k < len(small):
            r = small[k * max(1, len(small) // 40)] if k * max(1, len(small) // 40) < len(small) \
                else small[-1]
            if r.kind == "plate":
                sim = OraclePlate(r.spec.config, dt=r.spec.dt)
            else:
                sim = OracleBed(r.spec.config, r.spec.positions, r.spec.radii)
            for _ in range(5):
                sim.step()
            ps += 5 * r.n
            k += 1
        el = time.perf_counter() - t_c
        res["cpu_baseline"] = dict(value=ps / el, unit="particle-steps/s", cores=1, kind="port",
                                   sample=f"{k} requests across the size range, 5 steps each "
                                          f"(FP64 C oracles), {el:.1f} s")
    return res


def cpu_baseline(cfg, dt, budget_s=12.0):
    """C oracle (FP64, one core) on the same plate, run for ~budget_s."""
    from oracle.plate import OraclePlate
    sim = OraclePlate(cfg, dt=dt)
    sim.step()
    t0 = time.perf_counter()
    steps = 0
    while time.perf_counter() - t0 < budget_s:
        sim.step()
        steps += 1
    el = time.perf_counter() - t0
    n = cfg.nx * cfg.ny
    return dict(value=n * steps / el, unit="particle-steps/s", cores=1, kind="port",
                sample=f"{steps} steps of the {cfg.nx}x{cfg.ny} plate from rest "
                       f"(FP64 C oracle, cell grid rebuilt every step), {el:.1f} s")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps of the headline workload: the 1M-disk column collapse "
                         "(default 50000), or the plate with --workload plate (default 20000)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed warm-up steps (column: 1000 on a throw-away engine; "
                         "plate with --workload plate: 200)")
    ap.add_argument("--plate-steps", type=int, default=None,
                    help="timed plate steps in the full run (default 20000)")
    ap.add_argument("--plate-warmup", type=int, default=None)
    ap.add_argument("--chunk", type=int, default=100, help="plate steps per captured graph")
    ap.add_argument("--plain-steps", type=int, default=30)
    ap.add_argument("--event-launches", type=int, default=0,
                    help="steps of the kernel-timing pass (0: same as --steps)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--disk-steps", type=int, default=2000)
    ap.add_argument("--disk-warmup", type=int, default=200)
    ap.add_argument("--d3-steps", type=int, default=2048)
    ap.add_argument("--d3-warmup", type=int, default=128)
    ap.add_argument("--workload", choices=("all", "plate", "bed", "ensemble", "serve", "disk",
                                           "horizon", "dem3d"),
                    default="all",
                    help="plate: the 1024x128 plate only; bed: the 1M-disk column only; "
                         "ensemble: the 512-plate batch only; serve: the request loop only; "
                         "disk: the 1M-disk granular2d box only; horizon: FP32 broken-bond "
                         "horizon of the 1024x128 plate against fp32_margin and dt; dem3d: "
                         "the 47^3-sphere IceDEM3D lattice only")
    ap.add_argument("--serve-window", type=float, default=60.0)
    ap.add_argument("--serve-chunk", type=int, default=250)
    ap.add_argument("--serve-slots", type=int, default=64)
    ap.add_argument("--ens-steps", type=int, default=10000)
    ap.add_argument("--bed-plain-steps", type=int, default=60)
    ap.add_argument("--bed-from", type=int, default=0,
                    help="start the bed timing this many steps into the collapse")
    ap.add_argument("--bed-start", default=None,
                    help="start the bed timing from a bed restart file")
    ap.add_argument("--bed-save-start", default=None,
                    help="write the --bed-from start state to this restart file")
    ap.add_argument("--bed-windows", type=int, default=BED_WINDOWS,
                    help="timed windows spread over the collapse when fewer than 50000 "
                         "column steps are timed (1: one window from release)")
    ap.add_argument("--bed-save-phases", default=None,
                    help="settle the column, write its release state and the states 25000 "
                         "and 48000 steps into the collapse to this directory, and exit")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a 512 MiB device copy (for counter calibration)")
    args = ap.parse_args(argv)
    plate_only = args.workload == "plate"
    if args.plate_steps is None:
        args.plate_steps = args.steps if plate_only and args.steps is not None else 20000
    if args.plate_warmup is None:
        args.plate_warmup = args.warmup if plate_only and args.warmup is not None else 200
    if args.steps is None:
        args.steps = 50000
    if args.warmup is None:
        args.warmup = 1000
    if args.copy_calibration:
        copy_calibration()
        return 0
    if args.bed_save_phases:
        save_bed_phases(Path(args.bed_save_phases))
        return 0
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: every plate and request of this workload runs on a "
              "single MI355X and no multi-GPU path is built.", file=sys.stderr)
        return 2
    # host-baseline workers are started before this process touches the GPU
    args.cpu_pools = {}
    try:
        if not args.no_cpu and args.workload in ("all", "bed"):
            bcfg, _, r0 = c3_inputs()
            args.cpu_pools["bed"] = CpuPool("bed", bcfg.stable_dt(float(r0.min())),
                                            host_workers())
        if not args.no_cpu and args.workload in ("all", "plate"):
            args.cpu_pools["plate"] = CpuPool("plate", c2_config("cpu").degree_bound_dt(),
                                              host_workers())
        if not args.no_cpu and args.workload in ("all", "disk"):
            args.cpu_pools["disk"] = CpuPool("disk", disk_dt(), host_workers())
    except OSError as exc:
        print(f"bench.py: host-baseline workers not started ({exc})", file=sys.stderr)
    try:
        return run_workloads(args)
    finally:
        for pool in args.cpu_pools.values():
            pool.close()


def run_workloads(args):
    if args.workload == "horizon":
        print(json.dumps(bench_horizon(args)))
        return 0
    if args.workload in ("bed", "ensemble", "serve", "disk", "dem3d"):
        bed = {"bed": bench_bed, "ensemble": bench_ensemble, "serve": bench_serve,
               "disk": bench_disk, "dem3d": bench_dem3d}[args.workload](args)
        line = {"metric": METRIC, "value": bed.pop("value"), "unit": "particle-steps/s",
                "n_gpus": 1, "steps": bed["steps"], "warmup": bed.pop("warmup", 0),
                "ms_per_step": bed["ms_per_step"], "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": bed.pop("dtype"),
                "data": "synthetic: " + {"bed": "seeded loose column packing",
                                         "ensemble": "seeded parameter sweep of plates",
                                         "serve": "seeded Poisson request stream",
                                         "disk": "seeded jittered-grid disk box",
                                         "dem3d": "cubic bonded-sphere lattice"}[args.workload],
                "config": {"workload": bed.pop("workload")}}
        line.update(bed)
        print(json.dumps(line))
        return 0
    if args.workload == "plate":
        print(json.dumps(bench_plate(args)))
        return 0
    # headline: the 1M-disk column (the configuration the FP32-force /
    # FP64-state target is stated at); the other configurations ride along
    bed = bench_bed(args)
    progress("bed column done")
    line = {"metric": METRIC, "value": bed.pop("value"), "unit": "particle-steps/s",
            "n_gpus": 1, "steps": bed["steps"], "warmup": bed.pop("warmup"),
            "ms_per_step": bed["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": bed.pop("dtype"),
            "data": "synthetic: seeded loose column packing",
            "config": {"workload": bed.pop("workload"), "disks": BED_N}}
    line.update(bed)
    c2 = bench_plate(args)
    c2.pop("metric")
    line["c2_plate"] = c2
    progress("1024x128 plate done")
    line["c4_ensemble"] = bench_ensemble(args)
    progress("512-plate ensemble done")
    line["c5_serve"] = bench_serve(args)
    progress("request loop done")
    line["disk_system"] = bench_disk(args)
    progress("1M-disk box done")
    line["dem3d_scale"] = bench_dem3d(args)
    progress("47^3 sphere lattice done")
    print(json.dumps(line))
    return 0


def bench_plate(args):
    """The 1024 x 128 plate: FP32 and FP64 forces over graph-captured steps,
    kernel events on a second pass, the plain FP64 path, the CPU oracle."""
    import torch
    from tensordem_amd.engine import PlateEngine, PlateSpec

    cfg = c2_config()
    dt = cfg.degree_bound_dt()
    n = cfg.nx * cfg.ny
    chunk = max(2, args.chunk - args.chunk % 2)
    steps = max(chunk, args.plate_steps - args.plate_steps % chunk)

    def engine(fp32):
        return PlateEngine([PlateSpec(cfg, dt=dt)], fp32=fp32)

    results = {}
    for fp32 in (True, False):
        eng = engine(fp32)
        eng.step(args.plate_warmup - args.plate_warmup % 2)
        eng.graph(chunk // 2)
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0.record(stream)
        eng.run_graph(chunk // 2, repeats=steps // chunk)
        e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        gpu_s = e0.elapsed_time(e1) * 1e-3
        st = eng.status()[0]
        broken = int((~eng.alive).sum())
        eng.close()
        del eng
        # same workload and phase again, each step kernel timed by HIP events
        # attached to its dispatch (the ledger launch is not in the bracket)
        eng = engine(fp32)
        eng.step(args.plate_warmup - args.plate_warmup % 2)
        k_s = kernel_event_time(eng, args.event_launches if args.event_launches else steps)
        results["fp32" if fp32 else "fp64"] = dict(
            seconds=gpu_s, wall_s=wall, ms_per_step=gpu_s / steps * 1e3,
            rate=n * steps / gpu_s, kernel_s=k_s, status=int(st), broken=broken)
        eng.close()
        del eng
        progress(f"plate {'FP32' if fp32 else 'FP64'} forces done")
    r32, r64 = results["fp32"], results["fp64"]
    for name, r in results.items():
        if r["status"]:
            print(f"bench.py: {name} run ended with status bits {r['status']} (contact search "
                  "invalid or non-finite state); its rate is not a valid measurement",
                  file=sys.stderr)

    plain = None
    if not args.no_plain:
        ps = time_plain(cfg, dt, args.plain_steps, 3, pinv=True)
        ps2 = time_plain(cfg, dt, args.plain_steps, 3, pinv=False)
        plain = dict(value=n / min(ps, ps2), unit="particle-steps/s",
                     ms_per_step=min(ps, ps2) * 1e3, steps=args.plain_steps,
                     ms_per_step_reference_pinv=ps * 1e3,
                     ms_per_step_closed_form_inverse=ps2 * 1e3,
                     what="eager PyTorch FP64 restatement of IceDEM.step on the same GPU "
                          "(per-step sort-based cell list); the faster of the reference's "
                          "torch.linalg.pinv and a closed-form 2x2 inverse")
    bytes_per_launch = PLATE_BYTES_PER_DISK * n
    achieved = bytes_per_launch / r32["kernel_s"] / 1e9
    line = {
        "metric": METRIC, "value": r32["rate"], "unit": "particle-steps/s", "n_gpus": 1,
        "steps": steps, "warmup": args.plate_warmup, "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-force/f64-state", "data": "synthetic: deterministic reference lattice",
        "config": {"workload": "large single ice plate 1024x128 (131072 disks, 520834 bonds), "
                               "indenter r=0.5 m at 0.5 m/s, strains 0.01/0.02, "
                               f"degree-bounded dt={dt:.6g} s, 20000 steps",
                   "nx": 1024, "ny": 128, "dt_s": dt, "graph_chunk_steps": chunk,
                   "broken_bonds_end": r32["broken"], "status_bits": r32["status"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profile(),
                     "kernel": "plate_step_kernel<float, true, true>",
                     "kernel_us": r32["kernel_s"] * 1e6,
                     "bytes_per_launch": bytes_per_launch},
        "fp64_kernel": {"value": r64["rate"], "ms_per_step": r64["ms_per_step"],
                        "kernel_us": r64["kernel_s"] * 1e6, "broken_bonds_end": r64["broken"]},
    }
    if plain:
        line["plain_fp64_gpu"] = plain
        line["speedup_vs_plain_fp64"] = r32["rate"] / plain["value"]
    if not args.no_cpu:
        line["cpu_baseline"] = with_pool(cpu_baseline(c2_config("cpu"), dt, args.cpu_seconds),
                                         getattr(args, "cpu_pools", {}).get("plate"),
                                         args.cpu_seconds, "the 1024x128 plate from rest")
    return line


if __name__ == "__main__":
    sys.exit(main())

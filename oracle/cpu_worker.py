"""One CPU worker of the benchmark's host baseline (test infrastructure only).

Runs the FP64 C oracle on one of the benchmark workloads for a fixed wall
time and prints one JSON line {"particle_steps", "seconds", "steps"}.  The
benchmark starts several of these as child processes at once (one per host
core it uses) to time the host's aggregate throughput; the product path never
runs this module.

    python -m oracle.cpu_worker bed --dt DT --seconds T
    python -m oracle.cpu_worker plate --dt DT --wait
    python -m oracle.cpu_worker disk --dt DT --wait

`bed`: the 1 048 576-disk deposited column (radii U[0.8, 1.2] x 0.01 m,
seed 0) before settling; `plate`: the 1024 x 128 plate indentation run from rest (indenter
r = 0.5 m at 0.5 m/s, strains 0.01 / 0.02); `disk`: a 256 x 256 box of rigid
disks with rotation (radii U[0.9, 1.1] x 0.01 m on a jittered grid, walls,
gravity, mu = 0.5, mu_rolling = 0.05, seed 0).  With --wait, the worker prints
"ready" after its set-up and then reads one line "START_AT SECONDS" from
stdin: it sleeps until the wall-clock instant START_AT and runs for SECONDS,
so that concurrent workers time the same window.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def _bed(dt):
    from oracle.bed import OracleBed
    from tensordem_amd.bedconfig import BedConfig, column_deposit
    x, r, _ = column_deposit(1 << 20, 0.01, seed=0)
    return OracleBed(BedConfig(), x, r, dt=dt), len(r)


def _plate(dt):
    from oracle.plate import OraclePlate
    from tensordem_amd.config import DEMConfig
    cfg = DEMConfig(nx=1024, ny=128, radius=0.025, tool_radius=0.5, tool_speed=0.5,
                    breaking_strain=0.01, shear_breaking_strain=0.02, device="cpu")
    return OraclePlate(cfg, dt=dt), cfg.nx * cfg.ny


def _disk(dt):
    from oracle.disk import OracleDisk
    from tensordem_amd.granular2d import DiskConfig, jittered_box
    pos, r, vel, walls = jittered_box(256, 0.01, seed=0)
    sim = OracleDisk(pos, r, DiskConfig(mu=0.5, mu_rolling=0.05, gravity=(0.0, -9.81), dt=dt),
                     walls=walls)
    sim.vel[:] = vel
    return sim, len(r)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("kind", choices=("bed", "plate", "disk"))
    ap.add_argument("--dt", type=float, required=True)
    ap.add_argument("--seconds", type=float, default=10.0)
    ap.add_argument("--wait", action="store_true")
    args = ap.parse_args(argv)
    sim, n = {"bed": _bed, "plate": _plate, "disk": _disk}[args.kind](args.dt)
    sim.step()  # warm-up: first-touch of the oracle's buffers
    if args.wait:
        print("ready", flush=True)
        line = sys.stdin.readline().split()
        if len(line) != 2:
            return 2
        start_at, args.seconds = float(line[0]), float(line[1])
        time.sleep(max(0.0, start_at - time.time()))
    t0 = time.perf_counter()
    steps = 0
    while time.perf_counter() - t0 < args.seconds:
        sim.step()
        steps += 1
    el = time.perf_counter() - t0
    print(json.dumps({"particle_steps": n * steps, "seconds": el, "steps": steps}))
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Write `plate_c2_fp64_horizon.json`: the FP64 oracle's state digests along
the 1024 x 128 plate indentation run (131 072 disks, 520 834 bonds, indenter
radius 0.5 m at 0.5 m/s, strains 0.01 / 0.02, degree-bounded dt, 20 000
steps).

The oracle (`oracle/dem2d.c`, a restatement of src/tensordem/dem.py:226-382
pinned by the reference's committed fracture and ledger artifacts) needs
about half an hour of one host core for this run, too long for a test, so
its results are kept as digests: every `--every` steps the broken-bond count,
SHA-256 of the alive mask (uint8 per bond, reference bond order), of the
positions and of the velocities (little-endian float64, (N, 2) row-major),
and the ledger.  The GPU test replays the same run through the HIP FP64 step
and compares the digests bit for bit.

Usage: python tests/golden/make_plate_c2_horizon.py [--steps 20000] [--every 500]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main() -> int:
    from oracle.plate import OraclePlate
    from tensordem_amd import DEMConfig

    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--every", type=int, default=500)
    ap.add_argument("--out", default=str(Path(__file__).with_name("plate_c2_fp64_horizon.json")))
    args = ap.parse_args()

    cfg = DEMConfig(nx=1024, ny=128, radius=0.025, tool_radius=0.5, tool_speed=0.5,
                    breaking_strain=0.01, shear_breaking_strain=0.02, device="cpu")
    dt = cfg.degree_bound_dt()
    sim = OraclePlate(cfg, dt=dt)
    rows = []
    t0 = time.perf_counter()
    for k in range(1, args.steps + 1):
        sim.step()
        if k % args.every == 0:
            acc = sim.accounting
            rows.append(dict(
                step=k, broken=sim.broken_bonds,
                alive_sha256=digest(sim.alive.astype(np.uint8)),
                positions_sha256=digest(sim.positions.astype("<f8")),
                velocities_sha256=digest(sim.velocities.astype("<f8")),
                fracture_release_J=acc["fracture_release_J"],
                pair_contact_acquired_J=acc["pair_contact_acquired_J"],
                tool_work_J=acc["tool_work_J"],
                damping_dissipation_J=acc["damping_dissipation_J"]))
            print(f"step {k}: broken {sim.broken_bonds}, "
                  f"{time.perf_counter() - t0:.0f} s", flush=True)
    out = dict(
        what="FP64 oracle digests along the 1024x128 plate indentation run",
        config=dict(nx=cfg.nx, ny=cfg.ny, radius=cfg.radius, tool_radius=cfg.tool_radius,
                    tool_speed=cfg.tool_speed, breaking_strain=cfg.breaking_strain,
                    shear_breaking_strain=cfg.shear_breaking_strain, dt=dt,
                    dt_rule="DEMConfig.degree_bound_dt() (8 partners)"),
        steps=args.steps, every=args.every, generator="tests/golden/make_plate_c2_horizon.py",
        checkpoints=rows)
    Path(args.out).write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {args.out} ({len(rows)} checkpoints, {time.perf_counter() - t0:.0f} s)")
    return 0


if __name__ == "__main__":
    sys.exit(main())

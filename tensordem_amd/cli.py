"""Command line: the 2-D ice-breaking run of the reference CLI on the HIP path.

`python -m tensordem_amd dem-2d ...` takes the options of the reference's
`dem-2d` command (src/tensordem/cli.py:42-110, 261-285) and writes the same
three files: `history.csv` (one `IceDEM.diagnostics()` row per saved frame),
`trajectory.pt` (frames of positions, velocities, bond states and tool
positions) and `accounting.json` (the ledger plus bond and mass totals).
Between saved frames the plate advances with `IceDEM.run`, which does not
synchronise with the host per step, and frames reach the host through a ring
of pinned buffers (`FrameStream`), so saving a frame does not wait for the
device either.  The run needs the ROCm device
(`--device cuda`, the default here); `--precision fp32` and
`--dt-policy degree` are additions.
"""
from __future__ import annotations

import argparse
import csv
import json
import math
import sys
from dataclasses import asdict
from pathlib import Path

import torch


def read_hull_profile(path: Path) -> torch.Tensor:
    """Local hull vertices (m) from a CSV with x,y columns, validated as the
    reference validates them: at least two finite vertices, no zero-length
    segment."""
    try:
        with Path(path).open(newline="", encoding="utf-8-sig") as fh:
            rows = csv.DictReader(fh)
            if rows.fieldnames is None or not {"x", "y"} <= set(rows.fieldnames):
                raise ValueError("hull profile CSV must contain x,y column headers")
            pts = []
            for line, row in enumerate(rows, start=2):
                try:
                    pts.append((float(row["x"]), float(row["y"])))
                except (TypeError, ValueError) as exc:
                    raise ValueError(f"invalid x,y value in hull profile CSV line {line}") from exc
    except OSError as exc:
        raise ValueError(f"cannot read hull profile CSV: {exc}") from exc
    if len(pts) < 2:
        raise ValueError("hull profile CSV must contain at least two vertices")
    prof = torch.tensor(pts, dtype=torch.float64)
    if not bool(torch.isfinite(prof).all()):
        raise ValueError("hull profile coordinates must be finite")
    if bool((torch.linalg.vector_norm(prof[1:] - prof[:-1], dim=1) <= 1e-12).any()):
        raise ValueError("hull profile must not contain zero-length segments")
    return prof


def _make_sim(args):
    from .config import DEMConfig
    from .icedem import IceDEM

    hull = read_hull_profile(args.hull_profile) if args.hull_profile else None
    speed = max(args.speed, math.hypot(args.hull_velocity_x, args.hull_velocity_y), 1e-12)
    cfg = DEMConfig(nx=args.nx, ny=args.ny, dt=args.dt, tool_speed=speed,
                    breaking_strain=args.breaking_strain,
                    shear_breaking_strain=args.shear_breaking_strain, device=args.device)
    kw = dict(precision=args.precision, dt_policy=args.dt_policy)
    if hull is None:
        return IceDEM(cfg, **kw)
    return IceDEM(cfg, hull_profile=hull, hull_start=(args.hull_start_x, args.hull_start_y),
                  hull_velocity=(args.hull_velocity_x, args.hull_velocity_y), **kw)


class FrameStream:
    """Saved frames of an IceDEM run, copied device -> host without a host
    wait per frame.

    `record()` enqueues the diagnostics reductions and copies of positions,
    velocities and bond states into a ring of pinned host slots on the
    current stream and returns; the host waits for a slot only when the ring
    comes round to it again, `ring` frames later, and then hands the frame to
    `sink(row, positions, velocities, bond_alive, tool_position)`.  Status bits
    carried with each frame raise as `IceDEM` would (non-finite state,
    invalid contact search).  `close()` drains the ring in order."""

    def __init__(self, sim, sink, ring: int = 4):
        self.sim, self.sink = sim, sink
        n, nb = len(sim.positions), int(sim.bonded.sum().item())
        pin = dict(dtype=torch.float64, pin_memory=True)
        self.slots = [dict(vec=torch.empty(sim._DIAG_LEN, **pin),
                           pos=torch.empty((n, 2), **pin), vel=torch.empty((n, 2), **pin),
                           alive=torch.empty(nb, dtype=torch.bool, pin_memory=True),
                           event=torch.cuda.Event(), time=None) for _ in range(max(1, ring))]
        self.count = 0

    def record(self) -> None:
        slot = self.slots[self.count % len(self.slots)]
        if slot["time"] is not None:
            self._drain(slot)
        sim = self.sim
        vec = sim._diagnostics_device()
        slot["vec"].copy_(vec, non_blocking=True)
        slot["pos"].copy_(sim.positions, non_blocking=True)
        slot["vel"].copy_(sim.velocities, non_blocking=True)
        slot["alive"].copy_(sim.alive[sim.bonded], non_blocking=True)
        slot["event"].record()
        slot["time"] = sim.time
        self.count += 1

    def _drain(self, slot) -> None:
        slot["event"].synchronize()
        v = slot["vec"].tolist()
        if int(v[8]):
            self.sim._raise_status()
        self.sink(self.sim._diagnostics_row(v, slot["time"]), slot["pos"].clone(),
                  slot["vel"].clone(), slot["alive"].clone(), slot["vec"][:2].clone())
        slot["time"] = None

    def close(self) -> None:
        k = len(self.slots)
        for i in range(max(0, self.count - k), self.count):
            slot = self.slots[i % k]
            if slot["time"] is not None:
                self._drain(slot)


def dem_2d(args: argparse.Namespace) -> dict:
    if args.steps < 1 or args.save_every < 1:
        sys.exit("--steps and --save-every must be positive")
    try:
        sim = _make_sim(args)
    except (ValueError, RuntimeError, TypeError) as exc:
        sys.exit(str(exc))
    out = Path(args.output)
    out.mkdir(parents=True, exist_ok=True)
    frames = {"positions": [], "velocities": [], "bond_alive": [], "tool_positions": [],
              "times": []}

    with (out / "history.csv").open("w", newline="", encoding="utf-8") as fh:
        writer = csv.DictWriter(fh, fieldnames=list(sim._diagnostics_row(
            [0.0] * sim._DIAG_LEN, 0.0)))
        writer.writeheader()

        def sink(row, pos, vel, alive, tool):
            writer.writerow(row)
            frames["positions"].append(pos)
            frames["velocities"].append(vel)
            frames["bond_alive"].append(alive)
            frames["tool_positions"].append(tool)
            frames["times"].append(row["time"])

        stream = FrameStream(sim, sink)
        stream.record()
        done = 0
        while done < args.steps:
            n = min(args.save_every - done % args.save_every, args.steps - done)
            sim.run(n)
            done += n
            if done % args.save_every == 0 or done == args.steps:
                stream.record()
        stream.close()
    torch.save({
        "config": asdict(sim.config),
        "dt": sim.dt,
        "times": torch.tensor(frames["times"], dtype=torch.float64),
        "positions": torch.stack(frames["positions"]),
        "velocities": torch.stack(frames["velocities"]),
        "bond_pairs": sim.pairs[sim.bonded].cpu(),
        "bond_alive": torch.stack(frames["bond_alive"]),
        "fixed": sim.fixed.cpu(),
        "tool_positions": torch.stack(frames["tool_positions"]),
        "hull_profile": None if sim.hull_profile is None else sim.hull_profile.cpu(),
        "hull_velocity": sim.hull_velocity.cpu(),
    }, out / "trajectory.pt")
    acc = dict(sim.accounting)
    acc["total_bonds"] = int(sim.bonded.sum().item())
    acc["broken_bonds"] = sim.broken_bonds
    acc["failure_energy_sum_J"] = float(sim.failure_energy_J.sum().item())
    acc["total_mass_kg"] = float(sim.config.mass * len(sim.positions))
    acc["dt"] = sim.dt
    acc["steps"] = sim.step_count
    (out / "accounting.json").write_text(json.dumps(acc, indent=2, sort_keys=True,
                                                    allow_nan=False), encoding="utf-8")
    print(f"done: t={sim.time:.6f} s, dt={sim.dt:.6g} s, broken bonds "
          f"{sim.broken_bonds}/{acc['total_bonds']}; results in {out}")
    return acc


def parser() -> argparse.ArgumentParser:
    p = argparse.ArgumentParser(prog="tensordem_amd",
                                description="TensorDEM 2-D ice-breaking DEM on MI355X")
    sub = p.add_subparsers(dest="command")
    d = sub.add_parser("dem-2d", help="2-D ice-breaking DEM simulation")
    d.add_argument("--nx", type=int, default=21, help="lattice columns")
    d.add_argument("--ny", type=int, default=5, help="lattice rows")
    d.add_argument("--steps", type=int, default=2000, help="integration steps")
    d.add_argument("--dt", type=float, default=None, help="time step (s); default automatic")
    d.add_argument("--speed", type=float, default=1.0,
                   help="indenter speed downwards; with a hull, the dt speed floor (m/s)")
    d.add_argument("--hull-profile", type=Path, default=None,
                   help="hull polyline CSV in local coordinates, columns x,y (m)")
    d.add_argument("--hull-start-x", type=float, default=0.0)
    d.add_argument("--hull-start-y", type=float, default=0.0)
    d.add_argument("--hull-velocity-x", type=float, default=0.0)
    d.add_argument("--hull-velocity-y", type=float, default=0.0)
    d.add_argument("--breaking-strain", type=float, default=0.015)
    d.add_argument("--shear-breaking-strain", type=float, default=0.03)
    d.add_argument("--device", default="cuda", help="cuda (the HIP path has no CPU device)")
    d.add_argument("--precision", choices=("fp64", "fp32"), default="fp64")
    d.add_argument("--dt-policy", choices=("all_pairs", "degree"), default="all_pairs")
    d.add_argument("--output", type=Path, default=Path("results"))
    d.add_argument("--save-every", type=int, default=20, help="trajectory sampling interval")
    d.set_defaults(func=dem_2d)
    return p


def main(argv: list[str] | None = None) -> None:
    args = parser().parse_args(argv)
    if not getattr(args, "func", None):
        parser().print_help()
        sys.exit(2)
    args.func(args)

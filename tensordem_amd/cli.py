"""Command line: the 2-D ice-breaking run of the reference CLI on the HIP path.

`python -m tensordem_amd dem-2d ...` takes the options of the reference's
`dem-2d` command (src/tensordem/cli.py:42-110, 261-285) and writes the same
three files: `history.csv` (one `IceDEM.diagnostics()` row per saved frame),
`trajectory.pt` (frames of positions, velocities, bond states and tool
positions) and `accounting.json` (the ledger plus bond and mass totals).
Between saved frames the plate advances with `IceDEM.run`, which does not
synchronise with the host per step.  The run needs the ROCm device
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

    def save_frame(writer):
        writer.writerow(sim.diagnostics())
        frames["positions"].append(sim.positions.detach().cpu().clone())
        frames["velocities"].append(sim.velocities.detach().cpu().clone())
        frames["bond_alive"].append(sim.alive[sim.bonded].detach().cpu().clone())
        frames["tool_positions"].append(sim.tool_position.detach().cpu().clone())
        frames["times"].append(sim.time)

    with (out / "history.csv").open("w", newline="", encoding="utf-8") as fh:
        writer = csv.DictWriter(fh, fieldnames=list(sim.diagnostics()))
        writer.writeheader()
        save_frame(writer)
        done = 0
        while done < args.steps:
            n = min(args.save_every - done % args.save_every, args.steps - done)
            sim.run(n)
            done += n
            if done % args.save_every == 0 or done == args.steps:
                save_frame(writer)
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

"""Autotuner for batched plate launches: block shape and cell size.

For a given batch it times a small grid of launch configurations on the
device -- the per-member tile policy (tile width, hence the 256-disk block's
shape) and the contact-grid skin (hence the cell size 2r + skin) -- each over
a short graph-captured run from the batch's initial state, and keeps the
fastest configuration whose run ended with clean status bits.  Results are
cached per batch signature (member shapes, precision) in memory and,
optionally, in a JSON file.
"""
from __future__ import annotations

import hashlib
import json
import os
import time
from dataclasses import dataclass, field
from pathlib import Path

import torch

from .engine import PlateEngine, PlateSpec
from .ensemble import with_tiles

DEFAULT_CANDIDATES = (("default", 1.0), ("min_tiles", 1.0), ("min_tiles", 0.5),
                      ("min_tiles", 1.5), ("min_tiles_128", 1.0), ("min_tiles_128", 0.5),
                      (32, 1.0), (64, 1.0))

_memory: dict = {}


@dataclass
class TuneResult:
    tile_policy: object
    skin_factor: float
    us_per_step: float
    table: list = field(default_factory=list)
    seconds: float = 0.0
    cached: bool = False

    def apply(self, specs: list[PlateSpec]) -> list[PlateSpec]:
        return with_tiles(specs, self.tile_policy)

    def engine_kwargs(self) -> dict:
        return {"skin_factor": self.skin_factor}


def signature(specs: list[PlateSpec], fp32: bool) -> str:
    h = hashlib.sha1()
    h.update(b"fp32" if fp32 else b"fp64")
    for s in specs:
        c = s.config
        h.update(f"{c.nx},{c.ny},{c.radius},{s.hull_profile is None};".encode())
    return h.hexdigest()


def _time(specs, fp32, skin, steps, warmup, device):
    eng = PlateEngine(specs, device=device, fp32=fp32, skin_factor=skin)
    try:
        eng.step(warmup - warmup % 2)
        eng.graph(steps // 2)
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        eng.run_graph(steps // 2)
        e1.record(stream)
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / steps
        ok = not any(eng.status())
    finally:
        eng.close()
    return us, ok


def autotune_plates(specs: list[PlateSpec], *, fp32: bool = True, steps: int = 200,
                    warmup: int = 40, candidates=DEFAULT_CANDIDATES, device="cuda",
                    cache_file: str | os.PathLike | None = None) -> TuneResult:
    key = signature(specs, fp32)
    if key in _memory:
        r = _memory[key]
        return TuneResult(r.tile_policy, r.skin_factor, r.us_per_step, r.table, 0.0, True)
    if cache_file is not None and Path(cache_file).exists():
        data = json.loads(Path(cache_file).read_text())
        if key in data:
            d = data[key]
            r = TuneResult(d["tile_policy"], d["skin_factor"], d["us_per_step"], d["table"], 0.0, True)
            _memory[key] = r
            return r
    t0 = time.perf_counter()
    table = []
    for policy, skin in candidates:
        us, ok = _time(with_tiles(specs, policy), fp32, skin, steps, warmup, device)
        table.append({"tile_policy": policy, "skin_factor": skin, "us_per_step": us, "clean": ok})
    good = [t for t in table if t["clean"]]
    if not good:
        raise RuntimeError("autotune: every candidate ended with status bits set")
    best = min(good, key=lambda t: t["us_per_step"])
    r = TuneResult(best["tile_policy"], best["skin_factor"], best["us_per_step"], table,
                   time.perf_counter() - t0)
    _memory[key] = r
    if cache_file is not None:
        p = Path(cache_file)
        data = json.loads(p.read_text()) if p.exists() else {}
        data[key] = {"tile_policy": r.tile_policy, "skin_factor": r.skin_factor,
                     "us_per_step": r.us_per_step, "table": table}
        p.parent.mkdir(parents=True, exist_ok=True)
        p.write_text(json.dumps(data, indent=1))
    return r

"""Restart layouts of the bonded plate.

The reference's restart (`tensordem.restart/1`, src/tensordem/dem.py:486-601)
stores `alive` and the four `failure_*` arrays over all N(N-1)/2 particle
pairs in `torch.triu_indices` order (dem.py:116-137); only bonded pairs can
be alive or carry a failure record.  The device engine keeps them per bond.
This module converts between the two, so a plate restarts from a reference
checkpoint and writes checkpoints the reference loads.  For plates whose
pair arrays would be too large to write (N(N-1)/2 above
`MAX_REFERENCE_PAIRS`), the bond-indexed layout is written under a schema of
its own, `tensordem_amd.restart_bonds/1`, with otherwise identical fields.
"""
from __future__ import annotations

import numpy as np

REFERENCE_SCHEMA = "tensordem.restart/1"
BOND_SCHEMA = "tensordem_amd.restart_bonds/1"
PAIR_FIELDS = ("alive", "failure_mode", "failure_time_s", "failure_extension_m",
               "failure_energy_J")
# 2**22 pairs (a 2 896-disk plate): about 30 MB of JSON per pair field
MAX_REFERENCE_PAIRS = 1 << 22


def pair_count(n: int) -> int:
    return n * (n - 1) // 2


def triu_index(i, j, n: int) -> np.ndarray:
    """Position of pair (i, j), i < j, in torch.triu_indices(n, n, 1) order."""
    i = np.asarray(i, dtype=np.int64)
    j = np.asarray(j, dtype=np.int64)
    return i * n - (i * (i + 1)) // 2 + (j - i - 1)


def bonds_to_pairs(fields: dict, n: int, bond_i, bond_j) -> dict:
    """Bond-indexed `alive` / `failure_*` (numpy) -> all-pairs lists with the
    reference's values for non-bonded pairs (alive False, mode 0, time -1,
    extension 0, energy 0)."""
    p = pair_count(n)
    if p > MAX_REFERENCE_PAIRS:
        raise ValueError(f"{p} particle pairs exceed the all-pairs restart limit "
                         f"({MAX_REFERENCE_PAIRS}); write the bond-indexed layout")
    idx = triu_index(bond_i, bond_j, n)
    out = {}
    fill = {"alive": False, "failure_mode": 0, "failure_time_s": -1.0,
            "failure_extension_m": 0.0, "failure_energy_J": 0.0}
    dtypes = {"alive": np.bool_, "failure_mode": np.uint8}
    for key in PAIR_FIELDS:
        arr = np.full(p, fill[key], dtype=dtypes.get(key, np.float64))
        arr[idx] = np.asarray(fields[key]).astype(arr.dtype)
        out[key] = arr.tolist()
    return out


def pairs_to_bonds(state: dict, n: int, bond_i, bond_j) -> dict:
    """All-pairs lists of a reference checkpoint -> bond-indexed lists.

    Fails closed like dem.py:536-575: wrong lengths, non-boolean `alive`, a
    non-bonded pair that is alive or carries a failure mode, time, extension
    or energy."""
    p = pair_count(n)
    idx = triu_index(bond_i, bond_j, n)
    other = np.ones(p, dtype=bool)
    other[idx] = False
    out = {}
    for key in PAIR_FIELDS:
        values = state[key]
        if not isinstance(values, list) or len(values) != p:
            raise ValueError(f"invalid restart {key} shape")
        if key == "alive":
            if any(type(v) is not bool for v in values):
                raise ValueError("invalid boolean restart alive")
            arr = np.asarray(values, dtype=bool)
            if bool(arr[other].any()):
                raise ValueError("nonbonded pair cannot be alive")
            out[key] = arr[idx].tolist()
            continue
        arr = np.asarray(values, dtype=np.float64)
        if not np.isfinite(arr).all():
            raise ValueError(f"nonfinite restart {key}")
        rest = arr[other]
        expect = -1.0 if key == "failure_time_s" else 0.0
        if bool((rest != expect).any()):
            if key == "failure_mode":
                raise ValueError("invalid restart failure modes")
            raise ValueError(f"nonbonded pair has a failure record in {key}")
        sel = arr[idx]
        out[key] = sel.astype(np.int64).tolist() if key == "failure_mode" else sel.tolist()
    return out

"""Renumbering a bed into Morton order on the device (`BedEngine.renumber`)
against the FP64 oracle given the same permutation: a collapsing 2 000-disk
column, renumbered every 400 steps, stays bit-identical to the oracle
(positions, velocities, contact pairs and spring histories) across the
renumberings, and `ids` maps storage back to the original disks."""
import numpy as np
import pytest

from oracle.bed import OracleBed
from tensordem_amd.bed import BedEngine, BedSpec
from tensordem_amd.bedconfig import BedConfig, column_packing

pytestmark = pytest.mark.gpu


def _permute_oracle(ref: OracleBed, perm: np.ndarray) -> None:
    n = len(perm)
    inv = np.empty(n, np.int64)
    inv[perm] = np.arange(n)
    ref.positions = np.ascontiguousarray(ref.positions[perm])
    ref.velocities = np.ascontiguousarray(ref.velocities[perm])
    ref.radii = np.ascontiguousarray(ref.radii[perm])
    ref.xi_floor = np.ascontiguousarray(ref.xi_floor[perm])
    ref.xi_left = np.ascontiguousarray(ref.xi_left[perm])
    i, j = ref.hist_keys // n, ref.hist_keys % n
    a, b = inv[i], inv[j]
    xi = np.where((a > b)[:, None], -ref.hist_xi, ref.hist_xi)
    keys = np.minimum(a, b) * n + np.maximum(a, b)
    o = np.argsort(keys)
    ref.hist_keys = np.ascontiguousarray(keys[o])
    ref.hist_xi = np.ascontiguousarray(xi[o])


def _check(eng: BedEngine, ref: OracleBed, what: str) -> None:
    snap = eng.snapshot()
    n = len(ref.radii)
    keys = snap["pairs"][:, 0] * n + snap["pairs"][:, 1]
    assert np.array_equal(snap["positions"], ref.positions), f"{what}: positions"
    assert np.array_equal(snap["velocities"], ref.velocities), f"{what}: velocities"
    assert np.array_equal(keys, ref.hist_keys), f"{what}: contact pairs"
    assert np.array_equal(snap["pair_xi"], ref.hist_xi), f"{what}: spring histories"


def test_renumbered_column_matches_oracle_with_same_permutation():
    x, r = column_packing(2000, 0.01, seed=3)
    cfg = BedConfig()
    ref = OracleBed(cfg, x, r)
    v = np.random.default_rng(0).normal(scale=0.2, size=x.shape)
    ref.velocities[:] = v
    eng = BedEngine([BedSpec(cfg, x, r, dt=ref.dt, velocities=v)], fp32=False)
    ids = np.arange(len(r))
    for block in range(4):
        eng.step(400)
        for _ in range(400):
            ref.step()
        _check(eng, ref, f"before renumbering {block}")
        perm = eng.renumber().cpu().numpy()
        assert sorted(perm.tolist()) == list(range(len(r)))
        _permute_oracle(ref, perm)
        ids = ids[perm]
        assert np.array_equal(eng.ids.cpu().numpy(), ids)
        assert np.array_equal(eng.radius.cpu().numpy(), r[ids])
    eng.step(400)
    for _ in range(400):
        ref.step()
    _check(eng, ref, "after the last renumbering")
    assert len(ref.hist_keys) > 0
    assert int(eng.status()[0]) == 0

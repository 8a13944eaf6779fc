"""Bed restart files: written as .npz with a JSON header, read back without
unpickling, and rejected when incomplete or inconsistent (host logic only)."""
import numpy as np
import pytest

from tensordem_amd import _lib
from tensordem_amd.bed import load_bed_restart, save_bed_restart, validate_bed_restart
from tensordem_amd.bedconfig import BedConfig


def state(n=5):
    from dataclasses import asdict
    rng = np.random.default_rng(0)
    return {
        "schema": "tensordem_amd.bed_restart/1", "config": asdict(BedConfig()),
        "dt": 1e-5, "time": 3e-5, "step_count": 3,
        "positions": rng.random((n, 2)), "velocities": rng.random((n, 2)),
        "radii": np.full(n, 0.01), "accounting": np.zeros(_lib.BED_NACC),
        "pairs": np.array([[0, 1], [1, 3]]), "pair_xi": rng.normal(size=(2, 2)) * 1e-6,
        "wall_xi": np.zeros((n, 4)), "wall_active": np.array([1, 0, 2, 3, 0], np.uint8),
    }


def test_round_trip_is_exact(tmp_path):
    s = state()
    save_bed_restart(s, tmp_path / "r.npz")
    back = load_bed_restart(tmp_path / "r.npz")
    assert set(back) == set(s)
    for k, v in s.items():
        if isinstance(v, np.ndarray):
            assert np.array_equal(back[k], v) and back[k].dtype == v.dtype, k
        else:
            assert back[k] == v, k


@pytest.mark.parametrize("edit, msg", [
    (lambda s: s.update(schema="tensordem.restart/1"), "not a"),
    (lambda s: s.pop("wall_xi"), "missing"),
    (lambda s: s.update(velocities=np.zeros((4, 2))), "velocities shape"),
    (lambda s: s["positions"].__setitem__((2, 0), np.nan), "nonfinite"),
    (lambda s: s.update(pairs=np.array([[1, 0], [1, 3]])), "i < j"),
    (lambda s: s.update(pairs=np.array([[0, 5], [1, 3]])), "i < j"),
    (lambda s: s.update(pairs=np.array([[1, 3], [1, 3]])), "repeated"),
    (lambda s: s.update(pair_xi=np.zeros((3, 2))), "pair_xi shape"),
    (lambda s: s.update(wall_active=np.array([8, 0, 0, 0, 0])), "wall_active"),
    (lambda s: s.update(wall_active=np.array([4, 0, 0, 0, 0])), "right wall"),
    (lambda s: s.update(wall_xi=np.zeros((5, 5))), "wall_xi shape"),
    (lambda s: s.update(dt=0.0), "timestep"),
    (lambda s: s.update(step_count=-1), "step_count"),
])
def test_inconsistent_states_are_rejected(edit, msg):
    s = state()
    edit(s)
    with pytest.raises(ValueError, match=msg):
        validate_bed_restart(s)


def test_three_wall_states_validate():
    s = state()
    s["wall_xi"] = np.zeros((5, 6))
    s["wall_active"] = np.array([4, 5, 7, 0, 2], np.uint8)
    validate_bed_restart(s)

"""Host logic of the disk-assembly drop-in (no GPU): configuration and input
validation, boundary geometry helpers -- the reference's checks in
tests/test_granular2d_force_chain_jamming.py:112-123, 333-340 and
tests/test_granular2d_wall_velocity.py:32-45 -- and the device requirement."""
import math

import numpy as np
import pytest
import torch

from tensordem_amd.granular2d import (DiskConfig, DiskSystem, box_walls, polyline_segments,
                                      ship_segments)


def test_config_validation():
    DiskConfig()
    for bad in (dict(kn=-1.0), dict(kt_ratio=0.0), dict(density=float("nan")),
                dict(zeta=-0.1), dict(skin=-1.0), dict(dt=0.0), dict(gravity=(0.0,)),
                dict(drift_velocity=(0.0, float("inf")))):
        with pytest.raises(ValueError):
            DiskConfig(**bad)


def test_engine_rejects_invalid_input():
    with pytest.raises(ValueError):
        DiskSystem(np.zeros((2, 2)), np.array([0.01, -0.01]))
    with pytest.raises(ValueError):
        DiskSystem(np.zeros((2, 3)), np.array([0.01, 0.01]))
    with pytest.raises(ValueError):
        DiskSystem(np.zeros((2, 2)), np.array([0.01]))
    with pytest.raises(ValueError):
        DiskSystem(np.zeros((0, 2)), np.zeros(0))
    with pytest.raises(ValueError, match="stability limit"):
        DiskSystem(np.array([[0.3, 0.2], [0.3, 0.2]]), np.array([0.01, 0.01]),
                   DiskConfig(dt=1.0))
    walls = box_walls(0.2, 0.1, left=False, right=False)
    with pytest.raises(ValueError):
        DiskSystem([[0.1, 0.02]], [0.005], walls=walls, wall_velocity=[[0.5, 0.0]])
    with pytest.raises(ValueError):
        DiskSystem([[0.1, 0.02]], [0.005], walls=walls,
                   wall_velocity=[[0.5, np.nan], [0.5, 0.0]])
    with pytest.raises(ValueError):
        DiskSystem([[0.1, 0.02]], [0.005], walls=np.zeros((2, 3, 2)))
    with pytest.raises(ValueError):
        DiskSystem([[0.1, 0.02]], [0.005], periodic_x=-1.0)


@pytest.mark.skipif(torch.cuda.is_available(), reason="checks the CPU-only refusal")
def test_device_is_required():
    with pytest.raises(ValueError, match="ROCm device"):
        DiskSystem([[0.1, 0.02]], [0.005])


def test_ship_geometry_is_closed_and_sharp():
    segs = ship_segments(0.24, 0.16, (0.1, 0.05), bow=0.06)
    pts = segs[:, 0, :]
    assert np.allclose(segs[:, 1, :], np.roll(pts, -1, axis=0))
    assert np.isclose(float(np.ptp(pts[:, 0])), 0.24)
    assert np.isclose(float(np.ptp(pts[:, 1])), 0.16)
    # 3 starboard corners, 5 bow points each side, the stem, 2 port corners
    assert len(segs) == 3 + 5 + 1 + 5 + 2
    assert np.allclose(pts[np.argmax(pts[:, 0])], [0.34, 0.05])
    with pytest.raises(ValueError):
        ship_segments(0.1, 0.1, bow=0.2)
    blunt = ship_segments(1.0, 0.5)
    assert len(blunt) == 6


def test_box_walls_and_polyline():
    w = box_walls(2.0, 1.0, margin=0.1)
    assert w.shape == (4, 2, 2)
    np.testing.assert_array_equal(w[0], [[-0.1, 0.0], [2.1, 0.0]])      # bottom
    np.testing.assert_array_equal(w[1], [[-0.1, 1.0], [2.1, 1.0]])      # top
    np.testing.assert_array_equal(w[2], [[0.0, -0.1], [0.0, 1.1]])      # left
    np.testing.assert_array_equal(w[3], [[2.0, -0.1], [2.0, 1.1]])      # right
    assert box_walls(1.0, 1.0, left=False, right=False).shape == (2, 2, 2)
    with pytest.raises(ValueError):
        box_walls(0.0, 1.0)
    tri = polyline_segments([[0, 0], [1, 0], [0, 1]])
    np.testing.assert_array_equal(tri[2], [[0, 1], [0, 0]])
    with pytest.raises(ValueError):
        polyline_segments([[0, 0], [1, 0]])


def test_recommended_dt_formula():
    # dt = 0.1 / sqrt(2 kn / (m_min / 2)), m = rho pi r^2 t (engine.py:170, 231-235)
    cfg = DiskConfig()
    m = cfg.density * math.pi * 0.01**2 * cfg.thickness
    want = 0.1 / math.sqrt(2.0 * cfg.kn / (0.5 * m))
    try:
        s = DiskSystem([[0.0, 0.0], [1.0, 0.0]], [0.01, 0.02], cfg)
    except ValueError as e:           # no device here: the bound is checked first
        assert "ROCm" in str(e)
        return
    assert s.recommended_dt == want

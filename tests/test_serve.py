"""Serving: request sweep, footprint estimate and admission policy (host side)."""
import math

import numpy as np
import pytest

from tensordem_amd.serve import Server, device_bytes, sweep_requests


def test_sweep_is_poisson_half_plates_log_uniform():
    reqs = sweep_requests(rate=50.0, window=60.0, steps=5000, seed=0)
    assert abs(len(reqs) - 3000) < 4 * math.sqrt(3000)
    arr = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(arr) > 0) and arr[-1] <= 60.0
    gaps = np.diff(np.concatenate([[0.0], arr]))
    assert gaps.mean() == pytest.approx(1 / 50.0, rel=0.06)        # exponential, mean 1/rate
    assert np.std(gaps) == pytest.approx(1 / 50.0, rel=0.1)
    kinds = np.array([r.kind == "plate" for r in reqs])
    assert abs(kinds.mean() - 0.5) < 0.05
    n = np.array([r.n for r in reqs])
    assert n.min() >= 900 and n.max() <= 201_000
    assert abs(np.log(n).mean() - 0.5 * (math.log(1000) + math.log(200_000))) < 0.15
    again = sweep_requests(rate=50.0, window=60.0, steps=5000, seed=0)
    assert [(r.kind, r.n, r.arrival) for r in again] == [(r.kind, r.n, r.arrival) for r in reqs]


def test_footprint_grows_with_size():
    reqs = sweep_requests(rate=50.0, window=2.0, seed=1)
    for kind in ("plate", "bed"):
        rs = sorted([r for r in reqs if r.kind == kind], key=lambda r: r.n)
        b = [device_bytes(r) for r in rs]
        assert all(x > 0 for x in b)
        assert b[-1] > b[0]
        assert b[-1] / rs[-1].n > 100       # at least the state and lists


def test_admission_respects_budget_slots_and_order():
    reqs = sweep_requests(rate=50.0, window=1.0, seed=2)
    need = [device_bytes(r) for r in reqs]
    # the first two fit together, the third does not fit beside them (alone it would)
    srv = Server(device="cpu", hbm_budget=need[0] + need[1] + need[2] - 1, max_active=8)
    waiting = list(reqs)
    join = srv._admit(waiting, 0.0)
    admitted = join["plate"] + join["bed"]
    assert [r.rid for r in sorted(admitted, key=lambda r: r.rid)] == [0, 1]
    assert waiting[0].rid == 2 and srv.used == need[0] + need[1]
    # slots: max_active members per batch, FIFO (a blocked head blocks the rest)
    srv = Server(device="cpu", hbm_budget=1e15, max_active=1)
    waiting = list(reqs)
    join = srv._admit(waiting, 0.0)
    assert len(join["plate"]) <= 1 and len(join["bed"]) <= 1
    with pytest.raises(ValueError):
        Server(device="cpu", hbm_budget=10)._admit(list(reqs), 0.0)


def test_steps_must_be_whole_chunks():
    reqs = sweep_requests(rate=50.0, window=0.2, steps=300, seed=3)
    with pytest.raises(ValueError):
        Server(device="cpu", chunk=250).run(reqs)
    with pytest.raises(ValueError):
        Server(device="cpu", chunk=3)

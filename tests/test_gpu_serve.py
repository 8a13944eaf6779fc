"""Continuous batching on the device: every request's result is the same
alone and in the mix (bit for bit: members never interact, and repacking
carries the complete per-member state), admission follows the HBM budget,
and a short real-time window completes with sane latencies."""
import numpy as np
import pytest

from tensordem_amd.serve import (Server, bed_request, device_bytes, latency_summary,
                                 plate_request, run_alone, sweep_requests)

pytestmark = pytest.mark.gpu


def mixed_requests(steps=1000):
    rng = np.random.default_rng(7)
    reqs = []
    for rid, (kind, n, arrival) in enumerate([("plate", 1500, 0), ("bed", 1200, 0),
                                              ("bed", 2600, 1), ("plate", 3000, 1),
                                              ("plate", 800, 2), ("bed", 1800, 4),
                                              ("plate", 2000, 5), ("bed", 900, 6)]):
        make = plate_request if kind == "plate" else bed_request
        reqs.append(make(rid, n, rng, steps, float(arrival)))
    return reqs


@pytest.mark.parametrize("fp32", [True, False])
def test_requests_in_the_mix_equal_alone(fp32):
    reqs = mixed_requests()
    # two members per batch: later requests queue, join mid-flight, and every
    # join or finish repacks the batch around the members still running
    srv = Server(fp32=fp32, max_active=2, chunk=250, keep_state=True)
    done = srv.run(reqs, realtime=False)
    assert len(done) == len(reqs)
    assert max(p + b for _, p, b, _ in srv.log) == 4
    assert any(w > 0 for *_, w in srv.log)           # some requests had to wait
    for r in done:
        alone = run_alone(r, fp32=fp32, chunk=250)
        got = r.result
        assert got["step_count"] == alone["step_count"] == 1000
        assert got["status"] == 0
        assert np.array_equal(got["x"], alone["x"]), (r.rid, r.kind)
        assert np.array_equal(got["v"], alone["v"]), (r.rid, r.kind)
        assert np.array_equal(got["accounting"], alone["accounting"]), (r.rid, r.kind)
        if r.kind == "plate":
            assert np.array_equal(got["alive"], alone["alive"]), r.rid
        else:
            assert got["rebuilds"] >= 1


def test_hbm_budget_admits_one_at_a_time():
    reqs = mixed_requests(steps=500)[:4]
    budget = max(device_bytes(r) for r in reqs) * 1.01
    srv = Server(hbm_budget=budget, max_active=8, chunk=250)
    done = srv.run(reqs, realtime=False)
    assert len(done) == 4
    assert max(p + b for _, p, b, _ in srv.log) == 1
    assert srv.used == 0


def test_realtime_window_completes():
    reqs = sweep_requests(rate=20.0, window=1.5, steps=500, n_range=(1000, 6000), seed=3)
    srv = Server(chunk=250)
    done = srv.run(reqs, realtime=True)
    assert len(done) == len(reqs) > 10
    lat = latency_summary(done)
    assert 0 < lat["p50_s"] <= lat["p99_s"] <= lat["max_s"]
    assert all(r.admitted >= r.arrival for r in done)
    assert all(r.result["status"] == 0 for r in done)

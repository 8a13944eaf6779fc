"""The batched and served workloads at their full stated size.

* 512 independent bonded plates (nx log-uniform on [16, 256], ny uniform on
  [4, 32], tool speed U[0.1, 1] m/s, per-disk Weibull(5) strengths with means
  0.01 / 0.02, seed 0), 10 000 steps in one batch, FP64 and FP32 forces:
  no status bit, every 32nd member bit-identical to its solo run, and in
  FP64 the smallest member bit-identical to the FP64 oracle over all 10 000
  steps.
* The request loop over a Poisson window (50 requests/s for 60 s, seed 0),
  half plates half beds, 1k-200k disks log-uniform, 5 000 steps each, served
  in real time: every request completes with a clean status, and 16 requests
  sampled across the window are bit-identical to the same request served
  alone.
"""
import numpy as np
import pytest
import torch

from oracle.plate import OraclePlate
from tensordem_amd.engine import PlateEngine
from tensordem_amd.ensemble import random_ensemble, with_tiles
from tensordem_amd.serve import Server, latency_summary, run_alone, sweep_requests

pytestmark = pytest.mark.gpu

STEPS_C4 = 10_000


def np_(t):
    return t.detach().cpu().numpy()


def member(eng, k):
    m = eng.read_members()[k]
    n = m.nx * m.ny
    return (np_(eng.x[m.first:m.first + n]), np_(eng.v[m.first:m.first + n]),
            np_(eng.alive[m.bond_first:m.bond_first + m.nbonds]))


@pytest.mark.timeout(900)
@pytest.mark.parametrize("fp32", [False, True])
def test_512_plate_ensemble_full_run(fp32):
    specs = random_ensemble(512, seed=0, device="cuda")
    print(f"512-plate ensemble, fp32={fp32}: {sum(s.config.nx * s.config.ny for s in specs)} "
          f"disks, {STEPS_C4} steps", flush=True)
    batch = PlateEngine(with_tiles(specs, "min_tiles"), fp32=fp32)
    batch.run_graph(250, repeats=STEPS_C4 // 500)
    torch.cuda.synchronize()
    assert not any(batch.status())
    broken = int((~batch.alive).sum())
    assert broken > 100_000
    for k in range(0, 512, 32):
        solo = PlateEngine([specs[k]], fp32=fp32)
        solo.step(STEPS_C4)
        xb, vb, ab = member(batch, k)
        xs, vs, as_ = member(solo, 0)
        assert np.array_equal(xb, xs) and np.array_equal(vb, vs), k
        assert np.array_equal(ab, as_), k
        assert np.allclose(np_(batch.accounting[k]), np_(solo.accounting[0]), rtol=1e-12,
                           atol=1e-18), k
    if fp32:
        return
    k = min(range(512), key=lambda i: specs[i].config.nx * specs[i].config.ny)
    s = specs[k]
    ref = OraclePlate(s.config, dt=s.dt, eps_t_disk=s.eps_t_disk, eps_s_disk=s.eps_s_disk)
    for _ in range(STEPS_C4):
        ref.step()
    x, v, alive = member(batch, k)
    print(f"member {k} ({s.config.nx}x{s.config.ny}): {ref.broken_bonds} broken bonds in the "
          f"oracle", flush=True)
    assert np.array_equal(x, ref.positions)
    assert np.array_equal(v, ref.velocities)
    assert np.array_equal(alive, ref.alive.astype(bool))


@pytest.mark.timeout(900)
def test_poisson_window_served_in_real_time():
    reqs = sweep_requests(rate=50.0, window=60.0, steps=5000, seed=0)
    sample = set(range(0, len(reqs), len(reqs) // 16))
    print(f"Poisson window: {len(reqs)} requests, {len(sample)} sampled", flush=True)
    srv = Server(chunk=250, keep_state=sample)
    done = srv.run(reqs, realtime=True)
    assert len(done) == len(reqs) > 2500
    assert all(r.result["status"] == 0 for r in done)
    assert all(r.admitted >= r.arrival for r in done)
    assert all(r.result["step_count"] == 5000 for r in done)
    lat = latency_summary(done)
    print(f"p50 {lat['p50_s']:.3f} s, p99 {lat['p99_s']:.3f} s", flush=True)
    assert 0 < lat["p50_s"] <= lat["p99_s"] <= lat["max_s"]
    kinds = set()
    for r in done:
        if r.rid not in sample:
            continue
        kinds.add(r.kind)
        alone = run_alone(r, fp32=True, chunk=250)
        got = r.result
        assert np.array_equal(got["x"], alone["x"]), (r.rid, r.kind, r.n)
        assert np.array_equal(got["v"], alone["v"]), (r.rid, r.kind, r.n)
        assert np.array_equal(got["accounting"], alone["accounting"]), (r.rid, r.kind)
        if r.kind == "plate":
            assert np.array_equal(got["alive"], alone["alive"]), r.rid
    assert kinds == {"plate", "bed"}

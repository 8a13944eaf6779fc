"""The allocation-free fragment entry point (tdm_plate_fragments_noalloc,
static device scratch) against the workspace entry point
(tdm_plate_fragments), both called through the C ABI: bit-identical labels,
order, component starts, counts and sums on two calls in a row, the second
on a smaller plate (the scratch keeps the first call's contents), and sizes
above its limit rejected before anything runs."""
import ctypes

import numpy as np
import pytest
import torch

from tensordem_amd import DEMConfig, _lib
from tensordem_amd.icedem import IceDEM

pytestmark = pytest.mark.gpu

LIMIT = 1 << 18


def shattered(nx, ny, broken, seed):
    sim = IceDEM(DEMConfig(nx=nx, ny=ny, device="cuda"), dt_policy="degree")
    rng = np.random.default_rng(seed)
    sim.alive = torch.as_tensor(rng.random(sim.alive.numel()) >= broken, device="cuda")
    sim.velocities = torch.as_tensor(rng.normal(scale=0.1, size=(nx * ny, 2)))
    sim._sync_inputs()
    return sim


def outputs(n):
    i32 = dict(dtype=torch.int32, device="cuda")
    # filled with garbage so nothing unwritten can pass as equal by accident
    return [torch.full((n,), -7, **i32), torch.full((n,), -7, **i32),
            torch.full((n + 1,), -7, **i32), torch.full((1,), -7, **i32),
            torch.full((n, 4), float("nan"), dtype=torch.float64, device="cuda")]


def call(sim, allocation_free):
    e = sim._eng
    nx, ny = sim.config.nx, sim.config.ny
    n = nx * ny
    lib = _lib.load()
    outs = outputs(n)
    ptrs = [t.data_ptr() for t in outs]
    args = (e.links.data_ptr(), e.x.data_ptr(), e.v.data_ptr(), nx, ny, *ptrs)
    stream = torch.cuda.current_stream().cuda_stream
    if allocation_free:
        rc = lib.tdm_plate_fragments_noalloc(*args, stream)
    else:
        need = lib.tdm_fragments_workspace_bytes(n)
        ws = torch.empty(need + 256, dtype=torch.uint8, device="cuda")
        base = ws.data_ptr()
        aligned = (base + 255) & ~255
        rc = lib.tdm_plate_fragments(*args, ctypes.c_void_p(aligned),
                                     ctypes.c_size_t(ws.numel() - (aligned - base)), stream)
    _lib.check(rc)
    torch.cuda.synchronize()
    nc = int(outs[3].item())
    return nc, [t.cpu().numpy() for t in outs]


def test_allocation_free_fragments_equal_workspace_fragments():
    for nx, ny, broken, seed in ((512, 512, 0.6, 1), (200, 37, 0.4, 2)):
        sim = shattered(nx, ny, broken, seed)
        for rep in range(2):
            nc_a, a = call(sim, allocation_free=True)
            nc_b, b = call(sim, allocation_free=False)
            assert nc_a == nc_b > 1, (nx, ny, rep)
            n = nx * ny
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
            assert np.array_equal(a[2][:nc_a + 1], b[2][:nc_b + 1])
            assert a[2][nc_a] == n
            assert np.array_equal(a[4][:nc_a], b[4][:nc_b])          # bitwise, no NaN
            assert sorted(a[1].tolist()) == list(range(n))


def test_allocation_free_rejects_plates_above_its_limit():
    lib = _lib.load()
    d = torch.zeros(1, dtype=torch.float64, device="cuda").data_ptr()
    rc = lib.tdm_plate_fragments_noalloc(d, d, d, 513, 512, d, d, d, d, d, None)
    assert rc == _lib.TDM_EINVAL
    buf = ctypes.create_string_buffer(256)
    lib.tdm_last_error(buf, 256)
    assert b"262144" in buf.value
    assert 513 * 512 > LIMIT >= 512 * 512

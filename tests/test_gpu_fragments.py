"""IceDEM.fragments on the device (csrc/fragments.hip) against the reference's
algorithm (dem.py:462-484: breadth first from the lowest unvisited disk,
sorted particle ids, mass = count * m, centroid and mean velocity as means)
and against scipy's connected components on large shattered plates.

Record fields: ids and masses exact; centroids and mean velocities within
1e-12 relative of numpy's means (the sums are taken in a different, fixed
order)."""
import math

import numpy as np
import pytest
import torch
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

from tensordem_amd import DEMConfig
from tensordem_amd.fragments import label_plate
from tensordem_amd.icedem import IceDEM

pytestmark = pytest.mark.gpu


def np_(t):
    return t.detach().cpu().numpy()


def reference_fragments(sim):
    """dem.py:462-484 restated on host copies of the state."""
    n = sim.positions.shape[0]
    pos = sim.positions.cpu()
    vel = sim.velocities.cpu()
    nb = [set() for _ in range(n)]
    for i, j in np_(sim.pairs)[np_(sim.alive)].tolist():
        nb[i].add(j)
        nb[j].add(i)
    unseen, out = set(range(n)), []
    while unseen:
        todo = [min(unseen)]
        comp = []
        unseen.remove(todo[0])
        while todo:
            i = todo.pop()
            comp.append(i)
            for j in sorted(nb[i] & unseen):
                unseen.remove(j)
                todo.append(j)
        ids = sorted(comp)
        out.append(dict(particle_ids=ids, mass_kg=len(ids) * sim.config.mass,
                        centroid_m=pos[ids].mean(0).tolist(),
                        mean_velocity_m_s=vel[ids].mean(0).tolist()))
    return out


def same_records(got, want, scale):
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert g["particle_ids"] == w["particle_ids"]
        assert g["mass_kg"] == w["mass_kg"]
        for key in ("centroid_m", "mean_velocity_m_s"):
            for a, b in zip(g[key], w[key]):
                assert abs(a - b) <= 1e-12 * max(abs(b), scale[key]), (key, a, b)


def test_indentation_fragments_match_reference_algorithm():
    cfg = DEMConfig(nx=21, ny=5, breaking_strain=0.01, shear_breaking_strain=0.02,
                    device="cuda")
    sim = IceDEM(cfg)
    sim.run(2000)
    got = sim.fragments()
    want = reference_fragments(sim)
    assert len(want) > 1
    same_records(got, want, {"centroid_m": 0.5, "mean_velocity_m_s": 1e-3})
    again = sim.fragments()
    assert again == got                                    # deterministic


@pytest.mark.parametrize("nx,ny,broken", [(512, 512, 0.62), (300, 40, 0.45), (64, 4, 1.0),
                                          (33, 7, 0.0)])
def test_shattered_plates_match_scipy_components(nx, ny, broken):
    cfg = DEMConfig(nx=nx, ny=ny, device="cuda")
    sim = IceDEM(cfg, dt_policy="degree")
    rng = np.random.default_rng(nx * 1000 + ny)
    alive = rng.random(sim.alive.numel()) >= broken
    sim.alive = torch.as_tensor(alive, device="cuda")
    v = rng.normal(scale=0.1, size=(nx * ny, 2))
    sim.velocities = torch.as_tensor(v)
    x = np_(sim.positions)
    got = sim.fragments()
    n = nx * ny
    pairs = np_(sim.pairs)[alive]
    g = coo_matrix((np.ones(len(pairs)), (pairs[:, 0], pairs[:, 1])), shape=(n, n))
    nc, lab = connected_components(g, directed=False)
    # components in the order of their lowest disk
    low = np.full(nc, n)
    np.minimum.at(low, lab, np.arange(n))
    rank = np.empty(nc, dtype=np.int64)
    rank[np.argsort(low)] = np.arange(nc)
    comp = rank[lab]
    assert len(got) == nc
    order = np.argsort(comp, kind="stable")
    bounds = np.concatenate(([0], np.cumsum(np.bincount(comp, minlength=nc))))
    span = {"centroid_m": float(np.abs(x).max()), "mean_velocity_m_s": 0.1}
    for c in range(nc):
        ids = order[bounds[c]:bounds[c + 1]]
        f = got[c]
        assert f["particle_ids"] == ids.tolist()
        assert f["mass_kg"] == len(ids) * cfg.mass
        for key, arr in (("centroid_m", x), ("mean_velocity_m_s", v)):
            want = arr[ids].mean(0)
            for a, b in zip(f[key], want):
                assert abs(a - b) <= 1e-12 * max(abs(b), span[key]), (c, key)
    if broken == 1.0:
        assert nc == n
    if broken == 0.0:
        assert nc == 1
    # the device arrays directly: labels are the component numbers
    c = label_plate(sim._eng)
    assert np.array_equal(np_(c.labels), comp)
    assert int(c.ncomp.item()) == nc
    assert math.isfinite(float(c.sums[:nc].sum()))

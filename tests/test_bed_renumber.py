"""Renumbering of bed disks into Morton order (host logic, on CPU tensors):
the candidate-list remap against a direct per-disk restatement, and the
Morton order itself."""
import numpy as np
import torch

from tensordem_amd.bed import morton_keys, remap_lists, spatial_order


def _reference(cand, xi, mask, count, perm, inv, first):
    S, m = cand.shape
    oc = np.zeros_like(cand)
    ox = np.zeros_like(xi)
    om = np.zeros_like(mask)
    for j in range(m):
        p = perm[j]
        lp = p - first
        k = count[lp]
        me = first + j
        items = []
        for e in range(k):
            q = cand[e, lp]
            qn = inv[q]
            h = xi[e, lp].copy()
            if (p < q) != (me < qn):
                h = -h
            items.append((qn, h, (mask[lp] >> e) & 1))
        items.sort(key=lambda t: t[0])
        bits = mask[lp] & (3 << 16)
        for e, (qn, h, b) in enumerate(items):
            oc[e, j] = qn
            ox[e, j] = h
            bits |= b << e
        om[j] = bits
    return oc, ox, om


def test_remap_lists_matches_per_disk_restatement():
    rng = np.random.default_rng(3)
    S, m, first = 8, 300, 40
    N = first + m + 25
    cand = np.zeros((S, m), np.int32)
    xi = rng.normal(size=(S, m, 2))
    mask = np.zeros(m, np.int64)
    count = np.zeros(m, np.uint8)
    for j in range(m):
        k = int(rng.integers(0, S + 1))
        others = np.setdiff1d(np.arange(first, first + m), [first + j])
        q = np.sort(rng.choice(others, size=k, replace=False))
        cand[:k, j] = q
        cand[k:, j] = rng.integers(0, N, size=S - k)  # stale entries past the count
        count[j] = k
        mask[j] = int(rng.integers(0, 1 << max(k, 1))) & ((1 << k) - 1)
        mask[j] |= int(rng.integers(0, 4)) << 16
    perm_local = rng.permutation(m)
    perm = perm_local + first
    inv = np.arange(N)
    inv[perm] = np.arange(first, first + m)
    ref = _reference(cand, xi, mask, count, perm, inv, first)
    got = remap_lists(torch.from_numpy(cand), torch.from_numpy(xi),
                      torch.from_numpy(mask.astype(np.int32)), torch.from_numpy(count),
                      torch.from_numpy(perm), torch.from_numpy(inv), first)
    gc, gx, gm, gk = (t.numpy() for t in got)
    assert np.array_equal(gk, count[perm_local])
    for j in range(m):
        k = gk[j]
        assert np.array_equal(gc[:k, j], ref[0][:k, j])
        assert np.array_equal(gx[:k, j], ref[1][:k, j])
    assert np.array_equal(gm.astype(np.int64), ref[2])


def test_morton_order():
    # a 2x2 block of cells is visited (0,0), (1,0), (0,1), (1,1)
    k = morton_keys(torch.tensor([0, 1, 0, 1]), torch.tensor([0, 0, 1, 1]))
    assert k.tolist() == [0, 1, 2, 3]
    rng = np.random.default_rng(0)
    x = torch.from_numpy(rng.uniform(0, 1, size=(1000, 2)))
    p = spatial_order(x, (0.0, 0.0), 0.05, 20, 20)
    assert sorted(p.tolist()) == list(range(1000))
    cells = torch.floor(x[p] / 0.05).to(torch.int64)
    keys = morton_keys(cells[:, 0], cells[:, 1])
    assert bool((keys[1:] >= keys[:-1]).all())

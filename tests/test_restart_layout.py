"""Restart interoperability with the reference's `tensordem.restart/1`
(src/tensordem/dem.py:486-601): its `alive` and `failure_*` arrays run over
all N(N-1)/2 pairs in torch.triu_indices order; the drop-in keeps them per
bond.  A hand-built 7x3 reference-layout state with the field shapes of
dem.py:116-137 goes to the bond layout and back unchanged, and the
conversions fail closed where the reference's loader does."""
import numpy as np
import pytest
import torch

from tensordem_amd import lattice, restart

NX, NY, R = 7, 3, 0.025
N = NX * NY
P = N * (N - 1) // 2


def bonds():
    i, j, _, _ = lattice.bonds(NX, NY, R)
    return i, j


def reference_state():
    """The five pair fields as the reference writes them for a 7x3 plate in
    which bonds (0, 1), (8, 9) and (12, 20) broke: alive over all pairs is
    the bonded mask minus the broken bonds; failure records only there."""
    pos = torch.tensor(lattice.positions(NX, NY, R))
    pairs = torch.triu_indices(N, N, offset=1).T
    d = torch.linalg.vector_norm(pos[pairs[:, 1]] - pos[pairs[:, 0]], dim=1)
    bonded = d <= 2 * R * 2 ** 0.5 * (1 + 1e-10)
    alive = bonded.clone()
    mode = torch.zeros(P, dtype=torch.uint8)
    t = torch.full((P,), -1.0, dtype=torch.float64)
    ext = torch.zeros(P, dtype=torch.float64)
    en = torch.zeros(P, dtype=torch.float64)
    rows = {tuple(p): k for k, p in enumerate(pairs.tolist())}
    for (a, b), m, tt, e in (((0, 1), 1, 0.01, 6e-4), ((8, 9), 2, 0.02, 1e-5),
                             ((12, 20), 3, 0.03, 3e-4)):
        k = rows[(a, b)]
        assert bool(bonded[k])
        alive[k] = False
        mode[k], t[k], ext[k], en[k] = m, tt, e, 0.5 * 2000 * e * e
    return dict(alive=alive.tolist(), failure_mode=mode.tolist(), failure_time_s=t.tolist(),
                failure_extension_m=ext.tolist(), failure_energy_J=en.tolist()), bonded


def test_triu_index_matches_torch_order():
    pairs = torch.triu_indices(N, N, offset=1).T.numpy()
    idx = restart.triu_index(pairs[:, 0], pairs[:, 1], N)
    assert np.array_equal(idx, np.arange(P))
    i, j = bonds()
    _, bonded = reference_state()
    assert np.array_equal(np.sort(restart.triu_index(i, j, N)),
                          np.nonzero(bonded.numpy())[0])


def test_reference_layout_round_trip():
    ref, _ = reference_state()
    i, j = bonds()
    per_bond = restart.pairs_to_bonds(ref, N, i, j)
    assert len(per_bond["alive"]) == len(i) == 6 * 3 + 7 * 2 + 2 * 6 * 2
    assert sum(not a for a in per_bond["alive"]) == 3
    broken = [(int(i[b]), int(j[b])) for b, a in enumerate(per_bond["alive"]) if not a]
    assert broken == [(0, 1), (8, 9), (12, 20)]
    assert [per_bond["failure_mode"][b] for b, a in enumerate(per_bond["alive"]) if not a] \
        == [1, 2, 3]
    back = restart.bonds_to_pairs({k: np.asarray(v) for k, v in per_bond.items()}, N, i, j)
    for key in restart.PAIR_FIELDS:
        assert back[key] == ref[key], key
        assert type(back[key][0]) is type(ref[key][0]), key


@pytest.mark.parametrize("key,value,message", [
    ("alive", True, "nonbonded pair cannot be alive"),
    ("failure_mode", 1, "invalid restart failure modes"),
    ("failure_time_s", 0.5, "failure record"),
    ("failure_energy_J", 1e-6, "failure record"),
    ("failure_extension_m", 1e-6, "failure record"),
])
def test_nonbonded_records_are_rejected(key, value, message):
    ref, bonded = reference_state()
    k = int(np.nonzero(~bonded.numpy())[0][5])
    ref[key][k] = value
    i, j = bonds()
    with pytest.raises(ValueError, match=message):
        restart.pairs_to_bonds(ref, N, i, j)


def test_shapes_and_types_are_checked():
    ref, _ = reference_state()
    i, j = bonds()
    short = dict(ref, alive=ref["alive"][:-1])
    with pytest.raises(ValueError, match="shape"):
        restart.pairs_to_bonds(short, N, i, j)
    ints = dict(ref, alive=[int(a) for a in ref["alive"]])
    with pytest.raises(ValueError, match="boolean"):
        restart.pairs_to_bonds(ints, N, i, j)
    nan = dict(ref, failure_time_s=[float("nan")] + ref["failure_time_s"][1:])
    with pytest.raises(ValueError, match="nonfinite"):
        restart.pairs_to_bonds(nan, N, i, j)


def test_large_plates_use_the_bond_layout():
    assert restart.pair_count(2896) <= restart.MAX_REFERENCE_PAIRS < restart.pair_count(2897)
    with pytest.raises(ValueError, match="bond-indexed"):
        restart.bonds_to_pairs({}, 131072, np.zeros(1), np.ones(1))

"""Connected components of surviving bonds (IceDEM.fragments, dem.py:462-484).

Minimum-label propagation over the alive bonds on the device: every disk
starts with its own index as label and repeatedly takes the smallest label
among itself and its bonded partners until nothing changes; each component
then carries the index of its lowest disk, the same representative the
reference's breadth-first search starts from.
"""
from __future__ import annotations

import torch


def label_components(engine) -> torch.Tensor:
    n = engine.n
    dev = engine.x.device
    labels = torch.arange(n, device=dev, dtype=torch.int64)
    i = engine.bond_i[engine.alive]
    j = engine.bond_j[engine.alive]
    if i.numel() == 0:
        return labels
    while True:
        li, lj = labels[i], labels[j]
        low = torch.minimum(li, lj)
        new = labels.clone()
        new.scatter_reduce_(0, i, low, reduce="amin")
        new.scatter_reduce_(0, j, low, reduce="amin")
        # pointer jumping: follow labels to their own labels
        new = new[new]
        if torch.equal(new, labels):
            return labels
        labels = new

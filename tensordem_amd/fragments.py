"""Connected components of surviving bonds (IceDEM.fragments, dem.py:462-484).

`tdm_plate_fragments` (csrc/fragments.hip) labels the components of one
plate on the device: union-find over the lattice link bits with the lowest
disk of each component as its root, components numbered in the order of
their lowest disk, disk indices grouped by a stable radix sort, and
per-component sums of positions and velocities in a fixed order.  The host
only turns those arrays into the reference's list of dicts.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib


class Components:
    """Device outputs of one fragment labelling (see tdm_plate_fragments)."""

    def __init__(self, labels, order, comp_first, ncomp, sums):
        self.labels, self.order, self.comp_first = labels, order, comp_first
        self.ncomp, self.sums = ncomp, sums


NOALLOC_MAX_DISKS = 1 << 18


def label_plate(engine, k: int = 0, workspace: torch.Tensor | None = None, *,
                allocation_free: bool = False) -> Components:
    """Component labelling of member k of a PlateEngine at its current state.

    allocation_free: use tdm_plate_fragments_noalloc (the library's static
    scratch, at most NOALLOC_MAX_DISKS disks) instead of a workspace."""
    m = engine._members_host[k]
    nx, ny, first = int(m.nx), int(m.ny), int(m.first)
    n = nx * ny
    dev = engine.x.device
    i32 = dict(dtype=torch.int32, device=dev)
    labels = torch.empty(n, **i32)
    order = torch.empty(n, **i32)
    comp_first = torch.empty(n + 1, **i32)
    ncomp = torch.empty(1, **i32)
    sums = torch.empty((n, 4), dtype=torch.float64, device=dev)
    lib = engine.lib
    x = engine.x[first:first + n]
    v = engine.v[first:first + n]
    links = engine.links[first:first + n]
    outs = (labels.data_ptr(), order.data_ptr(), comp_first.data_ptr(), ncomp.data_ptr(),
            sums.data_ptr())
    if allocation_free:
        _lib.check(lib.tdm_plate_fragments_noalloc(links.data_ptr(), x.data_ptr(), v.data_ptr(),
                                                   nx, ny, *outs, engine.stream))
        return Components(labels, order, comp_first, ncomp, sums)
    need = int(lib.tdm_fragments_workspace_bytes(n))
    if workspace is None or workspace.numel() < need + 256:
        workspace = torch.empty(need + 256, dtype=torch.uint8, device=dev)
    base = workspace.data_ptr()
    ws = (base + 255) & ~255
    _lib.check(lib.tdm_plate_fragments(
        links.data_ptr(), x.data_ptr(), v.data_ptr(), nx, ny, *outs,
        ctypes.c_void_p(ws), ctypes.c_size_t(workspace.numel() - (ws - base)),
        engine.stream))
    return Components(labels, order, comp_first, ncomp, sums)


def fragment_records(c: Components, mass: float) -> list[dict]:
    """The reference's records: particle_ids (ascending), mass_kg,
    centroid_m, mean_velocity_m_s, fragments in the order of their lowest
    disk."""
    nc = int(c.ncomp.item())
    first = c.comp_first[:nc + 1].cpu().tolist()
    order = c.order.cpu().tolist()
    sums = c.sums[:nc].cpu().tolist()
    out = []
    for k in range(nc):
        ids = order[first[k]:first[k + 1]]
        cnt = len(ids)
        s = sums[k]
        out.append(dict(particle_ids=ids, mass_kg=cnt * mass,
                        centroid_m=[s[0] / cnt, s[1] / cnt],
                        mean_velocity_m_s=[s[2] / cnt, s[3] / cnt]))
    return out

"""Hash-cell capacity of the bed (host logic): the packing bound against
dense packings of the smallest disks, placed at every offset."""
import math

import numpy as np

from tensordem_amd.bed import cell_capacity


def _max_in_cell(h, r, offsets=24, overlap=0.0):
    """Most centres of a hexagonal packing of radius-r disks (the densest),
    neighbours overlapping by `overlap` (centre spacing 2r - overlap), that
    fall in one h x h cell, over shifts of the packing."""
    d = 2.0 * r - overlap
    best = 0
    for sx in np.linspace(0.0, d, offsets, endpoint=False):
        for sy in np.linspace(0.0, d * math.sqrt(3.0), offsets, endpoint=False):
            rows = int(h / (d * math.sqrt(3.0) / 2)) + 4
            cols = int(h / d) + 4
            iy, ix = np.mgrid[-2:rows, -2:cols]
            x = ix * d + (iy % 2) * (d / 2) + sx
            y = iy * d * math.sqrt(3.0) / 2 + sy
            inside = (x >= 0) & (x < h) & (y >= 0) & (y < h)
            best = max(best, int(inside.sum()))
    return best


def test_bound_holds_for_dense_packings():
    for h, r in ((0.028, 0.008), (0.034, 0.005), (0.05, 0.004), (0.02, 0.01)):
        assert _max_in_cell(h, r) <= cell_capacity(h, r, floor=0), (h, r)


def test_bound_holds_with_contact_overlaps():
    # every neighbour pair pressed together by 2 % of r (a settled bed's
    # Hertz overlaps are of order 1e-4 r)
    # (a 200-radius cell is where the bound without the 0.99 shrink fails)
    for h, r in ((0.028, 0.008), (0.034, 0.005), (0.05, 0.004), (0.02, 0.01), (0.4, 0.002)):
        assert _max_in_cell(h, r, overlap=0.02 * r) <= cell_capacity(h, r, floor=0), (h, r)


def test_column_keeps_eight_and_wide_spreads_grow():
    # the column collapse: radii U[0.8, 1.2] x 0.01, skin 0.4 x mean radius
    assert cell_capacity(2 * 0.012 + 0.004, 0.008) == 8
    # the floe bed's widest spread (polydispersity 1: radii in [0.5, 1.5] r0)
    assert cell_capacity(2 * 0.015 + 0.004, 0.005) == 22
    assert cell_capacity(0.01, 0.01) == 8          # never below the floor


def test_rectangular_cells():
    # (h + 2r')(hy + 2r') / (sqrt(12) r'^2), r' = 0.99 r, for an elongated cell
    h, hy, r = 0.03, 0.06, 0.005
    rr = 0.99 * r
    want = math.floor((h + 2 * rr) * (hy + 2 * rr) / (math.sqrt(12.0) * rr * rr))
    assert cell_capacity(h, r, floor=0, cell_size_y=hy) == want
    assert cell_capacity(h, r, floor=0, cell_size_y=h) == cell_capacity(h, r, floor=0)

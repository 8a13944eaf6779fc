"""Weibull strength scatter for bonded plates.

The draw follows the reference's `weibull_bond_strengths`
(src/tensordem/ice_bond_viscoelastic.py:300-322): u from a CPU
`torch.Generator` seeded with the member's seed, clamped below by the
float64 tiny, and F = scale * (-log1p(-u))^(1/m).  The reference fixes the
scale by the median (scale = median / ln2^(1/m)); a mean can be asked for
instead (scale = mean / Gamma(1 + 1/m)).  Draws are per disk: a bond then
breaks at the weaker of its two disks (the kernels and the oracle both take
the minimum of the endpoint thresholds).
"""
from __future__ import annotations

import math

import numpy as np
import torch


def weibull_scale(shape: float, *, mean: float | None = None, median: float | None = None) -> float:
    if shape <= 0 or not math.isfinite(shape):
        raise ValueError("shape must be finite and positive")
    if (mean is None) == (median is None):
        raise ValueError("give exactly one of mean, median")
    target = mean if mean is not None else median
    if target <= 0 or not math.isfinite(target):
        raise ValueError("mean / median must be finite and positive")
    if mean is not None:
        return mean / math.gamma(1.0 + 1.0 / shape)
    return median / math.log(2.0) ** (1.0 / shape)


def weibull_strengths(n: int, shape: float, *, seed: int, mean: float | None = None,
                      median: float | None = None) -> np.ndarray:
    """n Weibull(shape) thresholds with the given mean or median, float64."""
    if n <= 0:
        raise ValueError("n must be positive")
    scale = weibull_scale(shape, mean=mean, median=median)
    g = torch.Generator(device="cpu").manual_seed(int(seed))
    u = torch.rand(n, generator=g, dtype=torch.float64)
    u = u.clamp_min(torch.finfo(torch.float64).tiny)
    return (scale * (-torch.log1p(-u)).pow(1.0 / float(shape))).numpy()

"""FP64 numpy restatement of Bimage::fsc_dpr for two real cubic maps.

TEST INFRASTRUCTURE ONLY: the checker of bsx_fsc_dpr (tests/). Pinned in
tests/test_resolution.py by identities (equal maps, a negated map, a scaled
map) and by a brute-force per-voxel evaluation of a one-voxel shift.

Follows src/img/Bimage_resolution.cpp:93-240 (shell sums over x < N/2 with
linear interpolation between shells i and i+1, both inside maxrad; FSC
clamped to [-1, 1]; DPR = sqrt(sum (I1+I2) dphi^2 / sum (I1+I2)) in degrees)
and Bimage::check_resolution (src/img/Bimage.cpp:633-648).
"""
from __future__ import annotations

import math

import numpy as np


def check_resolution(res: float, n: int, sampling: float) -> float:
    if res > n * sampling:
        return n * sampling
    return max(res, 2 / math.sqrt(3 / sampling ** 2))


def fsc_dpr(a: np.ndarray, b: np.ndarray, sampling: float, hi_res: float = 0.0,
            ratio: float = 1.0):
    n = a.shape[0]
    hi_res = check_resolution(hi_res, n, sampling)
    rad_scale = n * sampling / ratio
    maxrad = int(2 + rad_scale / hi_res)
    F1 = np.fft.fftn(a.astype(np.float64))
    F2 = np.fft.fftn(b.astype(np.float64))
    f = np.arange(n)
    f = np.where(f > (n - 1) // 2, f - n, f).astype(np.float64)
    l, k, h = np.meshgrid(f, f, f[: n // 2], indexing="ij")      # (z, y, x < n/2)
    u, v = F1[:, :, : n // 2], F2[:, :, : n // 2]
    radius = np.sqrt(h * h + k * k + l * l) / ratio
    ir = radius.astype(np.int64)
    ok = ir + 1 < maxrad
    fr = (radius - ir)[ok]
    ir = ir[ok]
    u, v = u[ok], v[ok]
    I1, I2 = np.abs(u) ** 2, np.abs(v) ** 2
    dphi = np.remainder(np.angle(u) - np.angle(v) + np.pi, 2 * np.pi) - np.pi
    re = (u * np.conj(v)).real
    vals = [I1, I2, re, I1 + I2, (I1 + I2) * dphi ** 2]
    sums = np.zeros((5, maxrad))
    for c, x in enumerate(vals):
        np.add.at(sums[c], ir, (1 - fr) * x)
        np.add.at(sums[c], ir + 1, fr * x)
    s = np.arange(maxrad) / rad_scale
    fsc = sums[2].copy()
    dpr = np.zeros(maxrad)
    for i in range(1, maxrad):
        if sums[0, i] * sums[1, i] > 1e-30:
            fsc[i] /= math.sqrt(sums[0, i] * sums[1, i])
        fsc[i] = min(1.0, max(-1.0, fsc[i]))
        dpr[i] = math.sqrt(sums[4, i] / sums[3, i]) * 180 / math.pi if sums[3, i] > 1e-30 else 0
    fsc[0] = 1.0
    return s, fsc, dpr

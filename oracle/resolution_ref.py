"""FP64 numpy restatement of Bimage::fsc_dpr for two real cubic maps, and of
the shell statistics table of a weighed reconstruction
(fspace_reconstruction_stats, reconstruction_stats below).

TEST INFRASTRUCTURE ONLY: the checker of bsx_fsc_dpr (tests/). Pinned in
tests/test_resolution.py by identities (equal maps, a negated map, a scaled
map) and by a brute-force per-voxel evaluation of a one-voxel shift.

Follows src/img/Bimage_resolution.cpp:93-240 (shell sums over x < N/2 with
linear interpolation between shells i and i+1, both inside maxrad; FSC
clamped to [-1, 1]; DPR = sqrt(sum (I1+I2) dphi^2 / sum (I1+I2)) in degrees)
and Bimage::check_resolution (src/img/Bimage.cpp:633-648).

reconstruction_stats follows fspace_reconstruction_weigh
(src/img/Bimage_reconstruct.cpp:486-515), fspace_reconstruction_stats
(:526-599), fspace_radial_power (src/img/Bimage_fspace.cpp:848-925),
fspace_maximum_radius (:26-38) and Bimage::radial (src/img/Bimage_polar.cpp:
79-170). It is pinned in tests/test_resolution.py by the closed form of K
identical contributions per voxel (SNR = K (K - 2) in every shell) and by
the radial average of a constant.
"""
from __future__ import annotations

import math

import numpy as np


def check_resolution(res: float, n: int, sampling: float, ndim: int = 3) -> float:
    if res > n * sampling:
        return n * sampling
    return max(res, 2 / math.sqrt(ndim / sampling ** 2))


def fsc_dpr(a: np.ndarray, b: np.ndarray, sampling: float, hi_res: float = 0.0,
            ratio: float = 1.0):
    """Cubic maps (z, y, x) or square images (y, x)."""
    if a.ndim == 2:
        a, b = a[None], b[None]
    n, nz = a.shape[-1], a.shape[0]
    hi_res = check_resolution(hi_res, n, sampling, 3 if nz > 1 else 2)
    rad_scale = n * sampling / ratio
    maxrad = int(2 + rad_scale / hi_res)
    F1 = np.fft.fftn(a.astype(np.float64))
    F2 = np.fft.fftn(b.astype(np.float64))
    f = np.arange(n)
    f = np.where(f > (n - 1) // 2, f - n, f).astype(np.float64)
    fz = np.arange(nz)
    fz = np.where(fz > (nz - 1) // 2, fz - nz, fz).astype(np.float64)
    l, k, h = np.meshgrid(fz, f, f[: n // 2], indexing="ij")     # (z, y, x < n/2)
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


def _shell_split(radius: np.ndarray, values, nr: int):
    """Split each value between shells floor(r) and floor(r) + 1 (weights
    1 - f and f), both below nr: (counts, sums...) per shell."""
    ir = radius.astype(np.int64).ravel()
    f = (radius - np.floor(radius)).ravel()
    ok0 = ir < nr
    ok1 = ir + 1 < nr
    cnt = np.bincount(ir[ok0], weights=(1 - f)[ok0], minlength=nr)[:nr]
    cnt += np.bincount(ir[ok1] + 1, weights=f[ok1], minlength=nr + 1)[:nr]
    out = [cnt]
    for v in values:
        v = np.asarray(v, dtype=np.float64).ravel()
        a = np.bincount(ir[ok0], weights=((1 - f) * v)[ok0], minlength=nr)[:nr]
        a += np.bincount(ir[ok1] + 1, weights=(f * v)[ok1], minlength=nr + 1)[:nr]
        out.append(a)
    return out


def reconstruction_stats(acc: np.ndarray, sampling: float, resolution: float):
    """acc: (5, N, N, N) accumulator planes (z, y, x; (5, 1, N, N) for a 2D
    map), summed over the half
    sets: sum w re, sum w im, sum w, sum w |F|^2, sum w^2. Returns the table
    rows (radius, s, res, weight, power, signal, noise, SNR, FSC) for shells
    1 .. min(maxrad, N/2) - 1 and the (SNR(0.5), FSC(0.3)) resolutions."""
    acc = np.asarray(acc, dtype=np.float64)
    n, zn = acc.shape[-1], acc.shape[1]           # zn = 1: a 2D map
    re, im, W, P, W2 = acc
    full = W > 1e-30
    Ws = np.where(full, W, 1.0)
    w2 = np.where(full, W - W2 / Ws, 1e-30)
    fr = np.where(full, (re / Ws).astype(np.float32), 0).astype(np.float64)
    fi = np.where(full, (im / Ws).astype(np.float32), 0).astype(np.float64)
    P = np.where(full, np.where(w2 > 1, P / np.where(w2 > 1, w2, 1),
                                np.where(W > 1, P / Ws, P)), 0.0)
    P = np.maximum(P, 0)
    P.ravel()[0] = 1.0
    W = np.where(full, W, 1e-30)
    P, W, w2 = (x.astype(np.float32).astype(np.float64) for x in (P, W, w2))
    pw = fr * fr + fi * fi
    rs = n * sampling

    def maxrad(res):
        return int(rs / res) if res > 0.1 else int(n / 2.0)

    nr_stats = maxrad(resolution)
    nr1 = max(1, maxrad(check_resolution(resolution, n, sampling, 3 if zn > 1 else 2)))
    nr2 = min(nr_stats if nr_stats >= 1 else n // 2, n // 2, nr1)
    idx, iz = np.arange(n), np.arange(zn)
    hf = np.where(idx > (n - 1) // 2, idx - n, idx) / rs
    hz = np.where(iz > (zn - 1) // 2, iz - zn, iz) / rs
    z, y, x = np.meshgrid(hz, hf, hf, indexing="ij")
    r1 = rs * np.sqrt(x * x + y * y + z * z)
    c1, s1 = _shell_split(r1, [pw], nr1)
    d = np.where(idx > n // 2, idx - n, idx).astype(np.float64)
    dz = np.where(iz > zn // 2, iz - zn, iz).astype(np.float64)
    z, y, x = np.meshgrid(dz, d, d, indexing="ij")
    r2 = np.sqrt(x * x + y * y + z * z)
    c2, sp, sw, sw2 = _shell_split(r2, [P, W, w2], nr2)

    def avg(a, c, i):
        return a[i] / c[i] if c[i] != 0 else 0.0

    rows = []
    fscp, snrp, res_fsc, res_snr = 1.0, 10.0, 0.0, 0.0
    for i in range(1, max(0, min(nr_stats, nr2) - 1) + 1):
        sf, res = i / rs, rs / i
        prad, pfom, w = avg(s1, c1, i), avg(sp, c2, i), avg(sw, c2, i)
        fsc = snr = signal = noise = 0.0
        if w > 1e-3:
            w2i = avg(sw2, c2, i)
            noise = (pfom - prad) / w2i
            signal = (w * prad - pfom) / w2i
            if noise > 0:
                snr = signal / noise
            fsc = max(snr / (snr + 1), 0.0)
        if fscp >= 0.3 and fsc < 0.3:
            res_fsc = 1 / (sf - (0.3 - fsc) / ((fscp - fsc) * rs))
        if snrp >= 0.5 and snr < 0.5:
            res_snr = 1 / (sf - (0.5 - snr) / ((snrp - snr) * rs))
        fscp, snrp = fsc, snr
        rows.append([i, sf, res, w, prad, signal, noise, snr, fsc])
    return np.asarray(rows).reshape(-1, 9), (res_snr, res_fsc)

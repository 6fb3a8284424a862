"""FP64 restatement of the projection-matching scores (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the
bsx_search_* kernels; never by the product path.

Follows:
  img_recip_space_fom (fom_type 0)  src/mg/mg_refine.cpp:660-781
      per ring j = floor(|(h,k)|): sum Re(F_p conj(F_ref CTF)) /
      sqrt(sum |F_p|^2 sum |F_ref CTF|^2), mean over rings 0..max_rad-1
  phase_shift_to_origin             src/img/Bimage_complex.cpp:490-543
  CTFparam::calculate = sin(chi)    include/ctf.h:321-365 (via the C oracle)
  img_fit_in_recip_space_grid       src/mg/mg_refine.cpp:862-946 (fit_grid)
The half-plane form used by the kernels equals the full-plane sum because
the transforms of real images are Hermitian (checked in tests/test_oracle.py).
"""
from __future__ import annotations

import math

import numpy as np

from bsoft_amd.abi import Particle


def disk_points(radius: int):
    """Half-plane lattice points with floor(|(h,k)|) < radius: DC first, then
    ring by ring; returns (hk (Nd, 2), ring_start (R+1))."""
    hk = [(0, 0)]
    starts = [0]
    for j in range(1, radius):
        starts.append(len(hk))
        ring = []
        for k in range(0, j + 2):
            for h in range(-j - 1, j + 2):
                if k == 0 and h <= 0:
                    continue
                if int(math.isqrt(h * h + k * k)) == j:
                    ring.append((h, k))
        hk += ring
    starts.append(len(hk))
    return np.asarray(hk, dtype=np.int64), np.asarray(starts, dtype=np.int64)


def particle_disk(img: np.ndarray, part: Particle, hk: np.ndarray, oracle, ctf: bool = True):
    """Forward FFT (sign -1, x 1/N) of one particle, phase shift to its origin,
    values and CTF at the disk points."""
    n = img.shape[0]
    F = np.fft.fft2(np.asarray(img, dtype=np.float64)) / n
    h, k = hk[:, 0], hk[:, 1]
    v = F[k % n, h % n]
    phi = 2 * np.pi * (h * part.ori[0] + k * part.ori[1]) / n
    v = v * np.exp(1j * phi)
    c = np.ones(len(hk))
    if ctf and part.def_avg >= 1 and part.volt > 0:
        fs = 1.0 / (n * part.pixel_size)
        c = np.array([oracle.ctf_value(part, float(a) * fs, float(b) * fs) for a, b in hk])
    return v, c


def ring_of(ring_start: np.ndarray, nd: int) -> np.ndarray:
    r = np.zeros(nd, dtype=np.int64)
    for j in range(len(ring_start) - 1):
        r[ring_start[j]:ring_start[j + 1]] = j
    return r


def coarse_row(vals: np.ndarray, ring_start: np.ndarray, scale_rings: bool = True) -> np.ndarray:
    """Ring-normalised real row (Re, Im of points 1..Nd-1), x 1/(R-1) when
    scale_rings; rings with no power give zeros."""
    R = len(ring_start) - 1
    rg = ring_of(ring_start, len(vals))
    p = np.bincount(rg, weights=np.abs(vals) ** 2, minlength=R)
    with np.errstate(divide="ignore"):
        inv = np.where(p > 0, 1.0 / np.sqrt(p), 0.0)
    inv[0] = 0.0
    if scale_rings:
        inv /= (R - 1)
    a = vals[1:] * inv[rg[1:]]
    out = np.empty(2 * len(a))
    out[0::2], out[1::2] = a.real, a.imag
    return out


def coarse_score(pvals, pctf, svals, ring_start) -> float:
    """Mean over rings 1..R-1 of the ring correlation of F_p*CTF with F_ref."""
    return float(coarse_row(pvals * pctf, ring_start) @ coarse_row(svals, ring_start, False))


def frc_half(pvals, pctf, svals, ring_start) -> float:
    """img_recip_space_fom over the half-plane disk points (rings 0..R-1)."""
    R = len(ring_start) - 1
    rg = ring_of(ring_start, len(pvals))
    r = svals * pctf
    num = np.bincount(rg, weights=(pvals * np.conj(r)).real, minlength=R)
    dp = np.bincount(rg, weights=np.abs(pvals) ** 2, minlength=R)
    dr = np.bincount(rg, weights=np.abs(r) ** 2, minlength=R)
    den = dp * dr
    f = np.where(den > 0, num / np.sqrt(np.where(den > 0, den, 1)), 0.0)
    return float(f.sum() / R)


def frc_full(pft: np.ndarray, ref_ctf: np.ndarray, n: int, sampling: float, radius: int,
             oracle) -> float:
    """The reference formula on full N x N planes (the C oracle's
    img_recip_space_fom) with max_rad = R: hi_res is set a hair above
    N * sampling / R so that rounding cannot admit ring R."""
    f, _ = oracle.recip_space_fom(pft, ref_ctf, sampling, n * sampling / (radius - 1e-6))
    return f


def _ctf_at(part: Particle, hk: np.ndarray, n: int, def_avg: float, ctf_flag: bool, oracle):
    """CTF values at the disk points with the defocus replaced; ones unless
    the CTF flag is set and def >= 1 (img_recip_space_fom, mg_refine.cpp:687-688)."""
    if not (ctf_flag and def_avg >= 1 and part.volt > 0):
        return np.ones(len(hk))
    q = Particle.from_buffer_copy(part)
    q.def_avg = def_avg
    fs = 1.0 / (n * part.pixel_size)
    return np.array([oracle.ctf_value(q, float(a) * fs, float(b) * fs) for a, b in hk])


def fit_grid(v0: np.ndarray, part: Particle, svals: np.ndarray, hk: np.ndarray,
             ring_start: np.ndarray, n: int, oracle, step: float, accuracy: float,
             def_std: float, ctf_flag: bool = True):
    """img_fit_in_recip_space_grid on the disk points. v0: the particle's
    transform (sign -1, x 1/N) at hk before any phase shift; svals: the
    section. Defocus scan over def +- 5 def_std when def_std > 0 and either
    the CTF flag is set or the defocus is nonzero (def_std == 0 with the flag
    set would not end in the reference and is skipped); then 3 x 3 origin
    grids around the starting origin, the step halved when the centre wins
    or |prevfom - bestfom| < 1e-6, while step >= accuracy.
    Returns (origin (x, y), defocus, fom, number of FOM evaluations)."""
    h, k = hk[:, 0].astype(np.float64), hk[:, 1].astype(np.float64)

    def fom(ox, oy, c):
        v = v0 * np.exp(2j * np.pi * (h * ox + k * oy) / n)
        return frc_half(v, c, svals, ring_start)

    ox, oy = float(part.ori[0]), float(part.ori[1])
    bestdef, bestfom, number = float(part.def_avg), -1.0, 0
    c = _ctf_at(part, hk, n, bestdef, ctf_flag, oracle)
    if def_std > 0 and (ctf_flag or part.def_avg != 0):
        d, d1 = bestdef - 5 * def_std, bestdef + 5 * def_std
        while d <= d1:
            f = fom(ox, oy, _ctf_at(part, hk, n, d, ctf_flag, oracle))
            number += 1
            if bestfom < f:
                bestfom, bestdef = f, d
            d += def_std
        c = _ctf_at(part, hk, n, bestdef, ctf_flag, oracle)
    best = (ox, oy)
    while step >= accuracy:
        prevfom, bestfom = bestfom, -1.0
        i = ibest = 0
        y = oy - step
        while y <= oy + step:
            x = ox - step
            while x <= ox + step:
                f = fom(x, y, c)
                number += 1
                if bestfom < f:
                    bestfom, ibest, best = f, i, (x, y)
                x += step
                i += 1
            y += step
        if ibest == i // 2 or abs(prevfom - bestfom) < 1e-6:
            step /= 2
    return best, bestdef, bestfom, number

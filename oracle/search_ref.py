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


def fit_monte(v0: np.ndarray, part: Particle, svals: np.ndarray, hk: np.ndarray,
              ring_start: np.ndarray, n: int, oracle, max_iter: int, def_std: float,
              shift_std: float, draws: np.ndarray, ctf_flag: bool = True,
              def_start: float | None = None):
    """img_fit_in_recip_space_monte (mg_refine.cpp:783-860) on the disk
    points with caller-supplied random values, draws (max_iter, 5):
    (g, a, ux, uy, uniform) per iteration for random_gaussian(0, def_std),
    vector3_xy_random_gaussian(0, shift_std) = (ux, uy) * shift_std a and
    random() * irm. The defocus moves when the CTF flag is set or def_std > 0
    and the defocus is nonzero (apply_ctf, :787-792); predef and bestdef
    start at the defocus only in the latter case, else at 0. A move is kept
    when fom / prefom > uniform (prefom starts at -1), else the walk returns
    to the last kept origin and defocus. Returns (origin (x, y), defocus,
    fom) of the best FOM seen. def_start: the defocus the walk starts from
    (the particle's when None)."""
    h, k = hk[:, 0].astype(np.float64), hk[:, 1].astype(np.float64)
    ox0, oy0 = float(part.ori[0]), float(part.ori[1])

    def fom(ox, oy, d):
        v = v0 * np.exp(2j * np.pi * (h * ox + k * oy) / n)
        return frc_half(v, _ctf_at(part, hk, n, d, ctf_flag, oracle), svals, ring_start)

    d = float(part.def_avg if def_start is None else def_start)
    d_in = d
    dstep = def_std > 0 and d != 0
    apply = ctf_flag or dstep
    predef = bestdef = d if dstep else 0.0
    ox, oy = ox0, oy0
    pre, best = (ox, oy), (ox, oy)
    prefom = bestfom = -1.0
    for it in range(max_iter):
        g, a, ux, uy, u = (float(x) for x in draws[it])
        if it > 0:
            if apply:
                d = d + def_std * g
            amp = shift_std * a
            ox, oy = ox + ux * amp, oy + uy * amp
        f = fom(ox, oy, d)
        if bestfom < f:
            bestfom, best = f, (ox, oy)
            if apply:
                bestdef = d
        with np.errstate(divide="ignore", invalid="ignore"):
            keep = np.float64(f) / np.float64(prefom) > u
        if keep:
            if apply:
                predef = d
            pre, prefom = (ox, oy), f
        else:
            if apply:
                d = predef
            ox, oy = pre
    return best, (bestdef if apply else d_in), bestfom


def refine_monte(v0, part, section, hk, ring_start, n, oracle, view, max_iter, def_std,
                 shift_std, view_std, max_angle, draws, inner_draws, ctf_flag=True):
    """img_refine_monte (mg_refine.cpp:979-1060) without magnification
    steps for one particle. section(view) gives the section values at hk;
    draws (max_iter, 6): (g, ux, uy, uz, u_angle, u_accept); inner_draws
    (max_iter, max_iter, 5) for fit_monte. The defocus restarts from the
    particle's only when def_std, the defocus and the CTF flag are all
    nonzero (apply_ctf, :984-988); otherwise each fit starts from the
    defocus the previous one left. Returns (view, origin, defocus, fom)."""
    d0 = float(part.def_avg)
    apply = bool(def_std and d0 and ctf_flag)
    cur = d0
    view = _view_normalize(view)
    pre, prefom, bestfom = list(view), -1.0, -1.0
    best = (list(view), (0.0, 0.0), d0)
    for it in range(max_iter):
        g, ux, uy, uz, ua, uk = (float(x) for x in draws[it])
        if it > 0:
            amp = view_std * g
            view = _view_normalize([view[0] + ux * amp, view[1] + uy * amp,
                                    view[2] + uz * amp, view[3] + max_angle * (ua * 2 - 1)])
        start = d0 if apply else cur
        ori, cur, f = fit_monte(v0, part, section(view), hk, ring_start, n, oracle, max_iter,
                                def_std, shift_std, inner_draws[it], ctf_flag, def_start=start)
        if bestfom < f:
            bestfom = f
            best = (list(view), ori, cur if apply else best[2])
        with np.errstate(divide="ignore", invalid="ignore"):
            keep = np.float64(f) / np.float64(prefom) > uk
        if keep:
            pre, prefom = list(view), f
        else:
            view = list(pre)
    return best[0], best[1], (best[2] if apply else cur), bestfom


# ------------------------------------------------------------ local view grid
def _view_normalize(v):
    """View::normalize (include/View.h:110-116, check :34-38) and
    angle_set_negPI_to_PI (src/util/utilities.cpp:606-615)."""
    x, y, z, a = (float(c) for c in v)
    s = math.sqrt(x * x + y * y + z * z)
    if s < 1e-30:
        x, y, z = 0.0, 0.0, 1.0
    else:
        x, y, z = x / s, y / s, z / s
    x, y, z = (0.0 if abs(c) < 1e-30 else c for c in (x, y, z))
    if not math.isfinite(a):
        a = 0.0
    else:
        n = int(a / math.pi)                        # C's truncating conversion
        a -= (n + int(math.fmod(n, 2))) * math.pi    # C's remainder keeps the sign
    return [x, y, z, 0.0 if abs(a) < 1e-30 else a]


def _qmul(a, b):
    """Quaternion product (include/Quaternion.h:117-124); (s, x, y, z)."""
    s1, x1, y1, z1 = a
    s2, x2, y2, z2 = b
    return (s1 * s2 - (x1 * x2 + y1 * y2 + z1 * z2),
            s1 * x2 + x1 * s2 + y1 * z2 - z1 * y2,
            s1 * y2 + y1 * s2 + z1 * x2 - x1 * z2,
            s1 * z2 + z1 * s2 + x1 * y2 - y1 * x2)


def _qaxis(axis, angle):
    n = math.sqrt(sum(c * c for c in axis))
    sh = math.sin(angle / 2)
    q = (math.cos(angle / 2), axis[0] / n * sh, axis[1] / n * sh, axis[2] / n * sh)
    return tuple(0.0 if abs(c) < 1e-30 else c for c in q)


def views_for_refinement(view, step: float) -> np.ndarray:
    """views_for_refinement (src/util/View.cpp:371-436): 27 x (x, y, z, angle)."""
    v = _view_normalize(view)
    ax1 = [v[1], -v[0], 0.0]
    if math.sqrt(sum(c * c for c in ax1)) < 0.001:
        ax1[0] = 1.0
    n1 = math.sqrt(sum(c * c for c in ax1))
    ax1 = [c / n1 for c in ax1]
    ax2 = [v[1] * ax1[2] - v[2] * ax1[1], v[2] * ax1[0] - v[0] * ax1[2],
           v[0] * ax1[1] - v[1] * ax1[0]]
    if math.sqrt(sum(c * c for c in ax2)) < 0.001:
        ax2[1] = 1.0
    n2 = math.sqrt(sum(c * c for c in ax2))
    ax2 = [c / n2 for c in ax2]
    qv = (0.0, v[0], v[1], v[2])
    out = []
    steps = []
    a = -step
    while a <= step:                                # the reference's floating loop
        steps.append(a)
        a += step
    for a1 in steps:
        q1 = _qaxis(ax1, a1)
        for a2 in steps:
            q = _qmul(_qaxis(ax2, a2), q1)
            n2q = sum(c * c for c in q)
            qi = (q[0] / n2q, -q[1] / n2q, -q[2] / n2q, -q[3] / n2q)
            r = _qmul(_qmul(q, qv), qi)
            for a3 in steps:
                out.append(_view_normalize([r[1], r[2], r[3], v[3] + a3]))
    return np.asarray(out)


def refine_views(score, view, step: float, accuracy: float, max_iter: int = 10000):
    """The view loop of img_refine_grid (src/mg/mg_refine.cpp:1116-1140) with
    a scoring function score(views (27, 4)) -> (27,) FOMs: returns (best view,
    best FOM, passes)."""
    best_view = _view_normalize(view)
    bestfom = -1.0
    it = 0
    while step >= accuracy and it < max_iter:
        prefom = bestfom
        bestfom = -1.0
        grid = views_for_refinement(best_view, step)
        f = score(grid)
        ib = 0
        for i in range(len(grid)):
            if bestfom < f[i]:
                bestfom = float(f[i])
                ib = i
                best_view = list(grid[i])
        if ib == len(grid) // 2 or abs(prefom - bestfom) < 1e-6:
            step /= 2
        it += 1
    return np.asarray(best_view), bestfom, it

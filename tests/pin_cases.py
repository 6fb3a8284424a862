"""Hand-derived cases that pin the insertion geometry and the finish of the
reconstruction: each expected value below follows from the definitions in
the reference (Bimage_reconstruct.cpp:25-174 for the sample coordinates and
weights, :486-515 and :610-634 for the weighing and the SNR figure of merit,
Bimage_fspace.cpp:2024-2074 for the Friedel residual), worked out by hand
for one sample or one voxel, not by re-running the algorithm.

Shared by the oracle tests (CPU) and the library tests (GPU)."""
from __future__ import annotations

import math

import numpy as np

from bsoft_amd.synth import PARTICLE_DTYPE

# Relativistic electron wavelength at 1 kV (include/ctf.h:309-317 with the
# constants of include/utilities.h:55-58), written out: h / sqrt(2 m e V
# (1 + e V / (2 m c^2))) in Angstrom.
_H, _E, _M, _C = 6.626070e-34, 1.602177e-19, 9.109384e-31, 299792458.0
LAMBDA_1KV = _H * 1e10 / math.sqrt(2 * _M * _E * 1e3 * (1 + _E * 1e3 / (2 * _M * _C * _C)))

# Geometry shared by the insertion cases: identity view, ft = N = 2 box,
# ft pix = N pix_map = 16 A. A sample at Fourier index (h, k) has
# s = (h, k) / (16 mag) and lands at m = s N pix_map = (h, k) / mag, so the
# hand counts below hold for both geometries: the oracle's (N = 16 at 1 A)
# and the library's (N = 32 at 0.5 A: its insertion takes ft >= 32); only
# the periodic wrap of negative voxel indices (by N) differs, and the extra
# samples of the larger transform (|h| or |k| > 8) land away from every
# counted voxel.
class Geometry:
    def __init__(self, n: int, pix: float):
        self.N, self.BOX, self.FT, self.PIX = n, n // 2, n, pix


CPU = Geometry(16, 1.0)
GPU = Geometry(32, 0.5)
N, BOX, FT = CPU.N, CPU.BOX, CPU.FT
PAD_WEIGHT = 1.0 / (0.6 * (FT / BOX) ** 2)     # 1 / (0.6 (ft/box)^2), mg_reconstruct.cpp:192-199


def particle(g: Geometry = CPU, **kw) -> np.ndarray:
    p = np.zeros(1, dtype=PARTICLE_DTYPE)
    p["view"] = kw.get("view", (0.0, 0.0, 1.0, 0.0))
    p["pixel_size"] = g.PIX
    p["mag"] = kw.get("mag", 1.0)
    p["volt"] = kw.get("volt", 300e3)
    p["sel"] = kw.get("sel", 1.0)
    p["ori"] = (g.BOX // 2, g.BOX // 2)
    p["hdr_ori"] = (g.BOX // 2, g.BOX // 2)
    return p


def image(seed: int = 0, g: Geometry = CPU) -> np.ndarray:
    return np.random.default_rng(seed).normal(size=(1, g.BOX, g.BOX)).astype(np.float32)


# Each case: reconstruction parameters, the particle, and {(x, y, z): weight
# sum at that voxel} for voxels whose contributions were counted by hand
# (signed voxel coordinates are wrapped by N at the end).
def insertion_cases(g: Geometry = CPU):
    pw = PAD_WEIGHT
    cases = {}
    # Weighted nearest neighbour (interp 1): mag 1/0.7 puts sample (3, 0) at
    # m = (2.1, 0, 0). Voxel (2, 0, 0) receives only it: round(0.7 h) = 2
    # needs h = 3, round(0.7 k) = 0 needs k = 0. Weight pw (1 - |m - v|) =
    # 0.9 pw. Its mate (-3, 0) lands at (-2.1) -> voxel -2, also 0.9 pw.
    # Voxel (2, 1, 0) receives (3, 1) at (2.1, 0.7) and (3, 2) at (2.1, 1.4):
    # pw (2 - sqrt(0.01 + 0.09) - sqrt(0.01 + 0.16)).
    cases["weighted_nn"] = (dict(interp=1), particle(g, mag=1 / 0.7), {
        (2, 0, 0): 0.9 * pw, (-2, 0, 0): 0.9 * pw,
        (2, 1, 0): pw * (2 - math.sqrt(0.1) - math.sqrt(0.17)),
    })
    # Plain nearest neighbour at the same coordinates: each sample adds pw.
    # Voxel (2, 1, 0) gets 2 pw; voxel (1, 0, 0) gets h in {1, 2} (0.7, 1.4)
    # with k = 0: 2 pw.
    cases["nearest_mag"] = (dict(interp=0), particle(g, mag=1 / 0.7), {
        (2, 0, 0): pw, (2, 1, 0): 2 * pw, (1, 0, 0): 2 * pw, (0, 0, 0): pw,
    })
    # Magnification 1: sample (h, k) lands on voxel (h, k, 0) exactly.
    cases["nearest_mag1"] = (dict(interp=0), particle(g), {
        (3, 1, 0): pw, (7, 7, 0): pw, (-7, 0, 0): pw,
    })
    # Bootstrap (flags 4): the weight is the selection number times pw.
    cases["bootstrap"] = (dict(interp=0, flags=4), particle(g, sel=3.0), {
        (3, 1, 0): 3 * pw, (1, 0, 0): 3 * pw,
    })
    # Low-resolution limit: lo_res = 16 / sqrt(10.5) keeps s^2 >= 10.5 / 256,
    # i.e. h^2 + k^2 >= 10.5: (3, 1) (10) is cut, (3, 2) (13) kept, DC cut.
    cases["lo_res"] = (dict(interp=0, lo_res=16 / math.sqrt(10.5)), particle(g), {
        (3, 1, 0): 0.0, (3, 2, 0): pw, (0, 0, 0): 0.0, (1, 0, 0): 0.0,
    })
    # Resolution taper: hi_res = 16 / sqrt(50.5) (maxs2 = 50.5 / 256). Sample
    # (7, 1) has s^2 = 50 / 256: d = (maxs2 - s^2) N = 0.5 / 256 * 16 =
    # 1/32 < 1, weight pw sqrt(1/32). (5, 5) (50 as well): the same. (6, 4)
    # (52) is beyond the limit: 0. (1, 1): d = 48.5 / 16 > 1, weight pw.
    cases["taper"] = (dict(interp=0, hi_res=16 / math.sqrt(50.5)), particle(g), {
        (7, 1, 0): pw * math.sqrt(1 / 32), (5, 5, 0): pw * math.sqrt(1 / 32),
        (6, 4, 0): 0.0, (1, 1, 0): pw,
    })
    # Ewald sphere (flags 8) at 1 kV: each sample is inserted twice, at
    # s_z = +-(lambda/2) s^2, i.e. m_z = +-(lambda/2) (h^2 + k^2) / 16.
    # (7, 7): m_z = +-1.187 -> voxels z = 1 and z = -1, none at z = 0.
    # (2, 0): m_z = +-0.048 -> both at z = 0: 2 pw.
    lam2 = LAMBDA_1KV / 2
    assert round(lam2 * 98 / 16) == 1 and round(lam2 * 4 / 16) == 0
    cases["ewald"] = (dict(interp=0, flags=8), particle(g, volt=1e3), {
        (7, 7, 1): pw, (7, 7, -1): pw, (7, 7, 0): 0.0, (2, 0, 0): 2 * pw,
    })
    return {name: (params, part, {(x % g.N, y % g.N, z % g.N): w for (x, y, z), w in exp.items()})
            for name, (params, part, exp) in cases.items()}


def recon_params(g: Geometry = CPU, **kw) -> dict:
    d = dict(size=g.N, box=g.BOX, ft_size=g.FT, interp=0, sampling=g.PIX, hi_res=1.0, lo_res=0.0,
             ctf_action=0, flags=0, wiener=0.0, nhalves=1, batch=0, mem_cap=0)
    d.update(kw)
    return d


# ---------------------------------------------------------------- finish
def fom_accumulators(n: int, seed: int):
    """Hand-built sums of K contributions x_k with weights w_k per voxel, K in
    {0, 1, 2, 3, 5, 9} by voxel. Returns (planes (5, n, n, n), samples) where
    samples[i] = (x_k, w_k) of voxel i (flat index)."""
    rng = np.random.default_rng(seed)
    n3 = n ** 3
    Ks = np.array([0, 1, 2, 3, 5, 9])
    planes = np.zeros((5, n3))
    samples = []
    for i in range(n3):
        K = int(Ks[rng.integers(len(Ks))]) if i else 0
        if rng.random() < 0.5:
            w = np.ones(K)
        else:
            w = rng.uniform(0.2, 1.5, size=K)
        mean = rng.normal() + 1j * rng.normal()
        x = mean + rng.uniform(0.1, 2.0) * (rng.normal(size=K) + 1j * rng.normal(size=K))
        planes[0, i] = np.sum(w * x.real)
        planes[1, i] = np.sum(w * x.imag)
        planes[2, i] = np.sum(w)
        planes[3, i] = np.sum(w * np.abs(x) ** 2)
        planes[4, i] = np.sum(w * w)
        samples.append((x, w))
    return planes.reshape(5, n, n, n), samples


def fom_expected(x: np.ndarray, w: np.ndarray) -> float:
    """SNR figure of merit of one voxel from its contributions, in terms of
    sample statistics: the weighted mean m and the reliability-weighted
    variance v = sum w |x - m|^2 / (W - sum w^2 / W), W = sum w.

    fspace_reconstruction_weigh divides the power sum by W - sum w^2 / W =
    W (1 - sum w^2 / W^2) when that exceeds 1 (else by W when W > 1), so the
    power term is P = W |m|^2 / w2 + v (first case) and the noise estimate
    P - |m|^2; fspace_reconstruction_snr then reports (W |m|^2 - P) /
    max(noise, |m|^2) where W > 1, else 0."""
    W = float(np.sum(w))
    if W <= 1:
        return 0.0
    m = np.sum(w * x) / W
    sig = abs(m) ** 2
    w2 = W - float(np.sum(w * w)) / W
    ss = float(np.sum(w * np.abs(x - m) ** 2))
    if w2 > 1:
        P = W * sig / w2 + ss / w2
    else:
        P = (W * sig + ss) / W             # the power sum over W
    noise = max(P - sig, sig)
    return (W * sig - P) / noise


def friedel_volume(n: int = 8):
    """A transform (n^3, even n) that is Friedel-symmetric except one pair,
    and its residual R worked out by hand.

    Pair A: v at (1, 2, 3), conj(v) at its mate (n-1, n-2, n-3): no
    deviation, weight |v|^2. Pair B: u at (2, 0, 0), t at (n-2, 0, 0):
    deviation d = (Re u - Re t, Im u + Im t), weight (|u|^2 + |t|^2) / 2.
    The residual loop visits each of these pairs once (x < (n+1)/2), so
    R = sqrt(wB / (wA + wB)) |d|."""
    F = np.zeros((n, n, n), dtype=np.complex128)    # (z, y, x)
    v = 1.5 - 0.5j
    u, t = 2.0 + 1.0j, 1.0 + 0.25j
    F[3, 2, 1] = v
    F[n - 3, n - 2, n - 1] = np.conj(v)
    F[0, 0, 2] = u
    F[0, 0, n - 2] = t
    wA = abs(v) ** 2
    wB = (abs(u) ** 2 + abs(t) ** 2) / 2
    d = math.hypot(u.real - t.real, u.imag + t.imag)
    return F, math.sqrt(wB / (wA + wB)) * d

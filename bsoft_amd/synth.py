"""Seeded synthetic single-particle data: Gaussian-blob maps, uniform SO(3)
views, CTF-modulated noisy projections.

Conventions follow the reference:
  * projection-slice: the 2D transform of a particle equals the map transform
    on the plane R.(h,k,0) (Bimage::fspace_pack_2D, Bimage_reconstruct.cpp:150-153),
    so a blob at c (relative to the map origin) projects to (R^T c)_xy;
  * CTF phase aberration chi = -amp_shift + (abw22 cos2a + abw20 + abw2-2 sin2a) s^2
    + abw40 s^4 (include/ctf.h:321-335), CTF = sin(chi) (mg_ctf.cpp:102);
  * noise sigma = sigma_signal / sqrt(SNR) (src/bproject.cpp:429);
  * View <-> quaternion as include/View.h:52-63 / :131-153. Views are drawn
    uniformly on SO(3) with Shoemake's method (the reference's random_view,
    src/util/View.cpp:116-128, is not uniform).

Small cases run on the CPU (tests), large ones on the GPU (bench) through
torch; all random parameters come from numpy's seeded generator so both
devices see identical inputs.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np
import torch

from .abi import Particle

# numpy layout identical to struct bsx_particle (include/bsx.h)
PARTICLE_DTYPE = np.dtype([
    ("view", "<f8", (4,)), ("ori", "<f8", (2,)), ("hdr_ori", "<f8", (2,)),
    ("pixel_size", "<f8"), ("mag", "<f8"), ("volt", "<f8"), ("cs", "<f8"),
    ("amp_shift", "<f8"), ("def_avg", "<f8"), ("def_dev", "<f8"), ("ast_angle", "<f8"),
    ("sel", "<f8"), ("half", "<i4"), ("reserved", "<i4"),
], align=True)
assert PARTICLE_DTYPE.itemsize == C.sizeof(Particle)


def particles_ctypes(arr: np.ndarray):
    """View a PARTICLE_DTYPE array as a ctypes Particle pointer (no copy)."""
    assert arr.dtype == PARTICLE_DTYPE and arr.flags.c_contiguous
    return arr.ctypes.data_as(C.POINTER(Particle))


def particles_list(arr: np.ndarray):
    """Copy a PARTICLE_DTYPE array into a ctypes Particle array."""
    out = (Particle * len(arr))()
    C.memmove(out, arr.ctypes.data, arr.nbytes)
    return out


# ---------------------------------------------------------------- rotations
def quat_to_view(q: np.ndarray) -> np.ndarray:
    """View(Quaternion&), include/View.h:52-63. q: (..., 4) = (s, x, y, z)."""
    q = np.asarray(q, dtype=np.float64)
    ang = np.arctan2(q[..., 3], q[..., 0])
    ca, sa = np.cos(ang), np.sin(ang)
    q2 = q[..., 0] ** 2 + q[..., 3] ** 2
    f2 = 2 * np.sqrt(q2)
    ang = np.where(q2 < 1e-30, np.arctan2(q[..., 1], q[..., 2]), ang)
    v0 = f2 * (q[..., 2] * ca + q[..., 1] * sa)
    v1 = f2 * (q[..., 2] * sa - q[..., 1] * ca)
    v2 = 2 * q2 - 1
    nrm = np.sqrt(v0 ** 2 + v1 ** 2 + v2 ** 2)
    a = np.remainder(2 * ang + np.pi, 2 * np.pi) - np.pi
    return np.stack([v0 / nrm, v1 / nrm, v2 / nrm, a], axis=-1)


def view_to_matrix(view: np.ndarray) -> np.ndarray:
    """View::matrix (include/View.h:131-157, Matrix3.h:30-42), vectorised."""
    v = np.asarray(view, dtype=np.float64).reshape(-1, 4)
    x, y, z, a = v[:, 0].copy(), v[:, 1].copy(), v[:, 2].copy(), v[:, 3]
    s = np.sqrt(x * x + y * y + z * z)
    small = s < 1e-30
    s = np.where(small, 1, s)
    x, y, z = np.where(small, 0, x / s), np.where(small, 0, y / s), np.where(small, 1, z / s)
    a = np.remainder(a + np.pi, 2 * np.pi) - np.pi
    ca, sa = np.cos(a / 2), np.sin(a / 2)
    q = np.stack([ca, np.zeros_like(ca), np.zeros_like(ca), sa], -1)
    bottom = z <= 1e-30 - 1
    mid = (z < 1 - 1e-30) & ~bottom
    z1 = np.where(mid, z + 1, 1.0)
    f1, f2 = np.sqrt(z1 / 2), np.sqrt(1 / (2 * z1))
    qm = np.stack([f1 * ca, f2 * (x * sa - y * ca), f2 * (x * ca + y * sa), f1 * sa], -1)
    qb = np.stack([np.zeros_like(ca), sa, ca, np.zeros_like(ca)], -1)
    q = np.where(mid[:, None], qm, q)
    q = np.where(bottom[:, None], qb, q)
    q0, q1, q2, q3 = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    d = np.stack([
        q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3, 2 * q1 * q2 - 2 * q0 * q3, 2 * q1 * q3 + 2 * q0 * q2,
        2 * q1 * q2 + 2 * q0 * q3, q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3, 2 * q2 * q3 - 2 * q0 * q1,
        2 * q1 * q3 - 2 * q0 * q2, 2 * q2 * q3 + 2 * q0 * q1, q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3,
    ], -1)
    d = np.where(np.abs(d) < 1e-12, 0, d)
    return np.clip(d, -1, 1).reshape(-1, 3, 3)


def uniform_quaternions(n: int, rng: np.random.Generator) -> np.ndarray:
    """Shoemake (1992) uniform random unit quaternions, (s, x, y, z)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    return np.stack([b * np.cos(2 * np.pi * u3), a * np.sin(2 * np.pi * u2),
                     a * np.cos(2 * np.pi * u2), b * np.sin(2 * np.pi * u3)], -1)


def uniform_views(n: int, rng: np.random.Generator) -> np.ndarray:
    return quat_to_view(uniform_quaternions(n, rng))


# ---------------------------------------------------------------- CTF
def ctf_lambda(volt: float) -> float:
    """CTFparam::lambda (include/ctf.h:309-317), constants utilities.h:55-58."""
    h, e, m, c = 6.626070e-34, 1.602177e-19, 9.109384e-31, 299792458.0
    return h * 1e10 / math.sqrt(2.0 * m * e * volt * (1.0 + e * volt / (2.0 * m * c * c)))


def ctf_sin_chi(h, k, box, pix, volt, cs, amp_shift, def_avg, def_dev, ast):
    """sin(chi) on integer frequency grids h, k (torch tensors, broadcastable
    with per-particle parameter tensors of shape (P,1,1))."""
    wl = ctf_lambda(volt)
    t1, t2 = math.pi / 2 * wl ** 3, -math.pi * wl
    sx, sy = h / (box * pix), k / (box * pix)
    s2 = sx * sx + sy * sy
    a = torch.atan2(sy, sx)
    c20 = t2 * def_avg
    c22 = t2 * def_dev.abs() * torch.cos(2 * ast)
    c2m2 = t2 * def_dev.abs() * torch.sin(2 * ast)
    chi = -amp_shift + c2m2 * s2 * torch.sin(2 * a) + c20 * s2 + c22 * s2 * torch.cos(2 * a) \
        + t1 * cs * s2 * s2
    return torch.sin(chi)


# ---------------------------------------------------------------- data sets
@dataclass
class BlobMap:
    centers: np.ndarray   # (B, 3) voxels, relative to the map origin (N/2)
    sigmas: np.ndarray    # (B,)
    amps: np.ndarray      # (B,)
    n: int

    def volume(self, device="cpu", dtype=torch.float32) -> torch.Tensor:
        """Real-space map, origin at N/2 (z, y, x order)."""
        n = self.n
        g = torch.arange(n, device=device, dtype=torch.float64) - n // 2
        out = torch.zeros((n, n, n), device=device, dtype=torch.float64)
        for (cx, cy, cz), s, a in zip(self.centers, self.sigmas, self.amps):
            gx = torch.exp(-((g - cx) ** 2) / (2 * s * s))
            gy = torch.exp(-((g - cy) ** 2) / (2 * s * s))
            gz = torch.exp(-((g - cz) ** 2) / (2 * s * s))
            out += a * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
        return out.to(dtype)


def make_blob_map(n: int, nblobs: int, rng: np.random.Generator,
                  sigma=(2.0, 5.0), amp=(0.5, 1.0), radius=0.35) -> BlobMap:
    """Centres uniform inside a sphere of radius `radius`*N, sigma ~ U, amp ~ U."""
    d = rng.normal(size=(nblobs, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = radius * n * rng.random(nblobs) ** (1.0 / 3.0)
    return BlobMap(centers=d * r[:, None], sigmas=rng.uniform(*sigma, nblobs),
                   amps=rng.uniform(*amp, nblobs), n=n)


@dataclass
class ParticleSet:
    images: torch.Tensor      # (P, box, box) float32
    parts: np.ndarray         # PARTICLE_DTYPE (P,)
    clean: torch.Tensor | None = None


def make_particles(bm: BlobMap, nparts: int, box: int, rng: np.random.Generator, *,
                   pix=1.5, volt=300e3, cs=2.7e7, amp_contrast=0.07,
                   defocus=(1e4, 3e4), astig=(0.0, 0.0), snr=0.1, device="cpu",
                   seed=0, chunk=2048, keep_clean=False, views=None) -> ParticleSet:
    """Projections of a blob map at uniform SO(3) views with CTF and noise.

    Each particle has its own defocus (and astigmatism) — the per-particle
    values the reference honours through its micrograph CTF."""
    parts = np.zeros(nparts, dtype=PARTICLE_DTYPE)
    if views is None:
        views = uniform_views(nparts, rng)
    parts["view"] = views
    parts["ori"] = box // 2
    parts["hdr_ori"] = box // 2
    parts["pixel_size"] = pix
    parts["mag"] = 1.0
    parts["volt"] = volt
    parts["cs"] = cs
    parts["amp_shift"] = math.asin(amp_contrast)
    parts["def_avg"] = rng.uniform(*defocus, nparts)
    parts["def_dev"] = rng.uniform(*astig, nparts) if astig[1] > 0 else 0.0
    parts["ast_angle"] = rng.uniform(-math.pi, math.pi, nparts) if astig[1] > 0 else 0.0
    parts["sel"] = 1.0
    parts["half"] = np.arange(nparts) % 2
    mats = view_to_matrix(views)                     # (P,3,3)
    cen = bm.centers                                 # (B,3)
    # blob centre in the particle frame: (R^T c)_xy
    uv = np.einsum("pji,bj->pbi", mats, cen)[:, :, :2]  # (P,B,2)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    imgs = torch.empty((nparts, box, box), device=device, dtype=torch.float32)
    clean = torch.empty_like(imgs) if keep_clean else None
    g = torch.arange(box, device=device, dtype=torch.float32) - box // 2
    sig = torch.as_tensor(bm.sigmas, device=device, dtype=torch.float32)
    amp = torch.as_tensor(bm.amps * np.sqrt(2 * np.pi) * bm.sigmas, device=device, dtype=torch.float32)
    fh = torch.fft.fftfreq(box, d=1.0 / box, device=device).to(torch.float64)
    hh, kk = fh[None, None, :], fh[None, :, None]
    for s in range(0, nparts, chunk):
        e = min(nparts, s + chunk)
        u = torch.as_tensor(uv[s:e], device=device, dtype=torch.float32)   # (p,B,2)
        gx = torch.exp(-((g[None, None, :] - u[:, :, 0:1]) ** 2) / (2 * sig[None, :, None] ** 2))
        gy = torch.exp(-((g[None, None, :] - u[:, :, 1:2]) ** 2) / (2 * sig[None, :, None] ** 2))
        gy = (gy * amp[None, :, None]).transpose(1, 2).contiguous()          # (p,box,B)
        proj = torch.bmm(gy, gx.contiguous())                               # (p,box,box)
        P = lambda f: torch.as_tensor(parts[f][s:e], device=device, dtype=torch.float64)[:, None, None]
        ctf = ctf_sin_chi(hh, kk, box, pix, volt, cs, math.asin(amp_contrast),
                          P("def_avg"), P("def_dev"), P("ast_angle"))
        ft = torch.fft.fft2(proj.to(torch.float64))
        proj = torch.fft.ifft2(ft * ctf).real.to(torch.float32)
        if keep_clean:
            clean[s:e] = proj
        if snr is None or snr <= 0:
            imgs[s:e] = proj
            continue
        sd = proj.reshape(e - s, -1).std(dim=1)
        noise = torch.randn(proj.shape, generator=gen, device=device, dtype=torch.float32)
        imgs[s:e] = proj + noise * (sd / math.sqrt(snr))[:, None, None]
    return ParticleSet(images=imgs, parts=parts, clean=clean)

"""numpy bindings of the CPU oracle (oracle/bsx_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from bsoft_amd.abi import Particle, ReconParams, Stats

HERE = os.path.dirname(os.path.abspath(__file__))


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_libs: dict = {}
_vp = C.c_void_p
_dp = C.POINTER(C.c_double)


def _load(prec: int) -> C.CDLL:
    if prec in _libs:
        return _libs[prec]
    path = os.path.join(HERE, f"liboracle{prec}.so")
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    sig = {
        "or_real_bytes": ([], C.c_int),
        "or_part_ft_size": ([C.c_int, C.c_double, C.c_int], C.c_int),
        "or_find_next_power": ([C.c_int, C.c_int], C.c_int),
        "or_view_matrix": ([_dp, _dp], None),
        "or_lambda": ([C.c_double], C.c_double),
        "or_check_resolution": ([C.c_double, C.c_int, C.c_double, C.c_int], C.c_double),
        "or_ctf_value": ([C.POINTER(Particle), C.c_double, C.c_double, C.c_int, C.c_double], C.c_double),
        "or_fft": ([_vp, C.c_long, C.c_long, C.c_long, C.c_long, C.c_int, C.c_int], C.c_int),
        "or_background": ([_vp, C.c_long, _dp], C.c_double),
        "or_stage_particle": ([C.POINTER(ReconParams), C.POINTER(Particle), _vp, _vp, _dp], C.c_int),
        "or_recon_new": ([C.POINTER(ReconParams), C.c_int], _vp),
        "or_recon_free": ([_vp], None),
        "or_recon_insert": ([_vp, C.POINTER(Particle), _vp, C.c_long, C.c_int], C.c_int),
        "or_recon_partial_acc": ([_vp, C.c_int, _vp], C.c_int),
        "or_recon_finish": ([_vp, C.c_int, C.c_int, _vp, _vp, _vp, C.POINTER(Stats)], C.c_int),
        "or_fsc": ([_vp, _vp, C.c_long, C.c_double, C.c_double, _dp, C.c_int], C.c_int),
        "or_fsi_kernel": ([C.c_int, C.c_int, _dp], C.c_int),
        "or_map_transform": ([_vp, C.c_long, _dp, _vp], C.c_int),
        "or_central_section": ([_vp, C.c_long, C.c_double, _dp, C.c_double, C.c_int, C.c_int,
                                C.c_int, _vp], C.c_int),
        "or_recip_space_fom": ([_vp, _vp, C.c_long, C.c_double, C.c_double, C.c_double, _dp],
                               C.c_double),
        "or_asu_views_c1": ([C.c_double, C.c_double, C.c_int, _dp, C.c_long], C.c_long),
        "or_friedel": ([_vp, C.c_long, C.c_long, C.c_long], C.c_double),
        "or_recon_set_partial_acc": ([_vp, C.c_int, _vp, C.c_long], C.c_int),
        "or_rescale_to_avg_std": ([_vp, C.c_long, C.c_double, C.c_double], C.c_int),
        "or_rescale_f32": ([_vp, C.c_long, C.c_double, C.c_double], None),
        "or_filter_sinc": ([_vp, C.c_long, C.c_long, C.c_long, _dp], C.c_int),
    }
    for name, (a, r) in sig.items():
        f = getattr(L, name)
        f.argtypes = a
        f.restype = r
    _libs[prec] = L
    return L


class Oracle:
    """One precision of the oracle: prec=64 (parity) or 32 (CPU baseline)."""

    def __init__(self, prec: int = 64):
        self.prec = prec
        self.L = _load(prec)
        self.dt = np.float64 if prec == 64 else np.float32
        self.cdt = np.complex128 if prec == 64 else np.complex64

    # -- helpers -----------------------------------------------------------
    @staticmethod
    def _p(a: np.ndarray):
        return a.ctypes.data_as(_vp)

    @staticmethod
    def _d(a: np.ndarray):
        return a.ctypes.data_as(_dp)

    def part_ft_size(self, xsize, scale=1.0, pad=2):
        return self.L.or_part_ft_size(int(xsize), float(scale), int(pad))

    def view_matrix(self, view) -> np.ndarray:
        v = np.ascontiguousarray(view, dtype=np.float64)
        m = np.zeros(9)
        self.L.or_view_matrix(self._d(v), self._d(m))
        return m.reshape(3, 3)

    def lam(self, volt):
        return self.L.or_lambda(float(volt))

    def check_resolution(self, res, n, sampling, ndim=3):
        return self.L.or_check_resolution(float(res), int(n), float(sampling), int(ndim))

    def ctf_value(self, part: Particle, sx, sy, flip=False, wiener=0.0):
        return self.L.or_ctf_value(C.byref(part), float(sx), float(sy), int(flip), float(wiener))

    def fft(self, a: np.ndarray, dir=-1, norm=0) -> np.ndarray:
        """FFT of the last 1-3 axes (C order) of a complex array; batch = leading axes."""
        a = np.ascontiguousarray(a, dtype=self.cdt).copy()
        shp = a.shape
        nd = min(3, len(shp))
        dims = list(shp[-nd:])[::-1] + [1] * (3 - nd)
        batch = int(np.prod(shp[:-nd])) if len(shp) > nd else 1
        r = self.L.or_fft(self._p(a), dims[0], dims[1], dims[2], batch, int(dir), int(norm))
        assert r == 0
        return a

    def background(self, img: np.ndarray, hdr_ori) -> float:
        img = np.ascontiguousarray(img, dtype=np.float32)
        o = np.ascontiguousarray(hdr_ori, dtype=np.float64)
        return self.L.or_background(self._p(img), img.shape[0], self._d(o))

    def stage_particle(self, rp: ReconParams, part: Particle, img: np.ndarray):
        img = np.ascontiguousarray(img, dtype=np.float32)
        out = np.zeros((rp.ft_size, rp.ft_size), dtype=self.cdt)
        w = C.c_double()
        r = self.L.or_stage_particle(C.byref(rp), C.byref(part), self._p(img), self._p(out), C.byref(w))
        assert r == 0
        return out, w.value

    def fsc(self, a: np.ndarray, b: np.ndarray, sampling, hi_res) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=self.cdt)
        b = np.ascontiguousarray(b, dtype=self.cdt)
        n = a.shape[0]
        out = np.zeros(4 * n)
        k = self.L.or_fsc(self._p(a), self._p(b), n, float(sampling), float(hi_res), self._d(out), 4 * n)
        assert k >= 0
        return out[:k]

    def fsc_with_support(self, a: np.ndarray, b: np.ndarray, sampling, hi_res):
        """FSC plus a mask of the shells that hold power in both volumes (the
        FSC of an empty shell is undefined; the reference prints 0)."""
        f = self.fsc(a, b, sampling, hi_res)
        n = a.shape[0]
        h = np.fft.fftfreq(n, 1.0 / n)
        r = np.sqrt(h[:, None, None] ** 2 + h[None, :, None] ** 2 + h[None, None, :] ** 2)
        ir = np.floor(r).astype(np.int64)
        pa = np.bincount(ir.ravel(), weights=(np.abs(a) ** 2).ravel(), minlength=len(f) + 2)
        pb = np.bincount(ir.ravel(), weights=(np.abs(b) ** 2).ravel(), minlength=len(f) + 2)
        # shell i receives voxels with floor(r) = i and i-1 (linear split)
        sa = pa[: len(f)] + np.concatenate([[0], pa[: len(f) - 1]])
        sb = pb[: len(f)] + np.concatenate([[0], pb[: len(f) - 1]])
        return f, (sa > 0) & (sb > 0)

    def fsi_kernel(self, width, power) -> np.ndarray:
        w = 1 if power < 0 else width
        t = np.zeros(513 * w)
        self.L.or_fsi_kernel(int(width), int(power), self._d(t))
        return t.reshape(513, w)

    def map_transform(self, m: np.ndarray, origin) -> np.ndarray:
        m = np.ascontiguousarray(m, dtype=self.dt)
        n = m.shape[0]
        o = np.ascontiguousarray(origin, dtype=np.float64)
        out = np.zeros((n, n, n), dtype=self.cdt)
        assert self.L.or_map_transform(self._p(m), n, self._d(o), self._p(out)) == 0
        return out

    def central_section(self, ft: np.ndarray, sampling, mat, resolution, kw=8, kp=2, center=1):
        ft = np.ascontiguousarray(ft, dtype=self.cdt)
        n = ft.shape[0]
        mt = np.ascontiguousarray(mat, dtype=np.float64).reshape(9)
        out = np.zeros((n, n), dtype=self.cdt)
        assert self.L.or_central_section(self._p(ft), n, float(sampling), self._d(mt), float(resolution),
                                         int(kw), int(kp), int(center), self._p(out)) == 0
        return out

    def recip_space_fom(self, p: np.ndarray, ref: np.ndarray, sampling, hi_res, lo_res=0.0):
        p = np.ascontiguousarray(p, dtype=self.cdt)
        ref = np.ascontiguousarray(ref, dtype=self.cdt)
        cv = C.c_double()
        f = self.L.or_recip_space_fom(self._p(p), self._p(ref), p.shape[0], float(sampling),
                                      float(hi_res), float(lo_res), C.byref(cv))
        return f, cv.value

    def asu_views_c1(self, theta_step, phi_step, flag=1) -> np.ndarray:
        cap = 200000
        v = np.zeros(4 * cap)
        n = self.L.or_asu_views_c1(float(theta_step), float(phi_step), int(flag), self._d(v), cap)
        return v[: 4 * min(n, cap)].reshape(-1, 4)

    def rescale_to_avg_std(self, d: np.ndarray, avg: float, std: float) -> np.ndarray:
        """Bimage::rescale_to_avg_std of a copy of d (oracle precision)."""
        d = np.ascontiguousarray(d, dtype=self.dt).copy()
        assert self.L.or_rescale_to_avg_std(self._p(d), d.size, float(avg), float(std)) == 0
        return d

    def rescale_f32(self, img: np.ndarray, avg: float = 0.0, std: float = 1.0) -> np.ndarray:
        """The per-particle rescale of breconstruct -rescale (FP32 storage)."""
        img = np.ascontiguousarray(img, dtype=np.float32).copy()
        self.L.or_rescale_f32(self._p(img), img.size, float(avg), float(std))
        return img

    def filter_sinc(self, d: np.ndarray, origin) -> np.ndarray:
        """Bimage::filter_sinc of a copy of a (nz, ny, nx) map."""
        d = np.ascontiguousarray(d, dtype=self.dt).copy()
        nz, ny, nx = d.shape
        o = np.ascontiguousarray(origin, dtype=np.float64)
        assert self.L.or_filter_sinc(self._p(d), nx, ny, nz, self._d(o)) == 0
        return d

    def friedel(self, ft: np.ndarray) -> float:
        ft = np.ascontiguousarray(ft, dtype=self.cdt)
        nz, ny, nx = ft.shape
        return self.L.or_friedel(self._p(ft), nx, ny, nz)


class OracleRecon:
    """CPU reconstruction with T partial maps (breconstruct -threads T)."""

    def __init__(self, rp: ReconParams, npartial: int = 2, prec: int = 64):
        self.o = Oracle(prec)
        self.rp = rp
        self.h = self.o.L.or_recon_new(C.byref(rp), int(npartial))
        if not self.h:
            raise RuntimeError("or_recon_new failed")
        self.T = npartial

    def __del__(self):
        if getattr(self, "h", None):
            self.o.L.or_recon_free(self.h)
            self.h = None

    def insert(self, parts, imgs: np.ndarray, nthreads: int = 0) -> None:
        imgs = np.ascontiguousarray(imgs, dtype=np.float32)
        arr = (Particle * len(parts))(*parts) if not isinstance(parts, C.Array) else parts
        r = self.o.L.or_recon_insert(self.h, arr, self.o._p(imgs), len(parts), int(nthreads))
        assert r == 0

    def _shape(self):
        n = self.rp.size
        return (1 if self.rp.flags & 2 else n, n, n)     # flags & 2: a 2D map

    def set_partial_acc(self, t: int, planes: np.ndarray, count: int = 0) -> None:
        """Hand-built accumulators (5, *shape): sum w re, sum w im, sum w,
        sum w|F|^2, sum w^2."""
        planes = np.ascontiguousarray(planes, dtype=self.o.dt)
        assert planes.shape == (5, *self._shape())
        assert self.o.L.or_recon_set_partial_acc(self.h, int(t), self.o._p(planes), int(count)) == 0

    def partial_acc(self, t: int) -> np.ndarray:
        shp = self._shape()
        out = np.zeros(5 * int(np.prod(shp)), dtype=self.o.dt)
        assert self.o.L.or_recon_partial_acc(self.h, int(t), self.o._p(out)) == 0
        return out.reshape(5, *shp)

    def finish(self, imap=0, nmaps=1, want_ft=False, want_fom=False):
        shp = self._shape()
        m = np.zeros(shp, dtype=self.o.dt)
        fom = np.zeros(shp, dtype=self.o.dt) if want_fom else None
        ft = np.zeros(shp, dtype=self.o.cdt) if want_ft else None
        st = Stats()
        r = self.o.L.or_recon_finish(self.h, int(imap), int(nmaps), self.o._p(m),
                                     self.o._p(fom) if fom is not None else None,
                                     self.o._p(ft) if ft is not None else None, C.byref(st))
        assert r == 0
        return m, fom, ft, st

"""Host mirror of Bsoft's reconstruction / projection entry points, calling
libbsx.so (include/bsx.h) through ctypes.

Names and argument meanings follow the reference:
  part_ft_size                  src/mg/mg_reconstruct.cpp:38
  particle_reconstruct          src/mg/mg_reconstruct.cpp:100
  img_reconstruction_sum_weigh  src/mg/mg_reconstruct.cpp:266
  fft (fft_setup_plan + fftw)   src/util/fft.cpp:31-114, Bimage_fft.cpp:176
  project / central_section     src/img/Bimage_project.cpp:288-443
  fsc                           src/img/Bimage_resolution.cpp:330
  img_recip_space_fom           src/mg/mg_refine.cpp:660

Device arrays are torch tensors on a ROCm device; there is no CPU path: the
library raises when it is missing and every computation runs in HIP kernels.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import abi
from .abi import ReconParams, Stats, call
from .synth import PARTICLE_DTYPE, particles_ctypes, view_to_matrix


def _ptr(t: torch.Tensor | None):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("device tensor required (libbsx works on HBM-resident data)")
    if not t.is_contiguous():
        raise ValueError("contiguous tensor required")
    return C.c_void_p(t.data_ptr())


def _stream(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def part_ft_size(xsize: int, scale: float = 1.0, pad_factor: int = 2) -> int:
    """part_ft_size (mg_reconstruct.cpp:38-64): transform size of padded particles."""
    return abi.lib().bsx_part_ft_size(int(xsize), float(scale), int(pad_factor))


def check_resolution(res: float, n: int, sampling: float, ndim: int = 3) -> float:
    """Bimage::check_resolution (Bimage.cpp:633-648)."""
    if res > n * sampling:
        return n * sampling
    rm = 2 / np.sqrt(ndim / sampling ** 2)
    return max(res, rm)


@dataclass
class FinishResult:
    map: torch.Tensor
    fom: torch.Tensor | None
    ft: torch.Tensor | None
    coverage: int
    friedel: float
    particles: int
    weight_sum: float


class Reconstruction:
    """Device accumulators of one reconstruction (the `prec` Bimage chain of
    particle_reconstruct: complex sum, power, weight, weight^2)."""

    def __init__(self, size: int, box: int, ft_size: int, sampling: float, hi_res: float = 0.0,
                 lo_res: float = 0.0, interp: int = 0, ctf_action: int = 0, wiener: float = 0.0,
                 flags: int = 0, nhalves: int = 1, batch: int = 0, mem_cap: int = 0,
                 method: int = 0, order: int = 0, merge: int = 0):
        if ctf_action > 2 and wiener < 0.01:            # breconstruct.cpp:187
            wiener = 0.2
        if hi_res < 0.1:                                  # breconstruct.cpp:236
            hi_res = sampling
        self.params = ReconParams(size=size, box=box, ft_size=ft_size, interp=interp,
                                  sampling=sampling, hi_res=hi_res, lo_res=lo_res,
                                  ctf_action=ctf_action, flags=flags, wiener=wiener,
                                  nhalves=nhalves, batch=batch, mem_cap=mem_cap,
                                  method=method, order=order, merge=merge)
        self.h = C.c_void_p()
        call("bsx_recon_create", C.byref(self.h), C.byref(self.params))
        self.size = size

    def close(self):
        if self.h:
            abi.lib().bsx_recon_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self, stream=None):
        call("bsx_recon_reset", self.h, _stream(stream))

    def insert(self, imgs: torch.Tensor, parts: np.ndarray, stream=None):
        """Insert particles: imgs (n, box, box) float32 on the device."""
        parts = np.ascontiguousarray(parts, dtype=PARTICLE_DTYPE)
        n = len(parts)
        if n and (imgs.dtype != torch.float32 or tuple(imgs.shape[1:]) != (self.params.box,) * 2
                  or imgs.shape[0] < n):
            raise ValueError(f"images must be float32 (n, {self.params.box}, {self.params.box})")
        call("bsx_recon_insert", self.h, _ptr(imgs) if n else None,
             particles_ctypes(parts) if n else None, n, _stream(stream))

    def autotune(self, imgs: torch.Tensor, parts: np.ndarray, stream=None) -> int:
        """Pick the fastest insertion batch size on a sample (bsx_recon_autotune);
        the handle must be empty and is left empty."""
        parts = np.ascontiguousarray(parts, dtype=PARTICLE_DTYPE)
        b, ms = C.c_int32(), C.c_double()
        imgs = imgs.contiguous()
        call("bsx_recon_autotune", self.h, _ptr(imgs), particles_ctypes(parts),
             len(parts), C.byref(b), C.byref(ms), _stream(stream))
        self._batch = b.value
        return b.value

    def set_symmetry(self, sym: str | None):
        """Point-group symmetry of later inserts (fspace_pack_2D with a
        Bsymmetry): each particle is added at all its symmetry-related views."""
        call("bsx_recon_set_symmetry", self.h, (sym or "C1").encode())

    def batch(self) -> int:
        return self.tuning()[0]

    def set_tuning(self, batch: int = 0, piece: int = 0):
        """Pin the insertion batch and/or the tile-pass piece length (0 keeps)."""
        call("bsx_recon_set_tuning", self.h, int(batch), int(piece))

    def tuning(self):
        """(batch, tile-pass piece length) in use (bsx_recon_tuning)."""
        b, p = C.c_int32(), C.c_int32()
        call("bsx_recon_tuning", self.h, C.byref(b), C.byref(p))
        return b.value, p.value

    def graph_launches(self) -> int:
        n = C.c_int64()
        call("bsx_recon_graph_launches", self.h, C.byref(n))
        return n.value

    def last_timing(self):
        ms, st = C.c_double(), C.c_double()
        nl = C.c_int64()
        call("bsx_recon_last_timing", self.h, C.byref(ms), C.byref(nl), C.byref(st))
        return ms.value, nl.value, st.value

    KERNELS = ("background", "rows", "fallback", "scan", "fill", "tile", "direct", "entries")

    def kernel_times(self) -> dict:
        """Per-kernel milliseconds of the previous insert (BSX_FLAG_TIMING) and the
        number of bin entries written."""
        out = (C.c_double * 8)()
        call("bsx_recon_kernel_times", self.h, out, 8)
        return dict(zip(self.KERNELS, list(out)))

    def device_bytes(self) -> int:
        b = C.c_int64()
        call("bsx_recon_device_bytes", self.h, C.byref(b))
        return b.value

    def accumulators(self, half: int = 0, stream=None) -> torch.Tensor:
        n = self.size
        out = torch.empty((5, n, n, n), device="cuda", dtype=torch.float32)
        call("bsx_recon_accumulators", self.h, half, _ptr(out), _stream(stream))
        torch.cuda.synchronize()
        return out

    def set_accumulators(self, planes: torch.Tensor, half: int = 0, count: int = 0, stream=None):
        """Replace one half-set's sums with hand-built planes (5, N, N, N)."""
        n = self.size
        if planes.dtype != torch.float32 or tuple(planes.shape) != (5, n, n, n):
            raise ValueError(f"planes must be float32 (5, {n}, {n}, {n})")
        planes = planes.contiguous()
        call("bsx_recon_set_accumulators", self.h, half, _ptr(planes), int(count), _stream(stream))
        torch.cuda.synchronize()

    def fallbacks(self) -> int:
        """Atomic-fallback flags of the binned insertion seen so far
        (bsx_recon_fallbacks: 1 chunk table, 2 chunk pool, 4 wait bound)."""
        f = C.c_int32()
        call("bsx_recon_fallbacks", self.h, C.byref(f))
        return f.value

    def finish(self, imap: int = 0, nmaps: int = 1, want_fom: bool = False,
               want_ft: bool = False, want_map: bool = True, stream=None) -> FinishResult:
        n = self.size
        shp = (1 if self.params.flags & abi.BSX_FLAG_2D else n, n, n)   # 2D: one plane
        m = torch.empty(shp, device="cuda", dtype=torch.float32) if want_map else None
        fom = torch.empty(shp, device="cuda", dtype=torch.float32) if want_fom else None
        ft = torch.empty(shp, device="cuda", dtype=torch.complex64) if want_ft else None
        st = Stats()
        call("bsx_recon_finish", self.h, imap, nmaps, _ptr(m), _ptr(fom), _ptr(ft), C.byref(st),
             _stream(stream))
        return FinishResult(m, fom, ft, st.coverage, st.friedel, st.particles, st.weight_sum)

    def stats(self, imap: int = 0, nmaps: int = 1, resolution: float | None = None, stream=None):
        """fspace_reconstruction_stats: (rows (n, 9): radius, s, resolution,
        weight, power, signal, noise, SNR, FSC; (SNR(0.5), FSC(0.3)) in A).
        Needs the FOM sums (BSX_FLAG_FOM)."""
        res = self.params.hi_res if resolution is None else resolution
        cap = self.size
        rows = np.zeros((cap, 9))
        nrow = C.c_int32()
        est = np.zeros(2)
        call("bsx_recon_stats", self.h, imap, nmaps, float(res),
             rows.ctypes.data_as(C.POINTER(C.c_double)), cap, C.byref(nrow),
             est.ctypes.data_as(C.POINTER(C.c_double)), _stream(stream))
        return rows[: nrow.value], (float(est[0]), float(est[1]))


class ReconGroup:
    """Several reconstructions (one per particle box of a mixed-size data
    set) run as one graph: clears, insertion and finish of every member
    (bsx_recon_group_run). The reference runs breconstruct once per box
    (breconstruct.cpp:204-349); a run here gives each member's full map."""

    def __init__(self, recons: list[Reconstruction]):
        self.recons = list(recons)
        hs = (C.c_void_p * len(self.recons))(*[r.h for r in self.recons])
        self.g = C.c_void_p()
        call("bsx_recon_group_create", hs, len(self.recons), C.byref(self.g))

    def close(self):
        if self.g:
            abi.lib().bsx_recon_group_destroy(self.g)
            self.g = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, imgs: list[torch.Tensor], parts: list[np.ndarray],
            maps: list[torch.Tensor] | None = None, stream=None):
        """Reconstruct every member from its particles. Returns (maps, stats),
        stats a list of FinishResult without map copies. Pass the same `maps`
        (and image tensors) again for the run to replay the captured graph."""
        m = len(self.recons)
        if len(imgs) != m or len(parts) != m:
            raise ValueError(f"{m} image and particle sets expected")
        parts = [np.ascontiguousarray(p, dtype=PARTICLE_DTYPE) for p in parts]
        if maps is None:
            maps = []
            for r in self.recons:
                n = r.size
                nz = 1 if r.params.flags & abi.BSX_FLAG_2D else n
                maps.append(torch.empty((nz, n, n), device="cuda", dtype=torch.float32))
        for r, im, p, mp in zip(self.recons, imgs, parts, maps):
            b = r.params.box
            if (len(p) == 0 or im.dtype != torch.float32 or tuple(im.shape[1:]) != (b, b)
                    or im.shape[0] < len(p)):
                raise ValueError(f"images must be float32 (n, {b}, {b}) with n >= 1 particles")
            nz = 1 if r.params.flags & abi.BSX_FLAG_2D else r.size
            if tuple(mp.shape) != (nz, r.size, r.size) or mp.dtype != torch.float32:
                raise ValueError("map buffer of the member's size expected")
        ip = (C.c_void_p * m)(*[_ptr(t).value for t in imgs])
        pp = (C.c_void_p * m)(*[p.ctypes.data for p in parts])
        cnt = (C.c_int64 * m)(*[len(p) for p in parts])
        mp = (C.c_void_p * m)(*[_ptr(t).value for t in maps])
        st = (Stats * m)()
        call("bsx_recon_group_run", self.g, ip, pp, cnt, mp, C.cast(st, C.c_void_p),
             _stream(stream))
        res = [FinishResult(mm, None, None, s.coverage, s.friedel, s.particles, s.weight_sum)
               for mm, s in zip(maps, st)]
        return maps, res

    def graph_launches(self) -> int:
        n = C.c_int64()
        call("bsx_recon_group_graph_launches", self.g, C.byref(n))
        return n.value


def particle_reconstruct(imgs: torch.Tensor, parts: np.ndarray, size: int, sampling: float,
                         hi_res: float = 0.0, interp_type: int = 0, ctf_action: int = 0,
                         wiener: float = 0.0, flags: int = 0, pad_factor: int = 2,
                         nhalves: int = 1, batch: int = 0) -> Reconstruction:
    """particle_reconstruct (mg_reconstruct.cpp:100-264): all particles of a
    list into one set of accumulators, on the device."""
    box = imgs.shape[-1]
    ft = part_ft_size(max(box, size), 1, pad_factor)
    rec = Reconstruction(size, box, ft, sampling, hi_res=hi_res, interp=interp_type,
                         ctf_action=ctf_action, wiener=wiener, flags=flags, nhalves=nhalves,
                         batch=batch)
    rec.insert(imgs, parts)
    return rec


def rescale_to_avg_std(vol: torch.Tensor, avg: float, std: float, stream=None) -> torch.Tensor:
    """Bimage::rescale_to_avg_std (Bimage_rescale.cpp:163) of a float32
    device array, in place (breconstruct -rescale on each output map)."""
    if vol.dtype != torch.float32 or not vol.is_contiguous():
        raise ValueError("a contiguous float32 array is required")
    call("bsx_rescale_to_avg_std", _ptr(vol), vol.numel(), float(avg), float(std), _stream(stream))
    return vol


def filter_sinc(vol: torch.Tensor, origin=None, stream=None) -> torch.Tensor:
    """Bimage::filter_sinc (Bimage_filter.cpp:484-509) of a float32 (nz, ny,
    nx) device map, in place: the correction of trilinear insertion."""
    if vol.dtype != torch.float32 or not vol.is_contiguous() or vol.dim() != 3:
        raise ValueError("a contiguous float32 (nz, ny, nx) map is required")
    nz, ny, nx = vol.shape
    o = np.asarray(origin if origin is not None else (nx // 2, ny // 2, nz // 2), dtype=np.float64)
    call("bsx_filter_sinc", _ptr(vol), nx, ny, nz, o.ctypes.data_as(C.POINTER(C.c_double)),
         _stream(stream))
    return vol


def symmetry_views(sym: str, views: np.ndarray) -> np.ndarray:
    """symmetry_get_all_views (src/util/symmetry.cpp:277-319) of each
    (x, y, z, angle) row: (n, order, 4), the view itself first."""
    v = np.ascontiguousarray(views, dtype=np.float64).reshape(-1, 4)
    cnt = C.c_int64()
    call("bsx_symmetry_views", sym.encode(), v.ctypes.data_as(C.POINTER(C.c_double)), len(v),
         None, 0, C.byref(cnt))
    out = np.zeros((max(cnt.value, 0), 4))
    call("bsx_symmetry_views", sym.encode(), v.ctypes.data_as(C.POINTER(C.c_double)), len(v),
         out.ctypes.data_as(C.POINTER(C.c_double)), len(out), C.byref(cnt))
    return out.reshape(len(v), -1, 4)


def symmetrize(vol: torch.Tensor, sym: str, origin=None, stream=None) -> torch.Tensor:
    """Bimage::symmetrize (src/img/Bimage_symmetry.cpp:31-100) of a real
    N^3 map on the device, about `origin` (voxels, x y z; default N/2):
    breconstruct's symmetry mode 1, applied after the inverse transform."""
    n = int(vol.shape[-1])
    if vol.dtype != torch.float32:
        raise ValueError("a float32 map is required")
    if tuple(vol.shape) == (n, n, n) and n > 1:
        fn, dflt = "bsx_symmetrize", (n // 2,) * 3
    elif tuple(vol.shape) in ((n, n), (1, n, n)):        # a 2D image: cyclic groups
        fn, dflt = "bsx_symmetrize_2d", (n // 2,) * 2
    else:
        raise ValueError("a cubic map or a square image is required")
    vol = vol.contiguous()
    out = torch.empty_like(vol)
    o = np.asarray(origin if origin is not None else dflt, dtype=np.float64)
    call(fn, _ptr(vol), _ptr(out), n, o.ctypes.data_as(C.POINTER(C.c_double)),
         sym.encode(), _stream(stream))
    return out


def random_symmetric_views(sym: str, views: np.ndarray, rng: np.random.Generator) -> np.ndarray:
    """random_symmetric_view (src/util/symmetry.cpp:816-833) for each row:
    one of the order(sym) related views, index floor(order * u * 0.99999)
    with u uniform in [0, 1) (breconstruct's symmetry mode 2). The
    reference draws u from random(); this draws it from `rng`."""
    allv = symmetry_views(sym, views)
    n, order = allv.shape[:2]
    idx = (order * rng.random(n) * 0.99999).astype(np.int64)
    return allv[np.arange(n), idx]


def img_reconstruction_sum_weigh(rec: Reconstruction, imap: int = 0, nmaps: int = 1,
                                 **kw) -> FinishResult:
    """img_reconstruction_sum_weigh (mg_reconstruct.cpp:266-361) followed by
    fft_back and complex_to_real (breconstruct.cpp:340)."""
    return rec.finish(imap, nmaps, **kw)


def fft(data: torch.Tensor, dims: int, direction: int = -1, norm: int = 1, stream=None):
    """In-place C2C transform of the last `dims` axes of a complex64 tensor
    (fftw with the sign and norm_flag semantics of Bimage::fft)."""
    if data.dtype != torch.complex64:
        raise ValueError("complex64 tensor required")
    shp = list(data.shape)
    sizes = shp[-dims:][::-1] + [1] * (3 - dims)
    batch = int(np.prod(shp[:-dims])) if len(shp) > dims else 1
    call("bsx_fft", _ptr(data), sizes[0], sizes[1], sizes[2], batch, direction, norm,
         _stream(stream))
    return data


def map_transform(vol: torch.Tensor, origin=None, stream=None) -> torch.Tensor:
    """Forward FFT (x 1/sqrt(N^3)) and phase shift to the map origin — the
    prologue of Bimage::project (Bimage_project.cpp:373-375)."""
    n = vol.shape[0]
    if origin is None:
        origin = (n // 2, n // 2, n // 2)
    o = (C.c_double * 3)(*origin)
    out = torch.empty((n, n, n), device=vol.device, dtype=torch.complex64)
    vol = vol.contiguous()          # held until the call has returned
    call("bsx_map_transform", _ptr(vol), n, o, _ptr(out), _stream(stream))
    return out


def central_sections(ft: torch.Tensor, mats: np.ndarray, sampling: float, resolution: float,
                     kernel_width: int = 8, kernel_power: int = 2, center: bool = True,
                     stream=None) -> torch.Tensor:
    """Bimage::central_section for each rotation matrix (Bimage_project.cpp:288)
    with the centre shift of Bimage::project."""
    n = ft.shape[0]
    mats = np.ascontiguousarray(np.asarray(mats, dtype=np.float64).reshape(-1, 9))
    nv = mats.shape[0]
    out = torch.empty((nv, n, n), device=ft.device, dtype=torch.complex64)
    call("bsx_extract", _ptr(ft), n, float(sampling), mats.ctypes.data_as(C.POINTER(C.c_double)),
         nv, float(resolution), int(kernel_width), int(kernel_power), int(center), _ptr(out),
         _stream(stream))
    return out


def project(ft: torch.Tensor, views: np.ndarray, sampling: float, resolution: float,
            kernel_width: int = 8, kernel_power: int = 2) -> torch.Tensor:
    """Bimage::project over a list of views (x, y, z, angle)."""
    return central_sections(ft, view_to_matrix(views), sampling, resolution, kernel_width,
                            kernel_power, True)


def fsc(a: torch.Tensor, b: torch.Tensor, sampling: float, hi_res: float) -> np.ndarray:
    """Bimage::fsc (Bimage_resolution.cpp:330-406) of two complex volumes."""
    n = a.shape[0]
    cap = C.c_int(4 * n)
    out = np.zeros(4 * n)
    call("bsx_fsc", _ptr(a), _ptr(b), n, float(sampling), float(hi_res),
         out.ctypes.data_as(C.POINTER(C.c_double)), C.byref(cap), _stream(None))
    return out[: cap.value]


def img_recip_space_fom(parts: torch.Tensor, refs: torch.Tensor, pairs: np.ndarray,
                        sampling: float, hi_res: float, lo_res: float = 0.0) -> np.ndarray:
    """img_recip_space_fom, fom_type 0 (mg_refine.cpp:660-781), for
    (particle, reference) index pairs."""
    pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
    out = np.zeros(len(pairs))
    call("bsx_score_frc", _ptr(parts), parts.shape[0], _ptr(refs), refs.shape[0],
         parts.shape[-1], float(sampling),
         float(hi_res), float(lo_res), pairs.ctypes.data_as(C.POINTER(C.c_int32)), len(pairs),
         out.ctypes.data_as(C.POINTER(C.c_double)), _stream(None))
    return out

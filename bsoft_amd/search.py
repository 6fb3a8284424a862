"""Orientation search by projection matching on the GPU (host side of
bsx_search_* in include/bsx.h).

The reference's exhaustive search compares each particle transform with
central sections of the map over an asymmetric-unit view grid expanded
in-plane, scoring by the ring FRC of img_recip_space_fom
(src/mg/mg_refine.cpp:660-781; driver part_determine_orientation,
src/mg/mg_orient.cpp:545-798; grid asymmetric_unit_views,
src/util/symmetry.cpp:383-548, and view_list_expand_angles,
src/util/View.cpp:519-556). Here:

* ``Search`` wraps one device handle: the view library of a map, the staged
  particles, the BF16 MFMA GEMM with per-particle top-k, the FP32 FRC
  rescoring and the on-the-fly local refinement;
* ``views_c1`` / ``view_matrices`` give the view grid and its rotations;
* ``exhaustive_search`` is the plain pipeline: GEMM over every view -> FP32
  FRC of the top-k -> local refinement.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import abi
from .abi import SearchParams, call, lib
from .recon import _ptr as _dev_ptr, _stream
from .synth import PARTICLE_DTYPE, particles_ctypes


def views_c1(theta_step_deg: float, phi_step_deg: float | None = None, flag: int = 1,
             angle_step_deg: float = 0.0) -> np.ndarray:
    """asymmetric_unit_views(C1) expanded by in-plane angles q * angle_step in
    [0, 360): an (n, 4) array of (x, y, z, angle [rad])."""
    phi = theta_step_deg if phi_step_deg is None else phi_step_deg
    args = (math.radians(theta_step_deg), math.radians(phi), int(flag), math.radians(angle_step_deg))
    n = C.c_int64()
    call("bsx_views_c1", *args, None, 0, C.byref(n))
    out = np.zeros((n.value, 4))
    call("bsx_views_c1", *args, out.ctypes.data_as(C.POINTER(C.c_double)), n.value, C.byref(n))
    return out


def view_matrices(views: np.ndarray) -> np.ndarray:
    """View::matrix of each (x, y, z, angle) row: (n, 3, 3) FP64."""
    v = np.ascontiguousarray(views, dtype=np.float64).reshape(-1, 4)
    m = np.zeros((len(v), 9))
    call("bsx_view_matrices", v.ctypes.data_as(C.POINTER(C.c_double)), len(v),
         m.ctypes.data_as(C.POINTER(C.c_double)))
    return m.reshape(-1, 3, 3)


def matrix_views(mats) -> np.ndarray:
    """Views (x, y, z, angle) of rotation matrices (..., 3, 3): the inverse of
    view_matrices (Matrix3::quaternion, then View(Quaternion&))."""
    m = np.ascontiguousarray(np.asarray(mats, dtype=np.float64).reshape(-1, 9))
    v = np.zeros((len(m), 4))
    call("bsx_matrix_views", m.ctypes.data_as(C.POINTER(C.c_double)), len(m),
         v.ctypes.data_as(C.POINTER(C.c_double)))
    return v


def monte_draws(n: int, max_iter: int, seed: int = 0) -> np.ndarray:
    """Random values for Search.fit_monte, (n, max_iter, 5) float64 per
    iteration: random_gaussian(0, 1) for the defocus step and for the shift
    amplitude, the normalised (x, y) of two more (a unit direction,
    vector3_xy_random_gaussian, src/util/random_numbers.cpp:522-533), and a
    uniform value in [0, 1) for the acceptance test."""
    rng = np.random.default_rng(seed)
    d = np.empty((n, max_iter, 5), dtype=np.float64)
    d[..., 0] = rng.standard_normal((n, max_iter))
    d[..., 1] = rng.standard_normal((n, max_iter))
    xy = rng.standard_normal((n, max_iter, 2))
    r = np.hypot(xy[..., 0], xy[..., 1])
    r[r == 0] = 1.0
    d[..., 2] = xy[..., 0] / r
    d[..., 3] = xy[..., 1] / r
    d[..., 4] = rng.random((n, max_iter))
    return d


def monte_view_draws(n: int, max_iter: int, seed: int = 0) -> np.ndarray:
    """Random values for Search.refine_monte, (n, max_iter, 6) float64 per
    iteration: random_gaussian(0, 1) for the view step amplitude, a unit
    vector (x, y, z) (vector3_random_gaussian, src/util/random_numbers.cpp:
    502-510), a uniform value for the angle step and one for the acceptance
    test, both in [0, 1)."""
    rng = np.random.default_rng(seed)
    d = np.empty((n, max_iter, 6), dtype=np.float64)
    d[..., 0] = rng.standard_normal((n, max_iter))
    xyz = rng.standard_normal((n, max_iter, 3))
    r = np.sqrt((xyz * xyz).sum(-1))
    r[r == 0] = 1.0
    d[..., 1:4] = xyz / r[..., None]
    d[..., 4] = rng.random((n, max_iter))
    d[..., 5] = rng.random((n, max_iter))
    return d


def views_normalize(v: np.ndarray) -> np.ndarray:
    """View::normalize (include/View.h:110-116) of (n, 4) rows: the view
    vector to unit length ((0, 0, 1) when below 1e-30, components below 1e-30
    set to 0) and the angle into [-pi, pi] by angle_set_negPI_to_PI
    (src/util/utilities.cpp:606-615: n = trunc(a / pi), a -= (n + n % 2) pi)."""
    v = np.array(v, dtype=np.float64).reshape(-1, 4)
    s = np.sqrt(v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2])
    small = s < 1e-30
    v[~small, :3] = v[~small, :3] / s[~small, None]
    v[small, :3] = (0.0, 0.0, 1.0)
    xyz = v[:, :3]
    xyz[np.abs(xyz) < 1e-30] = 0.0
    a = v[:, 3]
    fin = np.isfinite(a)
    q = np.trunc(a[fin] / np.pi)
    a[fin] = a[fin] - (q + np.fmod(q, 2)) * np.pi
    a[~fin] = 0.0
    a[np.abs(a) < 1e-30] = 0.0
    return v


def rotation_angle_deg(a, b) -> np.ndarray:
    """Geodesic angle (degrees) between rotation matrices a and b (..., 3, 3)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    tr = np.einsum("...ij,...ij->...", a, b)
    return np.degrees(np.arccos(np.clip((tr - 1) / 2, -1.0, 1.0)))


class Search:
    """One search handle: a view library of one map and a staged particle set."""

    def __init__(self, box: int, radius: int, sampling: float, *, kernel_width: int = 8,
                 kernel_power: int = 2, topk: int = 8, ctf: bool = True, mem_cap: int = 0,
                 device: str | torch.device = "cuda"):
        self.device = torch.device(device)
        self.params = SearchParams(box=box, radius=radius, sampling=sampling,
                                   kernel_width=kernel_width, kernel_power=kernel_power,
                                   topk=topk, flags=abi.BSX_SEARCH_CTF if ctf else 0,
                                   mem_cap=mem_cap)
        self.h = C.c_void_p()
        call("bsx_search_create", C.byref(self.h), C.byref(self.params))
        nd, kd = C.c_int32(), C.c_int32()
        call("bsx_search_info", self.h, C.byref(nd), C.byref(kd), None, None)
        self.npoints, self.kdim = nd.value, kd.value
        self.box, self.radius, self.topk = box, radius, topk
        self.nviews = 0
        self.nparts = 0
        self._defocus = np.zeros(0)   # staged defocus per particle (refine_monte)
        self._ft = None

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            try:
                lib().bsx_search_destroy(h)
            except Exception:
                pass
            self.h = None

    def points(self):
        """Disk points (Nd, 2) as (h, k) and ring starts (R + 1)."""
        hk = np.zeros((self.npoints, 2), dtype=np.int32)
        rs = np.zeros(self.radius + 1, dtype=np.int32)
        call("bsx_search_points", self.h, hk.ctypes.data_as(C.POINTER(C.c_int32)),
             rs.ctypes.data_as(C.POINTER(C.c_int32)))
        return hk, rs

    def set_reference(self, ft: torch.Tensor, mats: np.ndarray) -> None:
        """ft: the N^3 complex64 map transform (bsx_map_transform); mats:
        (nviews, 3, 3) rotations. The transform is kept for local refinement."""
        assert ft.is_cuda and ft.dtype == torch.complex64 and ft.is_contiguous()
        assert ft.shape == (self.box,) * 3
        m = np.ascontiguousarray(mats, dtype=np.float64).reshape(-1, 9)
        call("bsx_search_set_reference", self.h, _dev_ptr(ft),
             m.ctypes.data_as(C.POINTER(C.c_double)), len(m), _stream())
        self.nviews = len(m)
        self._ft = ft

    def stage(self, images: torch.Tensor, parts) -> None:
        """images: (n, N, N) float32 on the device; parts: PARTICLE_DTYPE records."""
        arr = np.ascontiguousarray(parts, dtype=PARTICLE_DTYPE)
        n = len(arr)
        if n and (images.dtype != torch.float32 or tuple(images.shape) != (n, self.box, self.box)):
            raise ValueError(f"images must be float32 ({n}, {self.box}, {self.box})")
        call("bsx_search_stage", self.h, _dev_ptr(images) if n else None,
             particles_ctypes(arr) if n else None, n, _stream())
        self.nparts = n
        self._defocus = arr["def_avg"].astype(np.float64) if n else np.zeros(0)

    def exhaustive(self):
        """BF16 GEMM against every view: (cand, score), each (n, topk)."""
        cand = torch.empty((self.nparts, self.topk), dtype=torch.int32, device=self.device)
        score = torch.empty((self.nparts, self.topk), dtype=torch.float32, device=self.device)
        call("bsx_search_exhaustive", self.h, _dev_ptr(cand), _dev_ptr(score), _stream())
        return cand, score

    def exhaustive_rows(self, rows: torch.Tensor):
        """BF16 GEMM for a subset of the staged particles (rows: indices)."""
        rows = rows.to(device=self.device, dtype=torch.int32).contiguous()
        m = len(rows)
        cand = torch.empty((m, self.topk), dtype=torch.int32, device=self.device)
        score = torch.empty((m, self.topk), dtype=torch.float32, device=self.device)
        if m:
            call("bsx_search_exhaustive_rows", self.h, _dev_ptr(rows), m, _dev_ptr(cand),
                 _dev_ptr(score), _stream())
        return cand, score

    def set_classes(self, members: np.ndarray) -> None:
        """members: (nclasses, maxm) library view indices per class, -1 padded."""
        m = np.ascontiguousarray(members, dtype=np.int32)
        call("bsx_search_set_classes", self.h, m.ctypes.data_as(C.POINTER(C.c_int32)), m.shape[0],
             m.shape[1])

    def proposed(self, cls: torch.Tensor):
        """BF16 GEMM restricted to the members of each particle's proposed
        classes (cls: (n, k)): (cand, score), each (n, topk)."""
        cls = cls.to(device=self.device, dtype=torch.int32).contiguous()
        cand = torch.empty((self.nparts, self.topk), dtype=torch.int32, device=self.device)
        score = torch.empty((self.nparts, self.topk), dtype=torch.float32, device=self.device)
        call("bsx_search_proposed", self.h, _dev_ptr(cls) if self.nparts else None, cls.shape[1],
             _dev_ptr(cand), _dev_ptr(score), _stream())
        return cand, score

    def cnn_input(self, size: int) -> torch.Tensor:
        """(n, size, size) low-pass, CTF-flipped, normalised images of the staged particles."""
        out = torch.empty((self.nparts, size, size), dtype=torch.float32, device=self.device)
        call("bsx_search_cnn_input", self.h, int(size), _dev_ptr(out) if self.nparts else None,
             _stream())
        return out

    def rescore(self, cand: torch.Tensor):
        """FP32 FRC of candidate views (n, c): (frc (n, c), best view (n), best frc (n))."""
        cand = cand.to(device=self.device, dtype=torch.int32).contiguous()
        n, c = cand.shape
        assert n == self.nparts
        frc = torch.empty((n, c), dtype=torch.float32, device=self.device)
        best = torch.empty(n, dtype=torch.int32, device=self.device)
        bf = torch.empty(n, dtype=torch.float32, device=self.device)
        call("bsx_search_rescore", self.h, _dev_ptr(cand), c, _dev_ptr(frc), _dev_ptr(best),
             _dev_ptr(bf), _stream())
        return frc, best, bf

    def score_mats(self, mats: torch.Tensor, ft: torch.Tensor | None = None):
        """FP32 FRC at arbitrary rotations: mats (n, c, 3, 3) float64 -> (n, c)."""
        ft = self._ft if ft is None else ft
        mats = mats.to(device=self.device, dtype=torch.float64).contiguous()
        n, c = mats.shape[:2]
        assert n == self.nparts
        frc = torch.empty((n, c), dtype=torch.float32, device=self.device)
        call("bsx_search_score_mats", self.h, _dev_ptr(ft), _dev_ptr(mats), c, _dev_ptr(frc),
             _stream())
        return frc

    def view_mats(self, views: torch.Tensor) -> torch.Tensor:
        views = views.to(device=self.device, dtype=torch.int32).contiguous()
        out = torch.empty((len(views), 3, 3), dtype=torch.float64, device=self.device)
        call("bsx_search_view_mats", self.h, _dev_ptr(views), len(views), _dev_ptr(out), _stream())
        return out

    def refine(self, mats: torch.Tensor, steps_deg=(2.0, 1.0), ft: torch.Tensor | None = None):
        """Local 3x3x3 grid refinement at each step; returns (mats, frc)."""
        ft = self._ft if ft is None else ft
        mats = mats.to(device=self.device, dtype=torch.float64).contiguous().clone()
        frc = torch.empty(self.nparts, dtype=torch.float32, device=self.device)
        st = np.radians(np.asarray(steps_deg, dtype=np.float64))
        call("bsx_search_refine", self.h, _dev_ptr(ft), _dev_ptr(mats), _dev_ptr(frc),
             st.ctypes.data_as(C.POINTER(C.c_double)), len(st), _stream())
        return mats, frc

    def refine_views(self, mats: torch.Tensor, step_deg: float = 2.0, accuracy_deg: float = 0.25,
                     max_iter: int = 64, ft: torch.Tensor | None = None):
        """img_refine_grid's view loop (mg_refine.cpp:1116-1140) over the
        views_for_refinement grid (View.cpp:371-436), the origin kept:
        returns (mats (n, 3, 3), frc (n,), passes)."""
        ft = self._ft if ft is None else ft
        mats = mats.to(device=self.device, dtype=torch.float64).contiguous().clone()
        frc = torch.empty(self.nparts, dtype=torch.float32, device=self.device)
        it = C.c_int32()
        call("bsx_search_refine_views", self.h, _dev_ptr(ft), _dev_ptr(mats), _dev_ptr(frc),
             math.radians(step_deg), math.radians(accuracy_deg), int(max_iter), C.byref(it),
             _stream())
        return mats, frc, it.value

    def fit_origin(self, mats: torch.Tensor, step: float = 1.0, accuracy: float = 0.25,
                   def_std: float = 0.0, ft: torch.Tensor | None = None):
        """img_fit_in_recip_space_grid at rotations mats (n, 3, 3): the best
        origin (n, 2) float64 pixels, defocus (n,) float64 and FOM (n,)."""
        ft = self._ft if ft is None else ft
        mats = mats.to(device=self.device, dtype=torch.float64).contiguous()
        assert mats.shape[0] == self.nparts
        ori = torch.empty((self.nparts, 2), dtype=torch.float64, device=self.device)
        dfc = torch.empty(self.nparts, dtype=torch.float64, device=self.device)
        fom = torch.empty(self.nparts, dtype=torch.float32, device=self.device)
        call("bsx_search_fit_origin", self.h, _dev_ptr(ft), _dev_ptr(mats), float(step),
             float(accuracy), float(def_std), _dev_ptr(ori), _dev_ptr(dfc), _dev_ptr(fom),
             _stream())
        return ori, dfc, fom

    def fit_monte(self, mats: torch.Tensor, max_iter: int = 100, def_std: float = 0.0,
                  shift_std: float = 1.0, draws: torch.Tensor | None = None, seed: int = 0,
                  ft: torch.Tensor | None = None, def_start=None):
        """img_fit_in_recip_space_monte at rotations mats (n, 3, 3): the best
        origin (n, 2) float64 pixels, defocus (n,) float64 and FOM (n,).
        draws (n, max_iter, 5) float64: per iteration a standard normal
        defocus step, a standard normal shift amplitude, a unit shift
        direction (x, y) and a uniform acceptance value in [0, 1); made from
        a generator seeded with `seed` when not given (the reference draws
        from random(), seeded by the process id). def_start (n,): the defocus
        each walk starts from, the staged one when None."""
        ft = self._ft if ft is None else ft
        mats = mats.to(device=self.device, dtype=torch.float64).contiguous()
        assert mats.shape[0] == self.nparts
        if draws is None:
            draws = monte_draws(self.nparts, max_iter, seed)
        draws = torch.as_tensor(draws, dtype=torch.float64).to(self.device).contiguous()
        if tuple(draws.shape) != (self.nparts, max_iter, 5):
            raise ValueError(f"draws: shape {tuple(draws.shape)}, expected "
                             f"({self.nparts}, {max_iter}, 5)")
        d0 = None
        if def_start is not None:
            d0 = torch.as_tensor(def_start, dtype=torch.float64).to(self.device).contiguous()
            if tuple(d0.shape) != (self.nparts,):
                raise ValueError(f"def_start: shape {tuple(d0.shape)}, expected ({self.nparts},)")
        ori = torch.empty((self.nparts, 2), dtype=torch.float64, device=self.device)
        dfc = torch.empty(self.nparts, dtype=torch.float64, device=self.device)
        fom = torch.empty(self.nparts, dtype=torch.float32, device=self.device)
        call("bsx_search_fit_monte", self.h, _dev_ptr(ft), _dev_ptr(mats), int(max_iter),
             float(def_std), float(shift_std), _dev_ptr(draws), _dev_ptr(d0), _dev_ptr(ori),
             _dev_ptr(dfc), _dev_ptr(fom), _stream())
        return ori, dfc, fom

    def refine_monte(self, views, max_iter: int = 10, def_std: float = 0.0,
                     shift_std: float = 1.0, view_std: float = 0.01, max_angle: float = 0.01,
                     draws=None, inner_draws=None, seed: int = 0,
                     ft: torch.Tensor | None = None):
        """img_refine_monte (src/mg/mg_refine.cpp:979-1060) for the staged
        particles from views (n, 4), without the magnification steps:
        iteration i > 0 adds view_std g (ux, uy, uz) to the view vector and
        max_angle (2 u - 1) to the angle, normalises, then runs the Monte
        Carlo origin/defocus fit (fit_monte, max_iter steps) at that view
        from the staged origin (and the staged defocus when the CTF flag is
        set and def_std > 0 with a nonzero defocus; otherwise the defocus the
        previous fit left). The best view, its fit's origin and defocus and
        the FOM are kept; a view is kept for the next step when
        fom / prefom > uniform (prefom starts at -1), else the walk returns to
        the last kept view. draws (n, max_iter, 6) as monte_view_draws,
        inner_draws (n, max_iter, max_iter, 5) as monte_draws per iteration.
        Returns views (n, 4), origins (n, 2), defocus (n,), FOM (n,)."""
        n = self.nparts
        view = views_normalize(np.asarray(views, dtype=np.float64).reshape(n, 4))
        if draws is None:
            draws = monte_view_draws(n, max_iter, seed)
        if inner_draws is None:
            inner_draws = np.stack([monte_draws(n, max_iter, seed + 1 + i)
                                    for i in range(max_iter)], axis=1)
        draws, inner_draws = np.asarray(draws, np.float64), np.asarray(inner_draws, np.float64)
        if draws.shape != (n, max_iter, 6) or inner_draws.shape != (n, max_iter, max_iter, 5):
            raise ValueError("draws: (n, max_iter, 6) and inner_draws: (n, max_iter, max_iter, 5)")
        d0 = self._defocus.copy()
        apply = (def_std > 0) & (d0 != 0) & bool(self.params.flags & abi.BSX_SEARCH_CTF)
        cur = d0.copy()
        pre = view.copy()
        prefom, bestfom = np.full(n, -1.0), np.full(n, -1.0)
        bestview, bestori, bestdef = view.copy(), np.zeros((n, 2)), d0.copy()
        for it in range(max_iter):
            if it > 0:
                g = view_std * draws[:, it, 0]
                step = view.copy()
                step[:, :3] = view[:, :3] + draws[:, it, 1:4] * g[:, None]
                step[:, 3] = view[:, 3] + max_angle * (draws[:, it, 4] * 2 - 1)
                view = views_normalize(step)
            start = np.where(apply, d0, cur)
            mats = torch.as_tensor(view_matrices(view))
            ori, dfc, fom = self.fit_monte(mats, max_iter, def_std, shift_std,
                                           draws=inner_draws[:, it], ft=ft, def_start=start)
            ori, cur = ori.cpu().numpy(), dfc.cpu().numpy()
            fom = fom.cpu().numpy().astype(np.float64)
            better = bestfom < fom
            bestfom[better] = fom[better]
            bestview[better], bestori[better] = view[better], ori[better]
            bestdef[better & apply] = cur[better & apply]
            with np.errstate(divide="ignore", invalid="ignore"):
                keep = fom / prefom > draws[:, it, 5]
            pre[keep], prefom[keep] = view[keep], fom[keep]
            view[~keep] = pre[~keep]
        return bestview, bestori, np.where(apply, bestdef, cur), bestfom

    def particle_data(self):
        v = torch.empty((self.nparts, self.npoints), dtype=torch.complex64, device=self.device)
        c = torch.empty((self.nparts, self.npoints), dtype=torch.float32, device=self.device)
        r = torch.empty((self.nparts, self.kdim), dtype=torch.int16, device=self.device)
        call("bsx_search_particle_data", self.h, _dev_ptr(v), _dev_ptr(c), _dev_ptr(r), _stream())
        return v, c, r

    def library_data(self):
        v = torch.empty((self.nviews, self.npoints), dtype=torch.complex64, device=self.device)
        r = torch.empty((self.nviews, self.kdim), dtype=torch.int16, device=self.device)
        call("bsx_search_library_data", self.h, _dev_ptr(v), _dev_ptr(r), _stream())
        return v, r


@dataclass
class SearchResult:
    mats: torch.Tensor        # (n, 3, 3) float64 final rotations
    frc: torch.Tensor         # (n,) final FRC
    grid_view: torch.Tensor   # (n,) best library view before refinement
    grid_frc: torch.Tensor    # (n,) its FRC


def exhaustive_search(s: Search, steps_deg=(2.0, 1.0)) -> SearchResult:
    """Plain pipeline on the staged particles: BF16 GEMM over every library
    view -> FP32 FRC of the top-k -> local refinement."""
    cand, _ = s.exhaustive()
    _, best, bf = s.rescore(cand)
    mats = s.view_mats(best)
    if steps_deg:
        mats, frc = s.refine(mats, steps_deg)
    else:
        frc = bf
    return SearchResult(mats=mats, frc=frc, grid_view=best, grid_frc=bf)

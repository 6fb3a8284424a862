"""ctypes mirror of include/bsx.h and the loader for libbsx.so.

The library is built in-tree (bsoft_amd/libbsx.so) by ``bsoft_amd.build``.
Loading fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbsx.so")

BSX_OK = 0
BSX_ERR_ARG = -1
BSX_ERR_DEVICE = -2
BSX_ERR_NOMEM = -3
BSX_ERR_STATE = -4
BSX_ERR_NONFINITE = -5
BSX_NOALLOC_MAX_VIEWS = 4096     # include/bsx.h

BSX_CTF_NONE = 0
BSX_CTF_FLIP = 1
BSX_CTF_APPLY = 2
BSX_CTF_WIENER = 4

BSX_FLAG_RESCALE = 1
BSX_FLAG_2D = 2
BSX_FLAG_BOOTSTRAP = 4
BSX_FLAG_EWALD = 8
BSX_FLAG_FOM = 16
BSX_FLAG_CHECK_NAN = 32
BSX_FLAG_TIMING = 64
BSX_FLAG_GRAPH = 128

BSX_INTERP_NEAREST = 0
BSX_INTERP_WEIGHTED = 1
BSX_INTERP_TRILINEAR = 2


class ReconParams(C.Structure):
    _fields_ = [
        ("size", C.c_int32),
        ("box", C.c_int32),
        ("ft_size", C.c_int32),
        ("interp", C.c_int32),
        ("sampling", C.c_double),
        ("hi_res", C.c_double),
        ("lo_res", C.c_double),
        ("ctf_action", C.c_int32),
        ("flags", C.c_int32),
        ("wiener", C.c_double),
        ("nhalves", C.c_int32),
        ("batch", C.c_int32),
        ("mem_cap", C.c_int64),
        ("method", C.c_int32),
        ("order", C.c_int32),
        ("merge", C.c_int32),
        ("reserved", C.c_int32),
    ]


class Particle(C.Structure):
    _fields_ = [
        ("view", C.c_double * 4),
        ("ori", C.c_double * 2),
        ("hdr_ori", C.c_double * 2),
        ("pixel_size", C.c_double),
        ("mag", C.c_double),
        ("volt", C.c_double),
        ("cs", C.c_double),
        ("amp_shift", C.c_double),
        ("def_avg", C.c_double),
        ("def_dev", C.c_double),
        ("ast_angle", C.c_double),
        ("sel", C.c_double),
        ("half", C.c_int32),
        ("reserved", C.c_int32),
    ]


BSX_SEARCH_CTF = 1


class SearchParams(C.Structure):
    _fields_ = [
        ("box", C.c_int32),
        ("radius", C.c_int32),
        ("sampling", C.c_double),
        ("kernel_width", C.c_int32),
        ("kernel_power", C.c_int32),
        ("topk", C.c_int32),
        ("flags", C.c_int32),
        ("mem_cap", C.c_int64),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("coverage", C.c_int64),
        ("friedel", C.c_double),
        ("particles", C.c_int64),
        ("weight_sum", C.c_double),
    ]


# Exported symbols of include/bsx.h with their argument types.
_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
SIGNATURES = {
    "bsx_last_error": ([], C.c_char_p),
    "bsx_version": ([C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
    "bsx_device_info": ([C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int64)], C.c_int),
    "bsx_part_ft_size": ([C.c_int, C.c_double, C.c_int], C.c_int),
    "bsx_recon_create": ([C.POINTER(_vp), C.POINTER(ReconParams)], C.c_int),
    "bsx_recon_destroy": ([_vp], C.c_int),
    "bsx_recon_reset": ([_vp, _vp], C.c_int),
    "bsx_recon_insert": ([_vp, _vp, C.POINTER(Particle), C.c_int64, _vp], C.c_int),
    "bsx_recon_finish": ([_vp, C.c_int, C.c_int, _vp, _vp, _vp, C.POINTER(Stats), _vp], C.c_int),
    "bsx_recon_accumulators": ([_vp, C.c_int, _vp, _vp], C.c_int),
    "bsx_recon_set_accumulators": ([_vp, C.c_int, _vp, C.c_int64, _vp], C.c_int),
    "bsx_recon_device_bytes": ([_vp, C.POINTER(C.c_int64)], C.c_int),
    "bsx_recon_last_timing": ([_vp, _dp, C.POINTER(C.c_int64), _dp], C.c_int),
    "bsx_recon_autotune": ([_vp, _vp, C.POINTER(Particle), C.c_int64, C.POINTER(C.c_int32), _dp,
                            _vp], C.c_int),
    "bsx_recon_graph_launches": ([_vp, C.POINTER(C.c_int64)], C.c_int),
    "bsx_recon_fallbacks": ([_vp, C.POINTER(C.c_int32)], C.c_int),
    "bsx_recon_group_create": ([C.POINTER(_vp), C.c_int, C.POINTER(_vp)], C.c_int),
    "bsx_recon_group_run": ([_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(C.c_int64),
                             C.POINTER(_vp), _vp, _vp], C.c_int),
    "bsx_recon_group_graph_launches": ([_vp, C.POINTER(C.c_int64)], C.c_int),
    "bsx_recon_group_destroy": ([_vp], C.c_int),
    "bsx_rescale_to_avg_std": ([_vp, C.c_int64, C.c_double, C.c_double, _vp], C.c_int),
    "bsx_filter_sinc": ([_vp, C.c_int32, C.c_int32, C.c_int32, _dp, _vp], C.c_int),
    "bsx_recon_set_symmetry": ([_vp, C.c_char_p], C.c_int),
    "bsx_recon_tuning": ([_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)], C.c_int),
    "bsx_recon_set_tuning": ([_vp, C.c_int32, C.c_int32], C.c_int),
    "bsx_recon_stats": ([_vp, C.c_int, C.c_int, C.c_double, _dp, C.c_int32,
                         C.POINTER(C.c_int32), _dp, _vp], C.c_int),
    "bsx_symmetrize": ([_vp, _vp, C.c_int32, _dp, C.c_char_p, _vp], C.c_int),
    "bsx_symmetrize_2d": ([_vp, _vp, C.c_int32, _dp, C.c_char_p, _vp], C.c_int),
    "bsx_symmetry_views": ([C.c_char_p, _dp, C.c_int64, _dp, C.c_int64, C.POINTER(C.c_int64)],
                           C.c_int),
    "bsx_recon_kernel_times": ([_vp, _dp, C.c_int], C.c_int),
    "bsx_fft": ([_vp, C.c_int, C.c_int, C.c_int, C.c_int64, C.c_int, C.c_int, _vp], C.c_int),
    "bsx_map_transform": ([_vp, C.c_int, _dp, _vp, _vp], C.c_int),
    "bsx_extract": ([_vp, C.c_int, C.c_double, _dp, C.c_int64, C.c_double, C.c_int, C.c_int,
                     C.c_int, _vp, _vp], C.c_int),
    "bsx_extract_noalloc": ([_vp, C.c_int, C.c_double, _dp, C.c_int64, C.c_double, C.c_int,
                             C.c_int, C.c_int, _vp, _vp], C.c_int),
    "bsx_fsc": ([_vp, _vp, C.c_int, C.c_double, C.c_double, _dp, C.POINTER(C.c_int), _vp], C.c_int),
    "bsx_fsc_dpr": ([_vp, _vp, C.c_int, C.c_double, C.c_double, C.c_double, _dp,
                     C.POINTER(C.c_int), _vp], C.c_int),
    "bsx_fsc_dpr_2d": ([_vp, _vp, C.c_int, C.c_double, C.c_double, C.c_double, _dp,
                        C.POINTER(C.c_int), _vp], C.c_int),
    "bsx_score_frc": ([_vp, C.c_int64, _vp, C.c_int64, C.c_int, C.c_double, C.c_double,
                       C.c_double, C.POINTER(C.c_int32), C.c_int64, _dp, _vp], C.c_int),
    "bsx_search_create": ([C.POINTER(_vp), C.POINTER(SearchParams)], C.c_int),
    "bsx_search_destroy": ([_vp], C.c_int),
    "bsx_search_info": ([_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                         C.POINTER(C.c_int64)], C.c_int),
    "bsx_search_points": ([_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)], C.c_int),
    "bsx_search_set_reference": ([_vp, _vp, _dp, C.c_int64, _vp], C.c_int),
    "bsx_search_stage": ([_vp, _vp, C.POINTER(Particle), C.c_int64, _vp], C.c_int),
    "bsx_search_exhaustive": ([_vp, _vp, _vp, _vp], C.c_int),
    "bsx_search_exhaustive_rows": ([_vp, _vp, C.c_int64, _vp, _vp, _vp], C.c_int),
    "bsx_search_set_classes": ([_vp, C.POINTER(C.c_int32), C.c_int32, C.c_int32], C.c_int),
    "bsx_search_proposed": ([_vp, _vp, C.c_int32, _vp, _vp, _vp], C.c_int),
    "bsx_search_rescore": ([_vp, _vp, C.c_int32, _vp, _vp, _vp, _vp], C.c_int),
    "bsx_search_score_mats": ([_vp, _vp, _vp, C.c_int32, _vp, _vp], C.c_int),
    "bsx_search_particle_data": ([_vp, _vp, _vp, _vp, _vp], C.c_int),
    "bsx_search_library_data": ([_vp, _vp, _vp, _vp], C.c_int),
    "bsx_search_view_mats": ([_vp, _vp, C.c_int64, _vp, _vp], C.c_int),
    "bsx_search_refine": ([_vp, _vp, _vp, _vp, _dp, C.c_int32, _vp], C.c_int),
    "bsx_search_refine_views": ([_vp, _vp, _vp, _vp, C.c_double, C.c_double, C.c_int32,
                                 C.POINTER(C.c_int32), _vp], C.c_int),
    "bsx_views_for_refinement": ([_dp, C.c_double, _dp], C.c_int),
    "bsx_search_fit_origin": ([_vp, _vp, _vp, C.c_double, C.c_double, C.c_double, _vp, _vp, _vp,
                               _vp], C.c_int),
    "bsx_search_fit_monte": ([_vp, _vp, _vp, C.c_int32, C.c_double, C.c_double, _vp, _vp, _vp,
                              _vp, _vp, _vp], C.c_int),
    "bsx_search_cnn_input": ([_vp, C.c_int32, _vp, _vp], C.c_int),
    "bsx_cnn_create": ([C.POINTER(_vp), C.c_int32, C.POINTER(C.c_int32), C.c_int32, C.c_int32,
                        C.c_uint64], C.c_int),
    "bsx_cnn_destroy": ([_vp], C.c_int),
    "bsx_cnn_nparams": ([_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int),
    "bsx_cnn_params": ([_vp, _vp, _vp, _vp], C.c_int),
    "bsx_cnn_forward": ([_vp, _vp, C.c_int32, _vp, _vp], C.c_int),
    "bsx_cnn_gradients": ([_vp, _vp, _vp, C.c_int32, _vp, _vp, _vp], C.c_int),
    "bsx_cnn_train_step": ([_vp, _vp, _vp, C.c_int32, C.c_float, C.c_float, _vp, _vp], C.c_int),
    "bsx_cnn_embed": ([_vp, _vp, C.c_int64, _vp, _vp], C.c_int),
    "bsx_cnn_train_pairs": ([_vp, _vp, C.c_int32, C.c_float, C.c_float, C.c_float, _vp, _vp],
                            C.c_int),
    "bsx_cnn_pair_gradients": ([_vp, _vp, C.c_int32, C.c_float, _vp, _vp, _vp], C.c_int),
    "bsx_cnn_match_topk": ([_vp, _vp, C.c_int64, _vp, C.c_int32, C.c_int32, C.c_float, _vp, _vp,
                            _vp], C.c_int),
    "bsx_cnn_infer_topk": ([_vp, _vp, C.c_int64, C.c_int32, _vp, _vp, _vp], C.c_int),
    "bsx_matrix_views": ([_dp, C.c_int64, _dp], C.c_int),
    "bsx_view_matrices": ([_dp, C.c_int64, _dp], C.c_int),
    "bsx_views_c1": ([C.c_double, C.c_double, C.c_int32, C.c_double, _dp, C.c_int64,
                      C.POINTER(C.c_int64)], C.c_int),
}

_lib = None


class BsxError(RuntimeError):
    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed ({code}): {msg}")
        self.code = code


def lib() -> C.CDLL:
    """Load libbsx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libbsx.so not found at {LIB_PATH}; build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'` (no CPU fallback exists)")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(fn: str, code: int) -> None:
    if code != BSX_OK:
        msg = lib().bsx_last_error()
        raise BsxError(fn, code, msg.decode() if msg else "")


def call(fn: str, *args) -> int:
    code = getattr(lib(), fn)(*args)
    check(fn, code)
    return code

"""bsx_extract_noalloc (central sections with no device allocation) against
bsx_extract, called through the C-ABI: two calls in a row, the second with
a smaller map and fewer views, must each equal bsx_extract bit for bit; the
capacity limit is an argument error."""
import ctypes as C

import numpy as np
import pytest
import torch

from bsoft_amd import abi, search, synth
from bsoft_amd.recon import map_transform

_dp = C.POINTER(C.c_double)


def _extract(name, ft, mats, sampling, res, kw, kp, center):
    n = ft.shape[0]
    mats = np.ascontiguousarray(mats.reshape(-1, 9))
    out = torch.full((len(mats), n, n), float("nan"), device=ft.device, dtype=torch.complex64)
    r = getattr(abi.lib(), name)(C.c_void_p(ft.data_ptr()), n, sampling,
                                 mats.ctypes.data_as(_dp), len(mats), res, kw, kp, center,
                                 C.c_void_p(out.data_ptr()), None)
    assert r == abi.BSX_OK, abi.lib().bsx_last_error()
    torch.cuda.synchronize()
    return out


@pytest.mark.gpu
def test_noalloc_extract_equals_extract(gpu):
    rng = np.random.default_rng(11)
    # first call: 48^3 map, 300 views, width 8; second: 32^3 map, 7 views,
    # width 3 (the scratch keeps the first call's larger contents)
    for n, nv, kw, kp, center in [(48, 300, 8, 2, 1), (32, 7, 3, 2, 0)]:
        print(f"noalloc extract: n {n}, views {nv}, width {kw}", flush=True)
        bm = synth.make_blob_map(n, 8, rng)
        ft = map_transform(bm.volume(device=gpu))
        views = search.views_c1(12.0, angle_step_deg=30.0)[:nv]
        mats = search.view_matrices(views)
        a = _extract("bsx_extract", ft, mats, 1.5, 3.2, kw, kp, center)
        b = _extract("bsx_extract_noalloc", ft, mats, 1.5, 3.2, kw, kp, center)
        assert torch.isfinite(torch.view_as_real(b)).all()
        assert torch.equal(a, b)


@pytest.mark.gpu
def test_noalloc_extract_capacity(gpu):
    L = abi.lib()
    ft = torch.zeros((16, 16, 16), device=gpu, dtype=torch.complex64)
    out = torch.zeros((1, 16, 16), device=gpu, dtype=torch.complex64)
    nv = abi.BSX_NOALLOC_MAX_VIEWS + 1
    mats = np.tile(np.eye(3).ravel(), nv)
    r = L.bsx_extract_noalloc(C.c_void_p(ft.data_ptr()), 16, 1.5, mats.ctypes.data_as(_dp), nv,
                              3.0, 3, 2, 1, C.c_void_p(out.data_ptr()), None)
    assert r == abi.BSX_ERR_ARG
    assert b"capacity" in L.bsx_last_error()

"""GPU parity of the FFT, map transform, extraction, FSC and FRC kernels
against the FP64 CPU oracle / the textbook DFT, through the C-ABI."""
import math
import sys

import numpy as np
import pytest
import torch

from bsoft_amd import recon, synth
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("shape,dims", [((16,), 1), ((512,), 1), ((24,), 1), ((384,), 1),
                                        ((7, 128), 1), ((3, 64, 256), 2), ((2, 512, 512), 2),
                                        ((96, 32), 2), ((4, 192, 96), 2), ((32, 32, 32), 3),
                                        ((64, 64, 64), 3), ((2, 48, 48, 48), 3),
                                        ((128, 256, 64), 3)])
@pytest.mark.parametrize("direction", [-1, 1])
def test_fft_vs_textbook(gpu, shape, dims, direction):
    say(f"fft shape={shape} dims={dims} dir={direction}")
    rng = np.random.default_rng(sum(shape))
    x = (rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(np.complex64)
    t = torch.from_numpy(x).cuda()
    recon.fft(t, dims, direction, norm=1)
    axes = tuple(range(len(shape) - dims, len(shape)))
    xd = x.astype(np.complex128)
    n = int(np.prod([shape[a] for a in axes]))
    ref = np.fft.fftn(xd, axes=axes) if direction < 0 else np.fft.ifftn(xd, axes=axes) * n
    ref /= math.sqrt(n)
    err = np.abs(t.cpu().numpy() - ref).max() / np.abs(ref).max()
    say(f"  rel err {err:.2e}")
    assert err < 2e-6


def test_map_transform_vs_oracle(gpu):
    rng = np.random.default_rng(1)
    bm = synth.make_blob_map(64, 20, rng)
    vol = bm.volume()
    ft = recon.map_transform(vol.cuda(), (32, 32, 32))
    ref = Oracle(64).map_transform(vol.numpy().astype(np.float64), (32, 32, 32))
    err = np.abs(ft.cpu().numpy() - ref).max() / np.abs(ref).max()
    assert err < 2e-6


@pytest.mark.parametrize("kw,kp", [(8, 2), (10, 2), (4, 1)])
def test_extract_vs_oracle(gpu, kw, kp):
    rng = np.random.default_rng(2)
    n = 64
    bm = synth.make_blob_map(n, 20, rng)
    vol = bm.volume()
    o = Oracle(64)
    ft_ref = o.map_transform(vol.numpy().astype(np.float64), (32, 32, 32))
    ft = torch.from_numpy(ft_ref.astype(np.complex64)).cuda()
    views = np.vstack([synth.uniform_views(12, rng), [[0, 0, 1, 0], [0, 0, 1, math.pi / 2]]])
    mats = synth.view_to_matrix(views)
    secs = recon.central_sections(ft, mats, 1.5, 3.0, kw, kp, center=True).cpu().numpy()
    for i, m in enumerate(mats):
        ref = o.central_section(ft_ref, 1.5, m, 3.0, kw, kp, 1)
        err = np.abs(secs[i] - ref).max() / np.abs(ref).max()
        assert err < 1e-5, (i, err)


def test_fsc_vs_oracle(gpu):
    rng = np.random.default_rng(3)
    n = 48
    a = (rng.normal(size=(n, n, n)) + 1j * rng.normal(size=(n, n, n))).astype(np.complex64)
    b = (a + 0.7 * (rng.normal(size=(n, n, n)) + 1j * rng.normal(size=(n, n, n)))).astype(np.complex64)
    f = recon.fsc(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), 1.5, 1.5)
    ref = Oracle(64).fsc(a, b, 1.5, 1.5)
    assert len(f) == len(ref)
    np.testing.assert_allclose(f, ref, atol=1e-9)


def test_frc_score_vs_oracle(gpu):
    rng = np.random.default_rng(4)
    n = 64
    P = (rng.normal(size=(5, n, n)) + 1j * rng.normal(size=(5, n, n))).astype(np.complex64)
    R = (P[[1, 0, 3, 2, 4]] + 0.5 * (rng.normal(size=(5, n, n)) + 1j * rng.normal(size=(5, n, n)))).astype(np.complex64)
    pairs = np.array([[i, j] for i in range(5) for j in range(5)], np.int32)
    s = recon.img_recip_space_fom(torch.from_numpy(P).cuda(), torch.from_numpy(R).cuda(), pairs, 1.5, 4.0, 50.0)
    o = Oracle(64)
    ref = [o.recip_space_fom(P[i], R[j], 1.5, 4.0, 50.0)[0] for i, j in pairs]
    np.testing.assert_allclose(s, ref, atol=1e-9)
    with pytest.raises(Exception):
        recon.img_recip_space_fom(torch.from_numpy(P).cuda(), torch.from_numpy(R).cuda(),
                                  np.array([[0, 5]], np.int32), 1.5, 4.0)

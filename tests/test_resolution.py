"""FSC and DPR between two real maps (bsx_fsc_dpr, Bimage::fsc_dpr) and the
resolution read-outs of bresolve.

CPU tests pin the FP64 restatement (oracle/resolution_ref.py) by identities
and by a brute-force evaluation of a one-voxel shift; the GPU test compares
the library with it."""
import math

import numpy as np
import pytest

from bsoft_amd.bresolve import cut, information_volume_radius
from oracle import resolution_ref


def _map(n, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, n, n))


def test_identities():
    a = _map(16, 1)
    s, fsc, dpr = resolution_ref.fsc_dpr(a, a, 1.5)
    assert np.allclose(fsc, 1) and np.allclose(dpr, 0)
    s, fsc, dpr = resolution_ref.fsc_dpr(a, -a, 1.5)
    assert np.allclose(fsc[1:], -1) and np.allclose(dpr[1:], 180)
    s, fsc, dpr = resolution_ref.fsc_dpr(a, 2.5 * a, 1.5)
    assert np.allclose(fsc, 1) and np.allclose(dpr, 0, atol=1e-6)
    assert s[1] == pytest.approx(1 / (16 * 1.5))


def test_shift_against_brute_force():
    """map2 = map1 rolled by one voxel along x: F2 = F1 exp(-2 pi i h / N),
    so per voxel dphi = 2 pi h / N (wrapped) and Re F1 conj F2 = |F1|^2 cos dphi;
    the shell sums are evaluated voxel by voxel."""
    n, samp = 8, 2.0
    a = _map(n, 2)
    b = np.roll(a, 1, axis=2)
    s, fsc, dpr = resolution_ref.fsc_dpr(a, b, samp)
    F = np.fft.fftn(a)
    hi = resolution_ref.check_resolution(0, n, samp)
    maxrad = int(2 + n * samp / hi)
    num = np.zeros(maxrad); den = np.zeros(maxrad); dp = np.zeros(maxrad); ws = np.zeros(maxrad)
    for z in range(n):
        for y in range(n):
            for x in range(n // 2):
                h = x - n if x > (n - 1) // 2 else x
                k = y - n if y > (n - 1) // 2 else y
                l = z - n if z > (n - 1) // 2 else z
                r = math.sqrt(h * h + k * k + l * l)
                i = int(r)
                if i + 1 >= maxrad:
                    continue
                f = r - i
                I = abs(F[z, y, x]) ** 2
                ph = math.remainder(2 * math.pi * h / n, 2 * math.pi)
                for j, wt in ((i, 1 - f), (i + 1, f)):
                    num[j] += wt * I * math.cos(ph)
                    den[j] += wt * I
                    dp[j] += wt * 2 * I * ph * ph
                    ws[j] += wt * 2 * I
    for j in range(1, maxrad):
        if den[j] > 0:
            assert fsc[j] == pytest.approx(num[j] / den[j], abs=1e-9)
            assert dpr[j] == pytest.approx(math.degrees(math.sqrt(dp[j] / ws[j])), abs=1e-7)


def test_cut_and_ivr():
    s = np.arange(10) / 20.0
    y = 1 - np.arange(10) / 9.0                     # falls linearly from 1 to 0
    assert cut(s, y, 0.5, -1) == pytest.approx(s[4] + (s[5] - s[4]) * (0.5 - y[4]) / (y[5] - y[4]))
    assert cut(s, np.ones(10), 0.5, -1) == pytest.approx(s[-1])   # never crosses
    # a perfect FSC: the information volume radius is the shell sum's closed form
    ivr = information_volume_radius(s, np.ones(10))
    tot = sum(4 * math.pi * i * i for i in range(10))
    assert ivr == pytest.approx((4 * math.pi / (3 * tot)) ** (1 / 3) / s[1])


@pytest.mark.gpu
def test_gpu_fsc_dpr_matches_oracle():
    import torch
    from bsoft_amd.bresolve import fsc_dpr
    n = 32
    a = _map(n, 3).astype(np.float32)
    b = (a + 0.7 * _map(n, 4)).astype(np.float32)
    for ratio, hi in ((1.0, 0.0), (2.0, 6.0)):
        s, fsc, dpr = fsc_dpr(torch.as_tensor(a).cuda(), torch.as_tensor(b).cuda(), 1.5, hi, ratio)
        rs, rf, rd = resolution_ref.fsc_dpr(a, b, 1.5, hi, ratio)
        assert len(s) == len(rs)
        np.testing.assert_allclose(s, rs, rtol=1e-12)
        np.testing.assert_allclose(fsc, rf, atol=2e-5)
        np.testing.assert_allclose(dpr, rd, atol=2e-3)


@pytest.mark.gpu
def test_bresolve_cli(tmp_path, capsys):
    from bsoft_amd import bresolve, formats
    n = 32
    a = _map(n, 5).astype(np.float32)
    b = (a + 0.5 * _map(n, 6)).astype(np.float32)
    formats.write_mrc(str(tmp_path / "a.mrc"), a, (1.5,) * 3, (16, 16, 16))
    formats.write_mrc(str(tmp_path / "b.mrc"), b, (1.5,) * 3, (16, 16, 16))
    tsv = str(tmp_path / "c.tsv")
    assert bresolve.main(["-fsccut", "0.5,0.143", "-dprcut", "45", "-tsv", tsv,
                          str(tmp_path / "a.mrc"), str(tmp_path / "b.mrc")]) == 0
    out = capsys.readouterr().out
    assert "FSC(0.5):" in out and "DPR(45):" in out and "Information volume radius" in out
    c = np.loadtxt(tsv)
    rs, rf, rd = resolution_ref.fsc_dpr(a, b, 1.5)
    np.testing.assert_allclose(c[:, 1], rf, atol=2e-5)

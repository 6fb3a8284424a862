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
def test_gpu_fsc_dpr_non_contiguous_inputs():
    """Two strided views (transposed maps): each is copied to its own
    contiguous buffer for the call, so the curves are those of the two
    transposed maps, not of one map with itself."""
    import torch
    from bsoft_amd.bresolve import fsc_dpr
    n = 24
    a = _map(n, 7).astype(np.float32)
    b = (a + 0.8 * _map(n, 8)).astype(np.float32)
    ta = torch.as_tensor(a).cuda().permute(2, 1, 0)
    tb = torch.as_tensor(b).cuda().permute(2, 1, 0)
    assert not ta.is_contiguous() and not tb.is_contiguous()
    s, fsc, dpr = fsc_dpr(ta, tb, 1.5)
    rs, rf, rd = resolution_ref.fsc_dpr(np.ascontiguousarray(a.transpose(2, 1, 0)),
                                        np.ascontiguousarray(b.transpose(2, 1, 0)), 1.5)
    np.testing.assert_allclose(fsc, rf, atol=2e-5)
    np.testing.assert_allclose(dpr, rd, atol=2e-3)
    assert fsc[1:].min() < 0.99


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


# -------------------------------------- reconstruction shell statistics
def _identical_acc(n, K, seed):
    """Accumulators of K identical unit-weight contributions S per voxel."""
    rng = np.random.default_rng(seed)
    S = rng.normal(size=(n, n, n)) + 1j * rng.normal(size=(n, n, n))
    if np.isscalar(K):
        K = np.full((n, n, n), float(K))
    return np.stack([K * S.real, K * S.imag, K, K * np.abs(S) ** 2, K]), S


def test_stats_closed_form_identical_contributions():
    """K >= 3 identical contributions: F = S, power = K |S|^2 / (K - 1), weight K,
    weight2 = K - 1, so noise = <|S|^2> / (K - 1)^2, signal =
    K (K - 2) <|S|^2> / (K - 1)^2 and SNR = K (K - 2) in every shell."""
    n, K, samp = 16, 5, 2.0
    acc, S = _identical_acc(n, K, 3)
    rows, (r_snr, r_fsc) = resolution_ref.reconstruction_stats(acc, samp, samp)
    assert len(rows) == n // 2 - 1                       # shells 1 .. N/2 - 1
    np.testing.assert_allclose(rows[:, 0], np.arange(1, n // 2))
    np.testing.assert_allclose(rows[:, 1], rows[:, 0] / (n * samp))
    np.testing.assert_allclose(rows[:, 3], K)            # radial average of a constant
    np.testing.assert_allclose(rows[:, 7], K * (K - 2), rtol=1e-5)
    np.testing.assert_allclose(rows[:, 8], K * (K - 2) / (K * (K - 2) + 1), rtol=1e-6)
    np.testing.assert_allclose(rows[:, 6] * (K - 1) ** 2, rows[:, 4], rtol=1e-5)
    assert r_snr == 0 and r_fsc == 0                     # never below the cuts


def test_stats_resolution_limit_and_crossing():
    """Three unit-weight contributions S + n_k with sum n_k = 0 and
    sum |n_k|^2 = a |S|^2: power (3 + a) |S|^2 / 2, so SNR = (3 - a) / (1 + a):
    3 inside radius 4.5 (a = 0), 1/7 outside (a = 2.5); the FSC(0.3) and
    SNR(0.5) crossings fall between shells 4 and 6. A resolution limit
    shortens the table to rs / res - 1 rows."""
    n, samp = 16, 1.5
    idx = np.arange(n)
    d = np.where(idx > n // 2, idx - n, idx)
    z, y, x = np.meshgrid(d, d, d, indexing="ij")
    a = np.where(np.sqrt(x * x + y * y + z * z) < 4.5, 0.0, 2.5)
    acc, S = _identical_acc(n, 3, 4)
    acc[3] = (3 + a) * np.abs(S) ** 2
    rows, _ = resolution_ref.reconstruction_stats(acc, samp, n * samp / 6)
    assert len(rows) == 5
    rows, (r_snr, r_fsc) = resolution_ref.reconstruction_stats(acc, samp, samp)
    np.testing.assert_allclose(rows[:3, 7], 3, rtol=1e-5)
    np.testing.assert_allclose(rows[5:, 7], 1 / 7, rtol=1e-5)
    rs = n * samp
    assert rs / 6 < r_fsc < rs / 4 and rs / 6 < r_snr < rs / 4


@pytest.mark.gpu
def test_gpu_recon_stats_matches_oracle():
    """bsx_recon_stats on a reconstruction with the FOM sums against the
    restatement evaluated on the same accumulators (both halves)."""
    import torch
    from bsoft_amd import abi, synth
    from bsoft_amd.recon import Reconstruction
    n = 32
    rng = np.random.default_rng(9)
    bm = synth.make_blob_map(n, 12, rng)
    ps = synth.make_particles(bm, 300, n, rng, snr=0.5, seed=9)
    parts = ps.parts.copy()
    parts["half"] = np.arange(len(parts)) % 2
    rec = Reconstruction(n, n, 64, 1.5, ctf_action=abi.BSX_CTF_FLIP, flags=abi.BSX_FLAG_FOM,
                         nhalves=2)
    rec.insert(ps.images.cuda().contiguous(), parts)
    acc = (rec.accumulators(0).cpu().numpy().astype(np.float64)
           + rec.accumulators(1).cpu().numpy().astype(np.float64))
    for res in (1.5, 4.0):
        print(f"recon stats n={n} resolution={res}", flush=True)
        rows, est = rec.stats(0, 3, res)
        ref, rest = resolution_ref.reconstruction_stats(acc, 1.5, res)
        assert rows.shape == ref.shape and len(rows) > 3
        for c in range(9):
            np.testing.assert_allclose(rows[:, c], ref[:, c], rtol=1e-6,
                                       atol=1e-9 * max(1e-30, np.abs(ref[:, c]).max()))
        np.testing.assert_allclose(est, rest, rtol=1e-6)


# ------------------------------------------------------------- 2D images
def test_2d_identities_and_shift():
    """Square images: FSC(a, a) = 1, FSC(a, -a) = -1 with DPR 180; a
    one-pixel roll along x against a brute-force evaluation of the ring sums
    (resolution checked over two axes)."""
    n, samp = 16, 1.5
    rng = np.random.default_rng(21)
    a = rng.normal(size=(n, n))
    hi = resolution_ref.check_resolution(0, n, samp, 2)
    assert hi == pytest.approx(2 * samp / math.sqrt(2))
    maxrad = int(2 + n * samp / hi)
    # the last shell lies beyond the half-plane's corner (|(7, 8)| < 11): no
    # data, FSC 0 there as in the reference
    s, fsc, dpr = resolution_ref.fsc_dpr(a, a, samp)
    assert len(s) == maxrad and fsc[-1] == 0
    assert np.allclose(fsc[:-1], 1) and np.allclose(dpr, 0)
    s, fsc, dpr = resolution_ref.fsc_dpr(a, -a, samp)
    assert np.allclose(fsc[1:-1], -1) and np.allclose(dpr[1:-1], 180)
    b = np.roll(a, 1, axis=1)
    s, fsc, dpr = resolution_ref.fsc_dpr(a, b, samp)
    F = np.fft.fft2(a)
    num = np.zeros(maxrad); den = np.zeros(maxrad)
    for y in range(n):
        for x in range(n // 2):
            h = x - n if x > (n - 1) // 2 else x
            k = y - n if y > (n - 1) // 2 else y
            r = math.sqrt(h * h + k * k)
            i = int(r)
            if i + 1 >= maxrad:
                continue
            f = r - i
            I = abs(F[y, x]) ** 2
            for j, wt in ((i, 1 - f), (i + 1, f)):
                num[j] += wt * I * math.cos(2 * math.pi * h / n)
                den[j] += wt * I
    for j in range(1, maxrad):
        if den[j] > 0:
            assert fsc[j] == pytest.approx(num[j] / den[j], abs=1e-9)


@pytest.mark.gpu
def test_gpu_fsc_dpr_2d_matches_oracle(tmp_path, capsys):
    import torch
    from bsoft_amd import bresolve, formats
    from bsoft_amd.bresolve import fsc_dpr
    n = 64
    rng = np.random.default_rng(23)
    a = rng.normal(size=(n, n)).astype(np.float32)
    b = (a + 0.8 * rng.normal(size=(n, n))).astype(np.float32)
    for ratio, hi in ((1.0, 0.0), (2.0, 5.0)):
        s, fsc, dpr = fsc_dpr(torch.as_tensor(a).cuda(), torch.as_tensor(b).cuda(), 1.5, hi, ratio)
        rs, rf, rd = resolution_ref.fsc_dpr(a, b, 1.5, hi, ratio)
        assert len(s) == len(rs)
        np.testing.assert_allclose(fsc, rf, atol=2e-5)
        np.testing.assert_allclose(dpr, rd, atol=2e-3)
    formats.write_mrc(str(tmp_path / "a.mrc"), a[None], (1.5, 1.5, 1.0), (32, 32, 0))
    formats.write_mrc(str(tmp_path / "b.mrc"), b[None], (1.5, 1.5, 1.0), (32, 32, 0))
    assert bresolve.main([str(tmp_path / "a.mrc"), str(tmp_path / "b.mrc")]) == 0
    assert "FSC(0.3):" in capsys.readouterr().out


def test_stats_2d_closed_form():
    """A 2D map (one plane): the same closed form, SNR = K (K - 2), weight
    K, rows up to the in-plane Nyquist radius."""
    n, K, samp = 16, 4, 1.5
    rng = np.random.default_rng(31)
    S = rng.normal(size=(1, n, n)) + 1j * rng.normal(size=(1, n, n))
    acc = np.stack([K * S.real, K * S.imag, np.full(S.shape, float(K)), K * np.abs(S) ** 2,
                    np.full(S.shape, float(K))])
    rows, _ = resolution_ref.reconstruction_stats(acc, samp, samp)
    assert len(rows) == n // 2 - 1
    np.testing.assert_allclose(rows[:, 3], K)
    np.testing.assert_allclose(rows[:, 7], K * (K - 2), rtol=1e-5)


@pytest.mark.gpu
def test_gpu_recon_stats_2d_matches_oracle():
    import torch
    from bsoft_amd import abi, synth
    from bsoft_amd.recon import Reconstruction
    n = 32
    rng = np.random.default_rng(19)
    bm = synth.make_blob_map(n, 10, rng)
    ps = synth.make_particles(bm, 200, n, rng, snr=0.5, seed=19)
    parts = ps.parts.copy()
    ang = parts["view"][:, 3]
    parts["view"] = np.stack([0 * ang, 0 * ang, np.ones_like(ang), ang], axis=1)
    rec = Reconstruction(n, n, 64, 1.5, ctf_action=abi.BSX_CTF_FLIP,
                         flags=abi.BSX_FLAG_FOM | abi.BSX_FLAG_2D)
    rec.insert(ps.images.cuda().contiguous(), parts)
    acc = rec.accumulators(0).cpu().numpy().astype(np.float64)[:, :1]
    print("recon stats 2D n=32", flush=True)
    rows, est = rec.stats(0, 1)
    ref, rest = resolution_ref.reconstruction_stats(acc, 1.5, 1.5)
    assert rows.shape == ref.shape and len(rows) > 3
    for c in range(9):
        np.testing.assert_allclose(rows[:, c], ref[:, c], rtol=1e-6,
                                   atol=1e-9 * max(1e-30, np.abs(ref[:, c]).max()))
    np.testing.assert_allclose(est, rest, rtol=1e-6)

"""Pins of the CPU oracle (oracle/bsx_oracle.c) against closed forms,
textbook results, invariants and brute force — the reference ships no tests or
golden vectors for this path, so each oracle function is pinned here by a
check that a plausible mistake (dropped term, wrong sign, index or transposed
operand) would fail."""
import math

import numpy as np
import pytest

from bsoft_amd import synth
from bsoft_amd.abi import Particle, ReconParams
from oracle.oracle import Oracle, OracleRecon


def P(**kw):
    p = Particle()
    p.view[:] = kw.get("view", (0, 0, 1, 0))
    p.ori[:] = kw.get("ori", (0, 0))
    p.hdr_ori[:] = kw.get("hdr_ori", p.ori[:])
    p.pixel_size = kw.get("pix", 1.5)
    p.mag = kw.get("mag", 1.0)
    p.volt = kw.get("volt", 300e3)
    p.cs = kw.get("cs", 2.7e7)
    p.amp_shift = kw.get("amp_shift", math.asin(0.07))
    p.def_avg = kw.get("def_avg", 2e4)
    p.def_dev = kw.get("def_dev", 0.0)
    p.ast_angle = kw.get("ast", 0.0)
    p.sel = kw.get("sel", 1.0)
    p.half = kw.get("half", 0)
    return p


def RP(**kw):
    d = dict(size=16, box=16, ft_size=32, interp=0, sampling=1.5, hi_res=1.5, lo_res=0.0,
             ctf_action=0, flags=0, wiener=0.2, nhalves=1, batch=0, mem_cap=0)
    d.update(kw)
    return ReconParams(**d)


# ------------------------------------------------------------------ sizes
@pytest.mark.parametrize("x,scale,pad,ft", [
    (64, 1, 2, 128), (128, 1, 2, 256), (192, 1, 2, 512), (256, 1, 2, 512), (100, 1, 2, 256),
    (64, 1, 0, 64), (64, 1, 9, 512), (64, 1.5, 2, 256), (64, 1, 1, 64), (65, 1, 1, 128),
    (0, 1, 2, 0)])
def test_part_ft_size_table(oracle64, x, scale, pad, ft):
    assert oracle64.part_ft_size(x, scale, pad) == ft


def test_find_next_power(oracle64):
    L = oracle64.L
    assert [L.or_find_next_power(a, b) for a, b in [(100, 2), (128, 2), (1, 2), (5, 1), (5, 0), (10, 3)]] \
        == [128, 128, 2, 5, 0, 27]


def test_check_resolution(oracle64):
    assert oracle64.check_resolution(1.5, 64, 1.5) == pytest.approx(2 * 1.5 / math.sqrt(3))
    assert oracle64.check_resolution(500, 64, 1.5) == 96.0
    assert oracle64.check_resolution(10, 64, 1.5) == 10


# ------------------------------------------------------------------ FFT
def naive_dft(x, sign):
    n = x.shape[-1]
    j = np.arange(n)
    W = np.exp(sign * 2j * np.pi * np.outer(j, j) / n)
    return x @ W.T


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 8, 12, 16, 24, 30, 64, 96, 192, 256, 512])
def test_fft_1d_vs_textbook(oracle64, n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    f = oracle64.fft(x, dir=-1, norm=0)
    np.testing.assert_allclose(f, naive_dft(x, -1), rtol=0, atol=1e-9 * n)
    b = oracle64.fft(x, dir=+1, norm=0)
    np.testing.assert_allclose(b, naive_dft(x, +1), rtol=0, atol=1e-9 * n)


@pytest.mark.parametrize("shape", [(8, 8), (16, 32), (12, 8), (8, 8, 8), (6, 4, 16)])
def test_fft_nd_norms(oracle64, shape):
    rng = np.random.default_rng(1)
    x = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    n = x.size
    ref = np.fft.fftn(x)
    np.testing.assert_allclose(oracle64.fft(x, -1, 0), ref, atol=1e-9)
    np.testing.assert_allclose(oracle64.fft(x, -1, 1), ref / math.sqrt(n), atol=1e-9)
    np.testing.assert_allclose(oracle64.fft(x, -1, 2), ref / n, atol=1e-9)
    np.testing.assert_allclose(oracle64.fft(x, +1, 0), np.fft.ifftn(x) * n, atol=1e-9)
    # Parseval with the 1/sqrt(n) scaling
    y = oracle64.fft(x, -1, 1)
    assert np.sum(abs(y) ** 2) == pytest.approx(np.sum(abs(x) ** 2), rel=1e-12)


def test_fft_impulse_and_plane_wave(oracle64):
    x = np.zeros((16, 16), complex)
    x[0, 0] = 1
    np.testing.assert_allclose(oracle64.fft(x, -1, 0), np.ones((16, 16)), atol=1e-12)
    yy, xx = np.mgrid[:16, :16]
    w = np.exp(2j * np.pi * (3 * xx + 5 * yy) / 16)
    f = oracle64.fft(w, -1, 0)
    assert abs(f[5, 3]) == pytest.approx(256)
    f[5, 3] = 0
    assert np.abs(f).max() < 1e-9


# ------------------------------------------------------------------ rotations
def test_view_matrix_z_rotation(oracle64):
    for a in [0.0, 0.3, -1.2, math.pi / 2, 3.0]:
        R = oracle64.view_matrix([0, 0, 1, a])
        ref = np.array([[math.cos(a), -math.sin(a), 0], [math.sin(a), math.cos(a), 0], [0, 0, 1]])
        ref[np.abs(ref) < 1e-12] = 0
        np.testing.assert_allclose(R, ref, atol=1e-15)


def test_view_matrix_properties(oracle64):
    rng = np.random.default_rng(3)
    views = synth.uniform_views(200, rng)
    views = np.vstack([views, [[0, 0, -1, 0.4], [1, 0, 0, 0], [0, 1, 0, 1.0]]])
    for v in views:
        R = oracle64.view_matrix(v)
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-12)
        assert np.linalg.det(R) == pytest.approx(1, abs=1e-12)
        vn = np.asarray(v[:3]) / np.linalg.norm(v[:3])
        np.testing.assert_allclose(R[:, 2], vn, atol=1e-12)       # view vector = image normal
    # vectorised python conversion agrees
    np.testing.assert_allclose(synth.view_to_matrix(views[:200]),
                               np.stack([oracle64.view_matrix(v) for v in views[:200]]), atol=1e-14)


def test_quaternion_view_roundtrip(oracle64):
    rng = np.random.default_rng(4)
    q = synth.uniform_quaternions(100, rng)
    views = synth.quat_to_view(q)
    for qi, v in zip(q, views):
        s, x, y, z = qi
        Rq = np.array([[s*s+x*x-y*y-z*z, 2*x*y-2*s*z, 2*x*z+2*s*y],
                       [2*x*y+2*s*z, s*s-x*x+y*y-z*z, 2*y*z-2*s*x],
                       [2*x*z-2*s*y, 2*y*z+2*s*x, s*s-x*x-y*y+z*z]])
        np.testing.assert_allclose(oracle64.view_matrix(v), Rq, atol=1e-12)


# ------------------------------------------------------------------ CTF
def test_lambda_textbook(oracle64):
    # relativistic electron wavelength: 0.019687 A at 300 kV, 0.025079 A at 200 kV
    assert oracle64.lam(300e3) == pytest.approx(0.0196875, abs=2e-7)
    assert oracle64.lam(200e3) == pytest.approx(0.0250793, abs=2e-7)


def test_ctf_zeros_closed_form(oracle64):
    amp = math.asin(0.07)
    p = P(cs=0.0, def_avg=2e4, amp_shift=amp)
    lam = oracle64.lam(300e3)
    for n in range(1, 6):
        s = math.sqrt((n * math.pi - amp) / (math.pi * lam * 2e4))
        assert abs(oracle64.ctf_value(p, s, 0)) < 1e-9
        assert oracle64.ctf_value(p, s * 0.99, 0) * oracle64.ctf_value(p, s * 1.01, 0) < 0
    # value at s=0: sin(-amp)
    assert oracle64.ctf_value(p, 0, 0) == pytest.approx(-0.07)
    # flip maps to +-1 and 0 -> +1
    assert oracle64.ctf_value(p, 0, 0, flip=True) == -1
    p0 = P(cs=0.0, amp_shift=0.0)
    assert oracle64.ctf_value(p0, 0, 0, flip=True) == 1
    # Wiener (1+W) c / (c^2 + W)
    W = 0.2
    assert oracle64.ctf_value(p, 0, 0, wiener=W) == pytest.approx((1 + W) * -0.07 / (0.0049 + W))


def test_ctf_astigmatism_and_cs(oracle64):
    lam = oracle64.lam(300e3)
    ast = 0.7
    p = P(cs=0.0, amp_shift=0.0, def_avg=2e4, def_dev=3e3, ast=ast)
    s = 0.2
    for alpha, dfe in [(ast, 2.3e4), (ast + math.pi / 2, 1.7e4), (ast + math.pi / 4, 2e4)]:
        sx, sy = s * math.cos(alpha), s * math.sin(alpha)
        assert oracle64.ctf_value(p, sx, sy) == pytest.approx(math.sin(-math.pi * lam * dfe * s * s), abs=1e-9)
    pc = P(cs=2.7e7, amp_shift=0.0, def_avg=0.0)
    assert oracle64.ctf_value(pc, s, 0) == pytest.approx(math.sin(math.pi / 2 * lam ** 3 * 2.7e7 * s ** 4), abs=1e-12)


# ------------------------------------------------------------------ staging
def test_background_ellipse_bruteforce(oracle64):
    box = 16
    img = np.zeros((box, box), np.float32)
    yy, xx = np.mgrid[:box, :box]
    ori = (7.0, 9.0)
    out = ((xx - ori[0]) / 8) ** 2 + ((yy - ori[1]) / 8) ** 2 - 1 > 0
    img[out] = 3.25
    img[~out] = -7.0
    assert oracle64.background(img, ori) == pytest.approx(3.25)
    rng = np.random.default_rng(0)
    img = rng.normal(size=(box, box)).astype(np.float32)
    assert oracle64.background(img, ori) == pytest.approx(img[out].astype(np.float64).mean())


def test_stage_impulse_shift_theorem(oracle64):
    """An impulse at the particle origin transforms to a constant after the
    phase shift to origin (shift theorem), scaled by 1/ft; pad weight 1/(0.6 r^2)."""
    box, ft = 16, 32
    img = np.zeros((box, box), np.float32)
    img[5, 9] = 2.0
    rp = RP(box=box, ft_size=ft)
    F, w = oracle64.stage_particle(rp, P(ori=(9, 5), hdr_ori=(8, 8)), img)
    np.testing.assert_allclose(F, np.full((ft, ft), 2.0 / ft), atol=1e-12)
    assert w == pytest.approx(1 / (0.6 * 4))


def test_stage_padding_fills_background(oracle64):
    box, ft = 16, 32
    rng = np.random.default_rng(2)
    img = rng.normal(size=(box, box)).astype(np.float32)
    bkg = oracle64.background(img, (8, 8))
    pad = np.full((ft, ft), bkg, dtype=np.float64)
    pad[:box, :box] = img
    F, _ = oracle64.stage_particle(RP(box=box, ft_size=ft), P(ori=(0, 0), hdr_ori=(8, 8)), img)
    np.testing.assert_allclose(F, np.fft.fft2(pad.astype(np.float64)) / ft, atol=1e-9)


# ------------------------------------------------------------------ insertion / finish
def _identity_chain(o, view, N=16, interp=0):
    rng = np.random.default_rng(5)
    img = rng.normal(size=(N, N)).astype(np.float32)
    rp = RP(size=N, box=N, ft_size=N, interp=interp)
    r = OracleRecon(rp, 1)
    r.insert([P(view=view, ori=(N // 2, N // 2), hdr_ori=(N // 2, N // 2))], img[None])
    return img, r


def test_insertion_exact_plane(oracle64):
    """ft = N and the identity view: every sample lands on a grid point of the
    l = 0 plane; the plane holds the shifted 2D DFT / N, weight 1."""
    N = 16
    img, r = _identity_chain(oracle64, (0, 0, 1, 0), N)
    acc = r.partial_acc(0)
    h = np.fft.fftfreq(N, 1 / N)
    ph = np.exp(1j * np.pi * (h[None, :] + h[:, None]))     # phase shift to origin N/2
    F2 = np.fft.fft2(img.astype(np.float64)) / N * ph
    np.testing.assert_allclose(acc[0, 0] + 1j * acc[1, 0], F2, atol=1e-12)
    np.testing.assert_allclose(acc[2, 0], np.ones((N, N)))
    assert np.abs(acc[2, 1:]).max() == 0
    np.testing.assert_allclose(acc[3, 0], np.abs(F2) ** 2, atol=1e-12)   # sum w|F|^2
    np.testing.assert_allclose(acc[4, 0], np.ones((N, N)))               # sum w^2


def test_insertion_rotated_plane(oracle64):
    """A quarter turn about z sends sample (h, k) to voxel (-k, h, 0); a view
    along +x puts the section in the yz plane."""
    N = 16
    img, r = _identity_chain(oracle64, (0, 0, 1, math.pi / 2), N)
    acc = r.partial_acc(0)
    _, r0 = _identity_chain(oracle64, (0, 0, 1, 0), N)
    a0 = r0.partial_acc(0)
    F0 = a0[0, 0] + 1j * a0[1, 0]
    F1 = acc[0, 0] + 1j * acc[1, 0]
    for k in range(-N // 2, N // 2):
        for h in range(-N // 2, N // 2):
            assert F1[h % N, (-k) % N] == pytest.approx(F0[k % N, h % N], abs=1e-12)
    _, rx = _identity_chain(oracle64, (1, 0, 0, 0), N)
    ax = rx.partial_acc(0)
    assert ax[2].sum() == pytest.approx(N * N)
    assert np.count_nonzero(ax[2][:, :, 0]) == N * N      # the x = 0 plane holds all samples


def test_full_chain_closed_form(oracle64):
    """ft = N, identity view, one particle: the finished map equals img/sqrt(N)
    in every z slice (FFT, phase shift, insertion, weigh, centre shift,
    inverse FFT and real part compose to a known identity)."""
    N = 16
    img, r = _identity_chain(oracle64, (0, 0, 1, 0), N)
    m, _, _, st = r.finish()
    for z in range(N):
        np.testing.assert_allclose(m[z], img / math.sqrt(N), atol=1e-6)
    assert st.coverage == N * N
    assert st.friedel < 1e-12


def test_two_d_chain_closed_form(oracle64):
    """A 2D map (flags & 2: one z plane, z real size 1): one particle at the
    identity view with ft = N. Every in-band sample lands on its own voxel of
    the single plane with weight sqrt(maxs2 - s^2) (the taper with z real
    size 1 is below one everywhere); the finished transform is the particle's
    DFT / N inside the band and the map its band-limited image."""
    N, samp, hi = 16, 1.5, 4.0
    rng = np.random.default_rng(6)
    img = rng.normal(size=(N, N)).astype(np.float32)
    r = OracleRecon(RP(size=N, box=N, ft_size=N, sampling=samp, hi_res=hi, flags=2), 1)
    r.insert([P(view=(0, 0, 1, 0), ori=(N // 2, N // 2), hdr_ori=(N // 2, N // 2), pix=samp)],
             img[None])
    acc = r.partial_acc(0)
    assert acc.shape == (5, 1, N, N)
    h = np.fft.fftfreq(N, 1 / N)
    s2 = (h[None, :] ** 2 + h[:, None] ** 2) / (N * samp) ** 2
    band = s2 < 1 / hi ** 2
    np.testing.assert_allclose(acc[2, 0], np.where(band, np.sqrt(np.maximum(1 / hi ** 2 - s2, 0)), 0),
                               atol=1e-14)
    m, _, ft, st = r.finish(want_ft=True)
    F = np.fft.fft2(img.astype(np.float64))
    np.testing.assert_allclose(ft[0], np.where(band, F / N, 0), atol=1e-12)
    np.testing.assert_allclose(m[0], np.fft.ifft2(np.where(band, F, 0)).real, atol=1e-12)
    assert st.coverage == int(band.sum())


def test_weight_sum_counts_samples(oracle64):
    """Sum of weights = inserted samples x particle weight, for every
    interpolation (trilinear weights sum to one) and independent of the CTF."""
    rng = np.random.default_rng(7)
    N, box, ft = 16, 16, 32
    imgs = rng.normal(size=(5, box, box)).astype(np.float32)
    parts = [P(view=v, ori=(8, 8), hdr_ori=(8, 8)) for v in synth.uniform_views(5, rng)]
    wp = 1 / (0.6 * 4)
    res = {}
    for interp in (0, 2):
        for ctf in (0, 1):
            r = OracleRecon(RP(size=N, box=box, ft_size=ft, interp=interp, ctf_action=ctf), 1)
            r.insert(parts, imgs)
            a = r.partial_acc(0)
            assert a[2].sum() == pytest.approx(5 * ft * ft * wp, rel=1e-12)
            res[(interp, ctf)] = a[2]
    np.testing.assert_array_equal(res[(0, 0)], res[(0, 1)])
    np.testing.assert_array_equal(res[(2, 0)], res[(2, 1)])


def test_partial_map_additivity(oracle64):
    rng = np.random.default_rng(8)
    N, box, ft = 16, 16, 32
    imgs = rng.normal(size=(7, box, box)).astype(np.float32)
    parts = [P(view=v, ori=(8, 8), hdr_ori=(8, 8), half=i % 2) for i, v in enumerate(synth.uniform_views(7, rng))]
    rp = RP(size=N, box=box, ft_size=ft, ctf_action=1)
    r1 = OracleRecon(rp, 1)
    r1.insert(parts, imgs)
    r3 = OracleRecon(rp, 4)
    r3.insert(parts, imgs)
    m1, _, f1, _ = r1.finish(want_ft=True)
    m3, _, f3, _ = r3.finish(want_ft=True)
    np.testing.assert_allclose(m3, m1, atol=1e-12)
    # half maps (even/odd partial maps) recombine to the full map's sums
    a = sum(r3.partial_acc(t) for t in range(4))
    np.testing.assert_allclose(a, r1.partial_acc(0), atol=1e-12)


def test_projection_slice_reconstruction(oracle64):
    """Noise-free CTF-free projections of a blob map at uniform views
    reconstruct the map (projection-slice theorem end to end)."""
    rng = np.random.default_rng(9)
    N = 32
    bm = synth.make_blob_map(N, 8, rng)
    ps = synth.make_particles(bm, 600, N, rng, snr=None, amp_contrast=1.0, cs=0.0,
                              defocus=(0.0, 0.0))
    r = OracleRecon(RP(size=N, box=N, ft_size=64), 2)
    r.insert(synth.particles_list(ps.parts), ps.images.numpy())
    m, _, _, _ = r.finish()
    truth = bm.volume().numpy()
    # CTF with zero defocus and amp shift pi/2 is -1 everywhere: the map is -truth
    c = np.corrcoef(m.ravel(), truth.ravel())[0, 1]
    assert c < -0.97


# ------------------------------------------------------------------ FSC
def test_fsc_identities(oracle64):
    rng = np.random.default_rng(10)
    n = 16
    a = rng.normal(size=(n, n, n)) + 1j * rng.normal(size=(n, n, n))
    f = oracle64.fsc(a, a, 1.0, 1.0)
    assert len(f) == int(n * 1.0 / (2 / math.sqrt(3)))
    np.testing.assert_allclose(f, 1, atol=1e-12)
    g = oracle64.fsc(a, -a, 1.0, 1.0)
    assert g[0] == 1
    np.testing.assert_allclose(g[1:], -1, atol=1e-12)
    b = rng.normal(size=(n, n, n)) + 1j * rng.normal(size=(n, n, n))
    assert abs(np.mean(oracle64.fsc(a, b, 1.0, 1.0)[2:])) < 0.1


def test_fsc_shell_split(oracle64):
    """A single voxel at radius r contributes (1-f) to shell floor(r) and f to
    the next: with b = a except that voxel, shells respond exactly."""
    n = 16
    a = np.zeros((n, n, n), complex)
    a[0, 0, 3] = 1.0     # h = 3 -> radius 3
    a[0, 1, 2] = 1.0     # radius sqrt(5) = 2.236
    b = a.copy()
    b[0, 1, 2] = -1.0
    f = oracle64.fsc(a, b, 1.0, 2.0)
    r = math.sqrt(5)
    fr = r - 2
    # shell 2: only (0,1,2) with weight 1-fr -> -1 ; shell 3: (0,0,3)*1 and (0,1,2)*fr
    assert f[2] == pytest.approx(-1)
    assert f[3] == pytest.approx((1 - fr) / (1 + fr))


# ------------------------------------------------------------------ extraction
def test_fsi_kernel_closed_form(oracle64):
    for w, p in [(8, 2), (10, 2), (6, 3)]:
        K = oracle64.fsi_kernel(w, p)
        hw = (w - 1) // 2
        assert K[0, hw] == 1 and np.count_nonzero(K[0]) == 1
        assert K[512, hw + 1] == 1 and np.count_nonzero(K[512]) == 1
        for i in (1, 100, 256, 511):
            x = i / 512 + hw - np.arange(w)
            ref = np.sin(np.pi * x) / (w * np.sin(np.pi * x / w)) * np.cos(np.pi * x / w) ** p
            np.testing.assert_allclose(K[i], ref, atol=1e-12)
    assert np.all(oracle64.fsi_kernel(8, -1) == 1)


def test_central_section_grid_exact(oracle64):
    """On-grid sampling (identity and quarter-turn views) returns the volume's
    l = 0 plane exactly; the centre shift multiplies by (-1)^(h+k)."""
    rng = np.random.default_rng(11)
    n = 16
    ft = rng.normal(size=(n, n, n)) + 1j * rng.normal(size=(n, n, n))
    sec = oracle64.central_section(ft, 1.0, np.eye(3), 0.01, 8, 2, center=0)
    hh = np.fft.fftfreq(n, 1 / n)
    inband = (hh[None, :] ** 2 + hh[:, None] ** 2) <= (n / 2) ** 2 * 1e9
    np.testing.assert_allclose(sec[inband], ft[0][inband], atol=1e-12)
    secc = oracle64.central_section(ft, 1.0, np.eye(3), 0.01, 8, 2, center=1)
    sgn = (-1.0) ** (hh[None, :] + hh[:, None])
    np.testing.assert_allclose(secc, ft[0] * sgn, atol=1e-12)
    Rz = oracle64.view_matrix([0, 0, 1, math.pi / 2])
    s2 = oracle64.central_section(ft, 1.0, Rz, 0.01, 8, 2, center=0)
    for k in range(-n // 2, n // 2):
        for h in range(-n // 2, n // 2):
            assert s2[k % n, h % n] == pytest.approx(ft[0, h % n, (-k) % n], abs=1e-12)


def test_central_section_of_gaussian(oracle64):
    """The transform of a centred isotropic Gaussian is an isotropic Gaussian;
    every central section is the same 2D Gaussian (off-grid interpolation)."""
    n, sig = 32, 2.5
    g = np.arange(n) - n // 2
    zz, yy, xx = np.meshgrid(g, g, g, indexing="ij")
    m = np.exp(-(xx ** 2 + yy ** 2 + zz ** 2) / (2 * sig ** 2))
    ft = oracle64.map_transform(m, (n // 2,) * 3)
    h = np.fft.fftfreq(n, 1 / n)
    ref2d = ft[0]
    rng = np.random.default_rng(12)
    for v in synth.uniform_views(3, rng):
        sec = oracle64.central_section(ft, 1.0, oracle64.view_matrix(v), 0.01, 10, 2, center=0)
        r2 = h[None, :] ** 2 + h[:, None] ** 2
        sel = r2 <= 8 ** 2
        np.testing.assert_allclose(sec[sel], ref2d[sel], atol=2e-3 * np.abs(ref2d).max())


def test_map_transform_origin_phase(oracle64):
    """A delta at the origin (N/2) transforms to the constant 1/sqrt(N^3)."""
    n = 8
    m = np.zeros((n, n, n))
    m[4, 4, 4] = 1
    ft = oracle64.map_transform(m, (4, 4, 4))
    np.testing.assert_allclose(ft, np.full((n, n, n), 1 / math.sqrt(n ** 3)), atol=1e-12)


# ------------------------------------------------------------------ scoring / grids
def test_recip_space_fom_identities(oracle64):
    rng = np.random.default_rng(13)
    n = 32
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    assert oracle64.recip_space_fom(a, a, 1.0, 2.0)[0] == pytest.approx(1)
    assert oracle64.recip_space_fom(a, 3 * a, 1.0, 2.0)[0] == pytest.approx(1)
    assert oracle64.recip_space_fom(a, -a, 1.0, 2.0)[0] == pytest.approx(-1)
    b = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    assert abs(oracle64.recip_space_fom(a, b, 1.0, 2.0)[0]) < 0.15


def test_recip_space_fom_ring_bruteforce(oracle64):
    """Ring-averaged FRC equals a brute-force per-ring evaluation."""
    rng = np.random.default_rng(14)
    n, pix, hi = 16, 1.0, 4.0
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    b = a + 0.5 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    f, cv = oracle64.recip_space_fom(a, b, pix, hi)
    h = np.fft.fftfreq(n, 1 / n)
    d = np.sqrt(h[None, :] ** 2 + h[:, None] ** 2)
    rmax = n * pix / hi
    ring = np.floor(d).astype(int)
    vals = []
    for j in range(int(np.ceil(rmax))):
        s = ring == j
        num = np.sum((a[s] * np.conj(b[s])).real)
        den = np.sqrt(np.sum(abs(a[s]) ** 2) * np.sum(abs(b[s]) ** 2))
        vals.append(num / den if den > 0 else 0)
    assert f == pytest.approx(np.mean(vals), abs=1e-12)


@pytest.mark.parametrize("step,count", [(4, 2566), (2, 10326)])
def test_asu_views_c1_counts(oracle64, step, count):
    v = oracle64.asu_views_c1(math.radians(step), math.radians(step), 1)
    assert len(v) == count
    np.testing.assert_allclose(np.linalg.norm(v[:, :3], axis=1), 1, atol=1e-12)
    assert v[0, 2] == 1 and v[-1, 2] == -1

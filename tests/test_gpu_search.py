"""GPU parity of the projection-matching path (bsx_search_* through the
C-ABI) against the FP64 restatement in oracle/search_ref.py and the C oracle.

Tolerances: disk values and CTF values of the staging within FP32 rounding
(1e-5 of the maximum); BF16 rows within one BF16 rounding (2^-8 relative);
GEMM scores within FP32 accumulation of the exact BF16 products (1e-4
absolute); FP32 FRC within 1e-4 of the FP64 reference formula."""
import math
import sys

import numpy as np
import pytest
import torch

from bsoft_amd import search, synth
from bsoft_amd.recon import central_sections, map_transform
from oracle import search_ref as sr
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bf16_to_f32(x: torch.Tensor) -> np.ndarray:
    u = x.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
    return u.view(np.float32).astype(np.float64)


@pytest.fixture(scope="module")
def setup64():
    n, R = 64, 14
    rng = np.random.default_rng(21)
    bm = synth.make_blob_map(n, 30, rng)
    vol = bm.volume().cuda()
    ft = map_transform(vol)
    ps = synth.make_particles(bm, 40, n, rng, snr=0.5, seed=21, astig=(0.0, 1500.0))
    return n, R, bm, vol, ft, ps


def test_stage_matches_oracle(setup64):
    n, R, bm, vol, ft, ps = setup64
    say(f"stage n={n} R={R} particles={len(ps.parts)}")
    s = search.Search(n, R, 1.5)
    hk, rs = s.points()
    ohk, ors = sr.disk_points(R)
    assert np.array_equal(hk, ohk) and np.array_equal(rs, ors)
    s.stage(ps.images.cuda().contiguous(), ps.parts)
    v, c, rows = s.particle_data()
    v, c = v.cpu().numpy(), c.cpu().numpy()
    rows = bf16_to_f32(rows)
    o = Oracle(64)
    parts = synth.particles_list(ps.parts)
    for i in range(0, len(parts), 7):
        ov, oc = sr.particle_disk(ps.images[i].numpy(), parts[i], ohk, o)
        assert np.abs(v[i] - ov).max() <= 1e-5 * np.abs(ov).max()
        assert np.abs(c[i] - oc).max() <= 1e-5
        orow = sr.coarse_row(ov * oc, ors)
        k = len(orow)
        assert np.all(np.abs(rows[i, :k] - orow) <= 2.0 ** -8 * np.abs(orow) + 1e-7)
        assert np.all(rows[i, k:] == 0)


def test_library_matches_extraction_and_oracle(setup64):
    n, R, bm, vol, ft, ps = setup64
    views = search.views_c1(15.0, angle_step_deg=60.0)
    mats = search.view_matrices(views)
    say(f"library views={len(views)}")
    s = search.Search(n, R, 1.5, kernel_width=8, kernel_power=2)
    s.set_reference(ft, mats)
    lv, lrows = s.library_data()
    lv = lv.cpu().numpy()
    hk, rs = s.points()
    sec = central_sections(ft, mats.reshape(-1, 9), 1.5, 1.5, kernel_width=8, kernel_power=2,
                           center=False).cpu().numpy()
    # same interpolation code: identical values at the disk points
    ex = sec[:, hk[:, 1] % n, hk[:, 0] % n]
    assert np.array_equal(lv, ex.astype(np.complex64))
    o = Oracle(64)
    ftc = ft.cpu().numpy()
    for i in range(0, len(mats), 11):
        osec = o.central_section(ftc, 1.5, mats[i], 1.5, 8, 2, 0)
        ov = osec[hk[:, 1] % n, hk[:, 0] % n]
        assert np.abs(lv[i] - ov).max() <= 1e-5 * np.abs(ov).max()
        orow = sr.coarse_row(ov, rs, scale_rings=False)
        lr = bf16_to_f32(lrows[i])[: len(orow)]
        assert np.all(np.abs(lr - orow) <= 2.0 ** -8 * np.abs(orow) + 1e-7)


@pytest.mark.parametrize("topk", [8, 16, 3])
def test_gemm_topk_exact_on_bf16_rows(setup64, topk):
    """The GEMM scores are the FP32-accumulated dot products of the BF16 rows
    and the selected views are the top-k of those products."""
    n, R, bm, vol, ft, ps = setup64
    views = search.views_c1(6.0, angle_step_deg=30.0)   # 1154 x 12 views: ragged last tile
    mats = search.view_matrices(views)
    say(f"gemm views={len(mats)} topk={topk}")
    s = search.Search(n, R, 1.5, topk=topk)
    s.set_reference(ft, mats)
    s.stage(ps.images.cuda().contiguous(), ps.parts)
    cand, score = s.exhaustive()
    torch.cuda.synchronize()
    A = bf16_to_f32(s.particle_data()[2])
    B = bf16_to_f32(s.library_data()[1])
    full = A @ B.T                                   # exact products, FP64 sums
    cand, score = cand.cpu().numpy(), score.cpu().numpy()
    for i in range(len(A)):
        assert len(set(cand[i])) == topk and cand[i].min() >= 0 and cand[i].max() < len(mats)
        np.testing.assert_allclose(score[i], full[i, cand[i]], atol=1e-4)
        kth = np.sort(full[i])[::-1][topk - 1]
        assert np.all(full[i, cand[i]] >= kth - 2e-4)
        assert np.all(np.diff(score[i]) <= 0)


def test_gemm_topk_row_blocks_and_k_steps():
    """The GEMM over more than one 256-row block (300 particles) and disk
    radii whose K = 2(Nd - 1), rounded up to 64, gives both odd and even
    numbers of 64-deep K steps: every row's top-8 are the top-8 of the exact
    BF16 products."""
    n = 64
    rng = np.random.default_rng(29)
    bm = synth.make_blob_map(n, 30, rng)
    ft = map_transform(bm.volume().cuda())
    ps = synth.make_particles(bm, 300, n, rng, snr=0.5, seed=29)
    mats = search.view_matrices(search.views_c1(8.0, angle_step_deg=45.0))
    parities = set()
    for R in (10, 11, 14, 16):   # 5, 6, 10 and 13 K steps
        s = search.Search(n, R, 1.5, topk=8)
        s.set_reference(ft, mats)
        s.stage(ps.images.cuda().contiguous(), ps.parts)
        cand, score = s.exhaustive()
        torch.cuda.synchronize()
        nk = s.kdim // 64
        parities.add(nk % 2)
        say(f"gemm R={R} K={s.kdim} ({nk} steps) rows={len(ps.parts)} views={len(mats)}")
        A = bf16_to_f32(s.particle_data()[2])
        B = bf16_to_f32(s.library_data()[1])
        full = A @ B.T
        cand, score = cand.cpu().numpy(), score.cpu().numpy()
        for i in range(len(A)):
            np.testing.assert_allclose(score[i], full[i, cand[i]], atol=1e-4)
            kth = np.sort(full[i])[::-1][7]
            assert np.all(full[i, cand[i]] >= kth - 2e-4)
            assert len(set(cand[i])) == 8
    assert parities == {0, 1}


def test_rescore_and_score_mats_match_reference_frc(setup64):
    n, R, bm, vol, ft, ps = setup64
    views = search.views_c1(10.0, angle_step_deg=45.0)
    mats = search.view_matrices(views)
    s = search.Search(n, R, 1.5)
    s.set_reference(ft, mats)
    s.stage(ps.images.cuda().contiguous(), ps.parts)
    rng = np.random.default_rng(2)
    P, C = len(ps.parts), 12
    cand = rng.integers(0, len(mats), size=(P, C)).astype(np.int32)
    cand[3, 5] = -1
    frc, best, bf = s.rescore(torch.as_tensor(cand))
    fm = s.score_mats(torch.as_tensor(mats[np.maximum(cand, 0)]))
    torch.cuda.synchronize()
    frc, best, bf, fm = frc.cpu().numpy(), best.cpu().numpy(), bf.cpu().numpy(), fm.cpu().numpy()
    assert frc[3, 5] == -np.inf
    # on-the-fly sections use the library's interpolation arithmetic; the FRC
    # ring sums run in another order (per-lane partial sums), so the two
    # agree to FP32 summation rounding: 2e-6 absolute on FRC values <= 1
    ok = cand >= 0
    np.testing.assert_allclose(fm[ok], frc[ok], rtol=0, atol=2e-6)
    o = Oracle(64)
    hk, rs = s.points()
    parts = synth.particles_list(ps.parts)
    ftc = ft.cpu().numpy()
    h = np.fft.fftfreq(n, 1 / n)
    fs = 1.0 / (n * 1.5)
    for i in range(0, P, 9):
        pv, pc = sr.particle_disk(ps.images[i].numpy(), parts[i], hk, o)
        F = np.fft.fft2(ps.images[i].numpy().astype(np.float64)) / n
        ph = np.exp(2j * np.pi * (h[None, :] * parts[i].ori[0] + h[:, None] * parts[i].ori[1]) / n)
        ctf = np.array([[o.ctf_value(parts[i], x * fs, y * fs) for x in h] for y in h])
        for c in range(0, C, 4):
            if cand[i, c] < 0:
                continue
            osec = o.central_section(ftc, 1.5, mats[cand[i, c]], 1.5, 8, 2, 0)
            ref = sr.frc_full(F * ph, osec * ctf, n, 1.5, R, o)
            assert frc[i, c] == pytest.approx(ref, abs=1e-4)
        j = int(np.argmax(np.where(ok[i], frc[i], -np.inf)))
        assert best[i] == cand[i, j] and bf[i] == frc[i, j]


def test_search_recovers_ground_truth():
    """Exhaustive search (4-degree grid with in-plane angles) plus 2/1-degree
    local refinement: on noise-free particles it finds the grid view nearest
    to the generating orientation and refines to well under a degree; on
    noisy particles (SNR 1) the orientation it returns scores at least as high
    as the true one (the search finds the maximum of the score, which noise
    moves away from the truth); and the GPU's choice among 50 views matches
    the FP64 oracle's."""
    n, R = 64, 16
    views = search.views_c1(4.0, angle_step_deg=4.0)
    mats = search.view_matrices(views)
    say(f"recovery views={len(mats)}")
    rng = np.random.default_rng(33)
    bm = synth.make_blob_map(n, 40, rng)
    ft = map_transform(bm.volume().cuda())
    s = search.Search(n, R, 1.5, topk=8)
    s.set_reference(ft, mats)
    for snr in (None, 1.0):
        ps = synth.make_particles(bm, 48, n, rng, snr=snr, seed=33)
        s.stage(ps.images.cuda().contiguous(), ps.parts)
        res = search.exhaustive_search(s, steps_deg=(2.0, 1.0))
        truth = synth.view_to_matrix(ps.parts["view"])
        grid_err = search.rotation_angle_deg(mats[res.grid_view.cpu().numpy()], truth)
        err = search.rotation_angle_deg(res.mats.cpu().numpy(), truth)
        nearest = search.rotation_angle_deg(mats[np.einsum("vij,nij->nv", mats, truth).argmax(1)], truth)
        f_true = s.score_mats(torch.as_tensor(truth[:, None]).cuda())[:, 0].cpu().numpy()
        f_found = res.frc.cpu().numpy()
        say(f"  snr {snr}: grid median {np.median(grid_err):.2f} deg (nearest grid view "
            f"{np.median(nearest):.2f}), refined median {np.median(err):.2f}, max {err.max():.2f}; "
            f"FRC found >= true for {np.mean(f_found >= f_true - 1e-6):.2f}")
        if snr is None:
            assert np.median(grid_err) <= np.median(nearest) + 0.5
            assert np.median(err) < 1.0
        else:
            assert np.mean(f_found >= f_true - 1e-6) >= 0.9
    # the FP64 reference formula picks the same view among 50 candidates
    o = Oracle(64)
    hk, rs = s.points()
    parts = synth.particles_list(ps.parts)
    ftc = ft.cpu().numpy()
    for i in range(0, 48, 16):
        c50 = np.concatenate([[res.grid_view[i].item()], rng.integers(0, len(mats), 49)])
        frc, best, _ = s.rescore(torch.as_tensor(np.tile(c50, (48, 1)).astype(np.int32)))
        pv, pc = sr.particle_disk(ps.images[i].numpy(), parts[i], hk, o)
        of = []
        for v in c50:
            osec = o.central_section(ftc, 1.5, mats[v], 1.5, 8, 2, 0)
            of.append(sr.frc_half(pv, pc, osec[hk[:, 1] % n, hk[:, 0] % n], rs))
        assert c50[int(np.argmax(of))] == best[i].item()


def test_proposed_grouped_gemm_matches_restricted_topk(setup64):
    """The routed, grouped GEMM returns the top-k of the exact BF16 products
    over the member views of each particle's proposed classes."""
    from bsoft_amd.proposal import OrientationClasses
    n, R, bm, vol, ft, ps = setup64
    views = search.views_c1(6.0, angle_step_deg=6.0)
    mats = search.view_matrices(views)
    classes = OrientationClasses(20.0)
    members = classes.members(mats)
    say(f"proposed views={len(mats)} classes={classes.n} maxm={members.shape[1]}")
    s = search.Search(n, R, 1.5, topk=8)
    s.set_reference(ft, mats)
    s.set_classes(members)
    s.stage(ps.images.cuda().contiguous(), ps.parts)
    rng = np.random.default_rng(4)
    P, k = len(ps.parts), 5
    cls = rng.integers(0, classes.n, size=(P, k)).astype(np.int32)
    cls[0, 1] = -1                       # an empty slot
    cls[1, :] = cls[1, 0]                # the same class proposed five times
    cand, score = s.proposed(torch.as_tensor(cls))
    torch.cuda.synchronize()
    A = bf16_to_f32(s.particle_data()[2])
    B = bf16_to_f32(s.library_data()[1])
    cand, score = cand.cpu().numpy(), score.cpu().numpy()
    for i in range(P):
        allowed = np.unique(np.concatenate([members[c][members[c] >= 0] for c in cls[i] if c >= 0]))
        full = A[i] @ B[allowed].T
        kk = min(8, len(allowed))
        valid = cand[i] >= 0
        assert valid.sum() == kk
        assert set(cand[i][valid]) <= set(allowed)
        np.testing.assert_allclose(score[i][valid], A[i] @ B[cand[i][valid]].T, atol=1e-4)
        kth = np.sort(full)[::-1][kk - 1]
        assert np.all(A[i] @ B[cand[i][valid]].T >= kth - 2e-4)


def _fit_setup(snr, seed, ctf=True):
    n, R = 64, 16
    rng = np.random.default_rng(seed)
    bm = synth.make_blob_map(n, 30, rng)
    ft = map_transform(bm.volume().cuda())
    ps = synth.make_particles(bm, 24, n, rng, snr=snr, seed=seed)
    s = search.Search(n, R, 1.5, ctf=ctf)
    truth = synth.view_to_matrix(ps.parts["view"])
    return n, R, rng, ft, ps, s, truth


def test_fit_origin_recovers_shift_and_defocus():
    """Noise-free particles at their true views, staged half a pixel off in x
    and y and up to three defocus steps off: the defocus scan returns the true
    defocus and the 0.5-pixel grid the true origin (box / 2)."""
    n, R, rng, ft, ps, s, truth = _fit_setup(None, 41)
    sd = 800.0
    parts = ps.parts.copy()
    true_def = parts["def_avg"].copy()
    parts["ori"] += rng.choice([-0.5, 0.5], size=(len(parts), 2))
    parts["def_avg"] += rng.integers(-3, 4, len(parts)) * sd
    say(f"fit_origin recovery particles={len(parts)} step=1 accuracy=0.5 def_std={sd}")
    s.stage(ps.images.cuda().contiguous(), parts)
    ori, dfc, fom = s.fit_origin(torch.as_tensor(truth), 1.0, 0.5, sd, ft=ft)
    torch.cuda.synchronize()
    ori, dfc, fom = ori.cpu().numpy(), dfc.cpu().numpy(), fom.cpu().numpy()
    np.testing.assert_allclose(ori, n // 2, atol=1e-9)
    np.testing.assert_allclose(dfc, true_def, rtol=1e-9)
    # the section is interpolated from the map, the particle projected from
    # the blobs analytically: the FRC at the truth is high but below 1
    assert fom.min() > 0.8


@pytest.mark.parametrize("ctf", [True, False])
def test_fit_origin_matches_oracle(ctf):
    """Noisy particles (SNR 1) staged up to 1.2 pixels and two defocus steps
    off, step 1 down to accuracy 0.25: every decision (defocus, grid winner,
    step halving) equals the FP64 restatement of img_fit_in_recip_space_grid
    on the same particles and sections; the FOM agrees to 2e-5 (FP32 products
    against FP64). With the CTF flag off the scan keeps def - 5 def_std."""
    n, R, rng, ft, ps, s, truth = _fit_setup(1.0, 43, ctf)
    sd = 600.0
    parts = ps.parts.copy()
    parts["ori"] += rng.uniform(-1.2, 1.2, size=(len(parts), 2))
    parts["def_avg"] += rng.uniform(-2, 2, len(parts)) * sd
    say(f"fit_origin vs oracle ctf={ctf} particles={len(parts)}")
    s.stage(ps.images.cuda().contiguous(), parts)
    ori, dfc, fom = s.fit_origin(torch.as_tensor(truth), 1.0, 0.25, sd, ft=ft)
    torch.cuda.synchronize()
    ori, dfc, fom = ori.cpu().numpy(), dfc.cpu().numpy(), fom.cpu().numpy()
    if not ctf:
        np.testing.assert_array_equal(dfc, parts["def_avg"] - 5 * sd)
    o = Oracle(64)
    hk, rs = s.points()
    ftc = ft.cpu().numpy()
    plist = synth.particles_list(parts)
    for i in range(0, len(parts), 4):
        F = np.fft.fft2(ps.images[i].numpy().astype(np.float64)) / n
        v0 = F[hk[:, 1] % n, hk[:, 0] % n]
        osec = o.central_section(ftc, 1.5, truth[i], 1.5, 8, 2, 0)
        sv = osec[hk[:, 1] % n, hk[:, 0] % n]
        (ox, oy), d, f, _ = sr.fit_grid(v0, plist[i], sv, hk, rs, n, o, 1.0, 0.25, sd, ctf)
        assert ori[i, 0] == pytest.approx(ox, abs=1e-9) and ori[i, 1] == pytest.approx(oy, abs=1e-9)
        assert dfc[i] == pytest.approx(d, rel=1e-12)
        assert fom[i] == pytest.approx(f, abs=2e-5)


def test_chunked_staging_keeps_every_record():
    """A box without the fused staging kernel (48 = 3 x 16) and more
    particles than one 256-MB spectrum chunk holds (14563 at 48^2): the disk
    values of a particle in the second chunk match the oracle, and the
    origin/defocus fit, which reads every particle's record, matches the FP64
    restatement for particles in both chunks."""
    n, R, npart = 48, 12, 15000
    rng = np.random.default_rng(47)
    bm = synth.make_blob_map(n, 12, rng)
    ft = map_transform(bm.volume().cuda())
    ps = synth.make_particles(bm, npart, n, rng, snr=2.0, seed=47, device="cuda")
    parts = ps.parts.copy()
    parts["ori"] += rng.uniform(-1, 1, size=(npart, 2))
    say(f"chunked staging box={n} particles={npart}")
    s = search.Search(n, R, 1.5)
    s.stage(ps.images.contiguous(), parts)
    truth = synth.view_to_matrix(parts["view"])
    ori, dfc, fom = s.fit_origin(torch.as_tensor(truth), 1.0, 0.5, 400.0, ft=ft)
    v, c, _ = s.particle_data()
    torch.cuda.synchronize()
    o = Oracle(64)
    hk, rs = s.points()
    plist = synth.particles_list(parts)
    ftc = ft.cpu().numpy()
    for i in (3, npart - 2):
        img = ps.images[i].cpu().numpy()
        ov, oc = sr.particle_disk(img, plist[i], hk, o)
        assert np.abs(v[i].cpu().numpy() - ov).max() <= 1e-5 * np.abs(ov).max()
        F = np.fft.fft2(img.astype(np.float64)) / n
        osec = o.central_section(ftc, 1.5, truth[i], 1.5, 8, 2, 0)
        (ox, oy), d, f, _ = sr.fit_grid(F[hk[:, 1] % n, hk[:, 0] % n], plist[i],
                                        osec[hk[:, 1] % n, hk[:, 0] % n], hk, rs, n, o,
                                        1.0, 0.5, 400.0)
        assert ori[i, 0].item() == pytest.approx(ox, abs=1e-9)
        assert ori[i, 1].item() == pytest.approx(oy, abs=1e-9)
        assert dfc[i].item() == pytest.approx(d, rel=1e-12)
        assert fom[i].item() == pytest.approx(f, abs=2e-5)


@pytest.mark.parametrize("ctf", [True, False])
def test_fit_monte_matches_oracle(ctf):
    """img_fit_in_recip_space_monte with the same draws on both sides: noisy
    particles (SNR 1) staged up to 1.2 pixels and two defocus steps off, 24
    iterations of shift_std 0.6 pixels and def_std 300 A: the best origin and
    defocus equal the FP64 restatement's (every FOM comparison decides the
    same way), the FOM agrees to 2e-5. A second particle set with the section
    sign flipped would walk; here positive FOMs restart every move."""
    n, R, rng, ft, ps, s, truth = _fit_setup(1.0, 45, ctf)
    sd, sh, it = 300.0, 0.6, 24
    parts = ps.parts.copy()
    parts["ori"] += rng.uniform(-1.2, 1.2, size=(len(parts), 2))
    parts["def_avg"] += rng.uniform(-2, 2, len(parts)) * sd
    say(f"fit_monte vs oracle ctf={ctf} particles={len(parts)} iterations={it}")
    s.stage(ps.images.cuda().contiguous(), parts)
    draws = search.monte_draws(len(parts), it, seed=9)
    ori, dfc, fom = s.fit_monte(torch.as_tensor(truth), it, sd, sh, draws=draws, ft=ft)
    torch.cuda.synchronize()
    ori, dfc, fom = ori.cpu().numpy(), dfc.cpu().numpy(), fom.cpu().numpy()
    o = Oracle(64)
    hk, rs = s.points()
    ftc = ft.cpu().numpy()
    plist = synth.particles_list(parts)
    moved = 0
    for i in range(0, len(parts), 3):
        F = np.fft.fft2(ps.images[i].numpy().astype(np.float64)) / n
        v0 = F[hk[:, 1] % n, hk[:, 0] % n]
        osec = o.central_section(ftc, 1.5, truth[i], 1.5, 8, 2, 0)
        sv = osec[hk[:, 1] % n, hk[:, 0] % n]
        (ox, oy), d, f = sr.fit_monte(v0, plist[i], sv, hk, rs, n, o, it, sd, sh, draws[i], ctf)
        assert ori[i, 0] == pytest.approx(ox, abs=1e-9) and ori[i, 1] == pytest.approx(oy, abs=1e-9)
        assert dfc[i] == pytest.approx(d, rel=1e-12)
        assert fom[i] == pytest.approx(f, abs=2e-5)
        moved += (ox, oy) != (parts["ori"][i][0], parts["ori"][i][1])
    assert moved > 0


def test_refine_monte_matches_oracle():
    """img_refine_monte's view walk with the same draws on both sides: 6
    noisy particles started 1-2 degrees off their views and a pixel off, 5
    view steps of 5 origin/defocus steps each: the best view, origin and
    defocus equal the FP64 restatement's, the FOM agrees to 2e-5."""
    n, R, rng, ft, ps, s, truth = _fit_setup(1.0, 49, True)
    parts = ps.parts[:6].copy()
    parts["ori"] += rng.uniform(-1, 1, size=(len(parts), 2))
    start = parts["view"].copy()
    start[:, 3] += rng.uniform(-0.03, 0.03, len(parts))
    it, sd, sh, vs, ma = 5, 300.0, 0.5, 0.02, 0.03
    say(f"refine_monte vs oracle particles={len(parts)} iterations={it}x{it}")
    s.stage(ps.images[:6].cuda().contiguous(), parts)
    draws = search.monte_view_draws(len(parts), it, seed=11)
    inner = np.stack([search.monte_draws(len(parts), it, seed=20 + i) for i in range(it)], axis=1)
    bv, bo, bd, bf = s.refine_monte(start, it, sd, sh, vs, ma, draws=draws, inner_draws=inner,
                                    ft=ft)
    o = Oracle(64)
    hk, rs = s.points()
    ftc = ft.cpu().numpy()
    plist = synth.particles_list(parts)

    def section(view):
        m = synth.view_to_matrix(np.asarray(view, dtype=np.float64).reshape(1, 4))[0]
        osec = o.central_section(ftc, 1.5, m, 1.5, 8, 2, 0)
        return osec[hk[:, 1] % n, hk[:, 0] % n]

    for i in range(len(parts)):
        F = np.fft.fft2(ps.images[i].numpy().astype(np.float64)) / n
        v0 = F[hk[:, 1] % n, hk[:, 0] % n]
        v, (ox, oy), d, f = sr.refine_monte(v0, plist[i], section, hk, rs, n, o, start[i], it,
                                            sd, sh, vs, ma, draws[i], inner[i], True)
        np.testing.assert_allclose(bv[i], v, atol=1e-12)
        assert bo[i, 0] == pytest.approx(ox, abs=1e-9) and bo[i, 1] == pytest.approx(oy, abs=1e-9)
        assert bd[i] == pytest.approx(d, rel=1e-12)
        assert bf[i] == pytest.approx(f, abs=2e-5)


def test_fit_monte_edge_cases():
    """No staged particles: nothing to do. max_iter < 1, a negative standard
    deviation or draws of the wrong shape are argument errors. One iteration
    returns the staged origin and defocus."""
    from bsoft_amd.abi import BsxError
    n, R = 32, 8
    rng = np.random.default_rng(5)
    bm = synth.make_blob_map(n, 6, rng)
    ft = map_transform(bm.volume().cuda())
    s = search.Search(n, R, 1.5)
    s.stage(torch.empty((0, n, n), device="cuda"), synth.make_particles(bm, 1, n, rng).parts[:0])
    ori, dfc, fom = s.fit_monte(torch.empty((0, 3, 3), dtype=torch.float64), 4, ft=ft)
    assert ori.shape == (0, 2) and fom.shape == (0,)
    ps = synth.make_particles(bm, 4, n, rng, snr=None)
    s.stage(ps.images.cuda().contiguous(), ps.parts)
    mats = torch.as_tensor(synth.view_to_matrix(ps.parts["view"]))
    for it, dstd, sstd in ((0, 0.0, 1.0), (4, -1.0, 1.0), (4, 0.0, -1.0)):
        with pytest.raises(BsxError):
            s.fit_monte(mats, it, dstd, sstd, ft=ft)
    with pytest.raises(ValueError):
        s.fit_monte(mats, 4, draws=np.zeros((4, 3, 5)), ft=ft)
    ori, dfc, _ = s.fit_monte(mats, 1, 500.0, 1.0, ft=ft)
    torch.cuda.synchronize()
    np.testing.assert_allclose(ori.cpu().numpy(), ps.parts["ori"], atol=1e-9)
    np.testing.assert_array_equal(dfc.cpu().numpy(), ps.parts["def_avg"])


def test_fit_origin_edge_cases():
    """No staged particles: nothing to do. A step or accuracy that is not
    positive, or a step / accuracy ratio of 1e6 or more, is an argument error."""
    from bsoft_amd.abi import BsxError
    n, R = 32, 8
    rng = np.random.default_rng(3)
    bm = synth.make_blob_map(n, 6, rng)
    ft = map_transform(bm.volume().cuda())
    s = search.Search(n, R, 1.5)
    s.stage(torch.empty((0, n, n), device="cuda"), synth.make_particles(bm, 1, n, rng).parts[:0])
    ori, dfc, fom = s.fit_origin(torch.empty((0, 3, 3), dtype=torch.float64), ft=ft)
    assert ori.shape == (0, 2) and fom.shape == (0,)
    ps = synth.make_particles(bm, 4, n, rng, snr=None)
    s.stage(ps.images.cuda().contiguous(), ps.parts)
    mats = torch.as_tensor(synth.view_to_matrix(ps.parts["view"]))
    for step, acc in ((0.0, 0.25), (1.0, 0.0), (1.0, 1e-7)):
        with pytest.raises(BsxError):
            s.fit_origin(mats, step, acc, ft=ft)
    ori, _, _ = s.fit_origin(mats, 1.0, 1.0, ft=ft)
    torch.cuda.synchronize()
    assert np.all(np.abs(ori.cpu().numpy() - n // 2) <= 1.0)


def test_refine_views_matches_host_loop():
    """bsx_search_refine_views (img_refine_grid's view loop over the
    views_for_refinement grid) against the same loop driven from the host:
    grids from oracle/search_ref.views_for_refinement, scores from the
    library's FRC at those rotations, the reference's pick and contraction
    rules applied per particle. Both must take the same passes and end at
    the same views; on noise-free particles started from the best 4-degree
    grid view they end within a fraction of a degree of the truth."""
    n, R, P = 64, 16, 48
    rng = np.random.default_rng(61)
    bm = synth.make_blob_map(n, 40, rng)
    ft = map_transform(bm.volume().cuda())
    lm = search.view_matrices(search.views_c1(4.0, angle_step_deg=4.0))
    s = search.Search(n, R, 1.5, kernel_width=3, topk=8)
    s.set_reference(ft, lm)
    ps = synth.make_particles(bm, P, n, rng, snr=None, seed=61)
    s.stage(ps.images.cuda().contiguous(), ps.parts)
    start = s.view_mats(search.exhaustive_search(s, steps_deg=()).grid_view)
    step0, acc = math.radians(2.0), math.radians(0.25)
    say(f"refine_views: {P} particles, step 2 deg, accuracy 0.25 deg")
    mats, frc, passes = s.refine_views(start, 2.0, 0.25, max_iter=64)
    # the host-driven loop, all particles in lockstep like the library
    views = search.matrix_views(start.cpu().numpy())
    step = np.full(P, step0)
    fom = np.full(P, -1.0)
    active = step >= acc
    it = 0
    while active.any() and it < 64:
        grids = np.stack([sr.views_for_refinement(v, st) for v, st in zip(views, step)])
        M = search.view_matrices(grids.reshape(-1, 4)).reshape(P, 27, 3, 3)
        f = s.score_mats(torch.as_tensor(M)).cpu().numpy().astype(np.float64)
        for p in np.nonzero(active)[0]:
            prefom, best, ib = fom[p], -1.0, 0
            for c in range(27):
                if best < f[p, c]:
                    best, ib = f[p, c], c
            views[p] = grids[p, ib]
            fom[p] = best
            if ib == 13 or abs(prefom - best) < 1e-6:
                step[p] /= 2
            active[p] = step[p] >= acc
        it += 1
    host = synth.view_to_matrix(views)
    d = search.rotation_angle_deg(mats.cpu().numpy(), host)
    truth = synth.view_to_matrix(ps.parts["view"])
    err = search.rotation_angle_deg(mats.cpu().numpy(), truth)
    gerr = search.rotation_angle_deg(start.cpu().numpy(), truth)
    say(f"  {passes} passes (host {it}); max difference to the host loop {d.max():.2e} deg; "
        f"median error vs truth {np.median(gerr):.2f} -> {np.median(err):.2f} deg")
    assert passes == it
    # same views: angles below 1e-4 deg are the arccos floor of the angle
    # between equal rotations (about 2e-6 deg)
    same = d < 1e-4
    assert np.sum(~same) <= 1
    np.testing.assert_allclose(frc.cpu().numpy()[same], fom[same], atol=1e-6)
    assert np.median(err) < 0.5 and np.median(err) < np.median(gerr)

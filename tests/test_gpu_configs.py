"""The benchmark configurations as GPU tests:

* the learned pipeline's low-confidence fallback (particles whose top class
  probability is below the threshold also get the exhaustive candidates,
  bsx_search_exhaustive_rows) against the exhaustive search;
* batched projection matching at 128^3: 20,000 particles of 128^2 (SNR 0.1)
  against the 4-degree grid with 90 in-plane angles (230,940 views), BF16
  GEMM top-8 then FP32 FRC rescoring, checked against the FP64 scoring on a
  subset, against ground truth, and by properties over every particle;
* mixed particle boxes reconstructed as one graph (bsx_recon_group_run)
  against the FP64 oracle, and group members with half-sets or a 2D map
  against the handles' own insert + finish.
"""
import sys

import numpy as np
import pytest
import torch

from bsoft_amd import search, synth
from bsoft_amd.proposal import OrientationClasses, ProposalNet, learned_search
from bsoft_amd.recon import map_transform
from oracle import search_ref as sr
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_fallback_rows_and_learned_fallback(gpu):
    """bsx_search_exhaustive_rows equals the exhaustive GEMM's rows; with an
    untrained network and a confidence threshold that every particle misses,
    the learned pipeline sends every particle through the fallback, and with
    a threshold at the median probability about half of them; each particle
    that fell back ends at least as high in FP32 FRC as the exhaustive
    search's best grid view (its candidates include the exhaustive top-8)."""
    n, R = 64, 14
    rng = np.random.default_rng(41)
    bm = synth.make_blob_map(n, 30, rng)
    ft = map_transform(bm.volume().cuda())
    lm = search.view_matrices(search.views_c1(8.0, angle_step_deg=8.0))
    classes = OrientationClasses(20.0)
    s = search.Search(n, R, 1.5, topk=8)
    s.set_reference(ft, lm)
    s.set_classes(classes.members(lm))
    ps = synth.make_particles(bm, 96, n, rng, snr=0.5, seed=41)
    s.stage(ps.images.cuda().contiguous(), ps.parts)
    say(f"fallback: {len(lm)} views, {classes.n} classes, {len(ps.parts)} particles")
    cand, score = s.exhaustive()
    rows = torch.as_tensor(np.array([5, 0, 77, 31, 31, 95], dtype=np.int32))
    rc, rsc = s.exhaustive_rows(rows)
    assert torch.equal(rc, cand[rows.long().cuda()])
    assert torch.equal(rsc, score[rows.long().cuda()])
    _, ebest, ebf = s.rescore(cand)
    net = ProposalNet(32, classes.n, (16, 32, 64, 64), max_batch=128, seed=3)
    _, prob = net.topk(s.cnn_input(32), 16)
    p0 = prob[:, 0].cpu().numpy()
    for thr, expect in ((1.01, len(p0)), (float(np.median(p0)), int(np.sum(p0 < np.median(p0))))):
        res, st, _, _ = learned_search(s, net, k=16, min_prob=thr, steps_deg=())
        say(f"  threshold {thr:.5f}: {st.fallback} particles fell back, "
            f"{st.candidates} candidates per particle")
        assert st.fallback == expect and st.candidates == 16   # proposed 8 + fallback 8
        low = p0 < thr
        gf = res.grid_frc.cpu().numpy()
        assert np.all(gf[low] >= ebf.cpu().numpy()[low] - 1e-6)
        # where the exhaustive best wins, it is the view the pipeline returns
        same = low & (gf <= ebf.cpu().numpy() + 1e-7)
        assert np.array_equal(res.grid_view.cpu().numpy()[same], ebest.cpu().numpy()[same])


def test_projection_matching_128_full_grid(gpu):
    """128^3 map, 20,000 particles 128^2 at SNR 0.1, 230,940 library views."""
    n, R, P = 128, 16, 20000
    rng = np.random.default_rng(51)
    bm = synth.make_blob_map(n, 60, rng)
    ft = map_transform(bm.volume().cuda())
    lm = search.view_matrices(search.views_c1(4.0, angle_step_deg=4.0))
    assert len(lm) == 2566 * 90
    s = search.Search(n, R, 1.5, kernel_width=3, topk=8)
    s.set_reference(ft, lm)
    say(f"C3: {len(lm)} views, {s.npoints} disk points, GEMM K {s.kdim}; generating {P} particles")
    ps = synth.make_particles(bm, P, n, rng, snr=0.1, seed=51, device=torch.device("cuda"),
                              chunk=2048)
    s.stage(ps.images, ps.parts)
    cand, score = s.exhaustive()
    frc, best, bf = s.rescore(cand)
    torch.cuda.synchronize()
    c, sc, f = cand.cpu().numpy(), score.cpu().numpy(), frc.cpu().numpy()
    # properties over every particle
    assert np.all((c >= 0) & (c < len(lm)))
    assert np.all(np.diff(sc, axis=1) <= 1e-6)                  # top-8 in descending order
    assert np.all([len(set(r)) == 8 for r in c[::97]])           # distinct views
    assert np.all(np.abs(f) <= 1 + 1e-5)
    bi = np.argmax(f, axis=1)
    assert np.array_equal(best.cpu().numpy(), c[np.arange(P), bi])
    # against the FP64 scoring of the same candidates on a subset
    o = Oracle(64)
    hk, rs = s.points()
    ftc = np.ascontiguousarray(ft.cpu().numpy(), dtype=np.complex128)   # converted once
    parts = synth.particles_list(ps.parts[:16])
    imgs = ps.images[:16].cpu().numpy()
    for i in range(16):
        pv, pc = sr.particle_disk(imgs[i], parts[i], hk, o)
        ref = []
        for v in c[i]:
            osec = o.central_section(ftc, 1.5, lm[v], 1.5, 3, 2, 0)
            ref.append(sr.frc_half(pv, pc, osec[hk[:, 1] % n, hk[:, 0] % n], rs))
        np.testing.assert_allclose(f[i], ref, atol=1e-4)
        # the FP32 choice is the FP64 optimum up to the FP32 scoring error
        assert ref[int(bi[i])] >= max(ref) - 2e-4
    # angular error against the truth, grid and refined
    truth = synth.view_to_matrix(ps.parts["view"])
    gerr = search.rotation_angle_deg(lm[best.cpu().numpy()], truth)
    mats, _ = s.refine(s.view_mats(best), (2.0, 1.0))
    err = search.rotation_angle_deg(mats.cpu().numpy(), truth)
    say(f"  grid view median error {np.median(gerr):.2f} deg, refined {np.median(err):.2f} deg; "
        f"within 5 deg: {np.mean(err < 5):.3f}")
    assert np.median(gerr) <= 3.0
    assert np.median(err) <= 1.5


def test_mixed_boxes_one_graph_parity(gpu):
    """Mixed particle boxes (32, 48 and 64 in a 50/30/20 ratio) of one object,
    each box reconstructed into its own volume by one group run
    (bsx_recon_group_run): the first run launches directly, the second
    captures clears, staging, batches and finish of all three members into
    one graph, the third replays it, and a fourth replays it with new particle
    parameters. Every run's map equals the FP64 oracle's on the same
    particles (max relative voxel error 1e-4), and the sums each member keeps
    afterwards give transforms with FSC >= 0.9999 against the oracle's."""
    from bsoft_amd import abi
    from bsoft_amd.abi import ReconParams
    from bsoft_amd.recon import Reconstruction, ReconGroup, part_ft_size
    from oracle.oracle import OracleRecon
    rng = np.random.default_rng(71)
    base = synth.make_blob_map(64, 20, rng)
    total = 240
    sets = []
    for box, frac in ((32, 0.5), (48, 0.3), (64, 0.2)):
        f = box / 64
        bm = synth.BlobMap(centers=base.centers * f, sigmas=base.sigmas * f, amps=base.amps, n=box)
        pix = 96.0 / box
        n = int(round(total * frac))
        ps = synth.make_particles(bm, n, box, rng, pix=pix, snr=0.2, seed=box,
                                  astig=(0.0, 1500.0))
        sets.append((box, pix, ps))
    recs = [Reconstruction(box, box, part_ft_size(box, 1, 2), pix, hi_res=2 * pix,
                           ctf_action=abi.BSX_CTF_FLIP, batch=16)
            for box, pix, _ in sets]
    grp = ReconGroup(recs)
    imgs = [ps.images.cuda().contiguous() for _, _, ps in sets]
    maps = [torch.empty((box, box, box), device="cuda") for box, _, _ in sets]

    def oracle(box, pix, parts, images):
        rp = ReconParams(size=box, box=box, ft_size=part_ft_size(box, 1, 2), interp=0,
                         sampling=pix, hi_res=2 * pix, lo_res=0.0, ctf_action=abi.BSX_CTF_FLIP,
                         flags=0, wiener=0.0, nhalves=1, batch=0, mem_cap=0)
        o = OracleRecon(rp, 2, 64)
        o.insert(synth.particles_list(parts), images)
        return o.finish(want_ft=True)

    refs = [oracle(box, pix, ps.parts, ps.images.numpy()) for box, pix, ps in sets]
    new_parts = []
    for box, pix, ps in sets:
        q = ps.parts.copy()
        q["view"] = synth.uniform_views(len(q), np.random.default_rng(box + 5))
        new_parts.append(q)
    new_refs = [oracle(box, pix, q, ps.images.numpy())
                for (box, pix, ps), q in zip(sets, new_parts)]
    for run in range(4):
        parts = [ps.parts for _, _, ps in sets] if run < 3 else new_parts
        want = refs if run < 3 else new_refs
        say(f"group run {run}: boxes 32/48/64, {[len(p) for p in parts]} particles")
        _, stats = grp.run(imgs, parts, maps)
        torch.cuda.synchronize()
        for (box, pix, ps), mp, st, (m, _, ftc, ost) in zip(sets, maps, stats, want):
            gm = mp.cpu().numpy().astype(np.float64)
            err = np.abs(gm - m).max() / np.abs(m).max()
            say(f"  box {box}: max rel err {err:.2e}, coverage {st.coverage}/{ost.coverage}, "
                f"particles {st.particles}")
            assert err <= 1e-4
            assert st.coverage == ost.coverage and st.particles == ost.particles
            assert abs(st.weight_sum - ost.weight_sum) <= 1e-5 * ost.weight_sum
        assert grp.graph_launches() == max(0, run)
    for rec, (box, pix, ps), (m, _, ftc, ost) in zip(recs, sets, new_refs):
        g = rec.finish(want_ft=True)
        fsc, sup = Oracle(64).fsc_with_support(g.ft.cpu().numpy(), ftc, pix, 2 * pix)
        say(f"  box {box}: member sums after the runs, min FSC {fsc[sup].min():.6f}")
        assert fsc[sup].min() >= 0.9999
    grp.close()


def test_group_run_half_sets_and_2d_member(gpu):
    """A group of a 3D member with two half-sets, a 2D member (BSX_FLAG_2D)
    and a member with the FOM sums (BSX_FLAG_FOM): every group run (direct,
    captured, replayed) gives the maps of each handle's own insert + finish on
    the same particles, the half maps read from the sums a run leaves
    (bsx_recon_finish, nmaps 2) equal the direct path's half maps, and so
    does the FOM member's FOM map."""
    from bsoft_amd import abi
    from bsoft_amd.recon import Reconstruction, ReconGroup, part_ft_size
    rng = np.random.default_rng(83)
    bm = synth.make_blob_map(32, 12, rng)
    ps = synth.make_particles(bm, 60, 32, rng, pix=3.0, snr=0.3, seed=5)
    ps.parts["half"] = np.arange(len(ps.parts)) % 2
    ft = part_ft_size(32, 1, 2)
    kinds = [dict(nhalves=2, flags=0), dict(nhalves=1, flags=abi.BSX_FLAG_2D),
             dict(nhalves=1, flags=abi.BSX_FLAG_FOM)]

    def make(kind):
        return Reconstruction(32, 32, ft, 3.0, hi_res=6.0, ctf_action=abi.BSX_CTF_FLIP, batch=16,
                              **kind)

    imgs = ps.images.cuda().contiguous()
    direct = []
    fom_direct = None
    for kind in kinds:
        r = make(kind)
        r.insert(imgs, ps.parts)
        full = r.finish(0, 1).map.cpu().numpy()
        if kind["flags"] & abi.BSX_FLAG_FOM:
            fom_direct = r.finish(0, 1, want_fom=True).fom.cpu().numpy()
        halves = [r.finish(i, 2).map.cpu().numpy() for i in range(2)] if kind["nhalves"] == 2 else []
        direct.append((full, halves))
        r.close()
    recs = [make(kind) for kind in kinds]
    grp = ReconGroup(recs)
    maps = [torch.empty((32, 32, 32), device="cuda"), torch.empty((1, 32, 32), device="cuda"),
            torch.empty((32, 32, 32), device="cuda")]
    for run in range(3):
        say(f"group run {run}: 3D half sets + 2D member + FOM member, 60 particles each")
        _, stats = grp.run([imgs] * 3, [ps.parts] * 3, maps)
        torch.cuda.synchronize()
        for m, st, (full, _) in zip(maps, stats, direct):
            g = m.cpu().numpy()
            err = np.abs(g - full).max() / np.abs(full).max()
            say(f"  shape {tuple(g.shape)}: max rel err vs direct {err:.2e}, particles {st.particles}")
            assert err <= 1e-5 and st.particles == len(ps.parts)
        assert grp.graph_launches() == run
    for i in range(2):
        h = recs[0].finish(i, 2).map.cpu().numpy()
        ref = direct[0][1][i]
        err = np.abs(h - ref).max() / np.abs(ref).max()
        say(f"  half {i} after the group runs: max rel err {err:.2e}")
        assert err <= 1e-5
    fom = recs[2].finish(0, 1, want_fom=True).fom.cpu().numpy()
    err = np.abs(fom - fom_direct).max() / np.abs(fom_direct).max()
    say(f"  FOM map after the group runs: max rel err {err:.2e}")
    assert err <= 1e-4
    grp.close()


def test_group_run_rejects_bad_input(gpu):
    """bsx_recon_group_create / _run refuse a repeated handle, a member with
    BSX_FLAG_CHECK_NAN, an empty particle set and a map buffer of the wrong
    size with an error (no launch), and the group still runs afterwards."""
    from bsoft_amd import abi
    from bsoft_amd.abi import BsxError
    from bsoft_amd.recon import Reconstruction, ReconGroup, part_ft_size
    rng = np.random.default_rng(5)
    bm = synth.make_blob_map(32, 8, rng)
    ps = synth.make_particles(bm, 20, 32, rng, pix=3.0, snr=0.5, seed=3)
    imgs = ps.images.cuda().contiguous()
    ft = part_ft_size(32, 1, 2)
    a = Reconstruction(32, 32, ft, 3.0, hi_res=6.0, ctf_action=abi.BSX_CTF_FLIP)
    with pytest.raises(BsxError):
        ReconGroup([a, a])
    nan = Reconstruction(32, 32, ft, 3.0, hi_res=6.0, ctf_action=abi.BSX_CTF_FLIP,
                         flags=abi.BSX_FLAG_CHECK_NAN)
    g = ReconGroup([a, nan])
    with pytest.raises(BsxError):
        g.run([imgs, imgs], [ps.parts, ps.parts])
    g.close()
    g = ReconGroup([a])
    with pytest.raises(ValueError):
        g.run([imgs], [ps.parts[:0]])
    with pytest.raises(ValueError):
        g.run([imgs], [ps.parts], [torch.empty((16, 32, 32), device="cuda")])
    maps, st = g.run([imgs], [ps.parts])
    torch.cuda.synchronize()
    ref = Reconstruction(32, 32, ft, 3.0, hi_res=6.0, ctf_action=abi.BSX_CTF_FLIP)
    ref.insert(imgs, ps.parts)
    r = ref.finish()
    err = float((maps[0] - r.map).abs().max() / r.map.abs().max())
    say(f"  group run after refused calls: max rel err {err:.2e}, particles {st[0].particles}")
    assert err <= 1e-5 and st[0].particles == 20
    g.close()

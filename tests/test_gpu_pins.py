"""The hand-derived cases of tests/pin_cases.py run through the library:
insertion weights of single samples (weighted nearest neighbour,
magnification, bootstrap, resolution band and taper, Ewald sphere), the SNR
figure of merit and the weighed transform of hand-built accumulators, the
Friedel residual, the -rescale staging and the output rescale and sinc
correction — each against its closed form (and the last three against the
FP64 oracle on the same inputs)."""
import sys

import numpy as np
import pytest
import torch

import pin_cases as pc
from bsoft_amd import abi, synth
from bsoft_amd.abi import ReconParams
from bsoft_amd.recon import Reconstruction, filter_sinc, rescale_to_avg_std
from oracle.oracle import Oracle, OracleRecon

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_weights(params: dict, part: np.ndarray, img: np.ndarray, method: int = 0) -> np.ndarray:
    d = pc.recon_params(pc.GPU, **params)
    rec = Reconstruction(d["size"], d["box"], d["ft_size"], d["sampling"], hi_res=d["hi_res"],
                         lo_res=d["lo_res"], interp=d["interp"], flags=d["flags"], method=method)
    rec.insert(torch.as_tensor(img).cuda().contiguous(), part)
    acc = rec.accumulators(0).cpu().numpy().astype(np.float64)
    rec.close()
    return acc


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("name", sorted(pc.insertion_cases()))
def test_insertion_weights_by_hand(gpu, name, method):
    say(f"case {name} method {method}")
    params, part, expect = pc.insertion_cases(pc.GPU)[name]
    img = pc.image(g=pc.GPU)
    acc = gpu_weights(params, part, img, method)
    for (x, y, z), w in expect.items():
        assert acc[2, z, y, x] == pytest.approx(w, rel=2e-6, abs=1e-7), (name, (x, y, z))
    # the whole weight plane equals the FP64 oracle's (float sums)
    o = OracleRecon(ReconParams(**pc.recon_params(pc.GPU, **params)), 1, 64)
    o.insert(synth.particles_list(part), img)
    ref = o.partial_acc(0)
    np.testing.assert_allclose(acc[2], ref[2], rtol=2e-6, atol=1e-6)
    np.testing.assert_allclose(acc[:2], ref[:2], atol=2e-6 * np.abs(ref[:2]).max())


def _fom_handle(n):
    return Reconstruction(n, n, 2 * n, 1.0, flags=abi.BSX_FLAG_FOM)


def test_fom_and_weighed_transform_by_hand(gpu):
    """SNR FOM and weighed transform of hand-built sums against the closed
    forms in the sample statistics, compared voxel by voxel with a bound
    on the maximum error (FP32 sums)."""
    n = 16
    planes, samples = pc.fom_accumulators(n, seed=9)
    rec = _fom_handle(n)
    rec.set_accumulators(torch.as_tensor(planes, dtype=torch.float32).cuda(), 0, count=7)
    r = rec.finish(want_fom=True, want_ft=True)
    fom = r.fom.cpu().numpy().astype(np.float64).ravel()
    exp = np.array([pc.fom_expected(x, w) for x, w in samples])
    exp[0] = 1.0
    # well conditioned voxels: the FOM is a ratio of differences of sums, so
    # its FP32 error scales with the power sum over the noise estimate
    W = planes[2].ravel()
    P = planes[3].ravel()
    cond = np.array([np.sum(w * np.abs(x) ** 2) for x, w in samples]) / np.maximum(
        np.array([max(abs(np.sum(w * x) / max(np.sum(w), 1e-30)) ** 2, 1e-30) for x, w in samples]),
        1e-30)
    well = (W > 1) & (cond < 1e3)
    err = np.abs(fom - exp)
    say(f"  FOM: {well.sum()} well-conditioned voxels, max error {err[well].max():.2e}, "
        f"max |FOM| {np.abs(exp[well]).max():.2f}")
    assert well.sum() > 500
    assert err[well].max() <= 1e-4 * max(1.0, np.abs(exp[well]).max())
    assert np.all(fom[W <= 1][1:] == 0)
    sign = (-1.0) ** np.add.outer(np.add.outer(np.arange(n), np.arange(n)), np.arange(n))
    ft = (r.ft.cpu().numpy().astype(np.complex128) * sign).ravel()
    mean = np.array([np.sum(w * x) / np.sum(w) if np.sum(w) > 1e-30 else 0 for x, w in samples])
    assert np.abs(ft - mean).max() <= 1e-5 * np.abs(mean).max()
    assert r.coverage == np.count_nonzero(W > 1e-30)
    assert r.particles == 7
    assert P.min() >= 0


def test_friedel_residual_by_hand(gpu):
    """Unit weights at the voxels of the hand-built transform: the weighed
    transform is the transform itself and the residual its closed form."""
    F, R = pc.friedel_volume(16)
    n = 16
    planes = np.zeros((5, n, n, n))
    planes[0], planes[1] = F.real, F.imag
    planes[2] = (np.abs(F) > 0).astype(float)
    rec = Reconstruction(n, n, 2 * n, 1.0)
    rec.set_accumulators(torch.as_tensor(planes, dtype=torch.float32).cuda())
    r = rec.finish()
    say(f"  Friedel residual {r.friedel:.8f} expected {R:.8f}")
    assert r.friedel == pytest.approx(R, rel=1e-6)


def test_rescale_flag_matches_oracle(gpu):
    """breconstruct -rescale (flags bit 1): particles normalised before the
    background; the map equals the oracle's, and differs from the map
    without the flag."""
    rng = np.random.default_rng(12)
    bm = synth.make_blob_map(32, 10, rng)
    ps = synth.make_particles(bm, 120, 32, rng, snr=0.2, seed=12)
    imgs = (ps.images * 3.0 + 1.5).contiguous()
    out = {}
    for flags in (abi.BSX_FLAG_RESCALE, 0):
        rec = Reconstruction(32, 32, 64, 1.5, ctf_action=1, flags=flags)
        rec.insert(imgs.cuda(), ps.parts)
        out[flags] = rec.finish().map.cpu().numpy().astype(np.float64)
    rp = ReconParams(size=32, box=32, ft_size=64, interp=0, sampling=1.5, hi_res=1.5, lo_res=0.0,
                     ctf_action=1, flags=abi.BSX_FLAG_RESCALE, wiener=0.0, nhalves=1, batch=0,
                     mem_cap=0)
    o = OracleRecon(rp, 2, 64)
    o.insert(synth.particles_list(ps.parts), imgs.numpy())
    m, _, _, _ = o.finish()
    err = np.abs(out[abi.BSX_FLAG_RESCALE] - m).max() / np.abs(m).max()
    say(f"  rescale flag: max rel err {err:.2e}")
    assert err <= 1e-4
    assert np.abs(out[0] - m).max() / np.abs(m).max() > 0.01


def test_output_rescale_and_sinc(gpu):
    rng = np.random.default_rng(13)
    v = (rng.normal(size=(20, 24, 28)) * 0.7 + 3).astype(np.float32)
    t = torch.as_tensor(v).cuda()
    rescale_to_avg_std(t, 0.5, 2.0)
    got = t.cpu().numpy().astype(np.float64)
    ref = Oracle(64).rescale_to_avg_std(v.astype(np.float64), 0.5, 2.0)
    np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-6)
    assert got.mean() == pytest.approx(0.5, abs=1e-5) and got.std() == pytest.approx(2.0, rel=1e-5)
    for shape, org in (((20, 24, 28), (14, 12, 10)), ((1, 32, 32), (16, 16, 0))):
        w = rng.normal(size=shape).astype(np.float32)
        t = torch.as_tensor(w).cuda()
        filter_sinc(t, org)
        np.testing.assert_allclose(t.cpu().numpy(), Oracle(64).filter_sinc(w, org), rtol=2e-7,
                                   atol=1e-7)
    with pytest.raises(abi.BsxError):
        rescale_to_avg_std(torch.ones(8, device="cuda"), 0.0, -1.0)

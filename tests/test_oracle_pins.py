"""Pins of the oracle's insertion geometry, weighing, SNR figure of merit,
Friedel residual, per-particle and output rescaling and the trilinear sinc
correction against values worked out by hand (tests/pin_cases.py) and
against textbook statistics — one check per formula such that a plausible
mistake (a squared distance, a missing factor 1/2, a sign, the magnification
applied the wrong way) fails it."""
import math

import numpy as np
import pytest

import pin_cases as pc
from bsoft_amd import synth
from bsoft_amd.abi import ReconParams
from oracle.oracle import Oracle, OracleRecon


def oracle_weights(params: dict, part: np.ndarray, img: np.ndarray) -> np.ndarray:
    """Weight plane (z, y, x) of one particle inserted by the FP64 oracle."""
    o = OracleRecon(ReconParams(**pc.recon_params(**params)), 1, 64)
    o.insert(synth.particles_list(part), img)
    return o.partial_acc(0)[2]


@pytest.mark.parametrize("name", sorted(pc.insertion_cases()))
def test_insertion_weights_by_hand(name):
    params, part, expect = pc.insertion_cases()[name]
    W = oracle_weights(params, part, pc.image())
    for (x, y, z), w in expect.items():
        assert W[z, y, x] == pytest.approx(w, rel=1e-12, abs=1e-15), (name, (x, y, z))


def test_weighted_nn_total_is_below_nearest():
    """Weighted nearest neighbour only scales weights down (1 - |m - v| <= 1)
    and drops none that are positive: the totals differ by the sum of the
    distances, and each voxel of the weighted plane is at most the NN one."""
    cases = pc.insertion_cases()
    p_w, part, _ = cases["weighted_nn"]
    p_n, _, _ = cases["nearest_mag"]
    Ww = oracle_weights(p_w, part, pc.image())
    Wn = oracle_weights(p_n, part, pc.image())
    assert np.all(Ww <= Wn + 1e-15)
    assert Ww.sum() < Wn.sum()


def test_value_follows_weight():
    """The complex sums carry w * F for the same sample: at a voxel fed by
    one sample (weighted NN, voxel (2, 0, 0), sample (3, 0)) sum(wF) / sum(w)
    equals the staged transform value of that sample."""
    params, part, _ = pc.insertion_cases()["weighted_nn"]
    rp = ReconParams(**pc.recon_params(**params))
    img = pc.image(3)
    o = OracleRecon(rp, 1, 64)
    o.insert(synth.particles_list(part), img)
    acc = o.partial_acc(0)
    F, _ = Oracle(64).stage_particle(rp, synth.particles_list(part)[0], img[0])
    v = (acc[0, 0, 0, 2] + 1j * acc[1, 0, 0, 2]) / acc[2, 0, 0, 2]
    assert v == pytest.approx(F[0, 3], rel=1e-12)


# ------------------------------------------------------------------ finish
def test_fom_against_sample_statistics():
    """SNR FOM of hand-built sums equals the closed form in the weighted mean
    and reliability-weighted variance of each voxel's contributions; the DC
    voxel is 1 and voxels with total weight <= 1 are 0."""
    n = 8
    planes, samples = pc.fom_accumulators(n, seed=5)
    rp = ReconParams(**pc.recon_params(size=n, box=n, ft_size=2 * n, flags=16))
    o = OracleRecon(rp, 1, 64)
    o.set_partial_acc(0, planes)
    _, fom, _, st = o.finish(want_fom=True)
    fom = fom.ravel()
    exp = np.array([pc.fom_expected(x, w) for x, w in samples])
    exp[0] = 1.0
    np.testing.assert_allclose(fom, exp, rtol=1e-10, atol=1e-12)
    assert np.count_nonzero(planes[2].ravel() > 1) > 50        # the formula is exercised
    assert np.count_nonzero([np.sum(w) > 1 and np.sum(w) - np.sum(w * w) / np.sum(w) <= 1
                             for x, w in samples]) > 5          # and its W / w2 branch
    assert st.coverage == np.count_nonzero(planes[2] > 1e-30)


def test_weighed_transform_is_weighted_mean():
    """fspace_reconstruction_weigh: F = sum(wF) / sum(w) where sum(w) > 1e-30,
    else 0 (checked through the centre-shifted transform: (-1)^(h+k+l))."""
    n = 8
    planes, samples = pc.fom_accumulators(n, seed=6)
    rp = ReconParams(**pc.recon_params(size=n, box=n, ft_size=2 * n, flags=16))
    o = OracleRecon(rp, 1, 64)
    o.set_partial_acc(0, planes)
    _, _, ft, _ = o.finish(want_ft=True)
    sign = (-1.0) ** np.add.outer(np.add.outer(np.arange(n), np.arange(n)), np.arange(n))
    exp = np.array([np.sum(w * x) / np.sum(w) if np.sum(w) > 1e-30 else 0 for x, w in samples])
    np.testing.assert_allclose((ft * sign).ravel(), exp, rtol=1e-12, atol=1e-14)


def test_friedel_residual_by_hand(oracle64):
    F, R = pc.friedel_volume(8)
    assert oracle64.friedel(F) == pytest.approx(R, rel=1e-12)
    # one pair alone: R = |u - conj(t)|; a self-mate voxel (DC) alone: 2 |Im|
    G = np.zeros_like(F)
    G[0, 0, 2], G[0, 0, 6] = 2 + 1j, 1 + 0.25j
    assert oracle64.friedel(G) == pytest.approx(abs((2 + 1j) - np.conj(1 + 0.25j)), rel=1e-12)
    D = np.zeros_like(F)
    D[0, 0, 0] = 3 - 0.75j
    assert oracle64.friedel(D) == pytest.approx(1.5, rel=1e-12)
    # a Friedel-symmetric transform (of a real volume) has R = 0
    real = np.fft.fftn(np.random.default_rng(1).normal(size=(8, 8, 8)))
    assert oracle64.friedel(real) < 1e-12


# ------------------------------------------------------------ rescaling
def test_particle_rescale_statistics(oracle64):
    """-rescale: the particle becomes average 0, standard deviation 1 (FP32
    storage), an affine map of the input."""
    img = (np.random.default_rng(2).normal(size=(16, 16)) * 3.5 + 7).astype(np.float32)
    r = oracle64.rescale_f32(img).astype(np.float64)
    assert abs(r.mean()) < 1e-6 and r.std() == pytest.approx(1.0, abs=1e-6)
    a = np.polyfit(img.ravel().astype(np.float64), r.ravel(), 1)
    assert a[0] == pytest.approx(1 / img.astype(np.float64).std(), rel=1e-6)
    # a constant image has deviation 0: scale 1, shift -average
    c = oracle64.rescale_f32(np.full((4, 4), 2.5, np.float32))
    assert np.all(c == 0)


def test_rescaled_particle_stages_as_rescaled_image(oracle64):
    """Staging with the rescale flag equals staging the rescaled image
    without it (the background is taken after the rescale)."""
    rp = ReconParams(**pc.recon_params(flags=1))
    rp0 = ReconParams(**pc.recon_params())
    part = synth.particles_list(pc.particle())[0]
    img = (np.random.default_rng(4).normal(size=(8, 8)) * 2 + 3).astype(np.float32)
    a, _ = oracle64.stage_particle(rp, part, img)
    b, _ = oracle64.stage_particle(rp0, part, oracle64.rescale_f32(img))
    np.testing.assert_array_equal(a, b)
    c, _ = oracle64.stage_particle(rp0, part, img)
    assert np.abs(a - c).max() > 0.1


def test_output_rescale(oracle64):
    d = np.random.default_rng(3).normal(size=(6, 6, 6)) * 0.3 - 2
    r = oracle64.rescale_to_avg_std(d, 1.5, 0.25)
    assert r.mean() == pytest.approx(1.5, abs=1e-12) and r.std() == pytest.approx(0.25, rel=1e-12)
    assert np.corrcoef(d.ravel(), r.ravel())[0, 1] == pytest.approx(1.0, abs=1e-12)


def test_filter_sinc_closed_form(oracle64):
    """(r / sin r)^2 with r = pi |d|: at d = (1/2, 0, 0) the factor is
    (pi/2)^2; along a diagonal d = (1/4, 1/4, 1/4): r = pi sqrt(3)/4; the
    origin voxel is unchanged."""
    n = 8
    d = np.ones((n, n, n))
    out = oracle64.filter_sinc(d, (4, 4, 4))
    assert out[4, 4, 4] == 1.0
    assert out[4, 4, 0] == pytest.approx((math.pi / 2) ** 2, rel=1e-14)    # x = 0: d = -1/2
    r = math.pi * math.sqrt(3) / 4
    assert out[6, 6, 6] == pytest.approx((r / math.sin(r)) ** 2, rel=1e-14)
    # a 2D image: z extent 1, origin z 0
    img = np.full((1, n, n), 2.0)
    o2 = oracle64.filter_sinc(img, (4, 4, 0))
    assert o2[0, 4, 6] == pytest.approx(2 * ((math.pi / 4) / math.sin(math.pi / 4)) ** 2, rel=1e-14)

"""Pins of the FP64 projection-matching restatement (oracle/search_ref.py) and
of the host-side view grid of libbsx (no GPU needed)."""
import math

import numpy as np
import pytest

from bsoft_amd import search, synth
import ctypes as C

from bsoft_amd import abi
from bsoft_amd.abi import Particle
from oracle import search_ref as sr


def test_disk_points_bruteforce():
    for R in (2, 5, 16):
        hk, rs = sr.disk_points(R)
        assert rs[0] == 0 and rs[1] == 1 and tuple(hk[0]) == (0, 0)
        pts = {(h, k) for h in range(-R - 1, R + 2) for k in range(-R - 1, R + 2)
               if math.isqrt(h * h + k * k) < R and (k > 0 or (k == 0 and h > 0))}
        assert set(map(tuple, hk[1:])) == pts and len(hk) == len(pts) + 1
        for j in range(R):
            r = np.floor(np.hypot(hk[rs[j]:rs[j + 1], 0], hk[rs[j]:rs[j + 1], 1]))
            assert np.all(r == j)
        # every full-plane point of the disk is a half-plane point or its mate
        full = {(h, k) for h in range(-R, R + 1) for k in range(-R, R + 1)
                if math.isqrt(h * h + k * k) < R}
        assert full == pts | {(-h, -k) for h, k in pts} | {(0, 0)}


def _part(def_avg=1.5e4, def_dev=0.0, ast=0.0, ori=(0.0, 0.0)):
    p = Particle()
    p.view[:] = (0, 0, 1, 0)
    p.ori[:] = ori
    p.hdr_ori[:] = ori
    p.pixel_size, p.mag, p.volt, p.cs = 1.5, 1.0, 3e5, 2.7e7
    p.amp_shift = math.asin(0.07)
    p.def_avg, p.def_dev, p.ast_angle = def_avg, def_dev, ast
    p.sel = 1.0
    return p


@pytest.mark.parametrize("R", [6, 15])
def test_frc_half_plane_equals_reference_full_plane(oracle64, R):
    """The half-plane ring sums give the reference FRC of full planes (the C
    oracle's img_recip_space_fom with the CTF multiplied into the section)."""
    rng = np.random.default_rng(R)
    n, pix = 32, 1.5
    img = rng.normal(size=(n, n))
    ref = 0.7 * img + rng.normal(size=(n, n))
    part = _part(def_avg=1.2e4, def_dev=800.0, ast=0.4, ori=(n // 2, n // 2))
    hk, rs = sr.disk_points(R)
    pv, pc = sr.particle_disk(img, part, hk, oracle64)
    sv, _ = sr.particle_disk(ref, part, hk, oracle64, ctf=False)
    # full planes with the same phase shift, CTF on the section
    F = np.fft.fft2(img) / n
    G = np.fft.fft2(ref) / n
    h = np.fft.fftfreq(n, 1 / n)
    ph = np.exp(2j * np.pi * (h[None, :] * part.ori[0] + h[:, None] * part.ori[1]) / n)
    fs = 1.0 / (n * pix)
    ctf = np.array([[oracle64.ctf_value(part, x * fs, y * fs) for x in h] for y in h])
    f_full = sr.frc_full(F * ph, G * ph * ctf, n, pix, R, oracle64)
    f_half = sr.frc_half(pv, pc, sv, rs)
    assert f_half == pytest.approx(f_full, abs=1e-12)
    assert -1 <= f_half <= 1


def test_frc_and_coarse_self_match(oracle64):
    rng = np.random.default_rng(3)
    n, R = 32, 12
    hk, rs = sr.disk_points(R)
    img = rng.normal(size=(n, n))
    v, _ = sr.particle_disk(img, _part(), hk, oracle64, ctf=False)
    ones = np.ones(len(v))
    assert sr.frc_half(v, ones, v, rs) == pytest.approx(1.0, abs=1e-12)
    assert sr.coarse_score(v, ones, v, rs) == pytest.approx(1.0, abs=1e-12)
    assert sr.coarse_score(v, ones, -v, rs) == pytest.approx(-1.0, abs=1e-12)
    # with a CTF the coarse score of the CTF-modulated section against the
    # particle multiplied by the same CTF is |corr| of the products: still 1
    c = np.cos(np.arange(len(v)) * 0.37)
    assert sr.coarse_score(v, c, v * c, rs) == pytest.approx(1.0, abs=1e-12)
    # a sign flip in one ring lowers the mean by 2/(R-1)
    w = v.copy()
    w[rs[3]:rs[4]] *= -1
    assert sr.coarse_score(v, ones, w, rs) == pytest.approx(1 - 2 / (R - 1), abs=1e-12)


def test_particle_disk_shift_theorem(oracle64):
    """An impulse at the particle origin becomes a constant after the phase
    shift to the origin (phase_shift_to_origin)."""
    n = 32
    img = np.zeros((n, n))
    img[13, 9] = 1.0
    hk, _ = sr.disk_points(10)
    v, _ = sr.particle_disk(img, _part(ori=(9, 13)), hk, oracle64, ctf=False)
    np.testing.assert_allclose(v, np.full(len(v), 1.0 / n), atol=1e-14)


def test_views_c1_matches_oracle(oracle64):
    v = search.views_c1(4.0)
    ov = oracle64.asu_views_c1(math.radians(4), math.radians(4), 1)
    assert v.shape == ov.shape == (2566, 4)
    np.testing.assert_allclose(v, ov, atol=1e-15)
    # in-plane expansion (view_list_expand_angles): 90 angles of 4 degrees
    ve = search.views_c1(4.0, angle_step_deg=4.0)
    assert len(ve) == 2566 * 90
    np.testing.assert_allclose(ve[90 * 7 + 11, :3], v[7, :3])
    assert ve[90 * 7 + 11, 3] == pytest.approx(v[7, 3] + math.radians(44), abs=1e-12)
    m = search.view_matrices(v[::37])
    om = np.array([oracle64.view_matrix(x) for x in v[::37]])
    np.testing.assert_allclose(m, om, atol=1e-15)


def test_rotation_angle():
    rng = np.random.default_rng(5)
    q = synth.uniform_quaternions(20, rng)
    m = synth.view_to_matrix(synth.quat_to_view(q))
    assert np.allclose(search.rotation_angle_deg(m, m), 0, atol=1e-5)
    th = math.radians(3.0)
    rz = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    assert np.allclose(search.rotation_angle_deg(m, m @ rz), 3.0, atol=1e-9)


def test_cnn_oracle_gradient_finite_differences():
    """The numpy backward of the proposal network equals central finite
    differences of its loss (pins oracle/cnn_ref.py)."""
    from oracle import cnn_ref
    S, ch, ncls, B = 8, (4, 4, 8, 4), 5, 3
    off, n, feat, chans = cnn_ref.layout(S, ch, ncls)
    rng = np.random.default_rng(0)
    p = rng.normal(size=n) * 0.4
    x = rng.normal(size=(B, S, S))
    y = rng.integers(0, ncls, B)
    L, g, _ = cnn_ref.loss_and_grad(p, x, y, S, ch, ncls)
    idx = np.concatenate([rng.choice(np.arange(o, ob + c), size=6, replace=False)
                          for o, ob, c in off])
    h = 1e-6
    for i in idx:
        e = np.zeros(n)
        e[i] = h
        lp = cnn_ref.loss_and_grad(p + e, x, y, S, ch, ncls)[0]
        lm = cnn_ref.loss_and_grad(p - e, x, y, S, ch, ncls)[0]
        assert g[i] == pytest.approx((lp - lm) / (2 * h), rel=1e-5, abs=1e-8)


def test_cnn_oracle_pair_gradient_finite_differences():
    """The contrastive loss of oracle/cnn_ref.py: its gradient equals central
    finite differences, and on embeddings where every particle equals its
    reference and the pairs are orthogonal the loss is the closed form
    -log(e^(1/tau) / (e^(1/tau) + B - 1))."""
    from oracle import cnn_ref
    S, ch, d, B, tau = 8, (4, 4, 8, 4), 6, 3, 0.5
    off, n, feat, chans = cnn_ref.layout(S, ch, d)
    rng = np.random.default_rng(1)
    p = rng.normal(size=n) * 0.4
    x = rng.normal(size=(2 * B, S, S))
    L, g, _ = cnn_ref.pair_loss_and_grad(p, x, B, tau, S, ch, d)
    idx = np.concatenate([rng.choice(np.arange(o, ob + c), size=6, replace=False)
                          for o, ob, c in off])
    h = 1e-6
    for i in idx:
        e = np.zeros(n)
        e[i] = h
        lp = cnn_ref.pair_loss_and_grad(p + e, x, B, tau, S, ch, d)[0]
        lm = cnn_ref.pair_loss_and_grad(p - e, x, B, tau, S, ch, d)[0]
        assert g[i] == pytest.approx((lp - lm) / (2 * h), rel=1e-5, abs=1e-8)
    z = np.concatenate([np.eye(B, 5) * 3.0, np.eye(B, 5) * 0.5])   # same directions, other norms
    losses, dz = cnn_ref.pair_loss(z, B, tau)
    exact = -math.log(math.exp(1 / tau) / (math.exp(1 / tau) + B - 1))
    np.testing.assert_allclose(losses, exact, rtol=1e-12)


def test_matrix_view_round_trip():
    rng = np.random.default_rng(9)
    v = synth.uniform_views(500, rng)
    v = np.concatenate([v, [[0, 0, 1, 0.3], [0, 0, -1, -1.2], [1, 0, 0, 2.0]]])
    m = search.view_matrices(v)
    back = search.matrix_views(m)
    np.testing.assert_allclose(search.view_matrices(back), m, atol=1e-12)
    ok = np.abs(v[:, 2]) < 0.999
    np.testing.assert_allclose(back[ok], v[ok], atol=1e-9)


def _fit_case(oracle64, n=32, R=12, centre=(16.0, 15.0), d_true=1.5e4, ctf=True):
    """A section G (transform of a random image) and the particle disk values
    G * CTF(d_true) moved to `centre`: the FOM is exactly 1 at (centre, d_true)."""
    rng = np.random.default_rng(8)
    hk, rs = sr.disk_points(R)
    G = np.fft.fft2(rng.normal(size=(n, n))) / n
    sv = G[hk[:, 1] % n, hk[:, 0] % n]
    c = sr._ctf_at(_part(def_avg=d_true), hk, n, d_true, ctf, oracle64)
    h, k = hk[:, 0], hk[:, 1]
    v0 = sv * c * np.exp(-2j * np.pi * (h * centre[0] + k * centre[1]) / n)
    return hk, rs, sv, v0


def test_fit_grid_recovers_origin_and_defocus(oracle64):
    """Start half a pixel off in x and y and two defocus steps off: the scan
    lands on the true defocus and the 0.5-pixel grid on the true origin."""
    n, cx, cy, dt, sd = 32, 16.0, 15.0, 1.5e4, 700.0
    hk, rs, sv, v0 = _fit_case(oracle64, n, 12, (cx, cy), dt)
    part = _part(def_avg=dt + 2 * sd, ori=(cx + 0.5, cy - 0.5))
    (ox, oy), d, f, num = sr.fit_grid(v0, part, sv, hk, rs, n, oracle64, 1.0, 0.5, sd)
    assert d == pytest.approx(dt, rel=1e-12)
    assert (ox, oy) == (pytest.approx(cx, abs=1e-12), pytest.approx(cy, abs=1e-12))
    assert f == pytest.approx(1.0, abs=1e-9)
    assert num >= 11 + 2 * 9


def test_fit_grid_centre_never_moves(oracle64):
    """The grids stay centred on the starting origin and the result is the
    best of the last grid: from centre + (1, 0) with steps 1 then 0.5 the fit
    returns centre + (0.5, 0), after 4 grids (each step's grid twice: the best
    is not the centre, then the repeat leaves the FOM unchanged)."""
    n, cx, cy, dt = 32, 16.0, 15.0, 1.5e4
    hk, rs, sv, v0 = _fit_case(oracle64, n, 12, (cx, cy), dt)
    part = _part(def_avg=dt, ori=(cx + 1, cy))
    (ox, oy), d, f, num = sr.fit_grid(v0, part, sv, hk, rs, n, oracle64, 1.0, 0.5, 0.0)
    assert (ox, oy) == (cx + 0.5, cy) and d == dt and num == 4 * 9
    assert f < 1 - 1e-3


def test_fit_grid_defocus_scan_without_ctf(oracle64):
    """def_std > 0 and a nonzero defocus start the scan even with the CTF
    flag off; the FOM does not change with the defocus, so the first value,
    def - 5 def_std, is kept (strict comparison)."""
    n, cx, cy, dt, sd = 32, 16.0, 15.0, 1.5e4, 500.0
    hk, rs, sv, v0 = _fit_case(oracle64, n, 12, (cx, cy), dt, ctf=False)
    part = _part(def_avg=dt, ori=(cx, cy))
    _, d, _, num = sr.fit_grid(v0, part, sv, hk, rs, n, oracle64, 1.0, 1.0, sd, ctf_flag=False)
    assert d == dt - 5 * sd and num == 11 + 9


def _monte_draws(rows):
    """draws (max_iter, 5) from (g, a, ux, uy, uniform) rows."""
    return np.array(rows, dtype=np.float64)


def test_fit_monte_moves_start_from_the_kept_state(oracle64):
    """Positive FOMs are never kept (prefom starts at -1), so every move
    starts again from the staged origin and defocus: from (cx + 0.5, cy) at
    dt + sd, a first move by (+0.5, 0) and +sd, then one by (-0.5, 0) and
    -sd lands exactly on (cx, cy, dt), where the FOM is 1. A walk that kept
    the first move would end at (cx + 0.5, cy, dt + sd) instead."""
    n, cx, cy, dt, sd = 32, 16.0, 15.0, 1.5e4, 700.0
    hk, rs, sv, v0 = _fit_case(oracle64, n, 12, (cx, cy), dt)
    part = _part(def_avg=dt + sd, ori=(cx + 0.5, cy))
    dr = _monte_draws([(0, 0, 1, 0, 0.5), (1, 1, 1, 0, 0.5), (-1, 1, -1, 0, 0.5)])
    (ox, oy), d, f = sr.fit_monte(v0, part, sv, hk, rs, n, oracle64, 3, sd, 0.5, dr)
    assert (ox, oy) == (cx, cy) and d == dt
    assert f == pytest.approx(1.0, abs=1e-9)


def test_fit_monte_negative_fom_walk(oracle64):
    """With the section negated the FOM is negative, fom / prefom > 0 = the
    uniform value keeps every move, so the walk accumulates: the result is
    the point of the cumulative path with the highest (least negative) FOM,
    found here by scoring every point of the path directly."""
    n, cx, cy, dt = 32, 16.0, 15.0, 1.5e4
    hk, rs, sv, v0 = _fit_case(oracle64, n, 12, (cx, cy), dt, ctf=False)
    part = _part(def_avg=dt, ori=(cx, cy))
    rng = np.random.default_rng(3)
    m = 12
    dr = np.zeros((m, 5))
    dr[:, 1] = rng.standard_normal(m)
    ang = rng.uniform(0, 2 * np.pi, m)
    dr[:, 2], dr[:, 3] = np.cos(ang), np.sin(ang)
    (ox, oy), d, f = sr.fit_monte(v0, part, -sv, hk, rs, n, oracle64, m, 0.0, 0.3, dr,
                                  ctf_flag=False)
    h, k = hk[:, 0], hk[:, 1]
    path, x, y = [], cx, cy
    for i in range(m):
        if i > 0:
            amp = 0.3 * dr[i, 1]
            x, y = x + dr[i, 2] * amp, y + dr[i, 3] * amp
        v = v0 * np.exp(2j * np.pi * (h * x + k * y) / n)
        path.append((sr.frc_half(v, np.ones(len(hk)), -sv, rs), x, y))
    assert all(p[0] < 0 for p in path)
    best = max(path, key=lambda p: p[0])
    assert (ox, oy) == (best[1], best[2]) and f == best[0] and d == dt


def test_refine_monte_view_steps_start_from_the_kept_view(oracle64):
    """A section whose FOM falls as cos(5 (a - a_true)) with the view angle
    a: from a_true + 0.01, a first angle step of +0.01 (worse, not kept, as
    no positive FOM is) and a second of -0.01 from the staged view land on
    a_true exactly, FOM 1. A walk that kept the first step would end at
    a_true + 0.01."""
    n, cx, cy, dt, at = 32, 16.0, 15.0, 1.5e4, 0.25
    hk, rs, sv, v0 = _fit_case(oracle64, n, 12, (cx, cy), dt, ctf=False)
    part = _part(def_avg=dt, ori=(cx, cy))

    def section(view):
        return sv * np.exp(1j * 5 * (view[3] - at))

    dr = np.array([(0, 0, 0, 1, 0.5, 0.5), (0, 0, 0, 1, 1.0, 0.5), (0, 0, 0, 1, 0.0, 0.5)])
    inner = np.zeros((3, 3, 5))
    inner[..., 2] = 1.0
    view, (ox, oy), d, f = sr.refine_monte(v0, part, section, hk, rs, n, oracle64,
                                           [0, 0, 1, at + 0.01], 3, 0.0, 0.0, 0.0, 0.01, dr,
                                           inner, ctf_flag=False)
    assert view[3] == pytest.approx(at, abs=1e-15) and (ox, oy) == (cx, cy) and d == dt
    assert f == pytest.approx(1.0, abs=1e-9)


# ------------------------------------------------- views_for_refinement grid
def _grid_closed_form(view, step):
    """The 27 views from the geometry alone: with v the unit view vector,
    ax1 = unit(v x z) (x when v is along z) and ax2 = v x ax1, the turned
    vector is cos a1 cos a2 v + cos a1 sin a2 ax1 - sin a1 ax2 (a turn by a1
    about ax1, then by a2 about ax2) and the in-plane angle is angle + a3."""
    v = np.asarray(view[:3], dtype=float)
    v /= np.linalg.norm(v)
    ax1 = np.cross(v, [0.0, 0.0, 1.0])
    if np.linalg.norm(ax1) < 1e-3:
        ax1 = np.array([1.0, ax1[1], ax1[2]])
    ax1 /= np.linalg.norm(ax1)
    ax2 = np.cross(v, ax1)
    out = []
    for a1 in (-step, 0.0, step):
        for a2 in (-step, 0.0, step):
            vec = (math.cos(a1) * math.cos(a2) * v + math.cos(a1) * math.sin(a2) * ax1
                   - math.sin(a1) * ax2)
            for a3 in (-step, 0.0, step):
                out.append(np.concatenate([vec, [view[3] + a3]]))
    return np.asarray(out)


@pytest.mark.parametrize("view", [(0, 0, 1, 0.3), (0.3, -0.5, 0.8, 2.9), (0, 0, -1, -3.0),
                                  (0.6, 0.8, 0, -1.2), (-0.2, 0.1, 0.97, 3.1)])
@pytest.mark.parametrize("deg", [0.5, 2.0, 7.5])
def test_views_for_refinement_closed_form(view, deg):
    step = math.radians(deg)
    g = sr.views_for_refinement(view, step)
    assert g.shape == (27, 4)
    cf = _grid_closed_form(view, step)
    np.testing.assert_allclose(g[:, :3], cf[:, :3], atol=1e-12)
    # in-plane angles: angle + a3 brought to (-pi, pi]
    d = np.remainder(g[:, 3] - cf[:, 3] + math.pi, 2 * math.pi) - math.pi
    np.testing.assert_allclose(d, 0, atol=1e-12)
    assert np.all(np.abs(g[:, 3]) <= math.pi + 1e-12)
    # the centre is the normalised view; its vector has unit length
    v = np.asarray(view[:3], float) / np.linalg.norm(view[:3])
    np.testing.assert_allclose(g[13, :3], v, atol=1e-15)
    # the library's grid (bsx_views_for_refinement) equals the restatement
    out = np.zeros(108)
    assert abi.lib().bsx_views_for_refinement((C.c_double * 4)(*view), step,
                                              out.ctypes.data_as(C.POINTER(C.c_double))) == 0
    np.testing.assert_allclose(out.reshape(27, 4), g, atol=1e-15)


def test_views_for_refinement_along_z():
    """View (0, 0, 1): ax1 = x, ax2 = y, so entry (a1, a2) carries the vector
    (cos a1 sin a2, -sin a1, cos a1 cos a2)."""
    s = math.radians(3)
    g = sr.views_for_refinement((0, 0, 1, 0), s)
    assert np.allclose(g[0 * 9 + 2 * 3 + 1, :3], (math.cos(-s) * math.sin(s), math.sin(s),
                                                  math.cos(s) * math.cos(s)), atol=1e-15)
    assert np.allclose(g[2 * 9 + 1 * 3 + 0, :3], (0, -math.sin(s), math.cos(s)), atol=1e-15)
    assert g[2 * 9 + 1 * 3 + 0, 3] == pytest.approx(-s)


def _cos_score(target):
    """Score of views: cosine of the rotation angle to the target rotation."""
    T = synth.view_to_matrix(np.asarray(target, float)[None])[0]
    def f(views):
        M = synth.view_to_matrix(views)
        return (np.einsum("nij,ij->n", M, T) - 1) / 2
    return f


def test_refine_views_loop_rules():
    """img_refine_grid's view loop: a score peaked at the start view only
    contracts (2, 1, 0.5 degrees: three passes above a 0.4-degree accuracy);
    a smooth score with its peak 3 degrees away is climbed to within the
    accuracy."""
    start = (0.2, -0.3, 0.93, 0.7)
    v, f, it = sr.refine_views(_cos_score(start), start, math.radians(2), math.radians(0.4))
    assert it == 3 and f == pytest.approx(1.0)
    target = synth.view_to_matrix(np.asarray(start)[None])[0]
    # the peak: the start rotated by 3 degrees about a skew axis
    ax = np.array([0.3, 0.8, -0.5]) / np.linalg.norm([0.3, 0.8, -0.5])
    a = math.radians(3)
    K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    Rt = (np.eye(3) + math.sin(a) * K + (1 - math.cos(a)) * K @ K) @ target
    tv = search.matrix_views(Rt[None])[0]
    v, f, it = sr.refine_views(_cos_score(tv), start, math.radians(2), math.radians(0.1))
    err = search.rotation_angle_deg(synth.view_to_matrix(v[None]), Rt[None])[0]
    assert err < 0.2 and it > 3

"""Point-group symmetry views (bsx_symmetry_views): host code of libbsx, no GPU.

The generated views are pinned by group structure: with R_k the matrix of
the k-th related view of a base view R_0, the set {R_k R_0^T} must be a
rotation group of the point group's order that contains the identity and the
generator rotations, and is closed under products. The oracle restatement
(oracle/sym_ref.py) must agree with the library to 1e-12."""
import ctypes as C
import math

import numpy as np
import pytest

from bsoft_amd import abi, search
from bsoft_amd.recon import symmetry_views
from oracle import sym_ref

ORDERS = {"C1": 1, "C2": 2, "C7": 7, "D2": 4, "D5": 10, "T": 12, "O": 24, "I": 60, "I90": 60,
          "T-3": 4, "O-3": 8, "O-4": 6, "O-2": 12, "I-3": 20, "I-5": 12, "I-2": 30,
          "532": 60, "432": 24, "23": 12, "3": 3}


def _rot(axis, ang):
    a = np.asarray(axis, float)
    a = a / np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K


def _contains(mats, m, tol=1e-9):
    return np.min(np.abs(mats - m).reshape(len(mats), -1).max(axis=1)) < tol


@pytest.mark.parametrize("label,order", sorted(ORDERS.items()))
def test_counts_and_oracle(label, order):
    rng = np.random.default_rng(5)
    v = rng.normal(size=(3, 4))
    out = symmetry_views(label, v)
    assert out.shape == (3, order, 4)
    for i in range(3):
        ref = np.asarray(sym_ref.all_views(label, v[i]))
        np.testing.assert_allclose(out[i], ref, rtol=0, atol=1e-12)


@pytest.mark.parametrize("label", ["C4", "D3", "T", "O", "I", "I90"])
def test_views_form_the_point_group(label):
    # full point groups only: the variants that drop an operator (O-2, I-5,
    # ...) enumerate the reference's products of the remaining operators,
    # which need not be closed
    rng = np.random.default_rng(11)
    base = np.concatenate([rng.normal(size=3), [rng.uniform(-3, 3)]])
    views = symmetry_views(label, base[None])[0]
    R = search.view_matrices(views)
    R0 = R[0]
    # the first view is the input (normalised)
    n = base[:3] / np.linalg.norm(base[:3])
    np.testing.assert_allclose(views[0, :3], n, atol=1e-12)
    S = np.einsum("kij,lj->kil", R, R0)          # R_k R_0^T
    for s in S:
        np.testing.assert_allclose(s @ s.T, np.eye(3), atol=1e-9)
        assert abs(np.linalg.det(s) - 1) < 1e-9
    assert _contains(S, np.eye(3))
    # distinct elements, closed under products
    d = np.abs(S[:, None] - S[None]).reshape(len(S), len(S), -1).max(axis=2)
    assert np.all(d + np.eye(len(S)) * 10 > 1e-6)
    for a in S:
        for b in S:
            assert _contains(S, a @ b)
    # the generators of the label are in the group
    for axis, order in sym_ref.operators(label):
        assert _contains(S, _rot(axis, 2 * math.pi / order))


def test_cyclic_views_rotate_about_z():
    """C_n: the related planes are the base plane rotated about z, so their
    view vectors keep z and turn (x, y) by multiples of 2 pi / n."""
    base = np.array([0.3, -0.5, 0.8, 0.4])
    v = symmetry_views("C6", base[None])[0]
    n = base[:3] / np.linalg.norm(base[:3])
    for k in range(6):
        ang = math.atan2(v[k, 1], v[k, 0]) - math.atan2(n[1], n[0])
        assert abs(v[k, 2] - n[2]) < 1e-12
        assert min(abs(math.remainder(ang - j * math.pi / 3, 2 * math.pi)) for j in range(6)) < 1e-9


def test_bad_labels():
    for bad in ("H10,20", "X5", "C0"):
        with pytest.raises(abi.BsxError):
            symmetry_views(bad, np.array([[0, 0, 1, 0.0]]))


# ------------------------------------------------ real-space symmetrization
def _ball_map(n, seed, radius):
    """Random map, zero outside `radius` of the centre (so the background and
    the fill are 0 and no rotation reaches a nonzero voxel outside the map)."""
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, n, n)).astype(np.float32)
    z, y, x = np.meshgrid(*(np.arange(n) - n // 2,) * 3, indexing="ij")
    v[x * x + y * y + z * z > radius * radius] = 0
    return v


def _closure(gens):
    G = [np.eye(3, dtype=np.int64)]
    grew = True
    while grew:
        grew = False
        for a in list(G):
            for g in gens:
                p = g @ a
                if not any(np.array_equal(p, b) for b in G):
                    G.append(p)
                    grew = True
    return G


_Z4 = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]])
_Z2 = _Z4 @ _Z4
_X2 = np.diag([1, -1, -1])
_C3 = np.array([[0, 0, 1], [1, 0, 0], [0, 1, 0]])      # 3-fold about (1, 1, 1)
GRID_GROUPS = {"C2": [_Z2], "C4": [_Z4], "D2": [_Z2, _X2], "D4": [_Z4, _X2],
               "T": [_C3, _Z2, _X2], "O": [_C3, _Z4, _X2]}


@pytest.mark.parametrize("label", sorted(GRID_GROUPS))
def test_symmetrize_oracle_is_the_group_average(label):
    """Rotations that permute the grid: the symmetrized map is the exact
    average of f(g (x - o) + o) over the group's elements."""
    n = 16
    f = _ball_map(n, 4, 6)
    o = n // 2
    G = _closure(GRID_GROUPS[label])
    assert len(G) == {"C2": 2, "C4": 4, "D2": 4, "D4": 8, "T": 12, "O": 24}[label]
    out = sym_ref.symmetrize(f, label, (o, o, o))
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    u = np.stack([x - o, y - o, z - o])
    ref = np.zeros((n, n, n))
    for g in G:
        w = np.einsum("ij,jxyz->ixyz", g, u) + o
        ok = np.all((w >= 0) & (w < n), axis=0)
        ref += np.where(ok, f[np.clip(w[2], 0, n - 1), np.clip(w[1], 0, n - 1),
                              np.clip(w[0], 0, n - 1)], 0.0)
    ref /= len(G)
    np.testing.assert_allclose(out, ref, rtol=0, atol=1e-6 * np.abs(f).max())


@pytest.mark.parametrize("label", ["C3", "C5", "I"])
def test_symmetrize_oracle_linear_maps(label):
    """Trilinear interpolation reproduces a linear map exactly, so inside a
    ball that every rotation keeps in the map the result is b + a . P (x - o),
    P the average of the group's rotations: the projection on z for C_n, 0
    for I."""
    n, o = 24, 12
    a = np.array([0.3, -0.2, 0.5])
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    f = (1.5 + a[0] * (x - o) + a[1] * (y - o) + a[2] * (z - o)).astype(np.float32)
    out = sym_ref.symmetrize(f, label, (o, o, o))
    r2 = (x - o) ** 2 + (y - o) ** 2 + (z - o) ** 2
    inner = r2 <= 8 * 8
    ref = 1.5 + (a[2] * (z - o) if label[0] == "C" else 0.0)
    np.testing.assert_allclose(out[inner], np.broadcast_to(ref, out.shape)[inner], atol=2e-5)


def test_background_is_the_outside_mean():
    n = 10
    v = np.zeros((n, n, n), np.float32)
    v[0, 0, 0] = 8.0                 # a corner: outside the inscribed sphere
    v[5, 5, 5] = 100.0               # the centre: inside
    z, y, x = np.meshgrid(*(np.arange(n) - 5,) * 3, indexing="ij")
    nout = int(np.sum((x / 5.0) ** 2 + (y / 5.0) ** 2 + (z / 5.0) ** 2 > 1))
    assert sym_ref.background(v, (5, 5, 5)) == pytest.approx(8.0 / nout)


def test_random_symmetric_views():
    """Mode 2: each particle gets one of its related views; over many draws
    every related view is picked."""
    from bsoft_amd.recon import random_symmetric_views
    base = np.array([[0.2, 0.3, 0.9, 0.5]] * 400)
    allv = symmetry_views("D3", base[:1])[0]
    pick = random_symmetric_views("D3", base, np.random.default_rng(0))
    idx = [int(np.argmin(np.abs(allv - p).sum(axis=1))) for p in pick]
    for p, i in zip(pick, idx):
        np.testing.assert_allclose(p, allv[i], atol=1e-12)
    assert set(idx) == set(range(6))


@pytest.mark.gpu
@pytest.mark.parametrize("label", ["C1", "C3", "D5", "T", "O", "I", "I90"])
def test_gpu_symmetrize_matches_oracle(label):
    import torch
    from bsoft_amd.recon import symmetrize
    n = 24
    rng = np.random.default_rng(7)
    f = rng.normal(size=(n, n, n)).astype(np.float32) + 0.25   # nonzero background
    o = (n // 2, n // 2 - 1, n // 2 + 0.5)
    print(f"symmetrize {label} n={n} origin={o}", flush=True)
    out = symmetrize(torch.as_tensor(f).cuda(), label, o).cpu().numpy()
    ref = sym_ref.symmetrize(f, label, o)
    np.testing.assert_allclose(out, ref, rtol=0, atol=1e-6 * np.abs(f).max())


def test_symmetrize_oracle_2d_images():
    """A square image: C4 about the centre is the exact average of the four
    quarter turns (np.rot90 about pixel n/2 of an image that is zero near its
    border); C3 of a linear image leaves its mean b inside a disk that every
    rotation keeps in the image."""
    n, o = 16, 8
    rng = np.random.default_rng(12)
    f = rng.normal(size=(n, n)).astype(np.float32)
    y, x = np.meshgrid(np.arange(n) - o, np.arange(n) - o, indexing="ij")
    f[x * x + y * y > 36] = 0
    out = sym_ref.symmetrize(f, "C4", (o, o))
    g = np.zeros((n + 1, n + 1))
    g[:n, :n] = f                     # pad so that pixel n/2 is the rotation centre
    ref = sum(np.rot90(g[: 2 * o + 1, : 2 * o + 1], q) for q in range(4))[:n, :n] / 4
    np.testing.assert_allclose(out, ref, atol=1e-6 * np.abs(f).max())
    lin = (2.0 + 0.3 * x - 0.2 * y).astype(np.float32)
    out = sym_ref.symmetrize(lin, "C3", (o, o))
    inner = x * x + y * y <= 25
    np.testing.assert_allclose(out[inner], 2.0, atol=2e-5)


@pytest.mark.gpu
@pytest.mark.parametrize("label", ["C1", "C3", "C6"])
def test_gpu_symmetrize_2d_matches_oracle(label):
    import torch
    from bsoft_amd.recon import symmetrize
    n = 32
    f = np.random.default_rng(8).normal(size=(n, n)).astype(np.float32) + 0.5
    o = (n // 2, n // 2 - 0.5)
    out = symmetrize(torch.as_tensor(f[None]).cuda(), label, o).cpu().numpy()[0]
    np.testing.assert_allclose(out, sym_ref.symmetrize(f, label, o), rtol=0,
                               atol=1e-6 * np.abs(f).max())


def test_symmetrize_2d_rejects_non_cyclic_groups():
    L = abi.lib()
    org = (C.c_double * 2)(8, 8)
    buf = C.c_void_p(0x1000)
    assert L.bsx_symmetrize_2d(buf, C.c_void_p(0x2000), 16, org, b"D2", None) == abi.BSX_ERR_ARG
    assert b"cyclic" in L.bsx_last_error()

"""Point-group symmetry views (bsx_symmetry_views): host code of libbsx, no GPU.

The generated views are pinned by group structure: with R_k the matrix of
the k-th related view of a base view R_0, the set {R_k R_0^T} must be a
rotation group of the point group's order that contains the identity and the
generator rotations, and is closed under products. The oracle restatement
(oracle/sym_ref.py) must agree with the library to 1e-12."""
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

"""Plain-Python restatement of Bsoft's point-group symmetry views.

TEST INFRASTRUCTURE ONLY: the checker of bsx_symmetry_views (tests/). The
generated views are pinned by group properties in tests/test_symmetry.py
(closure, order, identity, generators), not by retyping this code.

Follows:
  Bsymmetry::Bsymmetry            src/util/symmetry.cpp:42-122 (operators)
  symmetry_get_all_views          src/util/symmetry.cpp:277-319
  View::quaternion                include/View.h:131-153
  View(Quaternion&)               include/View.h:52-63
  Quaternion(axis, angle), *      include/Quaternion.h:34-41, 117-124
  Bimage::symmetrize              src/img/Bimage_symmetry.cpp:31-100 (symmetrize)
  Bimage::transform               src/img/Bimage_transform.cpp:138-214
  Bimage::interpolate             src/img/Bimage.cpp:1436-1484
  Bimage::calculate_background    src/img/Bimage_background.cpp:27-80
  Matrix3(axis, angle)            include/Matrix3.h:21-42, 73-76
The real-space symmetrization is pinned in tests/test_symmetry.py by exact
group averages where the rotations permute the grid (C2, C4, D2, D4, T, O)
and by linear maps, whose trilinear interpolation is exact (C3, C5, I).
"""
from __future__ import annotations

import math

import numpy as np

GOLDEN = 1.61803398874989484820458683436563811772


def operators(label: str):
    """(unit axis, order) list of a point-group label (C<n>, D<n>, T, O, I
    and their -2/-3/-4/-5/90 variants)."""
    lbl = "".join(label.split()) or "C1"
    if lbl[0].isdigit():
        j = int("".join(ch for ch in lbl if ch.isdigit())[:9] or 0)
        if "532" in lbl and "90" in lbl:
            lbl = "I90"
        elif "532" in lbl:
            lbl = "I"
        elif "432" in lbl:
            lbl = "O"
        elif "23" in lbl:
            lbl = "T"
        else:
            lbl = f"C{j}" if j > 0 else "C1"
    lbl = lbl[0].upper() + lbl[1:]
    ops = []

    def add(x, y, z, order):
        n = math.sqrt(x * x + y * y + z * z)
        ops.append(((x / n, y / n, z / n), order))

    if lbl[0] == "C":
        add(0, 0, 1, 1 if lbl in ("Cs", "CS") else int(lbl[1:]))
    elif lbl[0] == "D":
        add(0, 0, 1, int(lbl[1:]))
        add(1, 0, 0, 2)
    elif lbl[0] == "T":
        if "-3" not in lbl:
            add(1, 1, 1, 3)
        add(0, 0, 1, 2)
        add(1, 0, 0, 2)
    elif lbl[0] == "O":
        if "-3" not in lbl:
            add(1, 1, 1, 3)
        if "-4" in lbl:
            add(1, -1, 0, 2)
        elif "-2" in lbl:
            add(0, 0, 1, 4)
        else:
            add(0, 0, 1, 4)
            add(1, 0, 0, 2)
    elif lbl[0] == "I":
        if "-3" not in lbl:
            add(1, 1, 1, 3)
        if "-5" in lbl:
            add(1, 0, 0, 2)
        elif "-2" in lbl:
            add(1, -1, 0, 2)
        else:
            add(1, 1 / GOLDEN, GOLDEN, 2)
        if "-5" not in lbl:
            add(1 / GOLDEN, 1, 0, 5)
        if "-2" not in lbl:
            add(0, 0, 1, 2)
        if "90" in lbl:
            ops = [((-a[1], a[0], a[2]), o) for a, o in ops]
    else:
        raise ValueError(f"not a point group: {label}")
    return ops


def _normalize(v):
    x, y, z, a = v
    s = math.sqrt(x * x + y * y + z * z)
    if s < 1e-30:
        x, y, z = 0.0, 0.0, 1.0
    else:
        x, y, z = x / s, y / s, z / s
    x, y, z = (0.0 if abs(c) < 1e-30 else c for c in (x, y, z))
    a = math.remainder(a, 2 * math.pi)
    return [x, y, z, 0.0 if abs(a) < 1e-30 else a]


def _quat(v):
    x, y, z, a = _normalize(v)
    ca, sa = math.cos(a / 2), math.sin(a / 2)
    q = [ca, 0.0, 0.0, sa]
    if z < 1 - 1e-30:
        if z <= 1e-30 - 1:
            q = [0.0, sa, ca, 0.0]
        else:
            f1, f2 = math.sqrt((z + 1) / 2), math.sqrt(1 / (2 * (z + 1)))
            q = [f1 * ca, f2 * (x * sa - y * ca), f2 * (x * ca + y * sa), f1 * sa]
    return [0.0 if abs(c) < 1e-30 else c for c in q]


def _view(q):
    ang = math.atan2(q[3], q[0])
    ca, sa = math.cos(ang), math.sin(ang)
    q2 = q[0] * q[0] + q[3] * q[3]
    f2 = 2 * math.sqrt(q2)
    if q2 < 1e-30:
        ang = math.atan2(q[1], q[2])
    return _normalize([f2 * (q[2] * ca + q[1] * sa), f2 * (q[2] * sa - q[1] * ca), 2 * q2 - 1,
                       2 * ang])


def _qmul(q, p):
    return [q[0] * p[0] - (q[1] * p[1] + q[2] * p[2] + q[3] * p[3]),
            q[0] * p[1] + q[1] * p[0] + q[2] * p[3] - q[3] * p[2],
            q[0] * p[2] + q[2] * p[0] + q[3] * p[1] - q[1] * p[3],
            q[0] * p[3] + q[3] * p[0] + q[1] * p[2] - q[2] * p[1]]


def all_views(label: str, view):
    """symmetry_get_all_views: the view first, then the rotated copies."""
    out = [_normalize(list(view))]
    nview = 1
    for (ax, order) in operators(label):
        for j in range(1, order):
            a = j * 2 * math.pi / order / 2
            q = [math.cos(a), ax[0] * math.sin(a), ax[1] * math.sin(a), ax[2] * math.sin(a)]
            q = [0.0 if abs(c) < 1e-30 else c for c in q]
            for k in range(nview):
                out.append(_view(_qmul(q, _quat(out[k]))))
        nview *= order
    return out


def axis_angle_matrix(axis, angle):
    """Matrix3(axis, angle): unit quaternion (components below 1e-30 zeroed),
    rotation matrix, entries below 1e-12 zeroed and clamped to [-1, 1]."""
    ax = np.asarray(axis, dtype=np.float64)
    ax = ax / math.sqrt(float(ax @ ax))
    a = angle / 2
    q = [math.cos(a)] + [float(c) * math.sin(a) for c in ax]
    q = [q[0]] + [0.0 if abs(c) < 1e-30 else c for c in q[1:]]
    M = np.array([
        [q[0] * q[0] + q[1] * q[1] - q[2] * q[2] - q[3] * q[3],
         2 * q[1] * q[2] - 2 * q[0] * q[3], 2 * q[1] * q[3] + 2 * q[0] * q[2]],
        [2 * q[1] * q[2] + 2 * q[0] * q[3],
         q[0] * q[0] - q[1] * q[1] + q[2] * q[2] - q[3] * q[3], 2 * q[2] * q[3] - 2 * q[0] * q[1]],
        [2 * q[1] * q[3] - 2 * q[0] * q[2], 2 * q[2] * q[3] + 2 * q[0] * q[1],
         q[0] * q[0] - q[1] * q[1] - q[2] * q[2] + q[3] * q[3]]])
    M[np.abs(M) < 1e-12] = 0
    return np.clip(M, -1, 1)


def background(vol: np.ndarray, origin) -> float:
    """Mean of the voxels outside the ellipsoid with semi-axes N/2 (integer)
    about origin (x, y, z); a single plane (z extent 1) has no z term."""
    zn, n = vol.shape[0], vol.shape[-1]
    h = float(n // 2)
    z, y, x = np.meshgrid(np.arange(zn), np.arange(n), np.arange(n), indexing="ij")
    zt = ((z - origin[2]) / float(zn // 2)) ** 2 if zn > 1 else 0.0
    r = ((x - origin[0]) / h) ** 2 + ((y - origin[1]) / h) ** 2 + zt - 1
    out = r > 0
    return float(vol[out].astype(np.float64).mean()) if out.any() else 0.0


def interpolate(vol: np.ndarray, xx, yy, zz, fill: float) -> np.ndarray:
    """Bimage::interpolate at arrays of positions: the (up to) 2 x 2 x 2
    neighbours inside the map, each axis weight toggled once per pass of its
    loop as the reference's loops do, normalised by the weight sum."""
    zn, n = vol.shape[0], vol.shape[-1]
    flat = vol.astype(np.float64).ravel()
    inside = (xx >= 0) & (xx < n) & (yy >= 0) & (yy < n) & (zz >= 0) & (zz < zn)
    xs, ys, zs = (np.where(inside, c, 0.0) for c in (xx, yy, zz))
    ix, iy, iz = xs.astype(np.int64), ys.astype(np.int64), zs.astype(np.int64)
    nx, ny = (np.where(n < i + 2, 1, 2) for i in (ix, iy))
    nz = np.where(zn < iz + 2, 1, 2)
    fx, fy, fz = xs - ix, ys - iy, zs - iz
    value = np.zeros_like(xs)
    ws = np.zeros_like(xs)
    for zk in range(2):
        zon = zk < nz
        fz = np.where(zon, 1.0 - fz, fz)
        for yk in range(2):
            yon = zon & (yk < ny)
            fy = np.where(yon, 1.0 - fy, fy)
            wyz = fy * fz
            for xk in range(2):
                on = yon & (xk < nx)
                fx = np.where(on, 1.0 - fx, fx)
                w = fx * wyz
                idx = ((iz + zk) * n + (iy + yk)) * n + (ix + xk)
                v = flat[np.where(on, idx, 0)]
                ws = np.where(on, ws + w, ws)
                value = np.where(on, value + v * w, value)
    with np.errstate(divide="ignore", invalid="ignore"):
        res = np.where(ws != 0, value / np.where(ws != 0, ws, 1), fill)
    return np.where(inside, res, fill)


def symmetrize(vol: np.ndarray, label: str, origin) -> np.ndarray:
    """Bimage::symmetrize(sym, View(), 1) of a real N^3 map (z, y, x) or
    N x N image (y, x; cyclic groups, origin (x, y)): per
    operator the map plus its order - 1 rotations about origin (each rotated
    map samples R^T (x - o) + o, rounded to FP32, added in FP32), with the
    background of the input as the fill; finally x 1/order in FP32."""
    if np.ndim(vol) == 2:                         # a 2D image
        return symmetrize(np.asarray(vol)[None], label, origin)[0]
    ops = operators(label)
    order = 1
    for _, m in ops:
        order *= m
    cur = np.asarray(vol, dtype=np.float32).copy()
    if all(m == 1 for _, m in ops):
        return cur
    zn, n = cur.shape[0], cur.shape[-1]           # zn = 1: a 2D image
    o = [float(c) for c in origin] + [0.0] * (3 - len(origin))
    fill = background(cur, o)
    z, y, x = np.meshgrid(np.arange(zn), np.arange(n), np.arange(n), indexing="ij")
    u = [x - o[0], y - o[1], z - o[2]]
    for axis, m in ops:
        if m == 1:
            continue
        acc = cur.copy()
        for j in range(1, m):
            M = axis_angle_matrix(axis, j * math.pi * 2.0 / m)
            old = [M[0, i] * u[0] + M[1, i] * u[1] + M[2, i] * u[2] + o[i] for i in range(3)]
            pj = interpolate(cur, old[0], old[1], old[2], fill).astype(np.float32)
            acc = (acc.astype(np.float64) + pj.astype(np.float64)).astype(np.float32)
        cur = acc
    return (cur.astype(np.float64) * (1.0 / order)).astype(np.float32)

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
"""
from __future__ import annotations

import math

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

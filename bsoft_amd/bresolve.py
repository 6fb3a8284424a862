"""bresolve on libbsx: FSC and DPR between two real maps (src/bresolve.cpp).

    python -m bsoft_amd.bresolve [options] map1.mrc map2.mrc

Options (a subset of bresolve's, same meaning):
  -resolution A          resolution limit of the curves (default: Nyquist)
  -sampling A            voxel size (default: from the first map)
  -fsccut a[,b,c,d]      FSC thresholds to report (default 0.3)
  -dprcut a[,b,c,d]      DPR thresholds to report (default: none)
  -ratio r               radial to Cartesian sampling ratio (default 1)
  -tsv file              write the curves (s, FSC, DPR) as tab-separated values

The curves come from bsx_fsc_dpr (Bimage::fsc_dpr, Bimage_resolution.cpp:
93-240); the thresholds and the information volume radius are read off them
as Bplot::resolution_display does (src/util/Bplot.cpp:181-231,
include/Bplot.h:269-289).
"""
from __future__ import annotations

import argparse
import ctypes as C
import math
import sys

import numpy as np


def fsc_dpr(map1, map2, sampling: float, hi_res: float = 0.0, sampling_ratio: float = 1.0):
    """Bimage::fsc_dpr of two real cubic maps or two square images (torch
    tensors on the device; an image may carry a leading axis of 1): arrays
    s (1/A), FSC and DPR (degrees), one row per shell."""
    from .abi import call
    from .recon import _ptr, _stream
    n = int(map1.shape[-1])
    if tuple(map1.shape) != tuple(map2.shape):
        raise ValueError("two maps of the same size are required")
    shp = tuple(map1.shape)
    if shp == (n, n, n) and n > 1:
        fn = "bsx_fsc_dpr"
    elif shp in ((n, n), (1, n, n)):
        fn = "bsx_fsc_dpr_2d"
    else:
        raise ValueError("cubic maps or square images are required")
    # contiguous copies bound to names so that both stay allocated (and
    # distinct) until the call has returned
    a, b = map1.contiguous(), map2.contiguous()
    cap = C.c_int(4 * n + 8)
    out = np.zeros((cap.value, 3))
    call(fn, _ptr(a), _ptr(b), n, float(sampling),
         float(hi_res), float(sampling_ratio), out.ctypes.data_as(C.POINTER(C.c_double)),
         C.byref(cap), _stream(None))
    out = out[: cap.value]
    return out[:, 0], out[:, 1], out[:, 2]


def cut(s: np.ndarray, y: np.ndarray, threshold: float, direction: int = 1) -> float:
    """Bplot::cut (include/Bplot.h:269-289): spatial frequency where the
    curve crosses the threshold (direction < 0: falling through it), by
    linear interpolation; the last crossing wins."""
    val = (s[0] + s[1]) / 2
    nr = len(s)
    for i in range(nr - 1):
        if direction < 0:
            if y[i] >= threshold and y[i + 1] < threshold:
                val = s[i] + (s[i + 1] - s[i]) * (threshold - y[i]) / (y[i + 1] - y[i])
        elif y[i] <= threshold and y[i + 1] > threshold:
            val = s[i] + (s[i + 1] - s[i]) * (threshold - y[i]) / (y[i + 1] - y[i])
    if val < s[1]:   # no crossing: the loop's final indices, i.e. the last row
        if direction < 0:
            if y[nr - 1] > threshold:
                val = s[nr - 1]
        elif y[nr - 1] < threshold:
            val = s[nr - 1]
    return val


def information_volume_radius(s: np.ndarray, fsc: np.ndarray, three_d: bool = True) -> float:
    """Bplot::resolution_display's information volume radius (Bplot.cpp:197-211)."""
    i = np.arange(len(s), dtype=np.float64)
    w = 2 * math.pi * i
    if three_d:
        w = w * 2.0 * i
    tot = float(np.sum(fsc * w))
    ivr = (4.0 * math.pi / (3.0 * tot)) ** (1.0 / 3.0) if three_d else math.sqrt(math.pi / tot)
    return ivr / s[1]


def _floats(text: str):
    return [float(v) for v in text.split(",") if v.strip()]


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="bresolve", prefix_chars="-")
    ap.add_argument("-resolution", type=float, default=0.0)
    ap.add_argument("-sampling", type=float, default=0.0)
    ap.add_argument("-fsccut", default="0.3")
    ap.add_argument("-dprcut", default="")
    ap.add_argument("-ratio", type=float, default=1.0)
    ap.add_argument("-tsv", default="")
    ap.add_argument("maps", nargs=2)
    a = ap.parse_args(sys.argv[1:] if argv is None else argv)
    import torch
    from .formats import read_mrc
    m1, m2 = read_mrc(a.maps[0]), read_mrc(a.maps[1])
    v1, v2 = np.asarray(m1.data[0], np.float32), np.asarray(m2.data[0], np.float32)
    if v1.shape != v2.shape:
        print("bresolve: the maps must have the same size", file=sys.stderr)
        return 1
    sampling = a.sampling if a.sampling > 0 else float(m1.sampling[0])
    s, fsc, dpr = fsc_dpr(torch.as_tensor(v1).cuda(), torch.as_tensor(v2).cuda(), sampling,
                          a.resolution, a.ratio)
    print("Resolution curves:")
    print("s(1/A)\tRes(A)\tFSC\tDPR")
    for i in range(len(s)):
        res = 1 / s[i] if i else 999.0
        print(f"{s[i]:7.5f}\t{res:7.3f}\t{fsc[i]:7.3f}\t{dpr[i]:7.3f}")
    print("\nResolution estimates:")
    for c in _floats(a.fsccut):
        if c:
            print(f"FSC({c:g}): {1 / cut(s, fsc, c, -1):g} A")
    for c in _floats(a.dprcut):
        if c:
            print(f"DPR({c:g}): {1 / cut(s, dpr, c, 1):g} A")
    print(f"Information volume radius: "
          f"{information_volume_radius(s, fsc, three_d=v1.shape[0] > 1):g} A")
    if a.tsv:
        np.savetxt(a.tsv, np.stack([s, fsc, dpr], axis=1), delimiter="\t")
    return 0


if __name__ == "__main__":
    sys.exit(main())

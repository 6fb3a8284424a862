"""Builds libbsx.so in-tree with hipcc for gfx950.

    python -m bsoft_amd.build            # build (incremental per source)
    python -m bsoft_amd.build --force    # rebuild everything
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(CSRC, "_build")
LIB = os.path.join(HERE, "libbsx.so")
SOURCES = ["bsx_fft.hip", "bsx_recon.hip", "bsx_extract.hip", "bsx_search.hip",
           "bsx_cnn.hip"]
HEADERS = ["bsx_internal.h", "bsx_fft_dev.h", "bsx_common.h"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("BSX_ARCH", "gfx950")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=off", "-Wall", "-Wno-unused-result",
]


def _newer(src: list[str], dst: str) -> bool:
    if not os.path.exists(dst):
        return True
    t = os.path.getmtime(dst)
    return any(os.path.getmtime(s) > t for s in src if os.path.exists(s))


def _compile(src: str, op: str, verbose: bool) -> None:
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", op]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {os.path.basename(src)}:\n{err}")
    if verbose and err.strip():
        print(err)


def build(force: bool = False, verbose: bool = False) -> str:
    """Compiles the HIP sources (in parallel, one hipcc per source) and links
    libbsx.so; without `force` only sources newer than their objects."""
    os.makedirs(OBJ, exist_ok=True)
    hdrs = [os.path.join(CSRC, h) for h in HEADERS] + [os.path.join(HERE, "..", "include", "bsx.h")]
    objs, todo = [], []
    for src in SOURCES:
        sp = os.path.join(CSRC, src)
        if not os.path.exists(sp):
            continue
        op = os.path.join(OBJ, src.replace(".hip", ".o"))
        objs.append(op)
        if force or _newer([sp] + hdrs, op):
            todo.append((sp, op))
    jobs = max(1, min(len(todo), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(_compile, sp, op, verbose) for sp, op in todo]:
            f.result()
    if force or _newer(objs, LIB):
        tmp = LIB + ".tmp"
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed:\n{r.stderr}")
        os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(force="--force" in sys.argv, verbose=True))

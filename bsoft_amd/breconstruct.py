"""breconstruct on libbsx: reconstruct a map from the particles of a Bsoft
STAR project (src/breconstruct.cpp:71-397 for the options and their
defaults).

    python -m bsoft_amd.breconstruct -reconstruction map.mrc [options] project.star

Options (breconstruct's, same meaning):
  -reconstruction file   output map (MRC)
  -resolution A          high resolution limit (default: the sampling)
  -sampling A            map voxel size (default: the particle pixel size)
  -size x[,y,z]          map size (default: the particle box)
  -pad n                 padding factor of the particle transforms (default 2)
  -CTF action            CTF treatment, parsed as the reference's
                         ctf_action (include/options.h:418-442): a number or
                         flip, apply, correct, wiener, baseline, baseflip,
                         basecorr, envelope, ...; flip (1) flips phases, every
                         other action applies the CTF, as a Wiener filter
                         when a Wiener factor is set (actions above 2 take
                         0.2 by default, breconstruct.cpp:187)
  -wiener w              Wiener factor (at least 0.01)
  -rescale avg,std       rescale each particle to average 0 and standard
                         deviation 1 before insertion, and each output map
                         to avg, std
  -interpolate nearest|weighted|trilinear
                         (trilinear maps get the sinc^2 correction,
                         Bimage::filter_sinc, before they are written)
  -symmetry label        point group (C<n>, D<n>, T, O, I, ...; default C1)
  -mode 0|1|2            symmetry mode: 0 insert each particle at all its
                         related views (default), 1 insert once and
                         symmetrize the map in real space, 2 insert once at a
                         randomly picked related view (-seed n, default 1)
  -bootstrap             weight particles by their selection number
  -ewald                 Ewald-sphere insertion
  -TwoD                  a 2D reconstruction (one plane, size x size x 1)
  -fullmap / -halfmaps   the full map, the two half maps, or both (both
                         flags); half maps are written as <name>_01, <name>_02
  -fom file              also write the figure-of-merit map of each output
                         map (file, file_01, file_02)
  -threads n             partial maps of the reference's partition (at least
                         2; made even for half maps): only the half sets
                         depend on it, and for an even count they alternate
                         over the selected particles
  -verbose               print the Fourier shell statistics of the full map
                         (fspace_reconstruction_stats); this accumulates the
                         power and squared-weight sums too, through the
                         slower direct-atomic insertion
"""
from __future__ import annotations

import argparse
import os
import sys

import numpy as np


def ctf_action_of(value: str) -> int:
    """Bsoft's option parser for -CTF (include/options.h:418-442): the number
    given, overridden by the last matching keyword."""
    v = value.strip()
    try:
        action = int(float(v))
    except ValueError:
        action = 0
    v = v.lower()
    for key, act in (("fli", 1), ("app", 2), ("cor", 3), ("wie", 4), ("bas", 5), ("baseline2", 6),
                     ("basef", 7), ("basec", 8), ("env", 9)):
        if key in v:
            action = act
    if "app" in v and "env" in v:
        action = 10
    if "prep" in v:
        action = 11
    if "fit" in v:
        action = 12
    if "prep" in v and "fit" in v:
        action = 13
    return action


def parse_args(argv):
    ap = argparse.ArgumentParser(prog="breconstruct", prefix_chars="-")
    ap.add_argument("-reconstruction", required=True)
    ap.add_argument("-resolution", type=float, default=0.0)
    ap.add_argument("-sampling", type=float, default=0.0)
    ap.add_argument("-size", default="")
    ap.add_argument("-pad", type=int, default=2)
    ap.add_argument("-CTF", default="0")
    ap.add_argument("-wiener", type=float, default=None)
    ap.add_argument("-rescale", default="")
    ap.add_argument("-interpolate", "-interpolation", dest="interpolate", default="nearest")
    ap.add_argument("-bootstrap", action="store_true")
    ap.add_argument("-ewald", action="store_true")
    ap.add_argument("-TwoD", action="store_true")
    ap.add_argument("-symmetry", default="C1")
    ap.add_argument("-mode", type=int, default=0, choices=[0, 1, 2])
    ap.add_argument("-seed", type=int, default=1)
    ap.add_argument("-fullmap", action="store_true")
    ap.add_argument("-halfmaps", action="store_true")
    ap.add_argument("-fom", default="")
    ap.add_argument("-threads", type=int, default=1)
    ap.add_argument("-verbose", action="store_true")
    ap.add_argument("project")
    return ap.parse_args(argv)


def _suffixed(path: str, imap: int) -> str:
    """<stem>_%02d.<ext> of breconstruct's output loop (breconstruct.cpp:360-374)."""
    stem, ext = os.path.splitext(path)
    return f"{stem}_{imap:02d}{ext or '.mrc'}" if imap else path


def main(argv=None) -> int:
    a = parse_args(sys.argv[1:] if argv is None else argv)
    import torch
    from . import abi
    from .formats import load_particle_images, project_particles, write_mrc
    from .recon import (Reconstruction, filter_sinc, part_ft_size, random_symmetric_views,
                        rescale_to_avg_std, symmetrize)

    # options (breconstruct.cpp:96-166)
    wiener = 0.0
    if a.wiener is not None:
        if a.wiener < 0.000001:
            print("-wiener: A Wiener factor must be specified!", file=sys.stderr)
        else:
            wiener = max(a.wiener, 0.01)
    rescale = None
    if a.rescale:
        vals = [float(v) for v in a.rescale.split(",") if v.strip()]
        if len(vals) < 2:
            print("-rescale: Both average and standard deviation must be specified!",
                  file=sys.stderr)
        elif vals[1] <= 0:
            print("-rescale: A positive standard deviation must be specified!", file=sys.stderr)
        else:
            rescale = (vals[0], vals[1])
    interp = {"n": 0, "w": 1, "t": 2}.get(a.interpolate[:1], 0)
    ctf = ctf_action_of(a.CTF)
    if ctf > 2 and wiener < 0.01:                # breconstruct.cpp:187
        wiener = 0.2
    nmaps = (1 if a.fullmap else 0) + (2 if a.halfmaps else 0)
    nmaps = min(max(nmaps, 1), 3)

    parts, srcs = project_particles(a.project)
    keep = parts["sel"] > 0                      # project_selected_partlist
    parts = parts[keep]
    srcs = [s for s, k in zip(srcs, keep) if k]
    if len(parts) == 0:
        print("Error: No particles selected!", file=sys.stderr)
        return 1
    imgs, hdr_ori = load_particle_images(srcs)
    box = imgs.shape[-1]
    if imgs.shape[-2] != box:
        print("breconstruct: particles must be square", file=sys.stderr)
        return 1
    size = [int(v) for v in a.size.split(",")] if a.size else [box]
    size = size[0]
    sampling = a.sampling if a.sampling > 0 else float(parts["pixel_size"][0])
    # the background ellipse uses the file-header origin; the phase origin is
    # the particle origin, else the image origin, else the map origin
    parts["hdr_ori"] = hdr_ori
    unset = parts["ori"][:, 0] <= 0
    parts["ori"][unset] = np.where(hdr_ori[unset, :1] > 0, hdr_ori[unset], size // 2)
    # Partition (breconstruct.cpp:204-205, mg_reconstruct.cpp:674, 741-744):
    # T >= 2 partial maps (at most half the particles), T made even when
    # half maps are wanted; the k-th selected particle goes to partial map
    # (k - 1) mod T, and half set j sums the partial maps of parity j, so
    # particle k is in half (k - 1) mod 2.
    T = max(a.threads, 2)
    T = min(T, max(len(parts) // 2, 1))
    if nmaps & 2 and T % 2:
        T += 1
    parts["half"] = (np.arange(len(parts)) % T) % 2
    flags = ((abi.BSX_FLAG_BOOTSTRAP if a.bootstrap else 0)
             | (abi.BSX_FLAG_EWALD if a.ewald else 0)
             | (abi.BSX_FLAG_RESCALE if rescale else 0)
             | (abi.BSX_FLAG_FOM if a.fom or a.verbose else 0)
             | (abi.BSX_FLAG_2D if a.TwoD else 0))
    ft = part_ft_size(max(box, size), 1.0, a.pad)
    rec = Reconstruction(size, box, ft, sampling, hi_res=a.resolution, interp=interp,
                         ctf_action=ctf, wiener=wiener, flags=flags,
                         nhalves=2 if nmaps > 1 else 1)
    # symmetry modes of particle_reconstruct (mg_reconstruct.cpp:175, 215-218)
    # and breconstruct.cpp:335
    if a.mode == 0:
        rec.set_symmetry(a.symmetry)
    elif a.mode == 2:
        parts["view"] = random_symmetric_views(a.symmetry, parts["view"],
                                               np.random.default_rng(a.seed))
    rec.insert(torch.as_tensor(imgs).cuda().contiguous(), parts)
    org = (size // 2, size // 2, 0 if a.TwoD else size // 2)
    for i in range(nmaps):
        # output map i: imap 1/2 are the half sets (breconstruct.cpp:355-374)
        imap = (i % 2 + 1) if nmaps == 2 else (i % 3 if nmaps == 3 else 0)
        if a.verbose and imap == 0:        # img_reconstruction_sum_weigh :346-347
            rows, (r_snr, r_fsc) = rec.stats(i, nmaps)
            print("Radius\ts(1/A)\tRes(A)\tWeight\tPower\tSignal\tNoise\tSNR\tFSC")
            for r in rows:
                print(f"{int(r[0])}\t{r[1]:.4g}\t{r[2]:.2f}\t{r[3]:.2f}\t{r[4]:.6g}\t"
                      f"{r[5]:.6g}\t{r[6]:.6g}\t{r[7]:.4g}\t{r[8]:.4g}")
            print(f"\nResolution estimates:\nSNR(0.5):                       {r_snr:g} A")
            print(f"FSC(0.3):                       {r_fsc:g} A\n")
        r = rec.finish(i, nmaps, want_fom=bool(a.fom))
        vol = r.map
        # after the inverse transform: symmetrize (mode 1), rescale, then the
        # sinc correction of trilinear maps (breconstruct.cpp:336-359)
        if a.mode == 1:
            vol = symmetrize(vol, a.symmetry, org[:2] if a.TwoD else org)
        if rescale:
            rescale_to_avg_std(vol, *rescale)
        if interp == 2:
            filter_sinc(vol, org)
        path = _suffixed(a.reconstruction, imap)
        write_mrc(path, vol.cpu().numpy(), (sampling,) * 3, org)
        if r.fom is not None:
            write_mrc(_suffixed(a.fom, imap), r.fom.cpu().numpy(), (sampling,) * 3, org)
        print(f"breconstruct: {path}: {r.particles} particles, coverage {r.coverage}, "
              f"Friedel residual {r.friedel:.4g}")
    return 0


if __name__ == "__main__":
    sys.exit(main())

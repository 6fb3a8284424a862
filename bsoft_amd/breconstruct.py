"""breconstruct on libbsx: reconstruct a map from the particles of a Bsoft
STAR project (src/breconstruct.cpp:71-397 for the options and their
defaults).

    python -m bsoft_amd.breconstruct -reconstruction map.mrc [options] project.star

Options (a subset of breconstruct's, same meaning):
  -reconstruction file   output map (MRC)
  -resolution A          high resolution limit (default: the sampling)
  -sampling A            map voxel size (default: the particle pixel size)
  -size x[,y,z]          map size (default: the particle box)
  -pad n                 padding factor of the particle transforms (default 2)
  -CTF flip|apply|wiener CTF treatment; -wiener w (default 0.2 for wiener)
  -interpolate nearest|weighted|trilinear
  -symmetry label        point group (C<n>, D<n>, T, O, I, ...; default C1)
  -mode 0|1|2            symmetry mode: 0 insert each particle at all its
                         related views (default), 1 insert once and
                         symmetrize the map in real space, 2 insert once at a
                         randomly picked related view (-seed n, default 1)
  -bootstrap             weight particles by their selection number
  -TwoD                  a 2D reconstruction (one plane, size x size x 1)
  -halfmaps              also write the two half maps (<name>_1, <name>_2)
  -fom                   also write the FOM map (<name>_fom)
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


def parse_args(argv):
    ap = argparse.ArgumentParser(prog="breconstruct", prefix_chars="-")
    ap.add_argument("-reconstruction", required=True)
    ap.add_argument("-resolution", type=float, default=0.0)
    ap.add_argument("-sampling", type=float, default=0.0)
    ap.add_argument("-size", default="")
    ap.add_argument("-pad", type=int, default=2)
    ap.add_argument("-CTF", default="none", choices=["none", "flip", "apply", "wiener"])
    ap.add_argument("-wiener", type=float, default=0.0)
    ap.add_argument("-interpolate", default="nearest", choices=["nearest", "weighted", "trilinear"])
    ap.add_argument("-bootstrap", action="store_true")
    ap.add_argument("-TwoD", action="store_true")
    ap.add_argument("-symmetry", default="C1")
    ap.add_argument("-mode", type=int, default=0, choices=[0, 1, 2])
    ap.add_argument("-seed", type=int, default=1)
    ap.add_argument("-halfmaps", action="store_true")
    ap.add_argument("-fom", action="store_true")
    ap.add_argument("-verbose", action="store_true")
    ap.add_argument("project")
    return ap.parse_args(argv)


def main(argv=None) -> int:
    a = parse_args(sys.argv[1:] if argv is None else argv)
    import torch
    from . import abi
    from .formats import load_particle_images, project_particles, write_mrc
    from .recon import Reconstruction, part_ft_size, random_symmetric_views, symmetrize

    parts, srcs = project_particles(a.project)
    keep = parts["sel"] > 0                      # project_selected_partlist
    parts = parts[keep]
    srcs = [s for s, k in zip(srcs, keep) if k]
    if len(parts) == 0:
        print("breconstruct: no selected particles", file=sys.stderr)
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
    # half sets alternate over the selected particles in project order, as
    # the round-robin partial maps of project_configure_for_reconstruction do
    parts["half"] = np.arange(len(parts)) % 2
    ctf = {"none": abi.BSX_CTF_NONE, "flip": abi.BSX_CTF_FLIP, "apply": abi.BSX_CTF_APPLY,
           "wiener": abi.BSX_CTF_WIENER}[a.CTF]
    interp = {"nearest": 0, "weighted": 1, "trilinear": 2}[a.interpolate]
    flags = ((abi.BSX_FLAG_BOOTSTRAP if a.bootstrap else 0)
             | (abi.BSX_FLAG_FOM if a.fom or a.verbose else 0)
             | (abi.BSX_FLAG_2D if a.TwoD else 0))
    ft = part_ft_size(max(box, size), 1.0, a.pad)
    rec = Reconstruction(size, box, ft, sampling, hi_res=a.resolution, interp=interp,
                         ctf_action=ctf, wiener=a.wiener, flags=flags,
                         nhalves=2 if a.halfmaps else 1)
    # symmetry modes of particle_reconstruct (mg_reconstruct.cpp:175, 215-218)
    # and breconstruct.cpp:335
    if a.mode == 0:
        rec.set_symmetry(a.symmetry)
    elif a.mode == 2:
        parts["view"] = random_symmetric_views(a.symmetry, parts["view"],
                                               np.random.default_rng(a.seed))
    rec.insert(torch.as_tensor(imgs).cuda().contiguous(), parts)
    stem, ext = os.path.splitext(a.reconstruction)
    outs = [(0, 3 if a.halfmaps else 1, a.reconstruction)]
    if a.halfmaps:
        outs += [(1, 3, f"{stem}_1{ext or '.mrc'}"), (2, 3, f"{stem}_2{ext or '.mrc'}")]
    for imap, nmaps, path in outs:
        if a.verbose and imap == 0:        # img_reconstruction_sum_weigh :346-347
            rows, (r_snr, r_fsc) = rec.stats(imap, nmaps)
            print("Radius\ts(1/A)\tRes(A)\tWeight\tPower\tSignal\tNoise\tSNR\tFSC")
            for r in rows:
                print(f"{int(r[0])}\t{r[1]:.4g}\t{r[2]:.2f}\t{r[3]:.2f}\t{r[4]:.6g}\t"
                      f"{r[5]:.6g}\t{r[6]:.6g}\t{r[7]:.4g}\t{r[8]:.4g}")
            print(f"\nResolution estimates:\nSNR(0.5):                       {r_snr:g} A")
            print(f"FSC(0.3):                       {r_fsc:g} A\n")
        r = rec.finish(imap, nmaps, want_fom=a.fom and imap == 0)
        org = (size // 2, size // 2, 0 if a.TwoD else size // 2)
        vol = r.map
        if a.mode == 1:
            vol = symmetrize(r.map, a.symmetry, org[:2] if a.TwoD else org)
        write_mrc(path, vol.cpu().numpy(), (sampling,) * 3, org)
        if r.fom is not None:
            write_mrc(f"{stem}_fom{ext or '.mrc'}", r.fom.cpu().numpy(), (sampling,) * 3, org)
        print(f"breconstruct: {path}: {r.particles} particles, coverage {r.coverage}, "
              f"Friedel residual {r.friedel:.4g}")
    return 0


if __name__ == "__main__":
    sys.exit(main())

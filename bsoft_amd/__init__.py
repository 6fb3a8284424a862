"""bsoft_amd — MI355X-native Fourier-space single-particle reconstruction.

A drop-in for Bsoft's breconstruct path (and its central-section extraction
mirror) built on hand-written gfx950 HIP kernels behind the C-ABI in
include/bsx.h (libbsx.so). See bsoft_amd.recon for the reference-named API.
"""
__version__ = "0.1.0"

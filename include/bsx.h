/*
 * bsx.h — C-ABI of the MI355X Fourier-space reconstruction library (libbsx.so).
 *
 * Every entry point returns an int status (BSX_OK or a negative BSX_ERR_*),
 * never throws, and leaves a message readable through bsx_last_error().
 * Pointers named *_dev are device (HBM) pointers; everything else is host.
 * Streams are hipStream_t passed as void* (NULL = default stream).
 *
 * Each entry point names the reference interface it replaces (paths are
 * relative to the Bsoft source tree, bsoft/).
 */
#ifndef BSX_H
#define BSX_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference returns NULL / negative int + error_show,
 *      src/util/utilities.cpp:745) ---- */
#define BSX_OK               0
#define BSX_ERR_ARG         -1   /* invalid argument / size */
#define BSX_ERR_DEVICE      -2   /* HIP runtime error */
#define BSX_ERR_NOMEM       -3   /* device allocation failed or over the cap */
#define BSX_ERR_STATE       -4   /* call out of order / unsupported mode */
#define BSX_ERR_NONFINITE   -5   /* non-finite particle data detected */

/* ---- CTF actions: include/options.h:418-442 (flip=1, apply=2, >2 Wiener) ---- */
#define BSX_CTF_NONE   0
#define BSX_CTF_FLIP   1
#define BSX_CTF_APPLY  2
#define BSX_CTF_WIENER 4

/* ---- reconstruction flags: src/mg/mg_reconstruct.cpp:83 (1 rescale,
 *      2 2D, 4 bootstrap, 8 Ewald) plus library-specific bits ---- */
#define BSX_FLAG_BOOTSTRAP   4
#define BSX_FLAG_EWALD       8
#define BSX_FLAG_FOM        16   /* also accumulate sum(w|F|^2) and sum(w^2) */
#define BSX_FLAG_CHECK_NAN  32   /* scan particles for non-finite values */
#define BSX_FLAG_TIMING     64   /* time the insertion kernel with HIP events
                                    (bsx_recon_insert then synchronises)   */

/* ---- interpolation: Bimage_reconstruct.cpp:25 (0 NN, 1 weighted NN, 2 trilinear) ---- */
#define BSX_INTERP_NEAREST   0
#define BSX_INTERP_WEIGHTED  1
#define BSX_INTERP_TRILINEAR 2

/* Reconstruction configuration (breconstruct options, src/breconstruct.cpp:71-266). */
typedef struct bsx_recon_params {
    int32_t size;          /* N: cubic map edge (voxels)                         */
    int32_t box;           /* particle image edge (pixels)                        */
    int32_t ft_size;       /* padded transform edge, see bsx_part_ft_size         */
    int32_t interp;        /* BSX_INTERP_*                                        */
    double  sampling;      /* map voxel size (A/voxel) = sam/scale                */
    double  hi_res;        /* high resolution limit (A); clamped as the reference */
    double  lo_res;        /* low resolution limit (A), 0 = none                  */
    int32_t ctf_action;    /* BSX_CTF_*                                           */
    int32_t flags;         /* BSX_FLAG_*                                          */
    double  wiener;        /* Wiener factor (used when ctf_action > 2)            */
    int32_t nhalves;       /* 1: one accumulator; 2: half-set accumulators        */
    int32_t batch;         /* particles per internal batch (0 = autotuned)        */
    int64_t mem_cap;       /* device-memory cap for this handle (bytes, 0 = none) */
    int32_t method;        /* 0: tile-binned insertion (FOM sums use 1); 1: direct atomics */
    int32_t order;         /* 0 auto, 1 input order, 2 plane-orientation order     */
    int32_t merge;         /* 0 auto, 1 merge equal-voxel runs before atomics, 2 off */
    int32_t reserved;
} bsx_recon_params;

/* One particle (Bparticle fields used on the path, include/mg_processing.h:101,
 * with the micrograph CTF already resolved per particle as particle_reconstruct
 * does, src/mg/mg_reconstruct.cpp:147,187). */
typedef struct bsx_particle {
    double  view[4];       /* view vector (x,y,z) and in-plane angle (rad)        */
    double  ori[2];        /* particle origin (pixels): phase origin              */
    double  hdr_ori[2];    /* image-header origin used for the background ellipse */
    double  pixel_size;    /* particle pixel size (A)                             */
    double  mag;           /* magnification (<=0 means 1)                         */
    double  volt;          /* acceleration voltage (V)                            */
    double  cs;            /* spherical aberration (A)                            */
    double  amp_shift;     /* amplitude-contrast phase shift (rad)                */
    double  def_avg;       /* defocus average (A, positive = underfocus)          */
    double  def_dev;       /* defocus deviation (A)                               */
    double  ast_angle;     /* astigmatism angle (rad)                             */
    double  sel;           /* selection number (bootstrap weight)                 */
    int32_t half;          /* half-set (0/1) when nhalves == 2                    */
    int32_t reserved;
} bsx_particle;

/* Finishing statistics (img_reconstruction_sum_weigh, mg_reconstruct.cpp:326-343). */
typedef struct bsx_stats {
    int64_t coverage;      /* voxels with sum(w) > 1e-30 (fspace_reconstruction_weigh) */
    double  friedel;       /* Friedel residual R (Bimage_fspace.cpp:2024)         */
    int64_t particles;     /* particles inserted into the returned map            */
    double  weight_sum;    /* total sum of weights (FP64)                         */
} bsx_stats;

typedef struct bsx_recon bsx_recon;   /* opaque handle: device accumulators */

/* Last error message of the calling thread (never NULL). */
const char* bsx_last_error(void);

/* Library / device information. */
int bsx_version(int* major, int* minor);
int bsx_device_info(char* name, int name_len, int* cu_count, int64_t* hbm_bytes);

/* ---- sizes ---- */
/* part_ft_size, src/mg/mg_reconstruct.cpp:38-64 (findNextPowerOf, utilities.cpp:447). */
int bsx_part_ft_size(int xsize, double scale, int pad_factor);

/* ---- reconstruction (breconstruct, src/breconstruct.cpp:204-349) ---- */
/* Allocates the FP32 accumulators (re, im, w [, sum w|F|^2, sum w^2]) in HBM.
 * Replaces `new Bimage(Float, TComplex, size, 1)` + check_resolution in
 * particle_reconstruct (mg_reconstruct.cpp:119-123). */
int bsx_recon_create(bsx_recon** h, const bsx_recon_params* p);
int bsx_recon_destroy(bsx_recon* h);
/* Zeroes the accumulators. */
int bsx_recon_reset(bsx_recon* h, void* stream);

/* Inserts n particle images (device, n x box x box FP32, row-major) into the
 * accumulators: background, pad, 2D FFT (x 1/ft), phase shift to origin, CTF
 * and central-section insertion — the per-particle loop of particle_reconstruct
 * (mg_reconstruct.cpp:143-229) with fspace_pack_2D / fspace_2D_interpolate
 * (Bimage_reconstruct.cpp:25-174). Asynchronous on `stream`. */
int bsx_recon_insert(bsx_recon* h, const float* imgs_dev, const bsx_particle* parts,
                     int64_t n, void* stream);

/* Sum, weigh (FP64 division), SNR figure-of-merit, centre shift, 3D inverse
 * FFT (x 1/sqrt(N^3)) and real part: img_reconstruction_sum_weigh
 * (mg_reconstruct.cpp:266-361) + fft_back (Bimage.h:903-908).
 * imap/nmaps follow the reference: nmaps 1 -> full; 2 -> imap 0/1 halves;
 * 3 -> imap 0 full, 1/2 halves. map_dev: N^3 FP32 (device). fom_dev: optional
 * N^3 FP32 FOM map (needs BSX_FLAG_FOM), may be NULL. ft_dev: optional N^3
 * complex64 copy of the weighed, centre-shifted transform, may be NULL.
 * stats may be NULL. Synchronises `stream`. */
int bsx_recon_finish(bsx_recon* h, int imap, int nmaps, float* map_dev, float* fom_dev,
                     float* ft_dev, bsx_stats* stats, void* stream);

/* Copies raw accumulator planes of one half-set to a device buffer laid out as
 * planes[5][N^3] FP32: sum(w re), sum(w im), sum(w), sum(w|F|^2), sum(w^2)
 * (planes 3-4 are zero without BSX_FLAG_FOM). For verification. */
int bsx_recon_accumulators(bsx_recon* h, int half, float* planes_dev, void* stream);

/* Device bytes held by a handle. */
int bsx_recon_device_bytes(const bsx_recon* h, int64_t* bytes);

/* Kernel timings of the previous bsx_recon_insert (needs BSX_FLAG_TIMING):
 * insertion kernels (count, scan, fill, tile accumulate or direct), staging
 * kernels (background, row transforms) in ms, and the number of batches. */
int bsx_recon_last_timing(const bsx_recon* h, double* insert_ms, int64_t* launches,
                          double* stage_ms);
/* Per-kernel totals of the previous insert (ms): background, rows, count,
 * scan, fill, tile accumulate, direct insertion; an 8th value receives the
 * number of bin entries written (n <= 8 values). */
int bsx_recon_kernel_times(const bsx_recon* h, double* ms, int n);

/* ---- FFT (fft_setup_plan / fftw, src/util/fft.cpp:31-114) ---- */
/* In-place batched C2C transform of `batch` contiguous nx*ny*nz complex64
 * arrays. dir -1 forward, +1 backward; norm 0 none, 1 x 1/sqrt(n), 2 x 1/n
 * (Bimage::fft norm_flag, Bimage_fft.cpp:218-223). Sizes: products of 2,3. */
int bsx_fft(float* data_dev, int nx, int ny, int nz, int64_t batch, int dir, int norm,
            void* stream);

/* ---- map transform for projection (Bimage::project prologue,
 *      src/img/Bimage_project.cpp:373-375): forward FFT x 1/sqrt(N^3) and
 *      phase shift to the map origin. map_dev N^3 FP32 -> ft_dev N^3 complex64. */
int bsx_map_transform(const float* map_dev, int n, const double origin[3],
                      float* ft_dev, void* stream);

/* ---- central-section extraction (Bimage::central_section,
 *      src/img/Bimage_project.cpp:288-338; fspace_interpolate,
 *      src/img/Bimage_fspace.cpp:85-147; FSI_Kernel, include/FSI_Kernel.h:28-68).
 * ft_dev: N^3 complex64 map transform. mats: nviews x 9 row-major rotation
 * matrices (host, FP64). sampling: map voxel size. resolution: high limit (A).
 * kernel_width/kernel_power: FSI kernel. center: 1 applies phase_shift_to_center
 * to each section (Bimage::project, Bimage_project.cpp:412-413).
 * out_dev: nviews x N x N complex64. */
int bsx_extract(const float* ft_dev, int n, double sampling, const double* mats,
                int64_t nviews, double resolution, int kernel_width, int kernel_power,
                int center, float* out_dev, void* stream);

/* ---- FSC (Bimage::fsc, src/img/Bimage_resolution.cpp:330-406 with
 *      fspace_radial_power, src/img/Bimage_fspace.cpp:848-958). FP64 shell sums
 *      on the device. a_dev/b_dev: N^3 complex64. fsc: host array of length
 *      *nshell (in: capacity, out: number of shells = (long)(N*pix/hi_res)). */
int bsx_fsc(const float* a_dev, const float* b_dev, int n, double sampling, double hi_res,
            double* fsc, int* nshell, void* stream);

/* ---- projection-matching score (img_recip_space_fom, fom_type 0,
 *      src/mg/mg_refine.cpp:660-781): ring-weighted FRC of particle transforms
 *      against CTF-free reference sections, FP32 on the device.
 *      part_dev: nparts x N x N complex64, ref_dev: nrefs x N x N complex64,
 *      pairs: npairs x 2 int32 (particle, reference) on the host, checked
 *      against nparts / nrefs; out: npairs FRC values (host). */
int bsx_score_frc(const float* part_dev, int64_t nparts, const float* ref_dev, int64_t nrefs,
                  int n, double sampling, double hi_res, double lo_res, const int32_t* pairs,
                  int64_t npairs, double* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* BSX_H */

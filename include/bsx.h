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

/* ---- CTF actions: include/options.h:418-442. particle_reconstruct passes
 *      flip = (action == 1) and the Wiener factor to img_ctf_calculate
 *      (mg_ctf.cpp:46-110): action 1 flips phases; any other nonzero action
 *      applies sin(chi) when wiener == 0 and the Wiener filter when wiener > 0
 *      (breconstruct sets wiener 0.2 for actions > 2 given no factor,
 *      breconstruct.cpp:187; the library takes the factor as given). ---- */
#define BSX_CTF_NONE   0
#define BSX_CTF_FLIP   1
#define BSX_CTF_APPLY  2
#define BSX_CTF_WIENER 4

/* ---- reconstruction flags: src/mg/mg_reconstruct.cpp:83 (1 rescale,
 *      2 2D, 4 bootstrap, 8 Ewald) plus library-specific bits ---- */
#define BSX_FLAG_RESCALE     1   /* rescale each particle image to average 0 and
                                    standard deviation 1 before its background
                                    (rescale_to_avg_std(0, 1),
                                    mg_reconstruct.cpp:160)                   */
#define BSX_FLAG_2D          2   /* 2D map (breconstruct -TwoD): one z plane,
                                    map / FOM / transform outputs N x N     */
#define BSX_FLAG_BOOTSTRAP   4
#define BSX_FLAG_EWALD       8
#define BSX_FLAG_FOM        16   /* also accumulate sum(w|F|^2) and sum(w^2) */
#define BSX_FLAG_CHECK_NAN  32   /* scan particles for non-finite values */
#define BSX_FLAG_TIMING     64   /* time the insertion kernel with HIP events
                                    (bsx_recon_insert then synchronises)   */
#define BSX_FLAG_GRAPH     128   /* replay the insertion kernels of a repeated
                                    insert (same images, count and batch) from
                                    a captured hipGraph; up to 8 such sets per
                                    handle (e.g. the chunks of a pipelined
                                    insert) are kept                         */

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
    double  wiener;        /* Wiener factor (> 0: Wiener for any action but flip) */
    int32_t nhalves;       /* 1: one accumulator; 2: half-set accumulators        */
    int32_t batch;         /* particles per internal batch (0 = autotuned)        */
    int64_t mem_cap;       /* device-memory cap for this handle (bytes, 0 = none) */
    int32_t method;        /* 0: tile-binned insertion; 1: direct atomics */
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

/* The inverse: replaces the accumulators of one half-set with planes_dev
 * (the same [5][N^3] layout; planes 3-4 are ignored without BSX_FLAG_FOM)
 * and sets its particle count. For verification of the finish on
 * hand-built sums (the reference's partial maps are plain images). */
int bsx_recon_set_accumulators(bsx_recon* h, int half, const float* planes_dev, int64_t count,
                               void* stream);

/* Device bytes held by a handle. */
int bsx_recon_device_bytes(const bsx_recon* h, int64_t* bytes);

/* Insertion autotuner (no counterpart in the reference): times the insertion
 * of n sample particles at the allocated batch and its halvings (down to 1/8,
 * at least 32), keeps the fastest, then times the tile-pass piece lengths
 * 8192, 16384 and 32768 at that batch and keeps the fastest. Requires an
 * empty handle and leaves it empty (accumulators reset). */
int bsx_recon_autotune(bsx_recon* h, const float* imgs_dev, const bsx_particle* parts, int64_t n,
                       int32_t* chosen_batch, double* best_ms, void* stream);
/* Current insertion batch (particles per launch sequence) and tile-pass
 * piece length (bin entries per workgroup), as set by bsx_recon_autotune. */
int bsx_recon_tuning(const bsx_recon* h, int32_t* batch, int32_t* piece);
/* Sets them by hand (0 keeps a value): batch up to the one the handle was
 * created for, piece a multiple of 4096 in [8192, 32768]. */
int bsx_recon_set_tuning(bsx_recon* h, int32_t batch, int32_t piece);
/* Number of inserts served by graph replay so far (BSX_FLAG_GRAPH). */
int bsx_recon_graph_launches(const bsx_recon* h, int64_t* n);
/* Capacity fallbacks of the tile-binned insertion (no counterpart in the
 * reference): the OR of the flags raised by the inserts read so far — by
 * each bsx_recon_finish and by each timed insert (BSX_FLAG_TIMING), graph
 * replays included: 1 a bin's chunk table was full, 2 the chunk pool was
 * exhausted, 4 a chunk wait reached its bound. The entries concerned were
 * added with float atomics: the sums are complete, the insert was slower. */
int bsx_recon_fallbacks(const bsx_recon* h, int32_t* flags);

/* ---- group runs (no counterpart in the reference) ----
 * Several reconstructions run as ONE hipGraph, e.g. one per particle box of a
 * mixed-size data set, where the reference runs breconstruct once per box
 * (breconstruct.cpp:204-349). A run of member i clears its accumulators,
 * inserts counts[i] particles (as bsx_recon_insert), then sums all half-sets,
 * weighs and inverse-transforms them into maps_dev[i] (size^3 FP32, size^2
 * with BSX_FLAG_2D), as bsx_recon_finish(h, 0, 1, ...) does. The first run
 * launches the kernels directly; the second identical run (same images,
 * counts, outputs and tuning; the particle parameters may change) captures
 * every member's clears, staging, batches and finish into one graph, and
 * later identical runs replay it. Only the particle records are uploaded
 * per run. Afterwards each member holds this run's sums, so bsx_recon_finish
 * can still give half maps. stats (array of count, may be NULL) get each
 * member's finishing statistics. BSX_FLAG_CHECK_NAN is refused (its host
 * check cannot sit inside one graph); BSX_FLAG_TIMING is ignored. Destroy a
 * group before its members. Synchronises `stream`. */
typedef struct bsx_recon_group bsx_recon_group;
int bsx_recon_group_create(bsx_recon* const* recons, int count, bsx_recon_group** out);
int bsx_recon_group_run(bsx_recon_group* g, const float* const* imgs_dev,
                        const bsx_particle* const* parts, const int64_t* counts,
                        float* const* maps_dev, bsx_stats* stats, void* stream);
/* Number of runs served by the group's graph so far. */
int bsx_recon_group_graph_launches(const bsx_recon_group* g, int64_t* n);
int bsx_recon_group_destroy(bsx_recon_group* g);

/* Fourier shell statistics of a weighed reconstruction, the table that
 * fspace_reconstruction_stats (src/img/Bimage_reconstruct.cpp:526-599) prints
 * from img_reconstruction_sum_weigh (src/mg/mg_reconstruct.cpp:346-347):
 * radial averages of |F|^2 (fspace_radial_power), of the weighed power, the
 * weight and w - sum w^2 / w (Bimage::radial), then per shell i >= 1 the
 * signal, noise, SNR and FSC = SNR / (SNR + 1); 2D maps (BSX_FLAG_2D) use
 * rings in the plane. Needs BSX_FLAG_FOM. imap /
 * nmaps select the half sets as bsx_recon_finish does. resolution: the
 * limit as passed to img_reconstruction_sum_weigh (<= 0.1: N/2 shells).
 * rows (cap x 9, optional): radius, s (1/A), resolution (A), weight, power,
 * signal, noise, SNR, FSC; *nrows = shells defined (those below N/2 and
 * the resolution limit). estimates (2, optional): the SNR(0.5) and FSC(0.3)
 * resolutions (A, 0 when not crossed). Synchronises the stream. */
int bsx_recon_stats(bsx_recon* h, int imap, int nmaps, double resolution, double* rows,
                    int32_t cap, int32_t* nrows, double* estimates, void* stream);

/* Kernel timings of the previous bsx_recon_insert (needs BSX_FLAG_TIMING):
 * insertion kernels (scan, fill, tile accumulate or direct) and staging
 * kernels (background, row transforms) in ms, and the number of batches. */
int bsx_recon_last_timing(const bsx_recon* h, double* insert_ms, int64_t* launches,
                          double* stage_ms);

/* Point-group symmetry of the insertion (fspace_pack_2D with a Bsymmetry,
 * Bimage_reconstruct.cpp:248-266): every later insert adds each particle at
 * all its symmetry-related views (symmetry_get_all_views,
 * src/util/symmetry.cpp:277-319). sym: "C1", "C<n>", "D<n>", "T", "O", "I",
 * with the reference's variants T-3, O-3, O-2, O-4, I-3, I-2, I-5, I90
 * (Bsymmetry::Bsymmetry, symmetry.cpp:42-122); NULL or "C1" switches it off. */
int bsx_recon_set_symmetry(bsx_recon* h, const char* sym);
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

/* bsx_extract without any device allocation. No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments and results as bsx_extract; the view matrices and the kernel
 * table go to a static per-GPU scratch array (about 315 KB, held until the
 * process ends), so nviews is limited to BSX_NOALLOC_MAX_VIEWS (larger counts
 * return BSX_ERR_ARG). Every call on one GPU shares that scratch: calls that
 * overlap on one GPU (other host threads, or other streams still running)
 * race. Synchronises `stream` before returning. */
#define BSX_NOALLOC_MAX_VIEWS 4096
int bsx_extract_noalloc(const float* ft_dev, int n, double sampling, const double* mats,
                        int64_t nviews, double resolution, int kernel_width, int kernel_power,
                        int center, float* out_dev, void* stream);

/* ---- FSC (Bimage::fsc, src/img/Bimage_resolution.cpp:330-406 with
 *      fspace_radial_power, src/img/Bimage_fspace.cpp:848-958). FP64 shell sums
 *      on the device. a_dev/b_dev: N^3 complex64. fsc: host array of length
 *      *nshell (in: capacity, out: number of shells = (long)(N*pix/hi_res)). */
int bsx_fsc(const float* a_dev, const float* b_dev, int n, double sampling, double hi_res,
            double* fsc, int* nshell, void* stream);

/* Bimage::fsc_dpr after resolution_prepare (src/img/Bimage_resolution.cpp:
 * 17-40, 93-240; bresolve src/bresolve.cpp:136-150) for two real n^3 maps on
 * the device: FSC and differential phase residual per shell out to
 * hi_res (checked as Bimage::check_resolution), shells every
 * 1/(n sampling / sampling_ratio). curves: rows of (s [1/A], FSC, DPR [deg]);
 * *nrows = capacity in, rows written out. */
int bsx_fsc_dpr(const float* map1_dev, const float* map2_dev, int n, double sampling,
                double hi_res, double sampling_ratio, double* curves, int* nrows, void* stream);

/* The same for two real n x n images (Bimage::fsc_dpr of 2D images, as
 * bresolve compares class averages or 2D reconstructions): shells in the
 * plane, the resolution checked over two axes. */
int bsx_fsc_dpr_2d(const float* img1_dev, const float* img2_dev, int n, double sampling,
                   double hi_res, double sampling_ratio, double* curves, int* nrows,
                   void* stream);

/* ---- projection-matching score (img_recip_space_fom, fom_type 0,
 *      src/mg/mg_refine.cpp:660-781): ring-weighted FRC of particle transforms
 *      against CTF-free reference sections, FP32 on the device.
 *      part_dev: nparts x N x N complex64, ref_dev: nrefs x N x N complex64,
 *      pairs: npairs x 2 int32 (particle, reference) on the host, checked
 *      against nparts / nrefs; out: npairs FRC values (host). */
int bsx_score_frc(const float* part_dev, int64_t nparts, const float* ref_dev, int64_t nrefs,
                  int n, double sampling, double hi_res, double lo_res, const int32_t* pairs,
                  int64_t npairs, double* out, void* stream);

/* ==== orientation search by projection matching ====
 * The reference compares a particle transform with central sections of the
 * map transform by img_recip_space_fom (src/mg/mg_refine.cpp:660-781,
 * fom_type 0, CTF applied to the section) over views of
 * asymmetric_unit_views (src/util/symmetry.cpp:383-548) expanded in-plane
 * (view_list_expand_angles, src/util/View.cpp:519-556); the exhaustive
 * driver is part_determine_orientation (src/mg/mg_orient.cpp:545-798).
 * Here comparisons use the rings 0..R-1 of a low-resolution disk (R =
 * radius in Fourier pixels, i.e. hi_res = box * sampling / R), held as
 * half-plane point lists: a BF16 MFMA GEMM ranks every library view by a
 * ring-normalised correlation, the top-k go to the FP32 FRC of the
 * reference formula, and bsx_search_score_mats evaluates the same FRC at
 * arbitrary rotations for local refinement. */
#define BSX_SEARCH_CTF 1   /* apply each particle's CTF (defocus >= 1 A) */

typedef struct bsx_search_params {
    int32_t box;           /* particle / section edge N (map N^3), 8..1024       */
    int32_t radius;        /* disk radius R (Fourier pixels), 2..min(64, N/2-1)   */
    double  sampling;      /* map voxel and particle pixel size (A)               */
    int32_t kernel_width;  /* FSI kernel width for sections (brefine: 8), 1..10   */
    int32_t kernel_power;  /* FSI kernel power (2); < 0 = nearest-neighbour table */
    int32_t topk;          /* candidates kept by the GEMM: 1..8 or 16             */
    int32_t flags;         /* BSX_SEARCH_*                                        */
    int64_t mem_cap;       /* device byte cap for the handle, 0 = none            */
} bsx_search_params;

typedef struct bsx_search bsx_search;  /* opaque handle: library + staged particles */

int bsx_search_create(bsx_search** h, const bsx_search_params* p);
int bsx_search_destroy(bsx_search* h);
/* Disk point count Nd (complex), GEMM row length (BF16 elements), views, particles. */
int bsx_search_info(const bsx_search* h, int32_t* npoints, int32_t* kdim, int64_t* nviews,
                    int64_t* nparts);
/* Host copies of the disk points (Nd x 2 int32: h, k) and ring starts (R+1). */
int bsx_search_points(const bsx_search* h, int32_t* hk, int32_t* ring_start);

/* Library from a map transform (bsx_map_transform output, N^3 complex64) at
 * nviews row-major FP64 rotation matrices (host): sections as
 * Bimage::central_section (Bimage_project.cpp:288-338) without the centre
 * shift, at the disk points. Replaces any previous library. */
int bsx_search_set_reference(bsx_search* h, const float* ft_dev, const double* mats,
                             int64_t nviews, void* stream);

/* Stage n particle images (n x N x N FP32, device): forward FFT x 1/N, phase
 * shift to each particle's origin `ori` (phase_shift_to_origin,
 * Bimage_complex.cpp:490-543), disk values, CTF values (sin chi of the
 * particle's parameters, include/ctf.h:321-365) and BF16 GEMM rows.
 * Replaces any previously staged set. */
int bsx_search_stage(bsx_search* h, const float* imgs_dev, const bsx_particle* parts, int64_t n,
                     void* stream);

/* Exhaustive coarse search: BF16 GEMM of every staged particle against every
 * library view with FP32 accumulation and per-particle top-k (k = topk).
 * cand_dev: n x topk int32 view indices, score_dev: n x topk FP32 coarse
 * scores, best first. Asynchronous on `stream`. */
int bsx_search_exhaustive(bsx_search* h, int32_t* cand_dev, float* score_dev, void* stream);

/* The same over a subset: rows_dev (m int32, device) staged-particle indices;
 * cand_dev / score_dev m x topk. */
int bsx_search_exhaustive_rows(bsx_search* h, const int32_t* rows_dev, int64_t m,
                               int32_t* cand_dev, float* score_dev, void* stream);

/* Proposal classes for bsx_search_proposed: members (host, nclasses x maxm)
 * holds each class's library view indices as a leading run, -1 padded. Call
 * after bsx_search_set_reference. No counterpart in the reference. */
int bsx_search_set_classes(bsx_search* h, const int32_t* members, int32_t nclasses, int32_t maxm);

/* Coarse search restricted to proposed classes: cls_dev (n x k int32,
 * device; negative = none) routes each (particle, proposal) to its class, and
 * a grouped BF16 MFMA GEMM scores the particle against that class's member
 * views only. Output as bsx_search_exhaustive: n x topk candidates and
 * scores. Synchronises `stream` once (work-item list). */
int bsx_search_proposed(bsx_search* h, const int32_t* cls_dev, int32_t k, int32_t* cand_dev,
                        float* score_dev, void* stream);

/* FP32 FRC (img_recip_space_fom, fom_type 0, CTF on the section, rings
 * 0..R-1) of candidate library views: cand_dev n x ncand int32 (negative =
 * empty slot). frc_dev (n x ncand, optional), best_dev / best_frc_dev (n,
 * optional): the best candidate's view index and FRC. Asynchronous. */
int bsx_search_rescore(bsx_search* h, const int32_t* cand_dev, int32_t ncand, float* frc_dev,
                       int32_t* best_dev, float* best_frc_dev, void* stream);

/* The same FRC at arbitrary rotations, sections interpolated on the fly from
 * ft_dev (N^3 complex64): mats_dev n x ncand x 9 FP64 (device), frc_dev
 * n x ncand. Asynchronous. */
int bsx_search_score_mats(bsx_search* h, const float* ft_dev, const double* mats_dev,
                          int32_t ncand, float* frc_dev, void* stream);

/* Rotation matrices (n x 9 FP64, device) of library views views_dev (n int32,
 * device; a negative index gives the identity). Asynchronous. */
int bsx_search_view_mats(const bsx_search* h, const int32_t* views_dev, int64_t n,
                         double* mats_dev, void* stream);

/* Local refinement of every staged particle's orientation, in the manner of
 * views_for_refinement (src/util/View.cpp:371-436) on a 3x3x3 grid: for each
 * step s (radians, host array), the 27 rotations R * Rot(s * (i,j,k)),
 * i,j,k in {-1,0,1}, are scored by the FRC of bsx_search_score_mats and the
 * best becomes R. mats_dev: n x 9 FP64 (device, in/out); frc_dev (n,
 * optional): the final FRC. Asynchronous. */
int bsx_search_refine(bsx_search* h, const float* ft_dev, double* mats_dev, float* frc_dev,
                      const double* steps, int32_t nsteps, void* stream);

/* img_refine_grid's view loop (src/mg/mg_refine.cpp:1116-1140) over the
 * views_for_refinement grid (src/util/View.cpp:371-436), for every staged
 * particle at once: starting from mats_dev (n x 9, FP64, device) with angular
 * step alpha_step (rad), each pass scores the 27 views around the best view
 * (the view vector tilted by -step, 0, +step about two axes perpendicular to
 * it, times the in-plane angle +-step) by the FP32 FRC of the staged
 * particle (as bsx_search_score_mats, the origin kept); the first highest
 * score becomes the best view, and the step halves when that is the centre
 * or the score changed by less than 1e-6; a particle stops once its step is
 * below `accuracy`. At most max_iter passes (the reference has no bound).
 * Out: mats_dev = the best views' matrices, frc_dev (n, may be NULL) their
 * scores (-1 where alpha_step < accuracy), *iterations = passes run.
 * Synchronises the stream once per pass. */
int bsx_search_refine_views(bsx_search* h, const float* ft_dev, double* mats_dev,
                            float* frc_dev, double alpha_step, double accuracy, int32_t max_iter,
                            int32_t* iterations, void* stream);
/* views_for_refinement (src/util/View.cpp:371-436) on the host: the 27 views
 * (x, y, z, angle; angle in rad) around `view` at angular step alpha_step > 0,
 * in the reference's order (centre = entry 13). out: 27 x 4 doubles. */
int bsx_views_for_refinement(const double view[4], double alpha_step, double* out);

/* Origin and defocus fit of every staged particle at its orientation, as
 * img_fit_in_recip_space_grid (src/mg/mg_refine.cpp:862-946) does: the
 * section at mats_dev (n x 9 FP64, device) is taken once; when def_std > 0
 * and either the search's CTF flag is set or the particle's defocus is
 * nonzero, the defocus is scanned from def - 5 def_std to def + 5 def_std in
 * steps of def_std (the CTF applied where the flag is set and the defocus is
 * >= 1); then 3 x 3 grids of origins around the staged origin are scored,
 * the step halved when the centre wins or the FOM stops changing, while
 * step >= accuracy. FOM: img_recip_space_fom, fom_type 0, rings 0..R-1, FP64
 * ring sums. Outputs, per particle: origin_dev n x 2 FP64 (pixels, the best
 * of the last grid), defocus_dev n FP64 (optional; the staged defocus when
 * not scanned), fom_dev n. def_std == 0 skips the scan (the reference's scan
 * loop does not end for it). Asynchronous. */
int bsx_search_fit_origin(bsx_search* h, const float* ft_dev, const double* mats_dev, double step,
                          double accuracy, double def_std, double* origin_dev, double* defocus_dev,
                          float* fom_dev, void* stream);

/* img_fit_in_recip_space_monte (src/mg/mg_refine.cpp:783-860) for the staged
 * particles at rotations mats_dev: a Metropolis walk of max_iter FOM
 * evaluations over origin and defocus, with the FOM of
 * bsx_search_fit_origin. Iteration i > 0 adds def_std * g to the defocus
 * (when the CTF flag is set or def_std > 0 and the defocus is nonzero) and
 * shift_std * a * (ux, uy) to the origin; the move is kept when
 * fom / previous fom > the uniform value, else the walk returns to the last
 * kept state (the previous FOM starts at -1, as in the reference). The
 * reference draws from random(); here the caller supplies the draws:
 * draws_dev is n x max_iter x 5 FP64, (g, a, ux, uy, uniform) with g and a
 * standard normal, (ux, uy) a unit vector (vector3_xy_random_gaussian,
 * src/util/random_numbers.cpp:522-533) and uniform in [0, 1); iteration 0
 * reads only the uniform value. defocus_start_dev (n FP64, optional): the
 * defocus each walk starts from instead of the staged one (img_refine_monte
 * leaves the previous fit's defocus in place when it does not reset it).
 * Outputs as bsx_search_fit_origin: the best origin, its defocus (the
 * starting one when the defocus does not move) and FOM. Asynchronous. */
int bsx_search_fit_monte(bsx_search* h, const float* ft_dev, const double* mats_dev,
                         int32_t max_iter, double def_std, double shift_std,
                         const double* draws_dev, const double* defocus_start_dev,
                         double* origin_dev, double* defocus_dev, float* fom_dev, void* stream);

/* Rotation matrices back to views: Matrix3::quaternion (include/Matrix3.h:
 * 129-166) then View(Quaternion&) (include/View.h:52-63). Host only. */
int bsx_matrix_views(const double* mats, int64_t n, double* views);

/* Input images of the orientation-proposal network for the staged particles:
 * the disk values phase-flipped by the sign of their CTF on an S x S grid
 * (centre at S/2), inverse transformed, real part normalised to zero mean and
 * unit variance. out_dev: n x S x S FP32. No counterpart in the reference. */
int bsx_search_cnn_input(bsx_search* h, int32_t S, float* out_dev, void* stream);

/* View::matrix (include/View.h:131-157, Matrix3.h:30-42) of n views (x, y,
 * z, angle) into n row-major 3x3 FP64 matrices. Host only. */
int bsx_view_matrices(const double* views, int64_t n, double* mats);

/* Device copies of the staged data (any pointer may be NULL): values n x Nd
 * complex64, CTF n x Nd FP32, GEMM rows n x kdim BF16 (raw uint16). */
int bsx_search_particle_data(const bsx_search* h, float* values_dev, float* ctf_dev,
                             uint16_t* rows_dev, void* stream);
int bsx_search_library_data(const bsx_search* h, float* values_dev, uint16_t* rows_dev,
                            void* stream);

/* ==== orientation-proposal network (no counterpart in the reference) ====
 * A 5-layer CNN: conv3x3 s1 (1->c1), conv3x3 s2 (c1->c2), conv3x3 s2 (c2->c3),
 * conv3x3 s2 (c3->c4), each + ReLU, then a fully connected layer to nclasses
 * coarse orientation classes with softmax cross-entropy. Trained with Adam on
 * FP32 weights by the library's own forward and backward kernels; inference
 * uses BF16 copies of the weights with FP32 accumulation. Inputs: batch x S x S
 * FP32 (bsx_search_cnn_input), S a multiple of 8. */
typedef struct bsx_cnn bsx_cnn;

int bsx_cnn_create(bsx_cnn** h, int32_t in_size, const int32_t channels[4], int32_t nclasses,
                   int32_t max_batch, uint64_t seed);
int bsx_cnn_destroy(bsx_cnn* h);
/* Parameter count (flat FP32: W1 b1 W2 b2 W3 b3 W4 b4 Wfc bfc) and FC input size. */
int bsx_cnn_nparams(const bsx_cnn* h, int64_t* n, int64_t* feature_size);
/* Copy the parameters out (get_dev) and/or in (set_dev), device arrays. */
int bsx_cnn_params(bsx_cnn* h, float* get_dev, const float* set_dev, void* stream);
/* FP32 forward: logits_dev batch x nclasses (optional). */
int bsx_cnn_forward(bsx_cnn* h, const float* x_dev, int32_t batch, float* logits_dev, void* stream);
/* Forward + backward without an update: grad_dev (nparams, optional) holds
 * the gradient of the mean cross-entropy, loss_dev (batch, optional) the
 * per-sample losses. */
int bsx_cnn_gradients(bsx_cnn* h, const float* x_dev, const int32_t* labels_dev, int32_t batch,
                      float* grad_dev, float* loss_dev, void* stream);
/* One Adam step (b1 0.9, b2 0.999, eps 1e-8; L2 weight decay). */
int bsx_cnn_train_step(bsx_cnn* h, const float* x_dev, const int32_t* labels_dev, int32_t batch,
                       float lr, float weight_decay, float* loss_dev, void* stream);
/* BF16-weight inference over n inputs (batched internally): the k <= 32 most
 * probable classes per input (cls_dev n x k) and their softmax probabilities. */
int bsx_cnn_infer_topk(bsx_cnn* h, const float* x_dev, int64_t n, int32_t k, int32_t* cls_dev,
                       float* prob_dev, void* stream);
/* The same network as an embedding (no counterpart in the reference): its
 * nclasses outputs are read as a d = nclasses vector, normalised to unit
 * length. A network trained with bsx_cnn_train_pairs maps a noisy,
 * CTF-modulated particle image and the clean projection of the same map at
 * the same orientation to nearby unit vectors, for maps it was not trained
 * on, so that a map's orientation classes are found by comparing a particle
 * with the map's own projections.
 * bsx_cnn_embed: BF16-weight inference of n inputs -> emb_dev (n x d, FP32). */
int bsx_cnn_embed(bsx_cnn* h, const float* x_dev, int64_t n, float* emb_dev, void* stream);
/* One Adam step of contrastive training: x_dev holds 2 * batch inputs, the
 * particles 0..batch-1 followed by their matching references. The loss is the
 * symmetric cross-entropy of the cosine similarities over temperature tau
 * (pair i matches particle i with reference i, the other references of the
 * batch are the negatives). 2 * batch <= max_batch. loss_dev (batch) may be
 * NULL. */
int bsx_cnn_train_pairs(bsx_cnn* h, const float* x_dev, int32_t batch, float tau, float lr,
                        float weight_decay, float* loss_dev, void* stream);
/* The gradients of that loss (grad_dev: the flat parameter vector's size,
 * may be NULL) and the per-pair losses (batch, may be NULL), no update. */
int bsx_cnn_pair_gradients(bsx_cnn* h, const float* x_dev, int32_t batch, float tau,
                           float* grad_dev, float* loss_dev, void* stream);
/* The k <= 32 best of m reference embeddings for each of n particle
 * embeddings (both unit rows of d = nclasses values, device FP32), by cosine;
 * prob_dev receives the softmax of cosine / tau over all m references. */
int bsx_cnn_match_topk(bsx_cnn* h, const float* p_dev, int64_t n, const float* r_dev, int32_t m,
                       int32_t k, float tau, int32_t* cls_dev, float* prob_dev, void* stream);

/* asymmetric_unit_views for C1 (src/util/symmetry.cpp:383-470; flag bit 0
 * full sphere, bit 1 no in-plane alignment) then view_list_expand_angles
 * (src/util/View.cpp:519-556) with angles q*angle_step in [0, 2 pi)
 * (angle_step 0: none). views: cap x 4 (x, y, z, angle); *count = total. */
int bsx_views_c1(double theta_step, double phi_step, int32_t flag, double angle_step,
                 double* views, int64_t cap, int64_t* count);

/* Real-space symmetrization of a reconstructed map, breconstruct's symmetry
 * mode 1 (src/breconstruct.cpp:335, Bimage::symmetrize with the default
 * reference view, src/img/Bimage_symmetry.cpp:31-100): for each operator
 * (axis, order m) of sym the map becomes itself plus its m - 1 rotations
 * about origin (voxels, x y z), trilinear as Bimage::interpolate with the
 * map's background (mean outside the inscribed ellipsoid) as the fill; the
 * result is multiplied by 1 / order. in_dev, out_dev: distinct N^3 FP32 maps
 * (z, y, x order) on the device. C1 copies. Synchronises the stream. */
int bsx_symmetrize(const float* in_dev, float* out_dev, int32_t n, const double* origin,
                   const char* sym, void* stream);

/* The same for an n x n image (breconstruct -TwoD -mode 1): cyclic point
 * groups only, about origin (x, y), as Bimage::symmetrize restricts 2D
 * images (Bimage_symmetry.cpp:40-46). Synchronises the stream. */
int bsx_symmetrize_2d(const float* in_dev, float* out_dev, int32_t n, const double* origin,
                      const char* sym, void* stream);

/* Bimage::rescale_to_avg_std (src/img/Bimage_rescale.cpp:163-183), as
 * breconstruct -rescale avg,std applies it to each output map
 * (src/breconstruct.cpp:336): v <- v * (std / s) + avg - a * std / s with
 * a, s the average and population standard deviation of the n values
 * (FP64 sums, Bimage::statistics) and the factor 1 when s <= 1e-30.
 * data_dev: n FP32 values (device), rescaled in place. std < 0 is an
 * error, as in the reference. Synchronises the stream. */
int bsx_rescale_to_avg_std(float* data_dev, int64_t n, double avg, double std, void* stream);

/* Bimage::filter_sinc (src/img/Bimage_filter.cpp:484-509), the correction
 * breconstruct applies to maps from trilinear insertion
 * (src/breconstruct.cpp:359): each voxel times (r / sin r)^2 with
 * r = pi |(x - ox) / nx, (y - oy) / ny, (z - oz) / nz|, the origin voxel
 * unchanged. data_dev: nz x ny x nx FP32 (device), in place; origin: 3
 * doubles (voxels). Asynchronous. */
int bsx_filter_sinc(float* data_dev, int32_t nx, int32_t ny, int32_t nz, const double* origin,
                    void* stream);

/* symmetry_get_all_views (src/util/symmetry.cpp:277-319) for each of n views
 * (x, y, z, angle): out holds, view after view, the order(sym) related views
 * (the view itself first); cap = room in out (views); *count = n * order. */
int bsx_symmetry_views(const char* sym, const double* views, int64_t n, double* out,
                       int64_t cap, int64_t* count);

#ifdef __cplusplus
}
#endif

#endif /* BSX_H */

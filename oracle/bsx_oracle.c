/*
 * bsx_oracle.c — CPU restatement of Bsoft's Fourier-space single-particle
 * reconstruction path (breconstruct) and its mirror (central-section
 * extraction), written from the reference's behaviour for use as the checker.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py. The product path (libbsx.so) never calls it.
 *
 * Two builds of this file (see oracle/Makefile):
 *   liboracle64.so  FP64 images and accumulators — the parity oracle;
 *   liboracle32.so  FP32 images and accumulators with double intermediates,
 *                   as the reference stores them, plus OpenMP over partial
 *                   maps like breconstruct -threads T — the CPU baseline.
 *
 * Reference paths below are relative to the Bsoft tree (bsoft/).
 * Pinning of each function is listed in tests/test_oracle.py; functions
 * without an independent pin are marked "parity unpinned" where they are
 * defined.
 */
#include "bsx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define SMALLFLOAT 1e-30            /* include/utilities.h:50 */

int or_real_bytes(void) { return (int)sizeof(orr); }

/* ------------------------------------------------------------------------ */
/* sizes                                                                     */
/* ------------------------------------------------------------------------ */

/* findNextPowerOf, src/util/utilities.cpp:447-462 */
int or_find_next_power(int start, int power)
{
    int r = power;
    if (power <= 0) return 0;
    if (power == 1) return start;
    while (r < start) r *= power;
    return r;
}

/* part_ft_size, src/mg/mg_reconstruct.cpp:38-64 */
int or_part_ft_size(int xsize, double scale, int pad_factor)
{
    if (xsize <= 0) return 0;
    int ft = xsize;
    if (scale > 1) ft = (int)(xsize * scale);
    if (pad_factor < 1) return ft;
    if (pad_factor > 8) pad_factor = 8;
    ft = or_find_next_power((int)(pad_factor * xsize * scale), 2);
    if (ft < xsize) ft = or_find_next_power(xsize, 2);
    return ft;
}

/* Bimage::check_resolution, src/img/Bimage.cpp:633-648 (cubic map, ndim axes
 * with size > 1). */
double or_check_resolution(double res, int n, double sampling, int ndim)
{
    if (res > n * sampling) return n * sampling;
    double invsum2 = ndim / (sampling * sampling);
    double rm = 2 / sqrt(invsum2);
    if (res < rm) res = rm;
    return res;
}

/* ------------------------------------------------------------------------ */
/* rotations: View::quaternion / View::matrix (include/View.h:110-157),      */
/* Matrix3::from_quaternion (include/Matrix3.h:30-42)                        */
/* ------------------------------------------------------------------------ */
void or_view_matrix(const double view[4], double d[9])
{
    double v0 = view[0], v1 = view[1], v2 = view[2], a = view[3];
    double size = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
    if (size < SMALLFLOAT) { v0 = v1 = 0; v2 = 1; }
    else { v0 /= size; v1 /= size; v2 /= size; }
    a = remainder(a, 2 * M_PI);                 /* angle_set_negPI_to_PI */
    double ang = a / 2.0, ca = cos(ang), sa = sin(ang);
    double q0 = ca, q1 = 0, q2 = 0, q3 = sa;
    if (v2 < 1.0 - SMALLFLOAT) {
        if (v2 <= SMALLFLOAT - 1.0) { q0 = 0; q1 = sa; q2 = ca; q3 = 0; }
        else {
            double z1 = v2 + 1, f1 = sqrt(z1 / 2.0), f2 = sqrt(1.0 / (2 * z1));
            q0 = f1 * ca;
            q1 = f2 * (v0 * sa - v1 * ca);
            q2 = f2 * (v0 * ca + v1 * sa);
            q3 = f1 * sa;
        }
    }
    d[0] = q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3;
    d[1] = 2 * q1 * q2 - 2 * q0 * q3;
    d[2] = 2 * q1 * q3 + 2 * q0 * q2;
    d[3] = 2 * q1 * q2 + 2 * q0 * q3;
    d[4] = q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3;
    d[5] = 2 * q2 * q3 - 2 * q0 * q1;
    d[6] = 2 * q1 * q3 - 2 * q0 * q2;
    d[7] = 2 * q2 * q3 + 2 * q0 * q1;
    d[8] = q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3;
    for (int i = 0; i < 9; i++) {               /* Matrix3::check */
        if (fabs(d[i]) < 1e-12) d[i] = 0;
        if (d[i] > 1) d[i] = 1;
        if (d[i] < -1) d[i] = -1;
    }
}

/* ------------------------------------------------------------------------ */
/* CTF: CTFparam (include/ctf.h:28-370), img_ctf_calculate                    */
/* (src/mg/mg_ctf.cpp:46-110)                                                 */
/* ------------------------------------------------------------------------ */

/* CTFparam::lambda, include/ctf.h:309-317; constants include/utilities.h:55-58 */
double or_lambda(double av)
{
    const double PLANCK = 6.626070E-34, ECHARGE = 1.602177E-19;
    const double EMASS = 9.109384E-31, LIGHTSPEED = 299792458.0;
    if (av <= 0) return 0;
    return PLANCK * 1e10 / sqrt(2.0 * EMASS * ECHARGE * av *
                                (1.0 + ECHARGE * av / (2.0 * EMASS * LIGHTSPEED * LIGHTSPEED)));
}

/* Aberration weights: amp_shift, defocus_average, astigmatism, Cs
 * (include/ctf.h:121-147, update_terms :79-83). */
static void ctf_coeffs(const bsx_particle* p, double c[5])
{
    double wl = or_lambda(p->volt);
    double t1 = M_PI_2 * wl * wl * wl, t2 = -M_PI * wl;
    double dev = fabs(p->def_dev);
    c[0] = -p->amp_shift;                   /* abw[0,0]  */
    c[1] = t2 * p->def_avg;                 /* abw[2,0]  */
    c[2] = t2 * dev * cos(2 * p->ast_angle);/* abw[2,2]  */
    c[3] = t2 * dev * sin(2 * p->ast_angle);/* abw[2,-2] */
    c[4] = t1 * p->cs;                      /* abw[4,0]  */
}

/* CTF multiplier at spatial frequency (sx, sy) (1/A), inside the resolution
 * band: calculate_aberration_even (include/ctf.h:321-335) summed in the same
 * order, sin, then flip / Wiener (mg_ctf.cpp:102-108). No odd aberrations. */
double or_ctf_value(const bsx_particle* p, double sx, double sy, int flip, double wiener)
{
    double c[5];
    ctf_coeffs(p, c);
    double s2 = sx * sx + sy * sy;
    double a = atan2(sy, sx);
    double dphi = c[0];
    dphi += c[3] * s2 * sin(2 * a);
    dphi += c[1] * s2;
    dphi += c[2] * s2 * cos(2 * a);
    dphi += c[4] * s2 * s2;
    double w = sin(dphi);
    if (flip) w = (w < 0) ? -1 : 1;
    else if (wiener) w = (wiener + 1) * w / (w * w + wiener);
    return w;
}

/* ------------------------------------------------------------------------ */
/* FFT: FFTW3f C2C semantics as used by Bimage::fft (Bimage_fft.cpp:176-234,  */
/* fft.cpp:31-114): unnormalised DFT with sign dir, then scaling by norm.     */
/* The algorithm (self-sorting Stockham, radices 4/2/3/5) is this oracle's   */
/* own; it is pinned against the textbook DFT in the tests.                   */
/* ------------------------------------------------------------------------ */
typedef struct { long n; int nf; int f[40]; double* tw; } fft1;

static int fft1_init(fft1* p, long n)
{
    p->n = n; p->nf = 0;
    long m = n;
    while (m % 4 == 0) { p->f[p->nf++] = 4; m /= 4; }
    while (m % 2 == 0) { p->f[p->nf++] = 2; m /= 2; }
    while (m % 3 == 0) { p->f[p->nf++] = 3; m /= 3; }
    while (m % 5 == 0) { p->f[p->nf++] = 5; m /= 5; }
    if (m != 1) return -1;
    p->tw = (double*)malloc(sizeof(double) * 2 * n);
    for (long i = 0; i < n; i++) {
        p->tw[2 * i] = cos(2 * M_PI * i / n);
        p->tw[2 * i + 1] = sin(2 * M_PI * i / n);
    }
    return 0;
}

static void fft1_free(fft1* p) { free(p->tw); p->tw = NULL; }

/* x: n interleaved complex values (contiguous); work: same size. */
static void fft1_exec(const fft1* p, orr* x, orr* work, int sign)
{
    long n = p->n;
    if (n == 1) return;
    orr* src = x;
    orr* dst = work;
    long Ns = 1;
    double vr[5], vi[5];
    double dr[25], di[25];   /* the R-point DFT matrix of a stage */
    for (int s = 0; s < p->nf; s++) {
        int R = p->f[s];
        long L = n / R;
        long tws = n / (Ns * R);
        for (int q = 0; q < R; q++)
            for (int r = 0; r < R; r++) {
                long a = (long)((q * r) % R) * (n / R);
                dr[q * R + r] = p->tw[2 * a];
                di[q * R + r] = sign * p->tw[2 * a + 1];
            }
        for (long j = 0; j < L; j++) {
            long k = j % Ns;
            for (int r = 0; r < R; r++) {
                double ar = src[2 * (j + r * L)], ai = src[2 * (j + r * L) + 1];
                long a = k * r * tws;            /* below n: k < Ns, r < R */
                double cr = p->tw[2 * a], ci = sign * p->tw[2 * a + 1];
                vr[r] = ar * cr - ai * ci;
                vi[r] = ar * ci + ai * cr;
            }
            long idxD = (j / Ns) * Ns * R + k;
            for (int q = 0; q < R; q++) {
                double sr = 0, si = 0;
                for (int r = 0; r < R; r++) {
                    double cr = dr[q * R + r], ci = di[q * R + r];
                    sr += vr[r] * cr - vi[r] * ci;
                    si += vr[r] * ci + vi[r] * cr;
                }
                dst[2 * (idxD + q * Ns)] = (orr)sr;
                dst[2 * (idxD + q * Ns) + 1] = (orr)si;
            }
        }
        Ns *= R;
        orr* t = src; src = dst; dst = t;
    }
    if (src != x) memcpy(x, src, sizeof(orr) * 2 * n);
}

/* Transform one nx*ny*nz complex array in place along each axis of size > 1. */
static int fft3(orr* d, long nx, long ny, long nz, int sign)
{
    long dims[3] = {nx, ny, nz};
    long maxn = nx; if (ny > maxn) maxn = ny; if (nz > maxn) maxn = nz;
    orr* line = (orr*)malloc(sizeof(orr) * 4 * maxn);
    orr* work = line + 2 * maxn;
    for (int ax = 0; ax < 3; ax++) {
        long n = dims[ax];
        if (n < 2) continue;
        fft1 p;
        if (fft1_init(&p, n)) { free(line); return -1; }
        long stride = (ax == 0) ? 1 : (ax == 1) ? nx : nx * ny;
        long nlines = nx * ny * nz / n;
        for (long li = 0; li < nlines; li++) {
            long base;
            if (ax == 0) base = li * nx;
            else if (ax == 1) base = (li / nx) * nx * ny + (li % nx);
            else base = li;
            if (stride == 1) {
                fft1_exec(&p, d + 2 * base, work, sign);
            } else {
                for (long i = 0; i < n; i++) {
                    line[2 * i] = d[2 * (base + i * stride)];
                    line[2 * i + 1] = d[2 * (base + i * stride) + 1];
                }
                fft1_exec(&p, line, work, sign);
                for (long i = 0; i < n; i++) {
                    d[2 * (base + i * stride)] = line[2 * i];
                    d[2 * (base + i * stride) + 1] = line[2 * i + 1];
                }
            }
        }
        fft1_free(&p);
    }
    free(line);
    return 0;
}

int or_fft(orr* data, long nx, long ny, long nz, long batch, int dir, int norm)
{
    long n = nx * ny * nz;
    double scale = 1.0 / n;
    if (norm == 1) scale = sqrt(scale);
    for (long b = 0; b < batch; b++) {
        orr* d = data + 2 * n * b;
        if (fft3(d, nx, ny, nz, dir < 0 ? -1 : 1)) return -1;
        if (norm)
            for (long i = 0; i < 2 * n; i++) d[i] = (orr)(d[i] * scale);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* staging: Bimage::calculate_background (Bimage_background.cpp:27-89,       */
/* flag 0), pad (Bimage_resize.cpp:347-358 -> resize :28-110),                */
/* phase_shift_to_origin (Bimage_complex.cpp:490-543), img_ctf_apply_complex  */
/* (mg_ctf.cpp:661-676), in the order of particle_reconstruct                 */
/* (mg_reconstruct.cpp:159-214).                                              */
/* ------------------------------------------------------------------------ */
/* Bimage::statistics (Bimage_stats.cpp:104-148): average and population
 * standard deviation from FP64 sums of v and v^2. */
static void moments(const double* d, const float* f, long n, double* avg, double* sd)
{
    double s1 = 0, s2 = 0;
    for (long i = 0; i < n; i++) {
        double v = d ? d[i] : (double)f[i];
        s1 += v;
        s2 += v * v;
    }
    *avg = s1 / n;
    double var = s2 / n - *avg * *avg;
    *sd = var > 0 ? sqrt(var) : 0;
}

/* Bimage::rescale_to_avg_std (Bimage_rescale.cpp:163-183) with
 * Bimage::rescale (:27-45): v <- v * scale + shift, scale = nustd / sd
 * (1 when sd <= 1e-30), shift = nuavg - avg * scale. */
static void rescale_coef(double avg, double sd, double nuavg, double nustd, double* sc, double* sh)
{
    *sc = sd > 1e-30 ? nustd / sd : 1;
    *sh = nuavg - avg * *sc;
}

void or_rescale_f32(float* d, long n, double nuavg, double nustd)
{
    double avg, sd, sc, sh;
    moments(NULL, d, n, &avg, &sd);
    rescale_coef(avg, sd, nuavg, nustd, &sc, &sh);
    for (long i = 0; i < n; i++) d[i] = (float)((double)d[i] * sc + sh);
}

int or_rescale_to_avg_std(orr* d, long n, double nuavg, double nustd)
{
    if (nustd < 0 || n < 1) return -1;
    double avg, sd, sc, sh;
#ifdef OR_FP32
    moments(NULL, d, n, &avg, &sd);
#else
    moments(d, NULL, n, &avg, &sd);
#endif
    rescale_coef(avg, sd, nuavg, nustd, &sc, &sh);
    for (long i = 0; i < n; i++) d[i] = (orr)((double)d[i] * sc + sh);
    return 0;
}

/* Bimage::filter_sinc (Bimage_filter.cpp:484-509): v <- v (r / sin r)^2,
 * r = pi |(x - o) / size| per axis, the origin voxel unchanged. */
int or_filter_sinc(orr* d, long nx, long ny, long nz, const double o[3])
{
    for (long z = 0, i = 0; z < nz; z++) {
        double dz = (z - o[2]) / nz;
        for (long y = 0; y < ny; y++) {
            double dy = (y - o[1]) / ny;
            for (long x = 0; x < nx; x++, i++) {
                double dx = (x - o[0]) / nx;
                double r = M_PI * sqrt(dx * dx + dy * dy + dz * dz);
                if (r != 0) {
                    double w = r / sin(r);
                    d[i] = (orr)(w * w * (double)d[i]);
                }
            }
        }
    }
    return 0;
}

double or_background(const float* img, long box, const double ori[2])
{
    long hx = box / 2, hy = box / 2, nm = 0;
    double bkg = 0;
    for (long yy = 0; yy < box; yy++) {
        double y2 = (yy - ori[1]) / hy;
        y2 *= y2;
        for (long xx = 0; xx < box; xx++) {
            double x2 = (xx - ori[0]) / hx;
            x2 *= x2;
            double r = x2 + y2 - 1;
            if (r > 0) { bkg += img[yy * box + xx]; nm++; }
        }
    }
    return nm ? bkg / nm : 0;
}

static inline long wrap_freq(long i, long n) { return (i > (n - 1) / 2) ? i - n : i; }

int or_stage_particle(const bsx_recon_params* rp, const bsx_particle* pp, const float* img,
                      orr* F, double* part_weight)
{
    long box = rp->box, ft = rp->ft_size;
    if (box < 2 || ft < box) return -1;
    /* a 2D map (flags & 2) has one z plane: its check_resolution counts two axes */
    double hi = or_check_resolution(rp->hi_res, rp->size, rp->sampling, (rp->flags & 2) ? 2 : 3);
    /* -rescale (flags bit 1): rescale_to_avg_std(0, 1) of the image before
     * its background (mg_reconstruct.cpp:160-161), the values stored FP32 */
    float* scaled = NULL;
    if (rp->flags & BSX_FLAG_RESCALE) {
        scaled = (float*)malloc(sizeof(float) * box * box);
        if (!scaled) return -1;
        for (long i = 0; i < box * box; i++) scaled[i] = img[i];
        or_rescale_f32(scaled, box * box, 0.0, 1.0);
        img = scaled;
    }
    double bkg = or_background(img, box, pp->hdr_ori);
    double w = 1;
    if (rp->flags & BSX_FLAG_BOOTSTRAP) w = pp->sel;
    double pad_ratio = (double)ft / box;
    if (pad_ratio > 1) { pad_ratio *= pad_ratio * 0.6; w /= pad_ratio; }
    if (part_weight) *part_weight = w;
    for (long y = 0; y < ft; y++)
        for (long x = 0; x < ft; x++) {
            orr v = (y < box && x < box) ? (orr)img[y * box + x] : (orr)bkg;
            F[2 * (y * ft + x)] = v;
            F[2 * (y * ft + x) + 1] = 0;
        }
    free(scaled);
    if (or_fft(F, ft, ft, 1, 1, -1, 1)) return -1;
    /* phase shift to origin: shift = origin/size * 2pi */
    double sx = pp->ori[0] / ft * 2 * M_PI, sy = pp->ori[1] / ft * 2 * M_PI;
    /* img_ctf_calculate(cp, flip = (action == 1), wiener, ...): a positive
     * factor makes any other action a Wiener filter (mg_ctf.cpp:100-107) */
    int flip = rp->ctf_action == BSX_CTF_FLIP;
    double wiener = rp->ctf_action >= 2 ? rp->wiener : 0;
    double pix = pp->pixel_size, shi = 1 / hi, shi2 = shi * shi;
    double fscale = 1.0 / (ft * pix);
    for (long y = 0; y < ft; y++) {
        long k = wrap_freq(y, ft);
        for (long x = 0; x < ft; x++) {
            long h = wrap_freq(x, ft);
            long i = y * ft + x;
            double phi = h * sx + k * sy;
            double cr = cos(phi), ci = sin(phi);
            double re = F[2 * i], im = F[2 * i + 1];
            double nr = re * cr - im * ci, ni = re * ci + im * cr;
            if (rp->ctf_action) {
                double fx = h * fscale, fy = k * fscale;
                double s2 = fx * fx + fy * fy;
                double c = 0;
                if (s2 <= shi2) c = or_ctf_value(pp, fx, fy, flip, wiener);
                nr *= c; ni *= c;
            }
            F[2 * i] = (orr)nr;
            F[2 * i + 1] = (orr)ni;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* accumulators and insertion                                                */
/* ------------------------------------------------------------------------ */
typedef struct {
    orr* F;      /* 2 N^3: sum w F            */
    orr* P;      /* N^3:   sum w |F|^2        */
    orr* W;      /* N^3:   sum w              */
    orr* W2;     /* N^3:   sum w^2            */
    long count;
} or_acc;

struct or_recon {
    bsx_recon_params p;
    double hi_res;
    long N, N3, Nz;   /* Nz = 1 for a 2D map (breconstruct -TwoD), else N */
    int T;
    or_acc* acc;
    long next;   /* running particle index (round-robin partition) */
};

static void acc_add(or_acc* a, long j, double cre, double cim, double w)
{
    /* add(j, cv*w): Complex<float>*double -> rounded to the data type, then
     * summed in double and stored (Bimage.h:325, Complex.h:84) */
    orr pr = (orr)(cre * w), pi = (orr)(cim * w);
    a->F[2 * j] = (orr)((double)a->F[2 * j] + (double)pr);
    a->F[2 * j + 1] = (orr)((double)a->F[2 * j + 1] + (double)pi);
    a->P[j] = (orr)((double)a->P[j] + w * (cre * cre + cim * cim));
    a->W[j] = (orr)((double)a->W[j] + w);
    a->W2[j] = (orr)((double)a->W2[j] + w * w);
}

/* Bimage::fspace_2D_interpolate, Bimage_reconstruct.cpp:25-88 */
static void interp_2d(or_acc* a, long N, long Nz, double cre, double cim, const double m[3],
                      double part_weight, int interp)
{
    if (interp < 2) {
        long c[3];
        double w = part_weight;
        for (int i = 0; i < 3; i++) c[i] = (long)floor(m[i] + 0.5);
        if (interp == 1) {
            double dx = m[0] - c[0], dy = m[1] - c[1], dz = m[2] - c[2];
            w *= 1 - sqrt(dx * dx + dy * dy + dz * dz);
        }
        if (w > 0) {
            for (int i = 0; i < 3; i++) {                 /* index_wrap, Bimage.h:418 */
                const long L = i == 2 ? Nz : N;
                while (c[i] < 0) c[i] += L;
                while (c[i] >= L) c[i] -= L;
            }
            acc_add(a, (c[2] * N + c[1]) * N + c[0], cre, cim, w);
        }
    } else {
        long c[3];
        double f[3];
        for (int i = 0; i < 3; i++) { c[i] = (long)floor(m[i]); f[i] = m[i] - c[i]; }
        for (int zz = 0; zz < 2; zz++) {
            long jz = c[2] + zz;
            while (jz < 0) jz += Nz;
            while (jz >= Nz) jz -= Nz;
            double wz = zz ? f[2] : 1 - f[2];
            for (int yy = 0; yy < 2; yy++) {
                long jy = c[1] + yy;
                if (jy < 0) jy += N; else if (jy >= N) jy -= N;
                double wy = yy ? f[1] : 1 - f[1];
                for (int xx = 0; xx < 2; xx++) {
                    long jx = c[0] + xx;
                    if (jx < 0) jx += N; else if (jx >= N) jx -= N;
                    double wx = xx ? f[0] : 1 - f[0];
                    double w = wx * wy * wz * part_weight;
                    acc_add(a, (jz * N + jy) * N + jx, cre, cim, w);
                }
            }
        }
    }
}

/* Bimage::fspace_pack_2D, Bimage_reconstruct.cpp:112-174 */
static void pack_2d(or_acc* a, const or_recon* r, const orr* F, const double R[9],
                    const bsx_particle* pp, double part_weight)
{
    long ft = r->p.ft_size, N = r->N;
    double hi = r->hi_res < 0.5 ? 0.5 : r->hi_res;
    double mins2 = r->p.lo_res > 0 ? 1.0 / r->p.lo_res : 0;
    double maxs2 = 1.0 / hi;
    mins2 *= mins2;
    maxs2 *= maxs2;
    double scale = pp->mag > 0 ? 1.0 / pp->mag : 1.0;
    double invscale = scale / (ft * pp->pixel_size);
    double mscale = N * r->p.sampling;
    /* real size along z: N sampling in 3D, 1 x 1 in 2D (particle_reconstruct
     * sets the z sampling to 1, mg_reconstruct.cpp:117) */
    double mscz = (r->p.flags & 2) ? 1.0 : mscale;
    double ew = (r->p.flags & BSX_FLAG_EWALD) ? or_lambda(pp->volt) / 2 : 0;
    long hx = (ft - 1) / 2;
    double s[3] = {0, 0, 0}, m[3];
    for (long yy = 0, i = 0; yy < ft; yy++) {
        s[1] = (double)yy;
        if (yy > hx) s[1] -= (double)ft;
        s[1] *= invscale;
        for (long xx = 0; xx < ft; xx++, i++) {
            s[0] = (double)xx;
            if (xx > hx) s[0] -= (double)ft;
            s[0] *= invscale;
            double s2 = s[0] * s[0] + s[1] * s[1];
            if (s2 >= mins2 && s2 <= maxs2) {
                double w = part_weight;
                double d = (maxs2 - s2) * mscz;
                if (d < 1) w *= sqrt(d);
                if (ew) s[2] = ew * s2;
                for (int k = 0; k < 3; k++)
                    m[k] = (s[0] * R[3 * k] + s[1] * R[3 * k + 1] + s[2] * R[3 * k + 2]) *
                           (k == 2 ? mscz : mscale);
                interp_2d(a, N, r->Nz, F[2 * i], F[2 * i + 1], m, w, r->p.interp);
                if (ew) {
                    s[2] = -s[2];
                    for (int k = 0; k < 3; k++)
                        m[k] = (s[0] * R[3 * k] + s[1] * R[3 * k + 1] + s[2] * R[3 * k + 2]) *
                               (k == 2 ? mscz : mscale);
                    interp_2d(a, N, r->Nz, F[2 * i], F[2 * i + 1], m, w, r->p.interp);
                }
            }
        }
    }
}

or_recon* or_recon_new(const bsx_recon_params* rp, int npartial)
{
    if (rp->size < 2 || rp->box < 2 || rp->ft_size < rp->box || npartial < 1) return NULL;
    or_recon* r = (or_recon*)calloc(1, sizeof(or_recon));
    r->p = *rp;
    r->N = rp->size;
    r->Nz = (rp->flags & 2) ? 1 : rp->size;
    r->N3 = r->N * r->N * r->Nz;
    r->T = npartial;
    r->hi_res = or_check_resolution(rp->hi_res, rp->size, rp->sampling, (rp->flags & 2) ? 2 : 3);
    r->acc = (or_acc*)calloc(npartial, sizeof(or_acc));
    for (int t = 0; t < npartial; t++) {
        r->acc[t].F = (orr*)calloc(2 * r->N3, sizeof(orr));
        r->acc[t].P = (orr*)calloc(r->N3, sizeof(orr));
        r->acc[t].W = (orr*)calloc(r->N3, sizeof(orr));
        r->acc[t].W2 = (orr*)calloc(r->N3, sizeof(orr));
        if (!r->acc[t].F || !r->acc[t].P || !r->acc[t].W || !r->acc[t].W2) {
            or_recon_free(r);
            return NULL;
        }
    }
    return r;
}

void or_recon_free(or_recon* r)
{
    if (!r) return;
    for (int t = 0; t < r->T; t++) {
        free(r->acc[t].F); free(r->acc[t].P); free(r->acc[t].W); free(r->acc[t].W2);
    }
    free(r->acc);
    free(r);
}

/* The per-particle loop of particle_reconstruct (mg_reconstruct.cpp:143-229),
 * with particles dealt round-robin to T partial maps as
 * project_configure_for_reconstruction does (mg_reconstruct.cpp:741-744) and
 * the partial maps processed in parallel as breconstruct does
 * (breconstruct.cpp:281-290). */
int or_recon_insert(or_recon* r, const bsx_particle* parts, const float* imgs, long n,
                    int nthreads)
{
    long ft = r->p.ft_size, box = r->p.box;
    long first = r->next;
    int T = r->T, err = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
#endif
    for (int t = 0; t < T; t++) {
        orr* F = (orr*)malloc(sizeof(orr) * 2 * ft * ft);
        for (long g = 0; g < n; g++) {
            if ((first + g) % T != t) continue;
            const bsx_particle* pp = parts + g;
            double w;
            if (or_stage_particle(&r->p, pp, imgs + g * box * box, F, &w)) { err |= 1; break; }
            double R[9];
            or_view_matrix(pp->view, R);
            pack_2d(&r->acc[t], r, F, R, pp, w);
            r->acc[t].count++;
        }
        free(F);
    }
    r->next += n;
    return err ? -1 : 0;
}

int or_recon_partial_acc(or_recon* r, int t, orr* planes)
{
    if (t < 0 || t >= r->T) return -1;
    long N3 = r->N3;
    for (long j = 0; j < N3; j++) {
        planes[j] = r->acc[t].F[2 * j];
        planes[N3 + j] = r->acc[t].F[2 * j + 1];
        planes[2 * N3 + j] = r->acc[t].W[j];
        planes[3 * N3 + j] = r->acc[t].P[j];
        planes[4 * N3 + j] = r->acc[t].W2[j];
    }
    return 0;
}

/* Replaces the accumulators of partial map t (planes as or_recon_partial_acc
 * returns them; the particle count is set to `count`): hand-built sums for
 * the tests of the finish. */
int or_recon_set_partial_acc(or_recon* r, int t, const orr* planes, long count)
{
    if (t < 0 || t >= r->T) return -1;
    long N3 = r->N3;
    for (long j = 0; j < N3; j++) {
        r->acc[t].F[2 * j] = planes[j];
        r->acc[t].F[2 * j + 1] = planes[N3 + j];
        r->acc[t].W[j] = planes[2 * N3 + j];
        r->acc[t].P[j] = planes[3 * N3 + j];
        r->acc[t].W2[j] = planes[4 * N3 + j];
    }
    r->acc[t].count = count;
    return 0;
}

/* Bimage::friedel_check, Bimage_fspace.cpp:2024-2074 (overall residual R) */
double or_friedel(const orr* d, long x, long y, long z)
{
    double s = 0, Rr = 0, Ri = 0;
    for (long zz = 0; zz < z; zz++) {
        long zf = zz ? z - zz : 0;
        for (long yy = 0; yy < y; yy++) {
            long yf = yy ? y - yy : 0;
            for (long xx = 0; xx < (x + 1) / 2; xx++) {
                long xf = xx ? x - xx : 0;
                long i = (zz * y + yy) * x + xx, j = (zf * y + yf) * x + xf;
                double vr = d[2 * i], vi = d[2 * i + 1], fr = d[2 * j], fi = d[2 * j + 1];
                double w = (vr * vr + vi * vi + fr * fr + fi * fi) / 2;
                double dd = vr - fr;
                Rr += dd * dd * w;
                dd = vi + fi;
                Ri += dd * dd * w;
                s += w;
            }
        }
    }
    if (s <= 0) return 0;
    Rr = sqrt(Rr / s);
    Ri = sqrt(Ri / s);
    return sqrt(Rr * Rr + Ri * Ri);
}

/* img_reconstruction_sum_weigh (mg_reconstruct.cpp:266-361) with
 * fspace_reconstruction_weigh (Bimage_reconstruct.cpp:486-515),
 * fspace_reconstruction_snr (:610-634), phase_shift_to_center
 * (Bimage_complex.cpp:545-551), then fft_back + complex_to_real
 * (Bimage.h:903-908, breconstruct.cpp:340). */
int or_recon_finish(or_recon* r, int imap, int nmaps, orr* map_out, orr* fom_out,
                    orr* ft_out, bsx_stats* st)
{
    long N = r->N, N3 = r->N3, Nz = r->Nz;
    int T = r->T;
    long s0 = 0, inc = 1;
    if (nmaps == 2) { s0 = imap % 2; inc = 2; }
    if (nmaps == 3 && imap % 3) { s0 = imap % 3 - 1; inc = 2; }
    orr* F = (orr*)calloc(2 * N3, sizeof(orr));
    orr* P = (orr*)calloc(N3, sizeof(orr));
    orr* W = (orr*)calloc(N3, sizeof(orr));
    orr* W2 = (orr*)calloc(N3, sizeof(orr));
    if (!F || !P || !W || !W2) { free(F); free(P); free(W); free(W2); return -1; }
    long npart = 0;
    for (long i = s0, j = 0; j < T; i += inc, j += inc) {
        or_acc* a = &r->acc[i];
        npart += a->count;
        for (long k = 0; k < N3; k++) {
            F[2 * k] = (orr)((double)F[2 * k] + a->F[2 * k]);
            F[2 * k + 1] = (orr)((double)F[2 * k + 1] + a->F[2 * k + 1]);
            P[k] += a->P[k];
            W[k] += a->W[k];
            W2[k] += a->W2[k];
        }
    }
    /* weigh */
    long cov = 0;
    double wsum = 0;
    for (long i = 0; i < N3; i++) {
        wsum += W[i];
        if (W[i] > SMALLFLOAT) {
            W2[i] = (orr)(W[i] - (double)W2[i] / W[i]);
            F[2 * i] = (orr)((double)F[2 * i] / W[i]);
            F[2 * i + 1] = (orr)((double)F[2 * i + 1] / W[i]);
            if (W2[i] > 1) P[i] = (orr)((double)P[i] / W2[i]);
            else if (W[i] > 1) P[i] = (orr)((double)P[i] / W[i]);
            if (P[i] < 0) P[i] = 0;
            cov++;
        } else {
            F[2 * i] = F[2 * i + 1] = 0;
            P[i] = 0;
            W[i] = W2[i] = (orr)SMALLFLOAT;
        }
    }
    P[0] = 1;
    double fr = or_friedel(F, N, N, Nz);
    /* SNR figure of merit */
    for (long i = 0; i < N3; i++) {
        double signal = (double)F[2 * i] * F[2 * i] + (double)F[2 * i + 1] * F[2 * i + 1];
        double noise = P[i] - signal;
        if (noise < signal) noise = signal;
        if (W[i] > 1 && signal > SMALLFLOAT && noise > SMALLFLOAT)
            P[i] = (orr)((W[i] * signal - P[i]) / noise);
        else
            P[i] = 0;
    }
    P[0] = 1;
    /* phase shift to centre: shift = size/2, t = shift/size * (-2pi) */
    double t = -2 * M_PI * (double)(N / 2) / N;
    for (long zz = 0, i = 0; zz < Nz; zz++) {
        long l = wrap_freq(zz, Nz);
        for (long yy = 0; yy < N; yy++) {
            long k = wrap_freq(yy, N);
            for (long xx = 0; xx < N; xx++, i++) {
                long h = wrap_freq(xx, N);
                double phi = h * t + k * t + l * t;
                double cr = cos(phi), ci = sin(phi);
                double re = F[2 * i], im = F[2 * i + 1];
                F[2 * i] = (orr)(re * cr - im * ci);
                F[2 * i + 1] = (orr)(re * ci + im * cr);
            }
        }
    }
    if (ft_out) memcpy(ft_out, F, sizeof(orr) * 2 * N3);
    if (fom_out) memcpy(fom_out, P, sizeof(orr) * N3);
    if (or_fft(F, N, N, Nz, 1, +1, 1)) { free(F); free(P); free(W); free(W2); return -1; }
    if (map_out)
        for (long i = 0; i < N3; i++) map_out[i] = F[2 * i];
    if (st) {
        st->coverage = cov;
        st->friedel = fr;
        st->particles = npart;
        st->weight_sum = wsum;
    }
    free(F); free(P); free(W); free(W2);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* resolution: Bimage::fsc (Bimage_resolution.cpp:330-406) with              */
/* fspace_radial_power (Bimage_fspace.cpp:848-958): linear split of each      */
/* voxel between shells floor(r) and floor(r)+1, r in voxel units.            */
/* ------------------------------------------------------------------------ */
int or_fsc(const orr* a, const orr* b, long n, double sampling, double hi_res,
           double* fsc, int cap)
{
    hi_res = or_check_resolution(hi_res, (int)n, sampling, 3);
    double rad_scale = n * sampling;
    long maxrad = (long)(rad_scale / hi_res);
    if (maxrad > cap) return -1;
    double* s1 = (double*)calloc(3 * maxrad, sizeof(double));
    double* s2 = s1 + maxrad;
    double* sr = s2 + maxrad;
    double fs = 1.0 / (n * sampling);
    for (long zz = 0, i = 0; zz < n; zz++) {
        double rz = wrap_freq(zz, n) * fs;
        rz *= rz;
        for (long yy = 0; yy < n; yy++) {
            double ry = wrap_freq(yy, n) * fs;
            ry *= ry;
            for (long xx = 0; xx < n; xx++, i++) {
                double rx = wrap_freq(xx, n) * fs;
                rx *= rx;
                double radius = rad_scale * sqrt(rx + ry + rz);
                long ir = (long)radius;
                if (ir >= maxrad) continue;
                double f = radius - ir, f1 = 1 - f;
                double p1 = (double)a[2 * i] * a[2 * i] + (double)a[2 * i + 1] * a[2 * i + 1];
                double p2 = (double)b[2 * i] * b[2 * i] + (double)b[2 * i + 1] * b[2 * i + 1];
                double re = (double)a[2 * i] * b[2 * i] + (double)a[2 * i + 1] * b[2 * i + 1];
                s1[ir] += f1 * p1; s2[ir] += f1 * p2; sr[ir] += f1 * re;
                if (ir + 1 < maxrad) { s1[ir + 1] += f * p1; s2[ir + 1] += f * p2; sr[ir + 1] += f * re; }
            }
        }
    }
    for (long i = 0; i < maxrad; i++) {
        double d = s1[i] * s2[i];
        fsc[i] = (d > 0) ? sr[i] / sqrt(d) : 0;
    }
    if (maxrad > 0) fsc[0] = 1;
    free(s1);
    return (int)maxrad;
}

/* ------------------------------------------------------------------------ */
/* extraction: FSI_Kernel (include/FSI_Kernel.h:28-68), central_section       */
/* (Bimage_project.cpp:288-338), fspace_interpolate (Bimage_fspace.cpp:85-147)*/
/* ------------------------------------------------------------------------ */
/* The kernel table of FSI_Kernel (include/FSI_Kernel.h:28-68) from its
 * closed form: row r (frac = r/512), tap j at distance x = frac + (wd-1)/2 - j,
 * K(x) = sin(pi x)/(wd sin(pi x/wd)) cos^p(pi x/wd), limit 1 at x = 0 and 0
 * at other integers; power < 0 gives one tap of weight 1. */
static double fsi_k(double x, int wd, int power)
{
    double r = nearbyint(x);
    if (fabs(x - r) < 1e-12) return r == 0 ? 1.0 : 0.0;
    double a = M_PI * x / wd;
    return sin(M_PI * x) / (wd * sin(a)) * pow(cos(a), power);
}

int or_fsi_kernel(int wd, int power, double* data)
{
    const int ndiv = 512;
    if (power < 0) {
        for (int r = 0; r <= ndiv; r++) data[r] = 1;
        return ndiv + 1;
    }
    int c = (wd - 1) / 2;
    for (int r = 0; r <= ndiv; r++)
        for (int j = 0; j < wd; j++) data[(long)r * wd + j] = fsi_k((double)r / ndiv + c - j, wd, power);
    return (ndiv + 1) * wd;
}

static void fsi_interp(const orr* ft, long n, const double m[3], const double* K, int wd,
                       double* vr, double* vi)
{
    int ndiv = 512, hw = (wd - 1) / 2;
    long ix = (long)floor(m[0]), iy = (long)floor(m[1]), iz = (long)floor(m[2]);
    long ikx = ((long)((m[0] - ix) * ndiv + 0.5)) * wd;
    long iky = ((long)((m[1] - iy) * ndiv + 0.5)) * wd;
    long ikz = ((long)((m[2] - iz) * ndiv + 0.5)) * wd;
    ix -= hw; iy -= hw; iz -= hw;
    double sr = 0, si = 0;
    /* single-step wrap of the reference, made total (identical whenever the
     * single step suffices) */
    for (int zz = 0; zz < wd; zz++) {
        long jz = (iz + zz) % n;
        if (jz < 0) jz += n;
        for (int yy = 0; yy < wd; yy++) {
            long jy = (iy + yy) % n;
            if (jy < 0) jy += n;
            double w = K[iky + yy] * K[ikz + zz];
            for (int xx = 0; xx < wd; xx++) {
                long jx = (ix + xx) % n;
                if (jx < 0) jx += n;
                long j = (jz * n + jy) * n + jx;
                double kw = K[ikx + xx] * w;
                sr += ft[2 * j] * kw;
                si += ft[2 * j + 1] * kw;
            }
        }
    }
    *vr = sr;
    *vi = si;
}

int or_map_transform(const orr* map, long n, const double origin[3], orr* ft)
{
    long n3 = n * n * n;
    for (long i = 0; i < n3; i++) { ft[2 * i] = map[i]; ft[2 * i + 1] = 0; }
    if (or_fft(ft, n, n, n, 1, -1, 1)) return -1;
    double sh[3];
    for (int k = 0; k < 3; k++) sh[k] = origin[k] / n * 2 * M_PI;
    for (long zz = 0, i = 0; zz < n; zz++) {
        long l = wrap_freq(zz, n);
        for (long yy = 0; yy < n; yy++) {
            long k = wrap_freq(yy, n);
            double skl = l * sh[2] + k * sh[1];
            for (long xx = 0; xx < n; xx++, i++) {
                long h = wrap_freq(xx, n);
                double phi = h * sh[0] + skl;
                double cr = cos(phi), ci = sin(phi), re = ft[2 * i], im = ft[2 * i + 1];
                ft[2 * i] = (orr)(re * cr - im * ci);
                ft[2 * i + 1] = (orr)(re * ci + im * cr);
            }
        }
    }
    return 0;
}

int or_central_section(const orr* ft, long n, double sampling, const double mat[9],
                       double resolution, int kw, int kp, int center, orr* out)
{
    if (resolution < 1e-10) resolution = sampling;
    double* K = (double*)malloc(sizeof(double) * 513 * (kw > 0 ? kw : 1));
    or_fsi_kernel(kw, kp, K);
    if (kp < 0) kw = 1;
    long hx = (n - 1) / 2;
    double inv = 1.0 / (n * sampling), rs = n * sampling;
    double maxs2 = 1.0 / (resolution * resolution);
    double iv[3] = {0, 0, 0}, m[3];
    double t = -2 * M_PI * (double)(n / 2) / n;
    for (long yy = 0, i = 0; yy < n; yy++) {
        iv[1] = (double)yy;
        if (iv[1] > hx) iv[1] -= n;
        iv[1] *= inv;
        double sy2 = iv[1] * iv[1];
        for (long xx = 0; xx < n; xx++, i++) {
            iv[0] = (double)xx;
            if (iv[0] > hx) iv[0] -= n;
            iv[0] *= inv;
            double s2 = iv[0] * iv[0] + sy2;
            double vr = 0, vi = 0;
            if (s2 <= maxs2) {
                for (int k = 0; k < 3; k++)
                    m[k] = (iv[0] * mat[3 * k] + iv[1] * mat[3 * k + 1] + iv[2] * mat[3 * k + 2]) * rs;
                fsi_interp(ft, n, m, K, kw, &vr, &vi);
            }
            if (center) {
                double phi = wrap_freq(xx, n) * t + wrap_freq(yy, n) * t;
                double cr = cos(phi), ci = sin(phi);
                double nr = vr * cr - vi * ci, ni = vr * ci + vi * cr;
                vr = nr; vi = ni;
            }
            out[2 * i] = (orr)vr;
            out[2 * i + 1] = (orr)vi;
        }
    }
    free(K);
    return 0;
}

/* img_recip_space_fom, fom_type 0, unit ring weights, no CTF
 * (src/mg/mg_refine.cpp:660-781). Returns the FRC score; *cv_out receives
 * the cross-validation score over [max_rad, 1.1 max_rad). */
double or_recip_space_fom(const orr* p, const orr* ref, long n, double sampling,
                          double hi_res, double lo_res, double* cv_out)
{
    if (lo_res > 0 && hi_res > lo_res) { double t = hi_res; hi_res = lo_res; lo_res = t; }
    hi_res = or_check_resolution(hi_res, (int)n, sampling, 2);
    double max_rad = n * sampling / hi_res;
    double max_rad_cv = 1.1 * max_rad;
    if (max_rad_cv <= max_rad) max_rad_cv = max_rad + 1;
    double max_rad_sq = max_rad * max_rad, max_rad_cv_sq = max_rad_cv * max_rad_cv;
    double min_rad_sq = 0;
    (void)max_rad_sq;
    if (lo_res > 0) { min_rad_sq = n * sampling / lo_res; min_rad_sq *= min_rad_sq; }
    double* fom = (double*)calloc(3 * n, sizeof(double));
    double* dsum = fom + n;
    double* msum = dsum + n;
    long hx = (n - 1) / 2;
    for (long yy = 0, i = 0; yy < n; yy++) {
        double iy = yy > hx ? (double)yy - n : (double)yy;
        for (long xx = 0; xx < n; xx++, i++) {
            double ix = xx > hx ? (double)xx - n : (double)xx;
            double d2 = ix * ix + iy * iy;
            if (d2 >= min_rad_sq && d2 <= max_rad_cv_sq) {
                long j = (long)sqrt(d2);
                if (j >= n) continue;
                double dr = p[2 * i], di = p[2 * i + 1], rr = ref[2 * i], ri = ref[2 * i + 1];
                fom[j] += dr * rr + di * ri;
                msum[j] += rr * rr + ri * ri;
                dsum[j] += dr * dr + di * di;
            }
        }
    }
    double f = 0, wsum = 0, cv = 0;
    for (long j = 0; j < max_rad && j < n; j++) {
        wsum += 1;
        double den = dsum[j] * msum[j];
        if (den > 0) f += fom[j] / sqrt(den);
    }
    f = wsum > 0 ? f / wsum : 0;
    wsum = 0;
    for (long j = (long)max_rad; j < max_rad_cv && j < n; j++) {
        wsum += 1;
        double den = dsum[j] * msum[j];
        if (den > 0) cv += fom[j] / sqrt(den);
    }
    cv = wsum > 0 ? cv / wsum : 0;
    if (cv_out) *cv_out = cv;
    free(fom);
    return f;
}

/* asymmetric_unit_views for C1 (point 101, order 1), symmetry.cpp:383-470.
 * views: cap x 4 doubles (x, y, z, angle). Returns the count. */
long or_asu_views_c1(double theta_step, double phi_step, int flag, double* views, long cap)
{
    if (theta_step < M_PI / 1800) theta_step = M_PI / 180;
    if (phi_step < M_PI / 1800) phi_step = M_PI / 180;
    int full = flag & 1, inplanerot = !(flag & 2);
    double max_theta = M_PI_2 + 1e-6;
    int nphi = (int)(M_PI * 2.0 / (1 * phi_step) + 0.5);
    if (full) max_theta = M_PI + 1e-6;
    if (nphi < 1) nphi = 1;
    long n = 0;
#define ADDV(X, Y, Z, A) do { if (n < cap) { double _x = (X), _y = (Y), _z = (Z); \
        double _s = sqrt(_x * _x + _y * _y + _z * _z); \
        views[4 * n] = _x / _s; views[4 * n + 1] = _y / _s; views[4 * n + 2] = _z / _s; \
        views[4 * n + 3] = (A); } n++; } while (0)
    ADDV(0, 0, 1, 0);
    double lastz = 1;
    for (double theta = theta_step; theta <= max_theta; theta += theta_step) {
        int nrphi = (int)(nphi * sin(theta) / 2 + 0.5);
        double phi = 0;
        int istart = 0;
        if (nrphi) { phi = M_PI * 1.0 / (1 * nrphi); istart = -nrphi; nrphi -= 1; }
        for (int j = istart; j <= nrphi; j++) {
            ADDV(sin(theta) * cos(j * phi), sin(theta) * sin(j * phi), cos(theta),
                 inplanerot ? j * phi : 0);
            lastz = cos(theta);
        }
    }
    if (full && lastz > -0.999999) ADDV(0, 0, -1, 0);
#undef ADDV
    return n;
}

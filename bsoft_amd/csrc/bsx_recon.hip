// bsx_recon.hip — Fourier-space reconstruction on gfx950.
//
// Per batch of particles (the per-particle loop of particle_reconstruct,
// src/mg/mg_reconstruct.cpp:143-229):
//   K1 k_background  background mean outside the inscribed ellipse
//                    (Bimage::calculate_background, Bimage_background.cpp:27-89)
//   K2 k_rows        pad with the background + x-direction FFT of the padded
//                    ft x ft image (Bimage::pad, Bimage_resize.cpp:347 and the
//                    first half of Bimage::fft, Bimage_fft.cpp:176-234). Two
//                    real rows share one complex FFT; only the half spectrum
//                    x = 0..ft/2 is kept, written transposed (x-major).
//   K3 k_cols_insert y-direction FFT of each kept column, x 1/ft, DC background
//                    term, phase shift to origin (Bimage_complex.cpp:490-543),
//                    CTF (mg_ctf.cpp:46-110, 661-676) and insertion of the
//                    column and of its Friedel mate into the FP32 volume
//                    accumulators (Bimage::fspace_pack_2D and
//                    fspace_2D_interpolate, Bimage_reconstruct.cpp:25-174).
// Finish (img_reconstruction_sum_weigh, mg_reconstruct.cpp:266-361):
//   K4 k_weigh       sum of half sets, FP64 weighing, SNR FOM, centre shift
//   K5 3 axis FFTs   inverse 3D transform, x 1/sqrt(N^3), real part.
#include <cmath>
#include <cstring>
#include <vector>

#include "bsx_fft_dev.h"
#include "bsx_internal.h"

namespace bsx {
int fft_axis_ex(float2* data, float* real_out, int nx, int ny, int nz, int64_t batch, int axis,
                int sign, float scale, hipStream_t s);
}

using namespace bsx;

// Device record of one particle, all derived quantities computed on the host
// in double precision.
struct alignas(16) PartDev {
    double R[9];        // View::matrix (row-major d[i][j])
    double invscale;    // (1/mag) / (ft * pixel_size)      fspace_pack_2D invscale
    double phx, phy;    // 2 pi ori / ft                      phase_shift_to_origin
    double c0, c20, c22, c2m2, c40;   // aberration weights   ctf.h:321-335
    double fscale;      // 1 / (ft * pixel_size)              img_ctf_calculate freq_scale
    double ew;          // Ewald offset lambda/2, 0 if off
    double hox, hoy;    // background-ellipse origin (image header)
    float wp;           // particle weight incl. pad ratio
    int half;           // accumulator set
};

struct ReconConst {
    int N, ft, box, stride, interp, ctf_mode, fom, nhalves;   // ctf_mode 0 none 1 flip 2 apply 3 wiener
    double mins2, maxs2, mscale, shi2, wiener;
};

// ------------------------------------------------------------------- K1
__global__ __launch_bounds__(256) void k_background(const float* __restrict__ imgs,
                                                    const PartDev* __restrict__ recs, int box,
                                                    float* __restrict__ bkg,
                                                    int* __restrict__ nonfinite) {
    const int p = blockIdx.x;
    const float* img = imgs + (size_t)p * box * box;
    const double ox = recs[p].hox, oy = recs[p].hoy;
    const long hx = box / 2, hy = box / 2;
    double sum = 0;
    int cnt = 0, bad = 0;
    for (int i = threadIdx.x; i < box * box; i += blockDim.x) {
        const int yy = i / box, xx = i - yy * box;
        const float v = img[i];
        if (!isfinite(v)) bad = 1;
        double y2 = (yy - oy) / hy;
        y2 *= y2;
        double x2 = (xx - ox) / hx;
        x2 *= x2;
        if (x2 + y2 - 1 > 0) { sum += v; ++cnt; }
    }
    __shared__ double ssum[256];
    __shared__ int scnt[256], sbad[256];
    ssum[threadIdx.x] = sum;
    scnt[threadIdx.x] = cnt;
    sbad[threadIdx.x] = bad;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            ssum[threadIdx.x] += ssum[threadIdx.x + o];
            scnt[threadIdx.x] += scnt[threadIdx.x + o];
            sbad[threadIdx.x] |= sbad[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        bkg[p] = scnt[0] ? (float)(ssum[0] / scnt[0]) : 0.0f;
        if (sbad[0] && nonfinite) atomicOr(nonfinite, 1);
    }
}

// ------------------------------------------------------------------- K2
// One workgroup: 2*LINES rows of one particle (rows y0 + 2j, y0 + 2j + 1 form
// complex line j). Output: inter[p][x][y], x in [0, FT/2], y in [0, box).
template <int FT, int LINES>
__global__ __launch_bounds__(256) void k_rows(const float* __restrict__ imgs,
                                              const float* __restrict__ bkg, int box,
                                              int tiles_per_part, float2* __restrict__ inter,
                                              const float2* __restrict__ twg) {
    constexpr int T = FftThreads<FT>::value;
    constexpr int G = 256 / T;
    constexpr int NT = G * T;
    constexpr int HX = FT / 2 + 1;
    __shared__ float2 lds[LINES * FT + FT];
    float2* tw = lds + LINES * FT;
    const int tid = threadIdx.x;
    const int p = blockIdx.x / tiles_per_part;
    const int y0 = (blockIdx.x % tiles_per_part) * 2 * LINES;
    const float* img = imgs + (size_t)p * box * box;
    const float b = bkg[p];
    for (int i = tid; i < FT; i += NT) tw[i] = twg[i];
    // load: line j = rows (y0+2j, y0+2j+1), x < box from the image minus the
    // background, zero elsewhere (the padded region equals the background).
    for (int idx = tid; idx < LINES * FT; idx += NT) {
        const int j = idx / FT, x = idx - j * FT;
        const int ya = y0 + 2 * j, yb = ya + 1;
        float a = 0.f, c = 0.f;
        if (x < box) {
            if (ya < box) a = img[(size_t)ya * box + x] - b;
            if (yb < box) c = img[(size_t)yb * box + x] - b;
        }
        lds[j * FT + x] = make_float2(a, c);
    }
    __syncthreads();
    const int g = tid / T, t = tid - g * T;
#pragma unroll 1
    for (int r0 = 0; r0 < LINES; r0 += G) FftLine<FT, -1, T>::run(lds + (r0 + g) * FT, tw, t);
    // store: for each x in [0, FT/2], the 2*LINES consecutive rows
    constexpr int ROWS = 2 * LINES;
    float2* out = inter + (size_t)p * HX * box;
    for (int idx = tid; idx < HX * ROWS; idx += NT) {
        const int x = idx / ROWS, r = idx - x * ROWS;
        const int y = y0 + r;
        if (y >= box) continue;
        const int j = r >> 1;
        const float2 z = lds[j * FT + x];
        const float2 zm = lds[j * FT + ((FT - x) & (FT - 1))];
        float2 v;
        if ((r & 1) == 0) v = make_float2(0.5f * (z.x + zm.x), 0.5f * (z.y - zm.y));   // A(x)
        else v = make_float2(0.5f * (z.y + zm.y), -0.5f * (z.x - zm.x));              // B(x)
        out[(size_t)x * box + y] = v;
    }
}

// ------------------------------------------------------------------- K3
__device__ __forceinline__ void acc_atomic(float* rec, float vr, float vi, double w, int fom) {
    const float wf = (float)w;
    unsafeAtomicAdd(rec + 0, (float)(vr * w));
    unsafeAtomicAdd(rec + 1, (float)(vi * w));
    unsafeAtomicAdd(rec + 2, wf);
    if (fom) {
        unsafeAtomicAdd(rec + 3, (float)(w * ((double)vr * vr + (double)vi * vi)));
        unsafeAtomicAdd(rec + 4, (float)(w * w));
    }
}

__device__ __forceinline__ long wrapi(long c, long n) {
    c %= n;
    return c < 0 ? c + n : c;
}

// fspace_2D_interpolate (Bimage_reconstruct.cpp:25-88) for one sample.
__device__ __forceinline__ void interpolate(float* __restrict__ acc, const ReconConst& K,
                                            const double m[3], float vr, float vi, double w) {
    const long N = K.N;
    if (K.interp < 2) {
        const double fx = floor(m[0] + 0.5), fy = floor(m[1] + 0.5), fz = floor(m[2] + 0.5);
        if (K.interp == 1) {
            const double dx = m[0] - fx, dy = m[1] - fy, dz = m[2] - fz;
            w *= 1 - sqrt(dx * dx + dy * dy + dz * dz);
        }
        if (w > 0) {
            const long cx = wrapi((long)fx, N), cy = wrapi((long)fy, N), cz = wrapi((long)fz, N);
            acc_atomic(acc + ((cz * N + cy) * N + cx) * K.stride, vr, vi, w, K.fom);
        }
    } else {
        const double fx = floor(m[0]), fy = floor(m[1]), fz = floor(m[2]);
        const double dx = m[0] - fx, dy = m[1] - fy, dz = m[2] - fz;
        const long cx = (long)fx, cy = (long)fy, cz = (long)fz;
#pragma unroll
        for (int zz = 0; zz < 2; ++zz) {
            const long jz = wrapi(cz + zz, N);
            const double wz = zz ? dz : 1 - dz;
#pragma unroll
            for (int yy = 0; yy < 2; ++yy) {
                const long jy = wrapi(cy + yy, N);
                const double wy = yy ? dy : 1 - dy;
#pragma unroll
                for (int xx = 0; xx < 2; ++xx) {
                    const long jx = wrapi(cx + xx, N);
                    const double wx = xx ? dx : 1 - dx;
                    const double ww = wx * wy * wz * w;
                    acc_atomic(acc + ((jz * N + jy) * N + jx) * K.stride, vr, vi, ww, K.fom);
                }
            }
        }
    }
}

// One Fourier sample (h, k) of a particle with raw value F (already x 1/ft,
// background DC included): phase shift, CTF, resolution band, taper,
// rotation and interpolation — fspace_pack_2D's loop body.
__device__ __forceinline__ void insert_sample(float* __restrict__ acc, const ReconConst& K,
                                              const PartDev& P, long h, long k, float2 F) {
    const double s0 = (double)h * P.invscale, s1 = (double)k * P.invscale;
    const double s2 = s0 * s0 + s1 * s1;
    if (!(s2 >= K.mins2 && s2 <= K.maxs2)) return;
    double w = P.wp;
    const double d = (K.maxs2 - s2) * K.mscale;
    if (d < 1) w *= sqrt(d);
    // phase shift to origin
    const double phi = (double)h * P.phx + (double)k * P.phy;
    double sn, cs;
    sincos(phi, &sn, &cs);
    double vr = F.x * cs - F.y * sn, vi = F.x * sn + F.y * cs;
    vr = (float)vr;   // stored as FP32 after the shift (Bimage_complex.cpp:536)
    vi = (float)vi;
    if (K.ctf_mode) {
        const double fx = (double)h * P.fscale, fy = (double)k * P.fscale;
        const double cs2 = fx * fx + fy * fy;
        double c = 0;
        if (cs2 <= K.shi2) {
            double c2a = 1, s2a = 0;
            if (cs2 > 0) {
                c2a = (fx * fx - fy * fy) / cs2;
                s2a = 2 * fx * fy / cs2;
            }
            double dphi = P.c0;
            dphi += P.c2m2 * cs2 * s2a;
            dphi += P.c20 * cs2;
            dphi += P.c22 * cs2 * c2a;
            dphi += P.c40 * cs2 * cs2;
            c = sin(dphi);
            if (K.ctf_mode == 1) c = (c < 0) ? -1 : 1;
            else if (K.ctf_mode == 3) c = (K.wiener + 1) * c / (c * c + K.wiener);
        }
        vr *= c;
        vi *= c;
    }
    const float fr = (float)vr, fi = (float)vi;
    double m[3];
    const double mz0 = P.ew * s2;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const double sz = e ? -mz0 : mz0;
        m[0] = (s0 * P.R[0] + s1 * P.R[1] + sz * P.R[2]) * K.mscale;
        m[1] = (s0 * P.R[3] + s1 * P.R[4] + sz * P.R[5]) * K.mscale;
        m[2] = (s0 * P.R[6] + s1 * P.R[7] + sz * P.R[8]) * K.mscale;
        interpolate(acc, K, m, fr, fi, w);
        if (P.ew == 0) break;
    }
}

// Column tasks: task = (particle, x) with x in [0, FT/2]; G tasks per WG.
template <int FT>
__global__ __launch_bounds__(256) void k_cols_insert(const float2* __restrict__ inter,
                                                     const float* __restrict__ bkg,
                                                     const PartDev* __restrict__ recs, int nb,
                                                     int box, ReconConst K,
                                                     float* __restrict__ acc,
                                                     const float2* __restrict__ twg) {
    constexpr int T = FftThreads<FT>::value;
    constexpr int G = 256 / T;
    constexpr int NT = G * T;
    constexpr int HX = FT / 2 + 1;
    __shared__ float2 lds[G * FT + FT];
    float2* tw = lds + G * FT;
    const int tid = threadIdx.x;
    for (int i = tid; i < FT; i += NT) tw[i] = twg[i];
    const int g = tid / T, t = tid - g * T;
    const long task = (long)blockIdx.x * G + g;
    const int p = (int)(task / HX);
    const int x = (int)(task - (long)p * HX);
    const bool valid = p < nb;
    float2* line = lds + g * FT;
    const float2* col = inter + ((size_t)p * HX + x) * box;
    for (int i = t; i < FT; i += T) line[i] = (valid && i < box) ? col[i] : make_float2(0.f, 0.f);
    __syncthreads();
    FftLine<FT, -1, T>::run(line, tw, t);
    if (!valid) return;
    const PartDev& P = recs[p];
    float* a = acc + (size_t)P.half * K.N * K.N * K.N * K.stride;
    const float inv = 1.0f / FT;
    const float b = bkg[p];
    const long h = (x <= (FT - 1) / 2) ? x : x - FT;
    const bool mirror = x >= 1 && x <= FT / 2 - 1;
#pragma unroll 1
    for (int r = 0; r < 8; ++r) {
        const int y = t + r * T;
        float2 F = line[y];
        F.x *= inv;
        F.y *= inv;
        if (x == 0 && y == 0) F.x += b * FT;
        const long k = (y <= (FT - 1) / 2) ? y : y - FT;
        insert_sample(a, K, P, h, k, F);
        if (mirror) {
            const int ym = (FT - y) & (FT - 1);
            const long km = (ym <= (FT - 1) / 2) ? ym : ym - FT;
            insert_sample(a, K, P, -h, km, make_float2(F.x, -F.y));
        }
    }
}

// ------------------------------------------------------------------- K4
// Sum of half sets, weighing (fspace_reconstruction_weigh,
// Bimage_reconstruct.cpp:486-515) in FP64, SNR FOM (:610-634), centre shift
// (phase_shift_to_center: x (-1)^(h+k+l) for even N, computed exactly via the
// phase for odd N), output complex volume.
__global__ __launch_bounds__(256) void k_weigh(const float* __restrict__ acc, long long n3, int N,
                                               int stride, int hmask, int nhalves, int fom,
                                               float2* __restrict__ out,
                                               float* __restrict__ fom_out,
                                               unsigned long long* __restrict__ cov,
                                               double* __restrict__ wsum) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double wloc = 0;
    int cloc = 0;
    for (; i < n3; i += (long long)gridDim.x * blockDim.x) {
        double fr = 0, fi = 0, W = 0, P = 0, W2 = 0;
        for (int hh = 0; hh < nhalves; ++hh) {
            if (!((hmask >> hh) & 1)) continue;
            const float* r = acc + ((size_t)hh * n3 + i) * stride;
            fr += r[0];
            fi += r[1];
            W += r[2];
            if (fom) { P += r[3]; W2 += r[4]; }
        }
        wloc += W;
        float ofr = 0.f, ofi = 0.f, oP = 0.f;
        if (W > 1e-30) {
            ofr = (float)(fr / W);
            ofi = (float)(fi / W);
            if (fom) {
                const double w2 = W - W2 / W;
                if (w2 > 1) P /= w2;
                else if (W > 1) P /= W;
                if (P < 0) P = 0;
            }
            ++cloc;
        } else {
            P = 0;
            W = 1e-30;
        }
        if (fom_out) {
            if (i == 0) P = 1;
            const double signal = (double)ofr * ofr + (double)ofi * ofi;
            double noise = P - signal;
            if (noise < signal) noise = signal;
            if (W > 1 && signal > 1e-30 && noise > 1e-30) oP = (float)((W * signal - P) / noise);
            else oP = 0.f;
            if (i == 0) oP = 1.f;
            fom_out[i] = oP;
        }
        const long long xx = i % N, yy = (i / N) % N, zz = i / ((long long)N * N);
        const long h = xx > (N - 1) / 2 ? xx - N : xx;
        const long k = yy > (N - 1) / 2 ? yy - N : yy;
        const long l = zz > (N - 1) / 2 ? zz - N : zz;
        if ((N & 1) == 0) {
            if ((h + k + l) & 1) { ofr = -ofr; ofi = -ofi; }
        } else {
            const double t = -2 * M_PI * (double)(N / 2) / N;
            double sn, cs;
            sincos(h * t + k * t + l * t, &sn, &cs);
            const double a = ofr, b = ofi;
            ofr = (float)(a * cs - b * sn);
            ofi = (float)(a * sn + b * cs);
        }
        out[i] = make_float2(ofr, ofi);
    }
    // block reductions of coverage and weight
    __shared__ double sw[256];
    __shared__ unsigned int sc[256];
    sw[threadIdx.x] = wloc;
    sc[threadIdx.x] = cloc;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) { sw[threadIdx.x] += sw[threadIdx.x + o]; sc[threadIdx.x] += sc[threadIdx.x + o]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(cov, (unsigned long long)sc[0]);
        atomicAdd(wsum, sw[0]);
    }
}

// Friedel residual sums (Bimage::friedel_check, Bimage_fspace.cpp:2024-2074).
__global__ __launch_bounds__(256) void k_friedel(const float2* __restrict__ v, int N,
                                                 double* __restrict__ sums) {
    const long long nx2 = (N + 1) / 2;
    const long long total = nx2 * N * N;
    double s = 0, rr = 0, ri = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const long long xx = i % nx2, yy = (i / nx2) % N, zz = i / (nx2 * N);
        const long long xf = xx ? N - xx : 0, yf = yy ? N - yy : 0, zf = zz ? N - zz : 0;
        const float2 a = v[(zz * N + yy) * N + xx];
        const float2 b = v[(zf * N + yf) * N + xf];
        const double w = ((double)a.x * a.x + (double)a.y * a.y + (double)b.x * b.x + (double)b.y * b.y) / 2;
        double d = (double)a.x - b.x;
        rr += d * d * w;
        d = (double)a.y + b.y;
        ri += d * d * w;
        s += w;
    }
    __shared__ double sh[3][256];
    sh[0][threadIdx.x] = s;
    sh[1][threadIdx.x] = rr;
    sh[2][threadIdx.x] = ri;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int q = 0; q < 3; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) atomicAdd(sums + q, sh[q][0]);
}

// Copy raw accumulator planes.
__global__ void k_planes(const float* __restrict__ acc, long long n3, int stride,
                         float* __restrict__ out) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
         i += (long long)gridDim.x * blockDim.x) {
        const float* r = acc + i * stride;
        for (int q = 0; q < 5; ++q) out[q * n3 + i] = q < stride ? r[q] : 0.f;
    }
}

// ================================================================== host side
struct bsx_recon {
    bsx_recon_params p;
    ReconConst K;
    double hi_res;
    int dev;
    float* acc = nullptr;
    size_t acc_elems = 0;
    float2* inter = nullptr;
    float* bkg = nullptr;
    PartDev* recs = nullptr;
    int* nonfinite = nullptr;
    int64_t recs_cap = 0;
    int batch = 0;
    int64_t inserted[2] = {0, 0};
    double last_insert_ms = 0, last_stage_ms = 0;
    int64_t last_launches = 0;
    std::vector<hipEvent_t> evs;
    size_t bytes = 0;
    PartDev* hrecs = nullptr;      // pinned staging of the records
    int64_t hrecs_cap = 0;
    hipEvent_t copied = nullptr;   // staging buffer reusable once this completes
};

static double check_resolution(double res, int n, double sampling) {
    // Bimage::check_resolution (Bimage.cpp:633-648), cubic 3D map
    if (res > n * sampling) return n * sampling;
    double rm = 2 / sqrt(3.0 / (sampling * sampling));
    return res < rm ? rm : res;
}

static double lambda_of(double av) {
    // CTFparam::lambda (include/ctf.h:309-317) with the constants of
    // include/utilities.h:55-58
    const double PLANCK = 6.626070E-34, ECHARGE = 1.602177E-19;
    const double EMASS = 9.109384E-31, LIGHTSPEED = 299792458.0;
    if (av <= 0) return 0;
    return PLANCK * 1e10 /
           sqrt(2.0 * EMASS * ECHARGE * av * (1.0 + ECHARGE * av / (2.0 * EMASS * LIGHTSPEED * LIGHTSPEED)));
}

// View::matrix (include/View.h:131-157; Matrix3.h:30-42)
static void view_matrix(const double view[4], double d[9]) {
    double v0 = view[0], v1 = view[1], v2 = view[2], a = view[3];
    double size = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
    if (size < 1e-30) { v0 = v1 = 0; v2 = 1; }
    else { v0 /= size; v1 /= size; v2 /= size; }
    a = remainder(a, 2 * M_PI);
    double ang = a / 2.0, ca = cos(ang), sa = sin(ang);
    double q0 = ca, q1 = 0, q2 = 0, q3 = sa;
    if (v2 < 1.0 - 1e-30) {
        if (v2 <= 1e-30 - 1.0) { q0 = 0; q1 = sa; q2 = ca; q3 = 0; }
        else {
            double z1 = v2 + 1, f1 = sqrt(z1 / 2.0), f2 = sqrt(1.0 / (2 * z1));
            q0 = f1 * ca; q1 = f2 * (v0 * sa - v1 * ca); q2 = f2 * (v0 * ca + v1 * sa); q3 = f1 * sa;
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
    for (int i = 0; i < 9; i++) {
        if (fabs(d[i]) < 1e-12) d[i] = 0;
        if (d[i] > 1) d[i] = 1;
        if (d[i] < -1) d[i] = -1;
    }
}

static void make_rec(const bsx_recon* h, const bsx_particle& q, PartDev& r) {
    const int ft = h->p.ft_size, box = h->p.box;
    view_matrix(q.view, r.R);
    double scale = q.mag > 0 ? 1.0 / q.mag : 1.0;
    r.invscale = scale / (ft * q.pixel_size);
    r.phx = q.ori[0] / ft * 2 * M_PI;
    r.phy = q.ori[1] / ft * 2 * M_PI;
    double wl = lambda_of(q.volt);
    double t1 = M_PI_2 * wl * wl * wl, t2 = -M_PI * wl, dev = fabs(q.def_dev);
    r.c0 = -q.amp_shift;
    r.c20 = t2 * q.def_avg;
    r.c22 = t2 * dev * cos(2 * q.ast_angle);
    r.c2m2 = t2 * dev * sin(2 * q.ast_angle);
    r.c40 = t1 * q.cs;
    r.fscale = 1.0 / (ft * q.pixel_size);
    r.ew = (h->p.flags & BSX_FLAG_EWALD) ? wl / 2 : 0;
    r.hox = q.hdr_ori[0];
    r.hoy = q.hdr_ori[1];
    double w = (h->p.flags & BSX_FLAG_BOOTSTRAP) ? q.sel : 1.0;
    double pad_ratio = (double)ft / box;
    if (pad_ratio > 1) { pad_ratio *= pad_ratio * 0.6; w /= pad_ratio; }
    r.wp = (float)w;
    r.half = h->p.nhalves > 1 ? (q.half & 1) : 0;
}

template <int FT>
static int launch_batch(bsx_recon* h, const float* imgs, int64_t p0, int nb, hipStream_t s,
                        hipEvent_t e0, hipEvent_t e1) {
    constexpr int T = FftThreads<FT>::value;
    constexpr int G = 256 / T;
    constexpr int LINES = G >= 8 ? G : 8;   // 16 rows per row-tile at FT = 512
    constexpr int NT = G * T;
    const int box = h->p.box;
    const float2* tw = twiddles(FT, s);
    BSX_REQUIRE(tw, BSX_ERR_DEVICE, "twiddles(%d) failed", FT);
    const int tiles = (box + 2 * LINES - 1) / (2 * LINES);
    hipLaunchKernelGGL((k_rows<FT, LINES>), dim3(nb * tiles), dim3(NT), 0, s,
                       imgs + (size_t)p0 * box * box, h->bkg + p0, box, tiles, h->inter, tw);
    BSX_HIP(hipGetLastError());
    const long ntask = (long)nb * (FT / 2 + 1);
    const long nblk = (ntask + G - 1) / G;
    if (e0) BSX_HIP(hipEventRecord(e0, s));
    hipLaunchKernelGGL((k_cols_insert<FT>), dim3((unsigned)nblk), dim3(NT), 0, s, h->inter,
                       h->bkg + p0, h->recs + p0, nb, box, h->K, h->acc, tw);
    if (e1) BSX_HIP(hipEventRecord(e1, s));
    BSX_HIP(hipGetLastError());
    return BSX_OK;
}

static int run_batch(bsx_recon* h, const float* imgs, int64_t p0, int nb, hipStream_t s,
                     hipEvent_t e0, hipEvent_t e1) {
    switch (h->p.ft_size) {
        case 32: return launch_batch<32>(h, imgs, p0, nb, s, e0, e1);
        case 64: return launch_batch<64>(h, imgs, p0, nb, s, e0, e1);
        case 128: return launch_batch<128>(h, imgs, p0, nb, s, e0, e1);
        case 256: return launch_batch<256>(h, imgs, p0, nb, s, e0, e1);
        case 512: return launch_batch<512>(h, imgs, p0, nb, s, e0, e1);
        default:
            set_error("ft_size %d not supported (32..512, power of 2)", h->p.ft_size);
            return BSX_ERR_ARG;
    }
}

extern "C" {

int bsx_version(int* major, int* minor) {
    if (major) *major = 0;
    if (minor) *minor = 1;
    return BSX_OK;
}

int bsx_device_info(char* name, int name_len, int* cu_count, int64_t* hbm_bytes) {
    int dev = 0;
    BSX_HIP(hipGetDevice(&dev));
    hipDeviceProp_t pr;
    BSX_HIP(hipGetDeviceProperties(&pr, dev));
    if (name && name_len > 0) {
        strncpy(name, pr.gcnArchName, name_len - 1);
        name[name_len - 1] = 0;
    }
    if (cu_count) *cu_count = pr.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = (int64_t)pr.totalGlobalMem;
    return BSX_OK;
}

// part_ft_size, src/mg/mg_reconstruct.cpp:38-64
int bsx_part_ft_size(int xsize, double scale, int pad_factor) {
    if (xsize <= 0) {
        set_error("part_ft_size: size not set");
        return 0;
    }
    auto next_pow2 = [](int start) { int r = 2; while (r < start) r *= 2; return r; };
    int ft = xsize;
    if (scale > 1) ft = (int)(xsize * scale);
    if (pad_factor < 1) return ft;
    if (pad_factor > 8) pad_factor = 8;
    ft = next_pow2((int)(pad_factor * xsize * scale));
    if (ft < xsize) ft = next_pow2(xsize);
    return ft;
}

int bsx_recon_create(bsx_recon** out, const bsx_recon_params* p) {
    BSX_REQUIRE(out && p, BSX_ERR_ARG, "bsx_recon_create: null argument");
    *out = nullptr;
    BSX_REQUIRE(p->size >= 2 && p->size <= 1024, BSX_ERR_ARG, "size %d out of range", p->size);
    BSX_REQUIRE(fft_length_supported(p->size), BSX_ERR_ARG,
                "map size %d: inverse FFT length not supported", p->size);
    BSX_REQUIRE(p->box >= 2 && p->ft_size >= p->box, BSX_ERR_ARG, "box %d / ft_size %d invalid",
                p->box, p->ft_size);
    BSX_REQUIRE(p->ft_size >= 32 && p->ft_size <= 512 && (p->ft_size & (p->ft_size - 1)) == 0,
                BSX_ERR_ARG, "ft_size %d not supported (32..512, power of 2)", p->ft_size);
    BSX_REQUIRE(p->interp >= 0 && p->interp <= 2, BSX_ERR_ARG, "interp %d invalid", p->interp);
    BSX_REQUIRE(p->nhalves == 1 || p->nhalves == 2, BSX_ERR_ARG, "nhalves must be 1 or 2");
    BSX_REQUIRE(p->sampling > 0, BSX_ERR_ARG, "sampling must be positive");
    bsx_recon* h = new bsx_recon();
    h->p = *p;
    BSX_HIP(hipGetDevice(&h->dev));
    h->hi_res = check_resolution(p->hi_res > 0 ? p->hi_res : p->sampling, p->size, p->sampling);
    ReconConst& K = h->K;
    K.N = p->size;
    K.ft = p->ft_size;
    K.box = p->box;
    K.fom = (p->flags & BSX_FLAG_FOM) ? 1 : 0;
    K.stride = K.fom ? 5 : 3;
    K.interp = p->interp;
    K.nhalves = p->nhalves;
    K.ctf_mode = p->ctf_action == 0 ? 0 : p->ctf_action == 1 ? 1
                 : (p->ctf_action > 2 && p->wiener > 0) ? 3 : 2;
    K.wiener = p->ctf_action > 2 ? p->wiener : 0;
    double hi = h->hi_res < 0.5 ? 0.5 : h->hi_res;   // fspace_pack_2D limit (:116)
    K.mins2 = p->lo_res > 0 ? 1.0 / (p->lo_res * p->lo_res) : 0;
    K.maxs2 = 1.0 / (hi * hi);
    K.mscale = p->size * p->sampling;
    K.shi2 = 1.0 / (h->hi_res * h->hi_res);          // img_ctf_calculate hires
    const size_t n3 = (size_t)p->size * p->size * p->size;
    h->acc_elems = n3 * K.stride * p->nhalves;
    int batch = p->batch;
    if (batch <= 0) {
        // intermediate half-spectra of ~128 MB per batch
        size_t per = (size_t)(p->ft_size / 2 + 1) * p->box * sizeof(float2);
        batch = (int)std::max<size_t>(1, std::min<size_t>(4096, (128u << 20) / per));
    }
    h->batch = batch;
    size_t inter_bytes = (size_t)batch * (p->ft_size / 2 + 1) * p->box * sizeof(float2);
    h->bytes = h->acc_elems * sizeof(float) + inter_bytes;
    if (p->mem_cap > 0 && (int64_t)h->bytes > p->mem_cap) {
        set_error("device memory %zu bytes exceeds the cap %lld", h->bytes, (long long)p->mem_cap);
        delete h;
        return BSX_ERR_NOMEM;
    }
    if (hipMalloc(&h->acc, h->acc_elems * sizeof(float)) != hipSuccess ||
        hipMalloc(&h->inter, inter_bytes) != hipSuccess ||
        hipMalloc(&h->nonfinite, sizeof(int)) != hipSuccess) {
        set_error("device allocation of %zu bytes failed", h->bytes);
        if (h->acc) (void)hipFree(h->acc);
        if (h->inter) (void)hipFree(h->inter);
        delete h;
        return BSX_ERR_NOMEM;
    }
    BSX_HIP(hipMemset(h->acc, 0, h->acc_elems * sizeof(float)));
    BSX_HIP(hipDeviceSynchronize());
    *out = h;
    return BSX_OK;
}

int bsx_recon_destroy(bsx_recon* h) {
    if (!h) return BSX_OK;
    if (h->acc) (void)hipFree(h->acc);
    if (h->inter) (void)hipFree(h->inter);
    if (h->bkg) (void)hipFree(h->bkg);
    if (h->recs) (void)hipFree(h->recs);
    if (h->nonfinite) (void)hipFree(h->nonfinite);
    for (auto e : h->evs) (void)hipEventDestroy(e);
    if (h->copied) {
        (void)hipEventSynchronize(h->copied);
        (void)hipEventDestroy(h->copied);
    }
    if (h->hrecs) (void)hipHostFree(h->hrecs);
    delete h;
    return BSX_OK;
}

int bsx_recon_reset(bsx_recon* h, void* stream) {
    BSX_REQUIRE(h, BSX_ERR_ARG, "null handle");
    BSX_HIP(hipMemsetAsync(h->acc, 0, h->acc_elems * sizeof(float), (hipStream_t)stream));
    h->inserted[0] = h->inserted[1] = 0;
    return BSX_OK;
}

int bsx_recon_device_bytes(const bsx_recon* h, int64_t* bytes) {
    BSX_REQUIRE(h && bytes, BSX_ERR_ARG, "null argument");
    *bytes = (int64_t)h->bytes + h->recs_cap * (int64_t)(sizeof(PartDev) + sizeof(float));
    return BSX_OK;
}

int bsx_recon_insert(bsx_recon* h, const float* imgs_dev, const bsx_particle* parts, int64_t n,
                     void* stream) {
    BSX_REQUIRE(h, BSX_ERR_ARG, "null handle");
    if (n == 0) return BSX_OK;
    BSX_REQUIRE(n > 0 && imgs_dev && parts, BSX_ERR_ARG, "bsx_recon_insert: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    BSX_HIP(hipSetDevice(h->dev));
    for (int64_t i = 0; i < n; ++i) {
        BSX_REQUIRE(parts[i].pixel_size > 0, BSX_ERR_ARG, "particle %lld: pixel size not set",
                    (long long)i);
        const bsx_particle& q = parts[i];
        const double vals[] = {q.view[0], q.view[1], q.view[2], q.view[3], q.ori[0], q.ori[1],
                               q.hdr_ori[0], q.hdr_ori[1], q.mag, q.volt, q.cs, q.amp_shift,
                               q.def_avg, q.def_dev, q.ast_angle, q.sel};
        for (double v : vals)
            BSX_REQUIRE(std::isfinite(v), BSX_ERR_ARG, "particle %lld: non-finite parameter",
                        (long long)i);
        BSX_REQUIRE(h->p.nhalves == 1 || parts[i].half == 0 || parts[i].half == 1, BSX_ERR_ARG,
                    "particle %lld: half must be 0 or 1", (long long)i);
    }
    if (n > h->recs_cap) {
        if (h->recs) (void)hipFree(h->recs);
        if (h->bkg) (void)hipFree(h->bkg);
        h->recs = nullptr;
        h->bkg = nullptr;
        int64_t extra = n * (int64_t)(sizeof(PartDev) + sizeof(float));
        if (h->p.mem_cap > 0 && (int64_t)h->bytes + extra > h->p.mem_cap) {
            h->recs_cap = 0;
            set_error("particle records (%lld bytes) exceed the device-memory cap", (long long)extra);
            return BSX_ERR_NOMEM;
        }
        if (hipMalloc(&h->recs, sizeof(PartDev) * n) != hipSuccess ||
            hipMalloc(&h->bkg, sizeof(float) * n) != hipSuccess) {
            h->recs_cap = 0;
            set_error("allocation of %lld particle records failed", (long long)n);
            return BSX_ERR_NOMEM;
        }
        h->recs_cap = n;
    }
    if (h->copied) BSX_HIP(hipEventSynchronize(h->copied));
    else BSX_HIP(hipEventCreateWithFlags(&h->copied, hipEventDisableTiming));
    if (n > h->hrecs_cap) {
        if (h->hrecs) BSX_HIP(hipHostFree(h->hrecs));
        h->hrecs = nullptr;
        h->hrecs_cap = 0;
        BSX_HIP(hipHostMalloc((void**)&h->hrecs, sizeof(PartDev) * n, hipHostMallocDefault));
        h->hrecs_cap = n;
    }
    for (int64_t i = 0; i < n; ++i) make_rec(h, parts[i], h->hrecs[i]);
    BSX_HIP(hipMemcpyAsync(h->recs, h->hrecs, sizeof(PartDev) * n, hipMemcpyHostToDevice, s));
    BSX_HIP(hipEventRecord(h->copied, s));
    const bool check_nan = h->p.flags & BSX_FLAG_CHECK_NAN;
    if (check_nan) BSX_HIP(hipMemsetAsync(h->nonfinite, 0, sizeof(int), s));
    const bool timing = h->p.flags & BSX_FLAG_TIMING;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (timing) {
        while (h->evs.size() < 2) {
            hipEvent_t e;
            BSX_HIP(hipEventCreate(&e));
            h->evs.push_back(e);
        }
        t0 = h->evs[0];
        t1 = h->evs[1];
        BSX_HIP(hipEventRecord(t0, s));
    }
    hipLaunchKernelGGL(k_background, dim3((unsigned)n), dim3(256), 0, s, imgs_dev, h->recs,
                       h->p.box, h->bkg, check_nan ? h->nonfinite : nullptr);
    BSX_HIP(hipGetLastError());
    if (check_nan) {
        int bad = 0;
        BSX_HIP(hipMemcpyAsync(&bad, h->nonfinite, sizeof(int), hipMemcpyDeviceToHost, s));
        BSX_HIP(hipStreamSynchronize(s));
        BSX_REQUIRE(!bad, BSX_ERR_NONFINITE, "non-finite values in the particle images");
    }
    // per-batch kernel timing of the insertion kernel
    const int64_t nbatches = (n + h->batch - 1) / h->batch;
    if (timing) {
        while ((int64_t)h->evs.size() < 2 + 2 * nbatches) {
            hipEvent_t e;
            BSX_HIP(hipEventCreate(&e));
            h->evs.push_back(e);
        }
    }
    for (int64_t b = 0; b < nbatches; ++b) {
        const int64_t p0 = b * h->batch;
        const int nb = (int)std::min<int64_t>(h->batch, n - p0);
        int r = run_batch(h, imgs_dev, p0, nb, s, timing ? h->evs[2 + 2 * b] : nullptr,
                          timing ? h->evs[3 + 2 * b] : nullptr);
        if (r != BSX_OK) return r;
    }
    for (int64_t i = 0; i < n; ++i) h->inserted[h->p.nhalves > 1 ? (parts[i].half & 1) : 0]++;
    if (timing) {
        BSX_HIP(hipEventRecord(t1, s));
        BSX_HIP(hipStreamSynchronize(s));
        float tot = 0, ins = 0;
        BSX_HIP(hipEventElapsedTime(&tot, t0, t1));
        for (int64_t b = 0; b < nbatches; ++b) {
            float ms = 0;
            BSX_HIP(hipEventElapsedTime(&ms, h->evs[2 + 2 * b], h->evs[3 + 2 * b]));
            ins += ms;
        }
        h->last_insert_ms = ins;
        h->last_stage_ms = tot - ins;
        h->last_launches = nbatches;
    }
    return BSX_OK;
}

int bsx_recon_last_timing(const bsx_recon* h, double* insert_ms, int64_t* launches,
                          double* stage_ms) {
    BSX_REQUIRE(h, BSX_ERR_ARG, "null handle");
    if (insert_ms) *insert_ms = h->last_insert_ms;
    if (launches) *launches = h->last_launches;
    if (stage_ms) *stage_ms = h->last_stage_ms;
    return BSX_OK;
}

int bsx_recon_accumulators(bsx_recon* h, int half, float* planes_dev, void* stream) {
    BSX_REQUIRE(h && planes_dev, BSX_ERR_ARG, "null argument");
    BSX_REQUIRE(half >= 0 && half < h->p.nhalves, BSX_ERR_ARG, "half %d out of range", half);
    const long long n3 = (long long)h->K.N * h->K.N * h->K.N;
    hipLaunchKernelGGL(k_planes, dim3(1024), dim3(256), 0, (hipStream_t)stream,
                       h->acc + (size_t)half * n3 * h->K.stride, n3, h->K.stride, planes_dev);
    BSX_HIP(hipGetLastError());
    return BSX_OK;
}

int bsx_recon_finish(bsx_recon* h, int imap, int nmaps, float* map_dev, float* fom_dev,
                     float* ft_dev, bsx_stats* stats, void* stream) {
    BSX_REQUIRE(h, BSX_ERR_ARG, "null handle");
    BSX_REQUIRE(nmaps >= 1 && nmaps <= 3 && imap >= 0 && imap < nmaps, BSX_ERR_ARG,
                "imap %d / nmaps %d invalid", imap, nmaps);
    BSX_REQUIRE(!fom_dev || h->K.fom, BSX_ERR_STATE, "FOM map requested without BSX_FLAG_FOM");
    int hmask = 3;
    if (nmaps == 2) hmask = 1 << (imap % 2);
    if (nmaps == 3 && imap % 3) hmask = 1 << (imap % 3 - 1);
    if (h->p.nhalves == 1) {
        BSX_REQUIRE(hmask == 3, BSX_ERR_STATE, "half maps need nhalves = 2");
        hmask = 1;
    }
    hipStream_t s = (hipStream_t)stream;
    BSX_HIP(hipSetDevice(h->dev));
    const int N = h->K.N;
    const long long n3 = (long long)N * N * N;
    float2* vol = (float2*)ft_dev;
    float2* tmp = nullptr;
    if (!vol) {
        BSX_HIP(hipMalloc((void**)&tmp, sizeof(float2) * n3));
        vol = tmp;
    }
    unsigned long long* dcov = nullptr;
    double* dsum = nullptr;
    BSX_HIP(hipMalloc((void**)&dcov, sizeof(unsigned long long) + 4 * sizeof(double)));
    dsum = (double*)(dcov + 1);
    BSX_HIP(hipMemsetAsync(dcov, 0, sizeof(unsigned long long) + 4 * sizeof(double), s));
    hipLaunchKernelGGL(k_weigh, dim3(2048), dim3(256), 0, s, h->acc, n3, N, h->K.stride, hmask,
                       h->p.nhalves, h->K.fom, vol, fom_dev, dcov, dsum);
    BSX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_friedel, dim3(1024), dim3(256), 0, s, vol, N, dsum + 1);
    BSX_HIP(hipGetLastError());
    unsigned long long hcov = 0;
    double hsum[4] = {0, 0, 0, 0};
    BSX_HIP(hipMemcpyAsync(&hcov, dcov, sizeof(hcov), hipMemcpyDeviceToHost, s));
    BSX_HIP(hipMemcpyAsync(hsum, dsum, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
    int r = BSX_OK;
    if (map_dev) {
        // inverse transform: work on a copy when the caller wants the transform back
        float2* work = vol;
        float2* wtmp = nullptr;
        if (ft_dev) {
            BSX_HIP(hipMalloc((void**)&wtmp, sizeof(float2) * n3));
            BSX_HIP(hipMemcpyAsync(wtmp, vol, sizeof(float2) * n3, hipMemcpyDeviceToDevice, s));
            work = wtmp;
        }
        const float sc = (float)(1.0 / sqrt((double)n3));
        r = fft_axis_ex(work, nullptr, N, N, N, 1, 0, +1, 1.0f, s);
        if (r == BSX_OK) r = fft_axis_ex(work, nullptr, N, N, N, 1, 1, +1, 1.0f, s);
        if (r == BSX_OK) r = fft_axis_ex(work, map_dev, N, N, N, 1, 2, +1, sc, s);
        BSX_HIP(hipStreamSynchronize(s));
        if (wtmp) BSX_HIP(hipFree(wtmp));
    }
    BSX_HIP(hipStreamSynchronize(s));
    if (tmp) BSX_HIP(hipFree(tmp));
    BSX_HIP(hipFree(dcov));
    if (r != BSX_OK) return r;
    if (stats) {
        stats->coverage = (int64_t)hcov;
        stats->weight_sum = hsum[0];
        double sw = hsum[1], rr = hsum[2], ri = hsum[3];
        double fr = 0;
        if (sw > 0) {
            rr = sqrt(rr / sw);
            ri = sqrt(ri / sw);
            fr = sqrt(rr * rr + ri * ri);
        }
        stats->friedel = fr;
        int64_t np = 0;
        for (int q = 0; q < h->p.nhalves; ++q)
            if ((hmask >> q) & 1) np += h->inserted[q];
        stats->particles = np;
    }
    return BSX_OK;
}

}  // extern "C"

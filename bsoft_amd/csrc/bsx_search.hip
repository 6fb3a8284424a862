// bsx_search.hip — orientation search by projection matching on gfx950.
//
// The reference scores a particle transform against central sections of the
// map transform with a ring-weighted Fourier ring correlation
// (img_recip_space_fom, src/mg/mg_refine.cpp:660-781, CTF applied to the
// section) over views from asymmetric_unit_views (src/util/symmetry.cpp:383-548)
// expanded in-plane (view_list_expand_angles, src/util/View.cpp:519-556);
// sections come from Bimage::central_section (src/img/Bimage_project.cpp:288-338).
//
// Here every comparison is restricted to a low-resolution disk of radius R
// Fourier pixels (rings 0..R-1), stored as a list of half-plane points (the
// transforms of real images are Hermitian, so the half plane carries every
// ring sum up to a factor 2 that cancels in each ring correlation):
//   k_part_stage   : particle FFT -> phase shift to its origin -> disk values,
//                    CTF values, and the BF16 row of the coarse score
//   k_lib_build<W> : library sections at the disk points by FSI-kernel
//                    interpolation (FP32 values) and their BF16 rows
//   k_gemm_topk4<T>: BF16 MFMA (v_mfma_f32_32x32x16_bf16) GEMM of particle rows
//                    against every library row with FP32 accumulation and a
//                    fused per-particle top-T selection (256 x 256 tiles)
//   k_topk_merge<T>: merge of the per-segment top-T lists
//   k_rescore      : FP32 FRC of the reference formula for candidate library
//                    views, best candidate per particle
//   k_score_mats<W>: FP32 FRC at arbitrary rotations with the section
//                    interpolated on the fly (local refinement)
//
// Coarse score (BF16 GEMM): with a = F_p*CTF and b = F_ref on ring j,
//   score = 1/(R-1) * sum_{j=1}^{R-1} sum_{i in j} Re(a_i conj b_i) / (|a|_j |b|_j),
// i.e. the mean ring correlation with the CTF moved from the section to the
// particle; it ranks candidates for the FP32 FRC, which decides.
#include <climits>
#include <cmath>
#include <cstring>
#include <algorithm>
#include <vector>

#include "bsx_common.h"
#include "bsx_fft_dev.h"

namespace bsx {
int fft_axis_ex(float2* data, float* real_out, int nx, int ny, int nz, int64_t batch, int axis,
                int sign, float scale, hipStream_t s);
}
using namespace bsx;

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

namespace {

constexpr int GM = 128;          // GEMM tile rows (particles)
constexpr int GN = 128;          // GEMM tile columns (library views)
constexpr int GK = 64;           // K step staged in LDS
constexpr int LROW = GK * 2 + 16;  // bytes per staged row (16-byte pad: conflict-free b128 reads)
constexpr int CROW = GN + 1;     // floats per epilogue row
constexpr int GEMM_LDS = 2 * (GM + GN) * LROW;   // 73,728 B, aliased by the 66,048-B epilogue tile
static_assert(GM * CROW * 4 <= GEMM_LDS, "epilogue tile must fit the staging buffers");

__device__ inline uint16_t to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// ---------------------------------------------------------------- staging
struct SPart {
    double phx, phy;     // 2 pi ori / N
    double fscale;       // 1 / (N * pixel_size)
    CtfCoef c;
    double ox, oy;       // particle origin used for the phase shift (pixels)
    double def_avg;      // defocus the CTF values were computed with
    double c20u;         // c20 per unit defocus (-pi lambda)
    int apply;           // CTF applied (flag and defocus >= 1, mg_refine.cpp:687-688)
    int flag;            // CTF flag set (and a voltage given)
};

// Disk values of one particle from its spectrum values sv[i] = F(h_i, k_i)
// (sign -1, x 1/N) in LDS: phase shift to the origin, CTF, then the BF16
// coarse row (F*CTF normalised per ring, x 1/(R-1)). sv is overwritten with
// F*CTF. Called by the whole workgroup.
__device__ __forceinline__ void stage_disk(float2* sv, float* rn, const SPart& P,
                                           const int2* __restrict__ hk,
                                           const int* __restrict__ ring_start, int R, int Nd,
                                           int Kp, int64_t p, float2* __restrict__ fp,
                                           float* __restrict__ ctf, uint16_t* __restrict__ arow) {
    for (int i = threadIdx.x; i < Nd; i += blockDim.x) {
        const int h = hk[i].x, k = hk[i].y;
        const float2 F = sv[i];
        double sn, cs;
        sincos((double)h * P.phx + (double)k * P.phy, &sn, &cs);
        const float vr = (float)(F.x * cs - F.y * sn), vi = (float)(F.x * sn + F.y * cs);
        float c = 1.f;
        if (P.apply) c = (float)sin(ctf_chi(P.c, h * P.fscale, k * P.fscale));
        fp[p * Nd + i] = make_float2(vr, vi);
        ctf[p * Nd + i] = c;
        sv[i] = make_float2(vr * c, vi * c);
    }
    __syncthreads();
    if (threadIdx.x < R) {
        const int j = threadIdx.x;
        float s = 0.f;
        for (int i = ring_start[j]; i < ring_start[j + 1]; ++i) s += sv[i].x * sv[i].x + sv[i].y * sv[i].y;
        rn[j] = (s > 0.f && j > 0) ? 1.f / (sqrtf(s) * (float)(R - 1)) : 0.f;
    }
    __syncthreads();
    uint16_t* A = arow + p * Kp;
    for (int q = threadIdx.x; q < Kp / 2; q += blockDim.x) {
        const int i = q + 1;
        uint32_t w = 0;
        if (i < Nd) {
            int j = 1;
            while (j < R && ring_start[j + 1] <= i) ++j;
            const float s = rn[j];
            w = (uint32_t)to_bf16(sv[i].x * s) | ((uint32_t)to_bf16(sv[i].y * s) << 16);
        }
        *(uint32_t*)(A + 2 * q) = w;
    }
}

// One workgroup per particle. spec: the particle's N x N forward transform
// (sign -1, x 1/N). Used where the fused kernel below does not fit.
__global__ __launch_bounds__(256) void k_part_stage(
    const float2* __restrict__ spec, int N, const SPart* __restrict__ sp,
    const int2* __restrict__ hk, const int* __restrict__ ring_start, int R, int Nd, int Kp,
    int64_t p0, float2* __restrict__ fp, float* __restrict__ ctf, uint16_t* __restrict__ arow) {
    extern __shared__ float2 sv[];          // Nd values
    __shared__ float rn[65];
    const int b = blockIdx.x;
    const float2* S = spec + (size_t)b * N * N;
    for (int i = threadIdx.x; i < Nd; i += blockDim.x) {
        const int h = hk[i].x, k = hk[i].y;
        const int xx = h < 0 ? h + N : h, yy = k < 0 ? k + N : k;
        sv[i] = S[(size_t)yy * N + xx];
    }
    __syncthreads();
    stage_disk(sv, rn, sp[b], hk, ring_start, R, Nd, Kp, p0 + b, fp, ctf, arow);
}

// Fused staging, one workgroup per particle: the image is read once and only
// the low frequencies the disk needs are transformed.
//  1. rows, two real rows per complex line (z = a + i b): FftLine<N> over the
//     lines, then X_a(h) = (Z(h) + conj Z(-h)) / 2 and
//     X_b(h) = (Z(h) - conj Z(-h)) / 2i for h = 0..R-1 into C[h][y];
//  2. columns: FftLine<N> over C[h] for h = 0..R-1, in place;
//  3. disk values F(h, k) = C[h][k] / N for h >= 0 and conj(C[-h][-k]) / N
//     for h < 0 (real image: F(-h, -k) = conj F(h, k)), then stage_disk.
// Dynamic LDS (float2): G lines x N | N twiddles | R x N columns | Nd values.
template <int N>
__global__ __launch_bounds__(256) void k_stage_fused(
    const float* __restrict__ imgs, const float2* __restrict__ twg, const SPart* __restrict__ sp,
    const int2* __restrict__ hk, const int* __restrict__ ring_start, int R, int Nd, int Kp,
    int64_t p0, float2* __restrict__ fp, float* __restrict__ ctf, uint16_t* __restrict__ arow) {
    constexpr int T = FftThreads<N>::value;
    constexpr int G = 256 / T;
    extern __shared__ float2 lds[];
    __shared__ float rn[65];
    float2* lines = lds;
    float2* tw = lines + G * N;
    float2* C = tw + N;
    float2* sv = C + (size_t)R * N;
    const int tid = threadIdx.x;
    const int g = tid / T, t = tid - g * T;
    const int b = blockIdx.x;
    const float* img = imgs + (size_t)b * N * N;
    for (int i = tid; i < N; i += blockDim.x) tw[i] = twg[i];
    float2* line = lines + g * N;
    // the rows of the next G lines are loaded into registers while the
    // current lines are transformed
    constexpr int PT = N / T;
    static_assert(N % T == 0, "whole values per thread");
    float2 pre[PT];
    auto load = [&](int l0) {
        const int l = l0 + g;
        const float* ra = img + (size_t)(2 * l) * N;
#pragma unroll
        for (int q = 0; q < PT; ++q) {
            const int x = t + q * T;
            pre[q] = (g < G && l < N / 2) ? make_float2(ra[x], ra[x + N]) : make_float2(0.f, 0.f);
        }
    };
    load(0);
    for (int l0 = 0; l0 < N / 2; l0 += G) {
        const int l = l0 + g;
        const bool ok = g < G && l < N / 2;
        if (ok) {
#pragma unroll
            for (int q = 0; q < PT; ++q) line[t + q * T] = pre[q];
        }
        __syncthreads();
        if (l0 + G < N / 2) load(l0 + G);
        FftLine<N, -1, T>::run(line, tw, t);
        if (ok) {
            for (int h = t; h < R; h += T) {
                const float2 zp = line[h], zm = line[(N - h) % N];
                C[(size_t)h * N + 2 * l] = make_float2(0.5f * (zp.x + zm.x), 0.5f * (zp.y - zm.y));
                C[(size_t)h * N + 2 * l + 1] = make_float2(0.5f * (zp.y + zm.y), 0.5f * (zm.x - zp.x));
            }
        }
        __syncthreads();
    }
    for (int c0 = 0; c0 < R; c0 += G) {
        const int h = c0 + g;
        float2* col = (g < G && h < R) ? C + (size_t)h * N : line;
        FftLine<N, -1, T>::run(col, tw, t);
    }
    const float inv = 1.0f / N;
    for (int i = tid; i < Nd; i += blockDim.x) {
        const int h = hk[i].x, k = hk[i].y;
        float2 F;
        if (h >= 0) {
            F = C[(size_t)h * N + k % N];
        } else {
            F = C[(size_t)(-h) * N + (N - k) % N];
            F.y = -F.y;
        }
        sv[i] = make_float2(F.x * inv, F.y * inv);
    }
    __syncthreads();
    stage_disk(sv, rn, sp[b], hk, ring_start, R, Nd, Kp, p0 + b, fp, ctf, arow);
}

// ---------------------------------------------------------------- library
// One workgroup per view. Section value at each disk point by FSI
// interpolation, with the same coordinate arithmetic as k_extract.
template <int WD>
__global__ __launch_bounds__(256) void k_lib_build(
    const float2* __restrict__ ft, int N, const double* __restrict__ mats, double inv, double rs,
    const float* __restrict__ ktab, const int2* __restrict__ hk, const int* __restrict__ ring_start,
    int R, int Nd, int Kp, int64_t v0, float2* __restrict__ lib, uint16_t* __restrict__ brow) {
    __shared__ float K[(FSI_NDIV + 1) * WD];
    __shared__ float rn[65];
    extern __shared__ float2 sv[];
    for (int i = threadIdx.x; i < (FSI_NDIV + 1) * WD; i += blockDim.x) K[i] = ktab[i];
    __syncthreads();
    const int64_t v = v0 + blockIdx.x;
    const double* M = mats + 9 * (size_t)blockIdx.x;
    for (int i = threadIdx.x; i < Nd; i += blockDim.x) {
        const double ivx = (double)hk[i].x * inv, ivy = (double)hk[i].y * inv;
        double m[3];
        for (int c = 0; c < 3; ++c) m[c] = (ivx * M[3 * c] + ivy * M[3 * c + 1] + 0.0 * M[3 * c + 2]) * rs;
        double vr, vi;
        fsi_sample<WD>(ft, N, m, K, vr, vi);
        const float2 val = make_float2((float)vr, (float)vi);
        lib[v * Nd + i] = val;
        sv[i] = val;
    }
    __syncthreads();
    if (threadIdx.x < R) {
        const int j = threadIdx.x;
        float s = 0.f;
        for (int i = ring_start[j]; i < ring_start[j + 1]; ++i) s += sv[i].x * sv[i].x + sv[i].y * sv[i].y;
        rn[j] = (s > 0.f && j > 0) ? 1.f / sqrtf(s) : 0.f;
    }
    __syncthreads();
    uint16_t* B = brow + v * Kp;
    for (int q = threadIdx.x; q < Kp / 2; q += blockDim.x) {
        const int i = q + 1;
        uint32_t w = 0;
        if (i < Nd) {
            int j = 1;
            while (j < R && ring_start[j + 1] <= i) ++j;
            const float s = rn[j];
            w = (uint32_t)to_bf16(sv[i].x * s) | ((uint32_t)to_bf16(sv[i].y * s) << 16);
        }
        *(uint32_t*)(B + 2 * q) = w;
    }
}

// ------------------------------------------------------------ GEMM + top-k
// Workgroup (512 threads, 8 waves as 4 x 2) = one 256-row block of particles
// x one segment of 256-column library tiles; each wave computes a 64 x 128
// block as 2 x 4 v_mfma_f32_32x32x16_bf16 blocks. At this tile size the
// operands cost 1 byte per 128 FLOP from L2 (128 x 128 tiles: 1 per 64, as
// much as the L2 -> CU path delivers at the MFMA rate). Operands go from
// global memory straight to LDS (global_load_lds_dwordx4) into unpadded,
// XOR-swizzled tiles: chunk c (16 B) of row r sits in slot c ^ (r & 7) of its
// 128-B row, so a wave's load instruction (1 KB, 8 rows in lane order) takes
// each lane's chunk from the slot it fills, and the b128 fragment reads of 8
// consecutive rows hit 8 distinct slots (all 32 banks); two stages, the
// loads of step k + 1 in flight while step k multiplies. The epilogue goes
// through LDS in two 128-row halves (rows of 260 floats, so b128 row reads
// stay conflict-free); thread t folds row t % 128 of the half, columns
// 64 (t / 128) .. + 63, into its running top-T list, 4 columns at a time
// against the list's T-th value. The top-T lists come out per (row,
// segment, column quarter).
constexpr int GM4 = 256, GN4 = 256;
constexpr int GSTAGE4 = (GM4 + GN4) * 128;        // bytes per pipeline stage
constexpr int CROW4 = GN4 + 4;
constexpr int GEMM4_LDS = (2 * GSTAGE4 > 128 * CROW4 * 4) ? 2 * GSTAGE4 : 128 * CROW4 * 4;

template <int T>
__global__ __launch_bounds__(512) void k_gemm_topk4(const uint16_t* __restrict__ A,
                                                    const uint16_t* __restrict__ B, int Kp,
                                                    int ncols, int tiles_per_seg, int ntiles,
                                                    int rblocks, int nseg,
                                                    float* __restrict__ pscore,
                                                    int* __restrict__ pidx) {
    __shared__ __attribute__((aligned(16))) char lds[GEMM4_LDS];
    typedef __attribute__((address_space(3))) void* lds_ptr;
    // XCD-aware order: hardware block b runs on XCD b % 8; logical block
    // L = (b % 8) * per + b / 8 gives each XCD a contiguous range, row blocks
    // fastest, so the workgroups resident on one XCD share library tiles
    // through its L2.
    const int nblk = rblocks * nseg;
    const int per = (nblk + 7) / 8;
    const int L = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (L >= nblk) return;
    const size_t row0 = (size_t)(L % rblocks) * GM4;
    const int seg = L / rblocks;
    const int t0 = seg * tiles_per_seg;
    const int t1 = min(ntiles, t0 + tiles_per_seg);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int my_row = tid & 127, my_q = tid >> 7;   // epilogue: row of a half, column quarter
    float best[2][T];
    int bidx[2][T];
#pragma unroll
    for (int hh = 0; hh < 2; ++hh)
#pragma unroll
        for (int q = 0; q < T; ++q) { best[hh][q] = -INFINITY; bidx[hh][q] = -1; }
    const int nk = Kp / GK;
    // this wave's load instructions: A rows 32 w .. + 31, B rows 32 w .. + 31
    const uint16_t* asrc[4];
    const uint16_t* bsrc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r = 8 * (wave * 4 + q) + (lane >> 3);
        asrc[q] = A + (row0 + r) * Kp + ((lane & 7) ^ (r & 7)) * 8;
        bsrc[q] = B + (size_t)r * Kp + ((lane & 7) ^ (r & 7)) * 8;
    }
    for (int t = t0; t < t1; ++t) {
        const size_t cofs = (size_t)t * GN4 * Kp;
        f32x16 acc[2][4];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;
        auto stage = [&](int buf, size_t k0) {
            char* base = lds + buf * GSTAGE4;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                __builtin_amdgcn_global_load_lds(asrc[q] + k0, (lds_ptr)(base + (wave * 4 + q) * 1024),
                                                 16, 0, 0);
                __builtin_amdgcn_global_load_lds(bsrc[q] + cofs + k0,
                                                 (lds_ptr)(base + GM4 * 128 + (wave * 4 + q) * 1024),
                                                 16, 0, 0);
            }
        };
        stage(0, 0);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        for (int kt = 0; kt < nk; ++kt) {
            // buffer (kt + 1) & 1 was read in step kt - 1, which every wave
            // has left (barrier at its end)
            if (kt + 1 < nk) stage((kt + 1) & 1, (size_t)(kt + 1) * GK);
            const char* abuf = lds + (kt & 1) * GSTAGE4;
            const char* bbuf = abuf + GM4 * 128;
#pragma unroll
            for (int ks = 0; ks < GK / 16; ++ks) {
                const int c = ks * 2 + (lane >> 5);
                bf16x8 fa[2], fb[4];
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    const int ra = wm * 64 + i * 32 + (lane & 31);
                    fa[i] = *(const bf16x8*)(abuf + ra * 128 + ((c ^ (ra & 7)) << 4));
                }
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int rb = wn * 128 + j * 32 + (lane & 31);
                    fb[j] = *(const bf16x8*)(bbuf + rb * 128 + ((c ^ (rb & 7)) << 4));
                }
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
            }
            __builtin_amdgcn_s_waitcnt(0);   // step kt + 1 has landed (and this step's reads)
            __syncthreads();
        }
        // epilogue, one 128-row half at a time
        float* C = (float*)lds;
#pragma unroll
        for (int hh = 0; hh < 2; ++hh) {
            if ((wm >> 1) == hh) {
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
#pragma unroll
                        for (int r = 0; r < 16; ++r) {
                            const int row = (wm & 1) * 64 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                            const int col = wn * 128 + j * 32 + (lane & 31);
                            C[row * CROW4 + col] = acc[i][j][r];
                        }
            }
            __syncthreads();
            const float* cr = C + my_row * CROW4 + my_q * 64;
            const int cbase = t * GN4 + my_q * 64;
            const int jmax = min(64, ncols - cbase);
            for (int j = 0; j < jmax; j += 4) {
                const float4 v4 = *(const float4*)(cr + j);
                if (fmaxf(fmaxf(v4.x, v4.y), fmaxf(v4.z, v4.w)) > best[hh][T - 1]) {
                    const float vv[4] = {v4.x, v4.y, v4.z, v4.w};
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const float v = vv[e];
                        if (j + e < jmax && v > best[hh][T - 1]) {
                            best[hh][T - 1] = v;
                            bidx[hh][T - 1] = cbase + j + e;
#pragma unroll
                            for (int q = T - 1; q > 0; --q) {
                                if (best[hh][q] > best[hh][q - 1]) {
                                    const float tv = best[hh][q]; best[hh][q] = best[hh][q - 1]; best[hh][q - 1] = tv;
                                    const int ti = bidx[hh][q]; bidx[hh][q] = bidx[hh][q - 1]; bidx[hh][q - 1] = ti;
                                }
                            }
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
        const size_t o = (((row0 + hh * 128 + my_row) * nseg + seg) * 4 + my_q) * T;
#pragma unroll
        for (int q = 0; q < T; ++q) {
            pscore[o + q] = best[hh][q];
            pidx[o + q] = bidx[hh][q];
        }
    }
}

// ------------------------------------------------- proposal-restricted GEMM
// Routing of (particle, proposal slot) pairs to their proposed class, so
// that one class's member views are scored against all particles that
// proposed it with one MFMA tile (a grouped GEMM).
__global__ void k_route_count(const int* __restrict__ cls, int64_t n, int k, int nclass,
                              int* __restrict__ counts) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * k) return;
    const int c = cls[t];
    if (c >= 0 && c < nclass) atomicAdd(counts + c, 1);
}

__global__ void k_route_fill(const int* __restrict__ cls, int64_t n, int k, int nclass,
                             int* __restrict__ cursors, int* __restrict__ pairs) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * k) return;
    const int c = cls[t];
    if (c >= 0 && c < nclass) pairs[atomicAdd(cursors + c, 1)] = (int)t;   // t = particle * k + slot
}

// Unrouted slots (no valid class) still need empty lists for the merge.
__global__ void k_route_empty(const int* __restrict__ cls, int64_t n, int k, int nclass, int T,
                              float* __restrict__ pscore, int* __restrict__ pidx) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * k) return;
    const int c = cls[t];
    if (c >= 0 && c < nclass) return;
    for (int q = 0; q < T; ++q) { pscore[t * T + q] = -INFINITY; pidx[t * T + q] = -1; }
}

struct GroupItem {
    int cls, p0, np, pad_;
};

// One workgroup per work item: up to 128 routed pairs of one class against
// the class's member views (128 per column tile). Rows are gathered by
// index (zero row for padding); the per-pair top-T over the members goes to
// pscore/pidx[(particle * k + slot) * T + q].
template <int T>
__global__ __launch_bounds__(256) void k_gemm_grouped(
    const uint16_t* __restrict__ A, const uint16_t* __restrict__ B, const uint16_t* __restrict__ zrow,
    int Kp, const GroupItem* __restrict__ items, const int* __restrict__ pairs, int k,
    const int* __restrict__ members, const int* __restrict__ mcount, int maxm,
    float* __restrict__ pscore, int* __restrict__ pidx) {
    __shared__ __attribute__((aligned(16))) char lds[GEMM_LDS];
    __shared__ int srow[GM];
    __shared__ int scol[GN];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const GroupItem it = items[blockIdx.x];
    const int my_row = tid & (GM - 1), my_half = tid >> 7;
    const int ncls_members = mcount[it.cls];
    if (tid < GM) srow[tid] = tid < it.np ? pairs[it.p0 + tid] : -1;
    float best[T];
    int bidx[T];
#pragma unroll
    for (int q = 0; q < T; ++q) { best[q] = -INFINITY; bidx[q] = -1; }
    const int nk = Kp / GK;
    int ld_row[4], ld_part[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int c = tid + 256 * q;
        ld_row[q] = c >> 3;
        ld_part[q] = c & 7;
    }
    const int ntile = (ncls_members + GN - 1) / GN;
    for (int t = 0; t < ntile; ++t) {
        __syncthreads();
        if (tid < GN) {
            const int j = t * GN + tid;
            scol[tid] = j < ncls_members ? members[(size_t)it.cls * maxm + j] : -1;
        }
        __syncthreads();
        const uint16_t* pa[4];
        const uint16_t* pb[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int pr = srow[ld_row[q]];
            pa[q] = pr >= 0 ? A + (size_t)(pr / k) * Kp : zrow;
            const int vc = scol[ld_row[q]];
            pb[q] = vc >= 0 ? B + (size_t)vc * Kp : zrow;
        }
        f32x16 acc[2][2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;
        uint4 ra0, ra1, ra2, ra3, rb0, rb1, rb2, rb3;
#define BSX_GGLOAD(K0)                                               \
    ra0 = *(const uint4*)(pa[0] + (K0) + ld_part[0] * 8);            \
    ra1 = *(const uint4*)(pa[1] + (K0) + ld_part[1] * 8);            \
    ra2 = *(const uint4*)(pa[2] + (K0) + ld_part[2] * 8);            \
    ra3 = *(const uint4*)(pa[3] + (K0) + ld_part[3] * 8);            \
    rb0 = *(const uint4*)(pb[0] + (K0) + ld_part[0] * 8);            \
    rb1 = *(const uint4*)(pb[1] + (K0) + ld_part[1] * 8);            \
    rb2 = *(const uint4*)(pb[2] + (K0) + ld_part[2] * 8);            \
    rb3 = *(const uint4*)(pb[3] + (K0) + ld_part[3] * 8);
#define BSX_GLSTORE(BUF)                                                            \
    *(uint4*)((BUF) + ld_row[0] * LROW + ld_part[0] * 16) = ra0;                    \
    *(uint4*)((BUF) + ld_row[1] * LROW + ld_part[1] * 16) = ra1;                    \
    *(uint4*)((BUF) + ld_row[2] * LROW + ld_part[2] * 16) = ra2;                    \
    *(uint4*)((BUF) + ld_row[3] * LROW + ld_part[3] * 16) = ra3;                    \
    *(uint4*)((BUF) + (GM + ld_row[0]) * LROW + ld_part[0] * 16) = rb0;             \
    *(uint4*)((BUF) + (GM + ld_row[1]) * LROW + ld_part[1] * 16) = rb1;             \
    *(uint4*)((BUF) + (GM + ld_row[2]) * LROW + ld_part[2] * 16) = rb2;             \
    *(uint4*)((BUF) + (GM + ld_row[3]) * LROW + ld_part[3] * 16) = rb3;
        BSX_GGLOAD(0)
        BSX_GLSTORE(lds)
        __syncthreads();
        for (int kt = 0; kt < nk; ++kt) {
            const bool more = kt + 1 < nk;
            if (more) {
                const size_t k0 = (size_t)(kt + 1) * GK;
                BSX_GGLOAD(k0)
            }
            const char* buf = lds + (kt & 1) * (GM + GN) * LROW;
#pragma unroll
            for (int ks = 0; ks < GK / 16; ++ks) {
                bf16x8 fa[2], fb[2];
                const int off = ks * 32 + (lane >> 5) * 16;
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    fa[i] = *(const bf16x8*)(buf + (wm * 64 + i * 32 + (lane & 31)) * LROW + off);
                    fb[i] = *(const bf16x8*)(buf + (GM + wn * 64 + i * 32 + (lane & 31)) * LROW + off);
                }
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
            }
            if (more) {
                char* nb = lds + ((kt + 1) & 1) * (GM + GN) * LROW;
                BSX_GLSTORE(nb)
            }
            __syncthreads();
        }
#undef BSX_GGLOAD
#undef BSX_GLSTORE
        float* C = (float*)lds;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int row = wm * 64 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                    const int col = wn * 64 + j * 32 + (lane & 31);
                    C[row * CROW + col] = acc[i][j][r];
                }
        __syncthreads();
        const float* cr = C + my_row * CROW + my_half * 64;
        for (int j = 0; j < 64; ++j) {
            const int vc = scol[my_half * 64 + j];
            if (vc < 0) continue;
            const float v = cr[j];
            if (v > best[T - 1]) {
                best[T - 1] = v;
                bidx[T - 1] = vc;
#pragma unroll
                for (int q = T - 1; q > 0; --q) {
                    if (best[q] > best[q - 1]) {
                        const float tv = best[q]; best[q] = best[q - 1]; best[q - 1] = tv;
                        const int ti = bidx[q]; bidx[q] = bidx[q - 1]; bidx[q - 1] = ti;
                    }
                }
            }
        }
    }
    // combine the two column halves of each row into the row's list
    __syncthreads();
    float* xs = (float*)lds;
    int* xi = (int*)(lds + GM * 2 * T * sizeof(float));
#pragma unroll
    for (int q = 0; q < T; ++q) {
        xs[(my_row * 2 + my_half) * T + q] = best[q];
        xi[(my_row * 2 + my_half) * T + q] = bidx[q];
    }
    __syncthreads();
    if (my_half == 0 && my_row < it.np) {
#pragma unroll
        for (int q = 0; q < T; ++q) {
            const float v = xs[(my_row * 2 + 1) * T + q];
            const int id = xi[(my_row * 2 + 1) * T + q];
            if (id >= 0 && v > best[T - 1]) {
                best[T - 1] = v;
                bidx[T - 1] = id;
#pragma unroll
                for (int z = T - 1; z > 0; --z) {
                    if (best[z] > best[z - 1]) {
                        const float tv = best[z]; best[z] = best[z - 1]; best[z - 1] = tv;
                        const int ti = bidx[z]; bidx[z] = bidx[z - 1]; bidx[z - 1] = ti;
                    }
                }
            }
        }
        const size_t o = (size_t)srow[my_row] * T;
#pragma unroll
        for (int q = 0; q < T; ++q) {
            pscore[o + q] = best[q];
            pidx[o + q] = bidx[q];
        }
    }
}

// One thread per particle: merge the 2*nseg sorted lists into the top T.
template <int T>
__global__ void k_topk_merge(const float* __restrict__ pscore, const int* __restrict__ pidx,
                             int nlist, int64_t n, int topk, int* __restrict__ cand,
                             float* __restrict__ cscore) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float best[T];
    int bidx[T];
#pragma unroll
    for (int q = 0; q < T; ++q) { best[q] = -INFINITY; bidx[q] = -1; }
    const float* S = pscore + (size_t)p * nlist * T;
    const int* I = pidx + (size_t)p * nlist * T;
    for (int e = 0; e < nlist * T; ++e) {
        const float v = S[e];
        const int id = I[e];
        if (id >= 0 && (v > best[T - 1] || (v == best[T - 1] && bidx[T - 1] >= 0 && id < bidx[T - 1]) ||
                        bidx[T - 1] < 0)) {
            best[T - 1] = v;
            bidx[T - 1] = id;
#pragma unroll
            for (int q = T - 1; q > 0; --q) {
                const bool sw = bidx[q - 1] < 0 || best[q] > best[q - 1] ||
                                (best[q] == best[q - 1] && bidx[q] < bidx[q - 1]);
                if (sw) {
                    const float tv = best[q]; best[q] = best[q - 1]; best[q - 1] = tv;
                    const int ti = bidx[q]; bidx[q] = bidx[q - 1]; bidx[q - 1] = ti;
                }
            }
        }
    }
    for (int q = 0; q < topk; ++q) {
        cand[p * topk + q] = bidx[q];
        cscore[p * topk + q] = best[q];
    }
}

// ------------------------------------------------------------ FP32 FRC
// Ring correlation of img_recip_space_fom (fom_type 0, unit ring weights,
// CTF on the section) over rings 0..R-1, one wavefront per comparison; ring
// sums reduced across the wavefront in a fixed order.
__device__ inline float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline float frc_wave(const float2* __restrict__ P, const float* __restrict__ Cf,
                                 const float2* __restrict__ S, const int* __restrict__ ring_start,
                                 int R, int lane) {
    float f = 0.f;
    for (int j = 0; j < R; ++j) {
        float num = 0.f, dp = 0.f, dr = 0.f;
        for (int i = ring_start[j] + lane; i < ring_start[j + 1]; i += 64) {
            const float2 a = P[i];
            const float c = Cf[i];
            const float2 r = make_float2(S[i].x * c, S[i].y * c);
            num += a.x * r.x + a.y * r.y;
            dp += a.x * a.x + a.y * a.y;
            dr += r.x * r.x + r.y * r.y;
        }
        num = wave_sum(num);
        dp = wave_sum(dp);
        dr = wave_sum(dr);
        const float den = dp * dr;
        if (den > 0.f) f += num / sqrtf(den);
    }
    return f / (float)R;
}

// One workgroup per particle; wave w scores candidates w, w+4, ...
// cand: n x ncand library indices (negative = unused slot).
__global__ __launch_bounds__(256) void k_rescore(const float2* __restrict__ fp,
                                                 const float* __restrict__ ctf,
                                                 const float2* __restrict__ lib, int64_t nviews,
                                                 const int* __restrict__ ring_start, int R, int Nd,
                                                 const int* __restrict__ cand, int ncand,
                                                 float* __restrict__ frc, int* __restrict__ best,
                                                 float* __restrict__ best_frc) {
    __shared__ float bs[4];
    __shared__ int bi[4];
    const int64_t p = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const float2* P = fp + p * Nd;
    const float* Cf = ctf + p * Nd;
    float wb = -INFINITY;
    int wi = -1;
    for (int c = wave; c < ncand; c += 4) {
        const int v = cand[p * ncand + c];
        float f = -INFINITY;
        if (v >= 0 && v < nviews) {
            f = frc_wave(P, Cf, lib + (size_t)v * Nd, ring_start, R, lane);
            if (f > wb) { wb = f; wi = c; }
        }
        if (frc && lane == 0) frc[p * ncand + c] = f;
    }
    if (lane == 0) { bs[wave] = wb; bi[wave] = wi; }
    __syncthreads();
    if (threadIdx.x == 0) {
        float b = -INFINITY;
        int ci = -1;
        for (int w = 0; w < 4; ++w)
            if (bi[w] >= 0 && (bs[w] > b || (bs[w] == b && bi[w] < ci))) { b = bs[w]; ci = bi[w]; }
        if (best) best[p] = ci >= 0 ? cand[p * ncand + ci] : -1;
        if (best_frc) best_frc[p] = b;
    }
}

// One workgroup per particle: FRC (the k_rescore formula) of the particle
// against central sections at ncand arbitrary rotations, each section value
// interpolated from the map transform with the FSI kernel on the fly.
// Lanes are (disk point, candidate) pairs: for ncand <= 32 a wavefront takes
// two points with one candidate per lane, so the lanes of one point gather
// from the same small neighbourhood of the map and a gather instruction
// touches few cache lines. Each lane keeps its candidate's ring sums in
// registers and folds them into LDS when the ring changes; candidates beyond
// 32 (64) run in further blocks.
__device__ __forceinline__ void lds_fadd(float* p, float x) {
    unsigned* u = reinterpret_cast<unsigned*>(p);
    unsigned old = *reinterpret_cast<volatile unsigned*>(u);
    while (true) {
        const unsigned assumed = old;
        old = atomicCAS(u, assumed, __float_as_uint(__uint_as_float(assumed) + x));
        if (old == assumed) break;
    }
}

// Copy of the map transform around the origin: cube[(c S + b) S + a] =
// ft[wrap(c - C)][wrap(b - C)][wrap(a - C)], C = S / 2, with the periodic
// wrap of fsi_sample, so a sample whose taps all lie within |i| <= C - 1 reads
// the same voxels from the cube without wrapping.
__global__ void k_center_cube(const float2* __restrict__ ft, int N, int S,
                              float2* __restrict__ cube) {
    const int64_t n = (int64_t)S * S * S;
    const int C = S / 2;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int a = (int)(i % S), b = (int)((i / S) % S), c = (int)(i / ((int64_t)S * S));
        const int x = wrap_index(a - C, N), y = wrap_index(b - C, N), z = wrap_index(c - C, N);
        cube[i] = ft[((int64_t)z * N + y) * N + x];
    }
}

// fsi_sample on the centred cube: the same taps, weights and summation order,
// with indices offset by C instead of wrapped. Returns false (nothing read)
// when a tap would leave the cube, e.g. for a rotation matrix that is not
// orthonormal; the caller then samples the full map.
template <int WD>
__device__ inline bool fsi_sample_cube(const float2* __restrict__ cube, int S, const double m[3],
                                       const float* K, double& vr, double& vi) {
    constexpr int HW = (WD - 1) / 2;
    const int C = S / 2;
    const double fx = floor(m[0]), fy = floor(m[1]), fz = floor(m[2]);
    const int ikx = ((int)((m[0] - fx) * FSI_NDIV + 0.5)) * WD;
    const int iky = ((int)((m[1] - fy) * FSI_NDIV + 0.5)) * WD;
    const int ikz = ((int)((m[2] - fz) * FSI_NDIV + 0.5)) * WD;
    const double lo = (double)(HW - C), hi = (double)(S - WD + HW - C);
    if (!(fx >= lo && fx <= hi && fy >= lo && fy <= hi && fz >= lo && fz <= hi)) return false;
    const int ix = (int)fx - HW + C, iy = (int)fy - HW + C, iz = (int)fz - HW + C;
    float kx[WD], ky[WD];
#pragma unroll
    for (int q = 0; q < WD; ++q) {
        kx[q] = K[ikx + q];
        ky[q] = K[iky + q];
    }
    vr = 0;
    vi = 0;
#pragma unroll 1
    for (int zz = 0; zz < WD; ++zz) {
        const float kzz = K[ikz + zz];
        float pr = 0.f, pi = 0.f;
#pragma unroll
        for (int yy2 = 0; yy2 < WD; ++yy2) {
            const float w = ky[yy2] * kzz;
            const float2* row = cube + ((int64_t)(iz + zz) * S + (iy + yy2)) * S + ix;
#pragma unroll
            for (int q = 0; q < WD; ++q) {
                const float kw = kx[q] * w;
                const float2 a = row[q];
                pr = fmaf(a.x, kw, pr);
                pi = fmaf(a.y, kw, pi);
            }
        }
        vr += pr;
        vi += pi;
    }
    return true;
}

template <int WD>
__global__ __launch_bounds__(256) void k_score_mats(
    const float2* __restrict__ ft, int N, const float2* __restrict__ cube, int S,
    const double* __restrict__ mats, int ncand, double inv,
    double rs, const float* __restrict__ ktab, const int2* __restrict__ hk,
    const int* __restrict__ ring_start, int R, int Nd, const float2* __restrict__ fp,
    const float* __restrict__ ctf, float* __restrict__ frc) {
    __shared__ float K[(FSI_NDIV + 1) * WD];
    __shared__ int rs_l[65];
    __shared__ float dp[64];
    __shared__ float anum[64 * 64], adr[64 * 64];   // [candidate in block][ring]
    for (int i = threadIdx.x; i < (FSI_NDIV + 1) * WD; i += blockDim.x) K[i] = ktab[i];
    for (int i = threadIdx.x; i <= R; i += blockDim.x) rs_l[i] = ring_start[i];
    const int64_t p = blockIdx.x;
    const float2* P = fp + p * Nd;
    const float* Cf = ctf + p * Nd;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // particle power per ring
    for (int j = wave; j < R; j += 4) {
        float s = 0.f;
        for (int i = ring_start[j] + lane; i < ring_start[j + 1]; i += 64) {
            const float2 a = P[i];
            s += a.x * a.x + a.y * a.y;
        }
        s = wave_sum(s);
        if (lane == 0) dp[j] = s;
    }
    const int cs = ncand <= 32 ? 32 : 64;   // lanes per point
    const int ppw = 64 / cs;                // points per wavefront step
    const int slot = lane / cs, cl = lane - slot * cs;
    for (int cb = 0; cb < ncand; cb += cs) {
        for (int i = threadIdx.x; i < cs * R; i += blockDim.x) {
            anum[i] = 0.f;
            adr[i] = 0.f;
        }
        __syncthreads();
        const int c = cb + cl;
        const bool act = c < ncand;
        double m0[3] = {0, 0, 0}, m1[3] = {0, 0, 0};
        if (act) {
            const double* M = mats + 9 * (p * ncand + c);
            for (int q = 0; q < 3; ++q) { m0[q] = M[3 * q]; m1[q] = M[3 * q + 1]; }
        }
        int cur = -1, j = 0;
        float num = 0.f, dr = 0.f;
        for (int i = wave * ppw + slot; i < Nd; i += 4 * ppw) {
            while (j < R - 1 && rs_l[j + 1] <= i) ++j;
            if (j != cur) {
                if (act && cur >= 0) {
                    lds_fadd(&anum[cl * R + cur], num);
                    lds_fadd(&adr[cl * R + cur], dr);
                }
                cur = j;
                num = 0.f;
                dr = 0.f;
            }
            if (act) {
                const double ivx = (double)hk[i].x * inv, ivy = (double)hk[i].y * inv;
                double m[3];
                for (int q = 0; q < 3; ++q) m[q] = (ivx * m0[q] + ivy * m1[q]) * rs;
                double vr, vi;
                if (!cube || !fsi_sample_cube<WD>(cube, S, m, K, vr, vi))
                    fsi_sample<WD>(ft, N, m, K, vr, vi);
                const float2 a = P[i];
                const float cf = Cf[i];
                const float rr = (float)vr * cf, ri = (float)vi * cf;
                num += a.x * rr + a.y * ri;
                dr += rr * rr + ri * ri;
            }
        }
        if (act && cur >= 0) {
            lds_fadd(&anum[cl * R + cur], num);
            lds_fadd(&adr[cl * R + cur], dr);
        }
        __syncthreads();
        for (int t = threadIdx.x; t < cs && cb + t < ncand; t += blockDim.x) {
            float f = 0.f;
            for (int jj = 0; jj < R; ++jj) {
                const float den = dp[jj] * adr[t * R + jj];
                if (den > 0.f) f += anum[t * R + jj] / sqrtf(den);
            }
            frc[p * ncand + cb + t] = f / (float)R;
        }
        __syncthreads();
    }
}

__device__ inline double wave_sum_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// img_recip_space_fom (fom_type 0; src/mg/mg_refine.cpp:660-781) of the
// particle moved to origin (ox + dx, oy + dy), i.e. its staged values times
// exp(2 pi i (h dx + k dy) / N) (ux = 2 pi dx / N), against the section S.
// With redef the CTF is evaluated at defocus def and applied only when the
// CTF flag is set and def >= 1 (mg_refine.cpp:687-688); otherwise the staged
// CTF is used. Products in FP32 as the reference's Complex<float> data, ring
// sums in FP64. Called by a whole wavefront.
__device__ inline double frc_fit(const float2* __restrict__ P, const float* __restrict__ Cf,
                                 const float2* __restrict__ S, const int2* __restrict__ hk,
                                 const int* __restrict__ ring_start, int R, int lane, double ux,
                                 double uy, const SPart& sp, bool redef, double def) {
    CtfCoef cc = sp.c;
    cc.c20 = sp.c20u * def;
    const bool ctf_on = sp.flag && def >= 1;
    double f = 0;
    for (int j = 0; j < R; ++j) {
        double num = 0, dp = 0, dr = 0;
        for (int i = ring_start[j] + lane; i < ring_start[j + 1]; i += 64) {
            const int h = hk[i].x, k = hk[i].y;
            double sn, cs;
            sincos((double)h * ux + (double)k * uy, &sn, &cs);
            const float2 a0 = P[i];
            const float ar = (float)(a0.x * cs - a0.y * sn), ai = (float)(a0.x * sn + a0.y * cs);
            float c = Cf[i];
            if (redef) c = ctf_on ? (float)sin(ctf_chi(cc, h * sp.fscale, k * sp.fscale)) : 1.f;
            const float rr = S[i].x * c, ri = S[i].y * c;
            num += (double)(ar * rr) + (double)(ai * ri);
            dp += (double)(ar * ar) + (double)(ai * ai);
            dr += (double)(rr * rr) + (double)(ri * ri);
        }
        num = wave_sum_d(num);
        dp = wave_sum_d(dp);
        dr = wave_sum_d(dr);
        const double den = dp * dr;
        if (den > 0) f += num / sqrt(den);
    }
    return f / R;
}

// img_fit_in_recip_space_grid (src/mg/mg_refine.cpp:862-946), one workgroup
// per particle at its current orientation. The central section at the disk
// points is interpolated once into LDS. A defocus scan from def - 5 def_std
// to def + 5 def_std in steps of def_std runs when def_std > 0 and either the
// CTF flag is set or the defocus is nonzero (the reference's apply_ctf); then
// 3 x 3 origin grids around the starting origin (the reference's floating
// loops, which never move the centre), the step halved when the best is the
// centre or the FOM stops changing, until the step falls below accuracy.
// The result is the best of the last grid: origin (pixels), defocus, FOM.
template <int WD>
__global__ __launch_bounds__(256) void k_fit_grid(
    const float2* __restrict__ ft, int N, const double* __restrict__ mats, double inv, double rs,
    const float* __restrict__ ktab, const int2* __restrict__ hk, const int* __restrict__ ring_start,
    int R, int Nd, const float2* __restrict__ fp, const float* __restrict__ ctf,
    const SPart* __restrict__ sps, double step0, double accuracy, double def_std,
    double* __restrict__ origin, double* __restrict__ defocus, float* __restrict__ fom) {
    __shared__ float K[(FSI_NDIV + 1) * WD];
    __shared__ double cf[16];
    extern __shared__ float2 sec[];   // Nd section values
    for (int i = threadIdx.x; i < (FSI_NDIV + 1) * WD; i += blockDim.x) K[i] = ktab[i];
    __syncthreads();
    const int64_t p = blockIdx.x;
    const SPart sp = sps[p];
    const double* M = mats + 9 * p;
    for (int i = threadIdx.x; i < Nd; i += blockDim.x) {
        const double ivx = (double)hk[i].x * inv, ivy = (double)hk[i].y * inv;
        double m[3];
        for (int q = 0; q < 3; ++q) m[q] = (ivx * M[3 * q] + ivy * M[3 * q + 1]) * rs;
        double vr, vi;
        fsi_sample<WD>(ft, N, m, K, vr, vi);
        sec[i] = make_float2((float)vr, (float)vi);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const float2* P = fp + p * Nd;
    const float* Cf = ctf + p * Nd;
    const double tpn = 2 * M_PI / N;
    double bestdef = sp.def_avg, bestfom = -1;
    const bool scan = def_std > 0 && (sp.flag || sp.def_avg != 0);
    if (scan) {
        const double d0 = bestdef - 5 * def_std, d1 = bestdef + 5 * def_std;
        double dv[16];
        int nb = 0;
        for (double d = d0; d <= d1 && nb < 16; d += def_std) dv[nb++] = d;
        for (int c0 = 0; c0 < nb; c0 += 4) {
            const int c = c0 + wave;
            if (c < nb) {
                const double f = frc_fit(P, Cf, sec, hk, ring_start, R, lane, 0, 0, sp, true, dv[c]);
                if (lane == 0) cf[c] = f;
            }
        }
        __syncthreads();
        for (int c = 0; c < nb; ++c)
            if (bestfom < cf[c]) { bestfom = cf[c]; bestdef = dv[c]; }
        __syncthreads();
    }
    // Each pass either halves the step or repeats the same grid, whose FOMs
    // are then equal, so the step halves at least every second pass; the
    // caller bounds step / accuracy, and the pass count is capped besides.
    double step = step0, bx = sp.ox, by = sp.oy;
    for (int pass = 0; step >= accuracy && pass < 128; ++pass) {
        const double prevfom = bestfom;
        bestfom = -1;
        double gx[9], gy[9];
        int n = 0;
        for (double y = sp.oy - step; y <= sp.oy + step && n < 9; y += step)
            for (double x = sp.ox - step; x <= sp.ox + step && n < 9; x += step) {
                gx[n] = x;
                gy[n] = y;
                ++n;
            }
        for (int c0 = 0; c0 < n; c0 += 4) {
            const int c = c0 + wave;
            if (c < n) {
                const double f = frc_fit(P, Cf, sec, hk, ring_start, R, lane, (gx[c] - sp.ox) * tpn,
                                         (gy[c] - sp.oy) * tpn, sp, scan, bestdef);
                if (lane == 0) cf[c] = f;
            }
        }
        __syncthreads();
        int ibest = 0;
        for (int c = 0; c < n; ++c)
            if (bestfom < cf[c]) { bestfom = cf[c]; ibest = c; bx = gx[c]; by = gy[c]; }
        __syncthreads();
        if (ibest == n / 2 || fabs(prevfom - bestfom) < 1e-6) step /= 2;
    }
    if (threadIdx.x == 0) {
        origin[2 * p] = bx;
        origin[2 * p + 1] = by;
        if (defocus) defocus[p] = bestdef;
        fom[p] = (float)bestfom;
    }
}

// img_fit_in_recip_space_monte (src/mg/mg_refine.cpp:783-860), one workgroup
// per particle at its current orientation: the section at the disk points is
// interpolated into LDS as in k_fit_grid, then one wavefront walks the chain.
// Iteration i > 0 moves the defocus by def_std g (when the CTF applies) and
// the origin by shift_std a u (u a unit direction), scores the FOM, keeps the
// best, and accepts the move when fom / prefom exceeds a uniform value, else
// returns to the last accepted state. prefom starts at -1, so a positive FOM
// is never accepted and every move starts from the staged origin, as in the
// reference. The random values come from the caller: draws holds max_iter x 5
// FP64 per particle, (g, a, ux, uy, uniform); iteration 0 reads only the
// uniform value. When the defocus is not stepped but the CTF flag is set, the
// first rejection resets the defocus to 0 (the reference's predef), as there.
template <int WD>
__global__ __launch_bounds__(256) void k_fit_monte(
    const float2* __restrict__ ft, int N, const double* __restrict__ mats, double inv, double rs,
    const float* __restrict__ ktab, const int2* __restrict__ hk, const int* __restrict__ ring_start,
    int R, int Nd, const float2* __restrict__ fp, const float* __restrict__ ctf,
    const SPart* __restrict__ sps, int max_iter, double def_std, double shift_std,
    const double* __restrict__ draws, const double* __restrict__ def0,
    double* __restrict__ origin, double* __restrict__ defocus, float* __restrict__ fom) {
    __shared__ float K[(FSI_NDIV + 1) * WD];
    extern __shared__ float2 sec[];   // Nd section values
    for (int i = threadIdx.x; i < (FSI_NDIV + 1) * WD; i += blockDim.x) K[i] = ktab[i];
    __syncthreads();
    const int64_t p = blockIdx.x;
    const SPart sp = sps[p];
    const double* M = mats + 9 * p;
    for (int i = threadIdx.x; i < Nd; i += blockDim.x) {
        const double ivx = (double)hk[i].x * inv, ivy = (double)hk[i].y * inv;
        double m[3];
        for (int q = 0; q < 3; ++q) m[q] = (ivx * M[3 * q] + ivy * M[3 * q + 1]) * rs;
        double vr, vi;
        fsi_sample<WD>(ft, N, m, K, vr, vi);
        sec[i] = make_float2((float)vr, (float)vi);
    }
    __syncthreads();
    if (threadIdx.x >= 64) return;   // the chain is sequential: one wavefront
    const int lane = threadIdx.x;
    const float2* P = fp + p * Nd;
    const float* Cf = ctf + p * Nd;
    const double tpn = 2 * M_PI / N;
    const double def_in = def0 ? def0[p] : sp.def_avg;   // the defocus the walk starts from
    const bool dstep = def_std > 0 && def_in != 0;
    const bool apply = sp.flag || dstep;
    double def = def_in, predef = dstep ? def : 0, bestdef = predef;
    double ox = sp.ox, oy = sp.oy, px = ox, py = oy, bx = ox, by = oy;
    double prefom = -1, bestfom = -1;
    const double* D = draws + (size_t)p * max_iter * 5;
    for (int it = 0; it < max_iter; ++it) {
        const double* d = D + 5 * it;
        if (it > 0) {
            if (apply) def = def + def_std * d[0];
            const double amp = shift_std * d[1];
            ox = ox + d[2] * amp;
            oy = oy + d[3] * amp;
        }
        const double f = frc_fit(P, Cf, sec, hk, ring_start, R, lane, (ox - sp.ox) * tpn,
                                 (oy - sp.oy) * tpn, sp, true, def);
        if (bestfom < f) {
            bestfom = f;
            if (apply) bestdef = def;
            bx = ox;
            by = oy;
        }
        if (f / prefom > d[4]) {
            if (apply) predef = def;
            px = ox;
            py = oy;
            prefom = f;
        } else {
            if (apply) def = predef;
            ox = px;
            oy = py;
        }
    }
    if (lane == 0) {
        origin[2 * p] = bx;
        origin[2 * p + 1] = by;
        if (defocus) defocus[p] = apply ? bestdef : def_in;
        fom[p] = (float)bestfom;
    }
}

// Local refinement grid: candidate c of particle p is R0 * Rot(delta) with
// delta = step * ((c%3)-1, (c/3%3)-1, (c/9)-1) a rotation vector (Rodrigues).
__global__ void k_local_mats(const double* __restrict__ r0, int64_t n, double step,
                             double* __restrict__ mats) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n * 27) return;
    const int64_t p = g / 27;
    const int c = (int)(g % 27);
    const double wx = step * (c % 3 - 1), wy = step * (c / 3 % 3 - 1), wz = step * (c / 9 - 1);
    const double th = sqrt(wx * wx + wy * wy + wz * wz);
    double D[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (th > 0) {
        const double ux = wx / th, uy = wy / th, uz = wz / th;
        const double cs = cos(th), sn = sin(th), t = 1 - cs;
        D[0] = cs + ux * ux * t;      D[1] = ux * uy * t - uz * sn; D[2] = ux * uz * t + uy * sn;
        D[3] = uy * ux * t + uz * sn; D[4] = cs + uy * uy * t;      D[5] = uy * uz * t - ux * sn;
        D[6] = uz * ux * t - uy * sn; D[7] = uz * uy * t + ux * sn; D[8] = cs + uz * uz * t;
    }
    const double* A = r0 + 9 * p;
    double* M = mats + 9 * g;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            M[3 * i + j] = A[3 * i] * D[j] + A[3 * i + 1] * D[3 + j] + A[3 * i + 2] * D[6 + j];
}

// Best of the 27 per particle (ties keep the centre, then the lowest index);
// the winner becomes the new centre.
__global__ void k_local_pick(const double* __restrict__ mats, const float* __restrict__ frc,
                             int64_t n, double* __restrict__ r0, float* __restrict__ best) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int bc = 13;
    float bf = frc[p * 27 + 13];
    for (int c = 0; c < 27; ++c) {
        const float f = frc[p * 27 + c];
        if (f > bf) { bf = f; bc = c; }
    }
    for (int i = 0; i < 9; ++i) r0[9 * p + i] = mats[(p * 27 + bc) * 9 + i];
    if (best) best[p] = bf;
}

// img_refine_grid's view loop (src/mg/mg_refine.cpp:1116-1140) on the GPU,
// one thread per particle. State per particle: the best view (x, y, z,
// angle), the angular step, the best FOM of the last pass (-1 before the
// first) and whether it still iterates (step >= accuracy).
__global__ void k_rv_init(const double* __restrict__ mats, int64_t n, double step,
                          double accuracy, double* __restrict__ view, double* __restrict__ st,
                          float* __restrict__ fom, int* __restrict__ active) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    matrix_view(mats + 9 * p, view + 4 * p);
    st[p] = step;
    fom[p] = -1.f;
    active[p] = step >= accuracy;
}

// The 27 rotation matrices of views_for_refinement around each iterating
// particle's best view.
__global__ void k_rv_mats(const double* __restrict__ view, const double* __restrict__ st,
                          const int* __restrict__ active, int64_t n, double* __restrict__ mats) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || !active[p]) return;
    double grid[27 * 4];
    refine_view_grid(view + 4 * p, st[p], grid);
    for (int c = 0; c < 27; ++c) view_matrix(grid + 4 * c, mats + (p * 27 + c) * 9);
}

// One pass's decision: the first view of the highest FOM (strict >, from
// -1) becomes the best view; the step halves when that is the centre
// (entry 13) or the FOM moved by less than 1e-6; the particle stops once
// the step is below the accuracy. `left` counts particles still iterating.
__global__ void k_rv_pick(const float* __restrict__ frc, int64_t n, double accuracy,
                          double* __restrict__ view, double* __restrict__ st,
                          float* __restrict__ fom, int* __restrict__ active,
                          int* __restrict__ left) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || !active[p]) return;
    const double prefom = fom[p];
    double best = -1;
    int ib = 0;
    for (int c = 0; c < 27; ++c) {
        const double f = frc[p * 27 + c];
        if (best < f) { best = f; ib = c; }
    }
    double grid[27 * 4];
    refine_view_grid(view + 4 * p, st[p], grid);
    for (int i = 0; i < 4; ++i) view[4 * p + i] = grid[4 * ib + i];
    fom[p] = (float)best;
    if (ib == 13 || fabs(prefom - best) < 1e-6) st[p] /= 2;
    active[p] = st[p] >= accuracy;
    if (active[p]) atomicAdd(left, 1);
}

// Matrices of the current view for particles that do not iterate (all 27
// entries), so that every matrix scored is a rotation.
__global__ void k_rv_fill(const double* __restrict__ view, const int* __restrict__ active,
                          int64_t n, double* __restrict__ mats) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || active[p]) return;
    double m[9];
    view_matrix(view + 4 * p, m);
    for (int c = 0; c < 27; ++c)
        for (int i = 0; i < 9; ++i) mats[(p * 27 + c) * 9 + i] = m[i];
}

__global__ void k_rv_finish(const double* __restrict__ view, const float* __restrict__ fom,
                            int64_t n, double* __restrict__ mats, float* __restrict__ frc) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    view_matrix(view + 4 * p, mats + 9 * p);
    if (frc) frc[p] = fom[p];
}

// Library view matrices of the chosen views (negative index: identity).
__global__ void k_view_mats(const double* __restrict__ lm, int64_t nviews,
                            const int* __restrict__ idx, int64_t n, double* __restrict__ out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int v = idx[p];
    for (int i = 0; i < 9; ++i)
        out[9 * p + i] = (v >= 0 && v < nviews) ? lm[9 * (int64_t)v + i] : (i % 4 == 0 ? 1.0 : 0.0);
}

// Input image of the proposal network: disk values phase-flipped by the
// sign of the CTF on an S x S grid (mates filled by Hermitian symmetry,
// centre shifted to S/2). gidx[g]: disk index, -(i+1) for the conjugate
// mate, INT_MIN outside the disk.
__global__ void k_cnn_grid(const float2* __restrict__ fp, const float* __restrict__ ctf, int Nd,
                           const int* __restrict__ gidx, int S, int64_t p0, int64_t nb,
                           float2* __restrict__ grid) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t S2 = (int64_t)S * S;
    if (t >= nb * S2) return;
    const int64_t p = p0 + t / S2;
    const int g = (int)(t % S2);
    const int code = gidx[g];
    float2 v = make_float2(0.f, 0.f);
    if (code != INT_MIN) {
        const int i = code >= 0 ? code : -code - 1;
        const float2 F = fp[p * Nd + i];
        const float sg = ctf[p * Nd + i] < 0.f ? -1.f : 1.f;
        v = make_float2(F.x * sg, code >= 0 ? F.y * sg : -F.y * sg);
        const int xx = g % S, yy = g / S;
        if ((xx + yy) & 1) { v.x = -v.x; v.y = -v.y; }
    }
    grid[t] = v;
}

// Real part of the inverse transform, normalised to zero mean and unit
// variance per image (one workgroup per image, FP64 sums).
__global__ __launch_bounds__(256) void k_cnn_norm(const float2* __restrict__ grid, int S,
                                                  int64_t p0, float* __restrict__ out) {
    __shared__ double rs[256], rq[256];
    const int64_t b = blockIdx.x;
    const int S2 = S * S;
    const float2* G = grid + b * S2;
    double s = 0, q = 0;
    for (int i = threadIdx.x; i < S2; i += blockDim.x) {
        const double v = G[i].x;
        s += v;
        q += v * v;
    }
    rs[threadIdx.x] = s;
    rq[threadIdx.x] = q;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            rs[threadIdx.x] += rs[threadIdx.x + st];
            rq[threadIdx.x] += rq[threadIdx.x + st];
        }
        __syncthreads();
    }
    const double mean = rs[0] / S2;
    const double var = rq[0] / S2 - mean * mean;
    const float inv = var > 0 ? (float)(1.0 / sqrt(var)) : 0.f;
    float* O = out + (p0 + b) * S2;
    for (int i = threadIdx.x; i < S2; i += blockDim.x) O[i] = (float)(G[i].x - mean) * inv;
}

__global__ void k_real_to_complex_b(const float* __restrict__ in, float2* __restrict__ out,
                                    long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        out[i] = make_float2(in[i], 0.f);
}

}  // namespace

// ================================================================ host side
struct bsx_search {
    bsx_search_params p{};
    int device = 0;
    int R = 0, Nd = 0, Kp = 0;
    std::vector<int2> hk;
    std::vector<int> ring_start;
    int2* d_hk = nullptr;
    int* d_ring = nullptr;
    // library
    int64_t nviews = 0, lib_cap = 0, lbf_cap = 0;
    float2* d_lib = nullptr;
    uint16_t* d_lbf = nullptr;
    // staged particles
    int64_t nparts = 0, fp_cap = 0, ctf_cap = 0, pbf_cap = 0;
    float2* d_fp = nullptr;
    float* d_ctf = nullptr;
    uint16_t* d_pbf = nullptr;
    // scratch
    float2* d_spec = nullptr;
    int64_t spec_cap = 0;
    float2* d_cube = nullptr;      // centred low-frequency cube of the map (refinement)
    int64_t cube_cap = 0;
    SPart* d_sp = nullptr;
    int64_t sp_cap = 0;
    double* d_libmats = nullptr;     // nviews x 9, kept for the chosen views
    int64_t libmats_cap = 0;
    double* d_lmats = nullptr;       // local refinement: n x 27 x 9
    int64_t lmats_cap = 0;
    float* d_lfrc = nullptr;
    int64_t lfrc_cap = 0;
    double* d_rv = nullptr;          // views_for_refinement loop: views (4n), steps (n)
    int64_t rv_cap = 0;
    int* d_rvi = nullptr;            // flags (n), FOMs (n floats), counter
    int64_t rvi_cap = 0;
    uint16_t* d_sub = nullptr;       // gathered rows for a partial exhaustive search
    int64_t sub_cap = 0;
    // proposal classes: member views per class
    int nclass = 0, maxm = 0;
    int* d_members = nullptr;
    int64_t members_cap = 0;
    int* d_mcount = nullptr;
    int64_t mcount_cap = 0;
    int* d_counts = nullptr;
    int64_t counts_cap = 0;
    int* d_cursors = nullptr;
    int64_t cursors_cap = 0;
    int* d_pairs = nullptr;
    int64_t pairs_cap = 0;
    GroupItem* d_items = nullptr;
    int64_t items_cap = 0;
    uint16_t* d_zrow = nullptr;
    int64_t zrow_cap = 0;
    float* d_ktab = nullptr;
    float* d_pscore = nullptr;
    int* d_pidx = nullptr;
    int64_t pscore_cap = 0, pidx_cap = 0;
    int64_t bytes = 0;
};

namespace {

template <class T>
int grow(bsx_search* h, T** ptr, int64_t* cap, int64_t need, const char* what) {
    if (need <= *cap && *ptr) return BSX_OK;
    const int64_t add = (int64_t)sizeof(T) * (need - *cap);
    BSX_REQUIRE(h->p.mem_cap <= 0 || h->bytes + add <= h->p.mem_cap, BSX_ERR_NOMEM,
                "bsx_search: %s needs %lld more bytes over the memory cap %lld", what,
                (long long)add, (long long)h->p.mem_cap);
    if (*ptr) {
        (void)hipFree(*ptr);
        h->bytes -= (int64_t)sizeof(T) * *cap;
        *ptr = nullptr;
        *cap = 0;
    }
    hipError_t e = hipMalloc((void**)ptr, sizeof(T) * (size_t)need);
    if (e != hipSuccess) {
        *ptr = nullptr;
        set_error("bsx_search: allocation of %lld bytes for %s failed: %s",
                  (long long)(sizeof(T) * need), what, hipGetErrorString(e));
        return BSX_ERR_NOMEM;
    }
    *cap = need;
    h->bytes += (int64_t)sizeof(T) * need;
    return BSX_OK;
}

inline int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

}  // namespace

namespace {

// Dynamic LDS of k_stage_fused for box N, disk radius R and Nd points; 0 when
// the box has no fused kernel or the columns do not fit.
size_t fused_stage_lds(int N, int R, int Nd) {
    if (N != 32 && N != 64 && N != 128 && N != 256 && N != 512) return 0;
    if (R < 1 || R > N / 2) return 0;
    const int G = 256 / (N / 8);
    const size_t bytes = sizeof(float2) * ((size_t)G * N + N + (size_t)R * N + Nd);
    return bytes <= 120 * 1024 ? bytes : 0;
}

template <int N>
int launch_fused(unsigned nb, size_t lds, hipStream_t s, const float* imgs, const float2* tw,
                 const SPart* sp, const int2* hk, const int* ring, int R, int Nd, int Kp,
                 float2* fp, float* ctf, uint16_t* arow) {
    static bool attr = false;
    if (!attr) {
        BSX_HIP(hipFuncSetAttribute((const void*)k_stage_fused<N>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, 120 * 1024));
        attr = true;
    }
    constexpr unsigned MAXB = 1u << 22;   // workgroups per launch
    for (unsigned p0 = 0; p0 < nb; p0 += MAXB) {
        const unsigned nl = std::min(MAXB, nb - p0);
        hipLaunchKernelGGL(k_stage_fused<N>, dim3(nl), dim3(256), lds, s,
                           imgs + (size_t)p0 * N * N, tw, sp + p0, hk, ring, R, Nd, Kp,
                           (int64_t)p0, fp, ctf, arow);
        BSX_HIP(hipGetLastError());
    }
    return BSX_OK;
}

int launch_stage_fused(int N, unsigned nb, size_t lds, hipStream_t s, const float* imgs,
                       const float2* tw, const SPart* sp, const int2* hk, const int* ring, int R,
                       int Nd, int Kp, float2* fp, float* ctf, uint16_t* arow) {
    switch (N) {
        case 32: return launch_fused<32>(nb, lds, s, imgs, tw, sp, hk, ring, R, Nd, Kp, fp, ctf, arow);
        case 64: return launch_fused<64>(nb, lds, s, imgs, tw, sp, hk, ring, R, Nd, Kp, fp, ctf, arow);
        case 128: return launch_fused<128>(nb, lds, s, imgs, tw, sp, hk, ring, R, Nd, Kp, fp, ctf, arow);
        case 256: return launch_fused<256>(nb, lds, s, imgs, tw, sp, hk, ring, R, Nd, Kp, fp, ctf, arow);
        case 512: return launch_fused<512>(nb, lds, s, imgs, tw, sp, hk, ring, R, Nd, Kp, fp, ctf, arow);
    }
    set_error("no fused staging kernel for box %d", N);
    return BSX_ERR_ARG;
}

}  // namespace

extern "C" {

int bsx_search_create(bsx_search** out, const bsx_search_params* prm) {
    BSX_REQUIRE(out && prm, BSX_ERR_ARG, "bsx_search_create: null argument");
    *out = nullptr;
    const bsx_search_params& p = *prm;
    BSX_REQUIRE(fft_length_supported(p.box) && p.box >= 8 && p.box <= 1024, BSX_ERR_ARG,
                "bsx_search_create: box %d unsupported", p.box);
    BSX_REQUIRE(p.radius >= 2 && p.radius <= 64 && 2 * p.radius + 2 <= p.box, BSX_ERR_ARG,
                "bsx_search_create: radius %d out of range 2..min(64, box/2-1)", p.radius);
    BSX_REQUIRE(p.sampling > 0, BSX_ERR_ARG, "bsx_search_create: sampling must be > 0");
    BSX_REQUIRE(p.kernel_width >= 1 && p.kernel_width <= 10, BSX_ERR_ARG,
                "bsx_search_create: kernel width %d out of range 1..10", p.kernel_width);
    BSX_REQUIRE(p.topk == 8 || p.topk == 16 || (p.topk >= 1 && p.topk < 8), BSX_ERR_ARG,
                "bsx_search_create: topk %d out of range 1..8 or 16", p.topk);
    bsx_search* h = new bsx_search();
    h->p = p;
    if (h->p.kernel_power < 0) h->p.kernel_width = 1;
    (void)hipGetDevice(&h->device);
    const int R = p.radius;
    h->R = R;
    // disk points: DC, then rings 1..R-1 of the half plane (k > 0, or k == 0
    // and h > 0), each ring j = floor(|(h,k)|) contiguous
    h->hk.push_back(make_int2(0, 0));
    h->ring_start.push_back(0);
    for (int j = 1; j < R; ++j) {
        h->ring_start.push_back((int)h->hk.size());
        for (int k = 0; k <= j + 1; ++k)
            for (int x = -j - 1; x <= j + 1; ++x) {
                if (k == 0 && x <= 0) continue;
                const int d2 = x * x + k * k;
                if (d2 >= j * j && d2 < (j + 1) * (j + 1)) h->hk.push_back(make_int2(x, k));
            }
    }
    h->ring_start.push_back((int)h->hk.size());
    h->Nd = (int)h->hk.size();
    h->Kp = (int)round_up(2 * (h->Nd - 1), GK);
    int rc = BSX_OK;
    int64_t c1 = 0, c2 = 0, c3 = 0;
    rc = grow(h, &h->d_hk, &c1, h->Nd, "disk table");
    if (rc == BSX_OK) rc = grow(h, &h->d_ring, &c2, R + 1, "ring table");
    std::vector<float> tab = fsi_table(h->p.kernel_width, h->p.kernel_power);
    if (rc == BSX_OK) rc = grow(h, &h->d_ktab, &c3, (int64_t)tab.size(), "kernel table");
    if (rc == BSX_OK &&
        (hipMemcpy(h->d_hk, h->hk.data(), sizeof(int2) * h->Nd, hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy(h->d_ring, h->ring_start.data(), sizeof(int) * (R + 1), hipMemcpyHostToDevice) != hipSuccess ||
         hipMemcpy(h->d_ktab, tab.data(), sizeof(float) * tab.size(), hipMemcpyHostToDevice) != hipSuccess)) {
        set_error("bsx_search_create: table upload failed");
        rc = BSX_ERR_DEVICE;
    }
    if (rc != BSX_OK) {
        bsx_search_destroy(h);
        return rc;
    }
    *out = h;
    return BSX_OK;
}

int bsx_search_destroy(bsx_search* h) {
    if (!h) return BSX_OK;
    void* ptrs[] = {h->d_hk,   h->d_ring,    h->d_lib,    h->d_lbf,    h->d_fp,
                    h->d_ctf,  h->d_pbf,     h->d_spec,   h->d_sp,
                    h->d_ktab, h->d_pscore,  h->d_pidx,   h->d_libmats, h->d_lmats,
                    h->d_lfrc, h->d_sub,     h->d_members, h->d_mcount, h->d_counts,
                    h->d_cursors, h->d_pairs, h->d_items, h->d_zrow, h->d_cube,
                    h->d_rv,   h->d_rvi};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    delete h;
    return BSX_OK;
}

int bsx_search_info(const bsx_search* h, int32_t* npoints, int32_t* kdim, int64_t* nviews,
                    int64_t* nparts) {
    BSX_REQUIRE(h, BSX_ERR_ARG, "bsx_search_info: null handle");
    if (npoints) *npoints = h->Nd;
    if (kdim) *kdim = h->Kp;
    if (nviews) *nviews = h->nviews;
    if (nparts) *nparts = h->nparts;
    return BSX_OK;
}

int bsx_search_points(const bsx_search* h, int32_t* hk, int32_t* ring_start) {
    BSX_REQUIRE(h, BSX_ERR_ARG, "bsx_search_points: null handle");
    if (hk)
        for (int i = 0; i < h->Nd; ++i) { hk[2 * i] = h->hk[i].x; hk[2 * i + 1] = h->hk[i].y; }
    if (ring_start)
        for (int j = 0; j <= h->R; ++j) ring_start[j] = h->ring_start[j];
    return BSX_OK;
}

int bsx_search_set_reference(bsx_search* h, const float* ft_dev, const double* mats,
                             int64_t nviews, void* stream) {
    BSX_REQUIRE(h && ft_dev && (mats || nviews == 0), BSX_ERR_ARG,
                "bsx_search_set_reference: null argument");
    BSX_REQUIRE(nviews >= 0 && nviews < (1ll << 31) - GN, BSX_ERR_ARG,
                "bsx_search_set_reference: nviews %lld out of range", (long long)nviews);
    for (int64_t i = 0; i < 9 * nviews; ++i)
        BSX_REQUIRE(std::isfinite(mats[i]) && fabs(mats[i]) <= 1.0 + 1e-9, BSX_ERR_ARG,
                    "bsx_search_set_reference: matrix %lld is not a rotation", (long long)(i / 9));
    hipStream_t s = (hipStream_t)stream;
    const int N = h->p.box, Nd = h->Nd, Kp = h->Kp;
    const int64_t vpad = round_up(std::max<int64_t>(nviews, 1), GN4);
    int rc = grow(h, &h->d_lib, &h->lib_cap, vpad * Nd, "library values");
    if (!rc) rc = grow(h, &h->d_lbf, &h->lbf_cap, vpad * Kp, "library rows");
    if (rc) return rc;
    // padded rows score zero and are never selected (column < nviews)
    BSX_HIP(hipMemsetAsync(h->d_lbf, 0, sizeof(uint16_t) * vpad * Kp, s));
    h->nviews = nviews;
    rc = grow(h, &h->d_libmats, &h->libmats_cap, 9 * std::max<int64_t>(nviews, 1), "view matrices");
    if (rc) return rc;
    if (nviews > 0)
        BSX_HIP(hipMemcpyAsync(h->d_libmats, mats, sizeof(double) * 9 * nviews, hipMemcpyHostToDevice, s));
    const int64_t chunk = 65536;
    const double inv = 1.0 / (N * h->p.sampling), rs = N * h->p.sampling;
    const size_t shm = sizeof(float2) * Nd;
    for (int64_t v0 = 0; v0 < nviews; v0 += chunk) {
        const int64_t nv = std::min(chunk, nviews - v0);
        switch (h->p.kernel_width) {
#define BSX_LIB(W) case W: hipLaunchKernelGGL((k_lib_build<W>), dim3((unsigned)nv), dim3(256), shm, s, \
            (const float2*)ft_dev, N, h->d_libmats + 9 * v0, inv, rs, h->d_ktab, h->d_hk, h->d_ring, h->R, \
            Nd, Kp, v0, h->d_lib, h->d_lbf); break;
            BSX_LIB(1) BSX_LIB(2) BSX_LIB(3) BSX_LIB(4) BSX_LIB(5)
            BSX_LIB(6) BSX_LIB(7) BSX_LIB(8) BSX_LIB(9) BSX_LIB(10)
#undef BSX_LIB
        }
        BSX_HIP(hipGetLastError());
    }
    BSX_HIP(hipStreamSynchronize(s));   // the host matrices may be released on return
    return BSX_OK;
}

int bsx_search_stage(bsx_search* h, const float* imgs_dev, const bsx_particle* parts, int64_t n,
                     void* stream) {
    BSX_REQUIRE(h && (n == 0 || (imgs_dev && parts)), BSX_ERR_ARG, "bsx_search_stage: null argument");
    BSX_REQUIRE(n >= 0 && n < (1ll << 31) - GM, BSX_ERR_ARG, "bsx_search_stage: n %lld out of range",
                (long long)n);
    const int N = h->p.box, Nd = h->Nd, Kp = h->Kp;
    for (int64_t i = 0; i < n; ++i) {
        const bsx_particle& q = parts[i];
        BSX_REQUIRE(std::isfinite(q.ori[0]) && std::isfinite(q.ori[1]) && q.pixel_size > 0 &&
                        std::isfinite(q.def_avg) && std::isfinite(q.def_dev) &&
                        std::isfinite(q.ast_angle) && std::isfinite(q.amp_shift),
                    BSX_ERR_ARG, "bsx_search_stage: particle %lld has invalid parameters", (long long)i);
    }
    // records (CTF coefficients, wavelength) built on host threads
    std::vector<SPart> sp((size_t)n);
    const bool ctf_flag = (h->p.flags & BSX_SEARCH_CTF) != 0;
    host_parallel(n, 4096, [&](int64_t b, int64_t e, unsigned) {
        for (int64_t i = b; i < e; ++i) {
            const bsx_particle& q = parts[i];
            SPart& r = sp[i];
            r.phx = q.ori[0] / N * 2 * M_PI;
            r.phy = q.ori[1] / N * 2 * M_PI;
            r.fscale = 1.0 / (N * q.pixel_size);
            r.c = ctf_coef(q);
            r.ox = q.ori[0];
            r.oy = q.ori[1];
            r.def_avg = q.def_avg;
            r.c20u = -M_PI * lambda_of(q.volt);
            r.flag = ctf_flag && q.volt > 0;
            r.apply = r.flag && q.def_avg >= 1;
        }
    });
    hipStream_t s = (hipStream_t)stream;
    const int64_t ppad = round_up(std::max<int64_t>(n, 1), GM4);
    int rc = grow(h, &h->d_fp, &h->fp_cap, ppad * Nd, "particle values");
    if (!rc) rc = grow(h, &h->d_ctf, &h->ctf_cap, ppad * Nd, "particle CTF");
    if (!rc) rc = grow(h, &h->d_pbf, &h->pbf_cap, ppad * Kp, "particle rows");
    if (rc) return rc;
    BSX_HIP(hipMemsetAsync(h->d_pbf, 0, sizeof(uint16_t) * ppad * Kp, s));
    h->nparts = n;
    if (n == 0) return BSX_OK;
    const int64_t per = (int64_t)N * N;
    // fused path: one kernel reads each image once
    const size_t fused_lds = fused_stage_lds(N, h->R, Nd);
    if (fused_lds) {
        rc = grow(h, &h->d_sp, &h->sp_cap, n, "particle records");
        if (rc) return rc;
        BSX_HIP(hipMemcpyAsync(h->d_sp, sp.data(), sizeof(SPart) * n, hipMemcpyHostToDevice, s));
        const float2* tw = twiddles(N, s);
        BSX_REQUIRE(tw, BSX_ERR_DEVICE, "twiddles(%d) failed", N);
        rc = launch_stage_fused(N, (unsigned)n, fused_lds, s, imgs_dev, tw, h->d_sp, h->d_hk,
                                h->d_ring, h->R, Nd, Kp, h->d_fp, h->d_ctf, h->d_pbf);
        if (rc) return rc;
        BSX_HIP(hipStreamSynchronize(s));   // the host records are released on return
        return BSX_OK;
    }
    // otherwise: full spectra, in chunks of at most 256 MB
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(n, (256ll << 20) / (8 * per)));
    rc = grow(h, &h->d_spec, &h->spec_cap, chunk * per, "spectra");
    if (!rc) rc = grow(h, &h->d_sp, &h->sp_cap, n, "particle records");
    if (rc) return rc;
    // all records stay on the device: the origin/defocus fit reads them later
    BSX_HIP(hipMemcpyAsync(h->d_sp, sp.data(), sizeof(SPart) * n, hipMemcpyHostToDevice, s));
    for (int64_t p0 = 0; p0 < n; p0 += chunk) {
        const int64_t nb = std::min(chunk, n - p0);
        hipLaunchKernelGGL(k_real_to_complex_b, dim3(2048), dim3(256), 0, s, imgs_dev + p0 * per,
                           h->d_spec, (long long)(nb * per));
        BSX_HIP(hipGetLastError());
        rc = fft_axis_ex(h->d_spec, nullptr, N, N, 1, nb, 0, -1, 1.0f, s);
        if (!rc) rc = fft_axis_ex(h->d_spec, nullptr, N, N, 1, nb, 1, -1, 1.0f / N, s);
        if (rc) return rc;
        hipLaunchKernelGGL(k_part_stage, dim3((unsigned)nb), dim3(256), sizeof(float2) * Nd, s,
                           h->d_spec, N, h->d_sp + p0, h->d_hk, h->d_ring, h->R, Nd, Kp, p0, h->d_fp,
                           h->d_ctf, h->d_pbf);
        BSX_HIP(hipGetLastError());
        BSX_HIP(hipStreamSynchronize(s));   // records and spectra are reused
    }
    return BSX_OK;
}

namespace {

// GEMM + top-k of `nrows` BF16 rows at A (allocated to a multiple of GM4 rows)
// against the whole library.
int run_exhaustive(bsx_search* h, const uint16_t* A, int64_t nrows, int32_t* cand_dev,
                   float* score_dev, hipStream_t s) {
    const int T = h->p.topk <= 8 ? 8 : 16;
    const int64_t rblocks = round_up(nrows, GM4) / GM4;
    const int ntiles = (int)(round_up(h->nviews, GN4) / GN4);
    // enough segments to give the chip >= 1024 workgroups (one per CU at a
    // time), at least 4 tiles each
    int nseg = (int)std::max<int64_t>(1, std::min<int64_t>(ntiles / 4 > 0 ? ntiles / 4 : 1,
                                                           (1024 + rblocks - 1) / rblocks));
    nseg = std::min(nseg, 64);
    const int tps = (ntiles + nseg - 1) / nseg;
    nseg = (ntiles + tps - 1) / tps;
    const int64_t need = rblocks * GM4 * nseg * 4 * T;
    int rc = grow(h, &h->d_pscore, &h->pscore_cap, need, "top-k partials");
    if (!rc) rc = grow(h, &h->d_pidx, &h->pidx_cap, need, "top-k partials");
    if (rc) return rc;
    BSX_REQUIRE(rblocks * nseg < (1ll << 31) - 8, BSX_ERR_ARG, "bsx_search_exhaustive: too many particles");
    const dim3 grid((unsigned)(8 * ((rblocks * nseg + 7) / 8)));
    if (T == 8)
        hipLaunchKernelGGL((k_gemm_topk4<8>), grid, dim3(512), 0, s, A, h->d_lbf, h->Kp,
                           (int)h->nviews, tps, ntiles, (int)rblocks, nseg, h->d_pscore, h->d_pidx);
    else
        hipLaunchKernelGGL((k_gemm_topk4<16>), grid, dim3(512), 0, s, A, h->d_lbf, h->Kp,
                           (int)h->nviews, tps, ntiles, (int)rblocks, nseg, h->d_pscore, h->d_pidx);
    BSX_HIP(hipGetLastError());
    const unsigned mg = (unsigned)((nrows + 255) / 256);
    if (T == 8)
        hipLaunchKernelGGL((k_topk_merge<8>), dim3(mg), dim3(256), 0, s, h->d_pscore, h->d_pidx,
                           4 * nseg, nrows, h->p.topk, cand_dev, score_dev);
    else
        hipLaunchKernelGGL((k_topk_merge<16>), dim3(mg), dim3(256), 0, s, h->d_pscore, h->d_pidx,
                           4 * nseg, nrows, h->p.topk, cand_dev, score_dev);
    BSX_HIP(hipGetLastError());
    return BSX_OK;
}

__global__ void k_gather_rows(const uint16_t* __restrict__ src, const int* __restrict__ rows,
                              int64_t m, int64_t nsrc, int Kp, uint16_t* __restrict__ dst) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = Kp / 8;
    if (t >= m * per) return;
    const int64_t r = t / per, c = t % per;
    const int64_t q = rows[r];
    uint4 v = make_uint4(0, 0, 0, 0);
    if (q >= 0 && q < nsrc) v = *(const uint4*)(src + q * Kp + c * 8);
    *(uint4*)(dst + r * Kp + c * 8) = v;
}

}  // namespace

int bsx_search_exhaustive(bsx_search* h, int32_t* cand_dev, float* score_dev, void* stream) {
    BSX_REQUIRE(h && cand_dev && score_dev, BSX_ERR_ARG, "bsx_search_exhaustive: null argument");
    BSX_REQUIRE(h->nviews > 0, BSX_ERR_STATE, "bsx_search_exhaustive: no reference views");
    if (h->nparts == 0) return BSX_OK;
    return run_exhaustive(h, h->d_pbf, h->nparts, cand_dev, score_dev, (hipStream_t)stream);
}

int bsx_search_exhaustive_rows(bsx_search* h, const int32_t* rows_dev, int64_t m,
                               int32_t* cand_dev, float* score_dev, void* stream) {
    BSX_REQUIRE(h && (m == 0 || (rows_dev && cand_dev && score_dev)), BSX_ERR_ARG,
                "bsx_search_exhaustive_rows: null argument");
    BSX_REQUIRE(h->nviews > 0, BSX_ERR_STATE, "bsx_search_exhaustive_rows: no reference views");
    BSX_REQUIRE(m >= 0 && m <= h->nparts, BSX_ERR_ARG,
                "bsx_search_exhaustive_rows: %lld rows of %lld staged", (long long)m,
                (long long)h->nparts);
    if (m == 0) return BSX_OK;
    hipStream_t s = (hipStream_t)stream;
    int rc = grow(h, &h->d_sub, &h->sub_cap, round_up(m, GM4) * h->Kp, "row subset");
    if (rc) return rc;
    const int64_t tot = m * (h->Kp / 8);
    hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, h->d_pbf,
                       rows_dev, m, h->nparts, h->Kp, h->d_sub);
    BSX_HIP(hipGetLastError());
    return run_exhaustive(h, h->d_sub, m, cand_dev, score_dev, s);
}

int bsx_search_set_classes(bsx_search* h, const int32_t* members, int32_t nclasses,
                           int32_t maxm) {
    BSX_REQUIRE(h && members, BSX_ERR_ARG, "bsx_search_set_classes: null argument");
    BSX_REQUIRE(nclasses >= 1 && maxm >= 1 && (int64_t)nclasses * maxm < (1ll << 31), BSX_ERR_ARG,
                "bsx_search_set_classes: %d classes x %d members out of range", nclasses, maxm);
    BSX_REQUIRE(h->nviews > 0, BSX_ERR_STATE, "bsx_search_set_classes: set the reference views first");
    std::vector<int> cnt(nclasses);
    for (int c = 0; c < nclasses; ++c) {
        int m = 0;
        while (m < maxm && members[(size_t)c * maxm + m] >= 0) ++m;
        for (int j = 0; j < maxm; ++j) {
            const int v = members[(size_t)c * maxm + j];
            BSX_REQUIRE(j < m ? v < h->nviews : v < 0, BSX_ERR_ARG,
                        "bsx_search_set_classes: class %d entry %d (%d) is not a view index of a "
                        "leading run", c, j, v);
        }
        cnt[c] = m;
    }
    int rc = grow(h, &h->d_members, &h->members_cap, (int64_t)nclasses * maxm, "class members");
    if (!rc) rc = grow(h, &h->d_mcount, &h->mcount_cap, nclasses, "class sizes");
    if (!rc) rc = grow(h, &h->d_zrow, &h->zrow_cap, h->Kp, "zero row");
    if (rc) return rc;
    BSX_HIP(hipMemcpy(h->d_members, members, sizeof(int) * nclasses * maxm, hipMemcpyHostToDevice));
    BSX_HIP(hipMemcpy(h->d_mcount, cnt.data(), sizeof(int) * nclasses, hipMemcpyHostToDevice));
    BSX_HIP(hipMemset(h->d_zrow, 0, sizeof(uint16_t) * h->Kp));
    h->nclass = nclasses;
    h->maxm = maxm;
    return BSX_OK;
}

int bsx_search_proposed(bsx_search* h, const int32_t* cls_dev, int32_t k, int32_t* cand_dev,
                        float* score_dev, void* stream) {
    BSX_REQUIRE(h && (h->nparts == 0 || (cls_dev && cand_dev && score_dev)), BSX_ERR_ARG,
                "bsx_search_proposed: null argument");
    BSX_REQUIRE(h->nclass > 0, BSX_ERR_STATE, "bsx_search_proposed: no classes (bsx_search_set_classes)");
    BSX_REQUIRE(k >= 1 && k <= 64, BSX_ERR_ARG, "bsx_search_proposed: k %d out of range 1..64", k);
    const int64_t n = h->nparts;
    if (n == 0) return BSX_OK;
    BSX_REQUIRE(n * k < (1ll << 31), BSX_ERR_ARG, "bsx_search_proposed: too many particles");
    hipStream_t s = (hipStream_t)stream;
    const int T = h->p.topk <= 8 ? 8 : 16;
    int rc = grow(h, &h->d_counts, &h->counts_cap, h->nclass, "routing counts");
    if (!rc) rc = grow(h, &h->d_cursors, &h->cursors_cap, h->nclass, "routing cursors");
    if (!rc) rc = grow(h, &h->d_pairs, &h->pairs_cap, n * k, "routing pairs");
    if (!rc) rc = grow(h, &h->d_pscore, &h->pscore_cap, n * k * T, "proposal partials");
    if (!rc) rc = grow(h, &h->d_pidx, &h->pidx_cap, n * k * T, "proposal partials");
    if (rc) return rc;
    const unsigned g = (unsigned)((n * k + 255) / 256);
    BSX_HIP(hipMemsetAsync(h->d_counts, 0, sizeof(int) * h->nclass, s));
    hipLaunchKernelGGL(k_route_count, dim3(g), dim3(256), 0, s, cls_dev, n, k, h->nclass, h->d_counts);
    BSX_HIP(hipGetLastError());
    std::vector<int> cnt(h->nclass), off(h->nclass);
    BSX_HIP(hipMemcpyAsync(cnt.data(), h->d_counts, sizeof(int) * h->nclass, hipMemcpyDeviceToHost, s));
    BSX_HIP(hipStreamSynchronize(s));
    std::vector<GroupItem> items;
    int acc = 0;
    for (int c = 0; c < h->nclass; ++c) {
        off[c] = acc;
        for (int p0 = 0; p0 < cnt[c]; p0 += GM)
            items.push_back(GroupItem{c, acc + p0, std::min(GM, cnt[c] - p0), 0});
        acc += cnt[c];
    }
    rc = grow(h, &h->d_items, &h->items_cap, std::max<int64_t>(1, (int64_t)items.size()), "work items");
    if (rc) return rc;
    BSX_HIP(hipMemcpyAsync(h->d_cursors, off.data(), sizeof(int) * h->nclass, hipMemcpyHostToDevice, s));
    if (!items.empty())
        BSX_HIP(hipMemcpyAsync(h->d_items, items.data(), sizeof(GroupItem) * items.size(),
                               hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_route_fill, dim3(g), dim3(256), 0, s, cls_dev, n, k, h->nclass, h->d_cursors,
                       h->d_pairs);
    BSX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_route_empty, dim3(g), dim3(256), 0, s, cls_dev, n, k, h->nclass, T,
                       h->d_pscore, h->d_pidx);
    BSX_HIP(hipGetLastError());
    if (!items.empty()) {
        if (T == 8)
            hipLaunchKernelGGL((k_gemm_grouped<8>), dim3((unsigned)items.size()), dim3(256), 0, s,
                               h->d_pbf, h->d_lbf, h->d_zrow, h->Kp, h->d_items, h->d_pairs, k,
                               h->d_members, h->d_mcount, h->maxm, h->d_pscore, h->d_pidx);
        else
            hipLaunchKernelGGL((k_gemm_grouped<16>), dim3((unsigned)items.size()), dim3(256), 0, s,
                               h->d_pbf, h->d_lbf, h->d_zrow, h->Kp, h->d_items, h->d_pairs, k,
                               h->d_members, h->d_mcount, h->maxm, h->d_pscore, h->d_pidx);
        BSX_HIP(hipGetLastError());
    }
    const unsigned mg = (unsigned)((n + 255) / 256);
    if (T == 8)
        hipLaunchKernelGGL((k_topk_merge<8>), dim3(mg), dim3(256), 0, s, h->d_pscore, h->d_pidx, k,
                           n, h->p.topk, cand_dev, score_dev);
    else
        hipLaunchKernelGGL((k_topk_merge<16>), dim3(mg), dim3(256), 0, s, h->d_pscore, h->d_pidx, k,
                           n, h->p.topk, cand_dev, score_dev);
    BSX_HIP(hipGetLastError());
    // the host work-item list is released on return
    BSX_HIP(hipStreamSynchronize(s));
    return BSX_OK;
}

int bsx_search_rescore(bsx_search* h, const int32_t* cand_dev, int32_t ncand, float* frc_dev,
                       int32_t* best_dev, float* best_frc_dev, void* stream) {
    BSX_REQUIRE(h && cand_dev, BSX_ERR_ARG, "bsx_search_rescore: null argument");
    BSX_REQUIRE(ncand >= 1 && ncand <= 4096, BSX_ERR_ARG, "bsx_search_rescore: ncand %d out of range",
                ncand);
    BSX_REQUIRE(h->nviews > 0, BSX_ERR_STATE, "bsx_search_rescore: no reference views");
    if (h->nparts == 0) return BSX_OK;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_rescore, dim3((unsigned)h->nparts), dim3(256), 0, s, h->d_fp, h->d_ctf,
                       h->d_lib, h->nviews, h->d_ring, h->R, h->Nd, cand_dev, ncand, frc_dev,
                       best_dev, best_frc_dev);
    BSX_HIP(hipGetLastError());
    return BSX_OK;
}

int bsx_search_score_mats(bsx_search* h, const float* ft_dev, const double* mats_dev,
                          int32_t ncand, float* frc_dev, void* stream) {
    BSX_REQUIRE(h && ft_dev && mats_dev && frc_dev, BSX_ERR_ARG,
                "bsx_search_score_mats: null argument");
    BSX_REQUIRE(ncand >= 1 && (int64_t)ncand * std::max<int64_t>(h->nparts, 1) < (1ll << 31),
                BSX_ERR_ARG, "bsx_search_score_mats: ncand %d out of range", ncand);
    if (h->nparts == 0) return BSX_OK;
    hipStream_t s = (hipStream_t)stream;
    const int N = h->p.box;
    const double inv = 1.0 / (N * h->p.sampling), rs = N * h->p.sampling;
    const unsigned g = (unsigned)h->nparts;
    // sections at the disk points only reach |coordinate| < R: gather from a
    // centred cube of the map around the origin when it is smaller than the map
    const int S = 2 * (h->R + h->p.kernel_width + 2);
    const float2* cube = nullptr;
    if (S < N) {
        int rc = grow(h, &h->d_cube, &h->cube_cap, (int64_t)S * S * S, "centred cube");
        if (rc) return rc;
        hipLaunchKernelGGL(k_center_cube, dim3(1024), dim3(256), 0, s, (const float2*)ft_dev, N,
                           S, h->d_cube);
        BSX_HIP(hipGetLastError());
        cube = h->d_cube;
    }
    switch (h->p.kernel_width) {
#define BSX_SM(W) case W: hipLaunchKernelGGL((k_score_mats<W>), dim3(g), dim3(256), 0, s, \
        (const float2*)ft_dev, N, cube, S, mats_dev, ncand, inv, rs, h->d_ktab, h->d_hk, h->d_ring, \
        h->R, h->Nd, h->d_fp, h->d_ctf, frc_dev); break;
        BSX_SM(1) BSX_SM(2) BSX_SM(3) BSX_SM(4) BSX_SM(5)
        BSX_SM(6) BSX_SM(7) BSX_SM(8) BSX_SM(9) BSX_SM(10)
#undef BSX_SM
    }
    BSX_HIP(hipGetLastError());
    return BSX_OK;
}

int bsx_search_fit_origin(bsx_search* h, const float* ft_dev, const double* mats_dev, double step,
                          double accuracy, double def_std, double* origin_dev, double* defocus_dev,
                          float* fom_dev, void* stream) {
    BSX_REQUIRE(h, BSX_ERR_ARG, "bsx_search_fit_origin: null handle");
    BSX_REQUIRE(h->nparts == 0 || (ft_dev && mats_dev && origin_dev && fom_dev), BSX_ERR_ARG,
                "bsx_search_fit_origin: null argument");
    BSX_REQUIRE(step > 0 && accuracy > 0 && std::isfinite(step) && std::isfinite(accuracy) &&
                    step / accuracy < 1e6 && def_std >= 0 && std::isfinite(def_std),
                BSX_ERR_ARG, "bsx_search_fit_origin: bad step %g / accuracy %g / def_std %g", step,
                accuracy, def_std);
    if (h->nparts == 0) return BSX_OK;
    BSX_REQUIRE(h->sp_cap >= h->nparts, BSX_ERR_STATE,
                "bsx_search_fit_origin: particle records missing (stage with the fused path)");
    hipStream_t s = (hipStream_t)stream;
    const int N = h->p.box;
    const double inv = 1.0 / (N * h->p.sampling), rs = N * h->p.sampling;
    const size_t shm = sizeof(float2) * h->Nd;
    switch (h->p.kernel_width) {
#define BSX_FG(W) case W: \
        BSX_HIP(hipFuncSetAttribute((const void*)k_fit_grid<W>, \
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm)); \
        hipLaunchKernelGGL((k_fit_grid<W>), dim3((unsigned)h->nparts), dim3(256), \
        shm, s, (const float2*)ft_dev, N, mats_dev, inv, rs, h->d_ktab, h->d_hk, h->d_ring, h->R, \
        h->Nd, h->d_fp, h->d_ctf, h->d_sp, step, accuracy, def_std, origin_dev, defocus_dev, \
        fom_dev); break;
        BSX_FG(1) BSX_FG(2) BSX_FG(3) BSX_FG(4) BSX_FG(5)
        BSX_FG(6) BSX_FG(7) BSX_FG(8) BSX_FG(9) BSX_FG(10)
#undef BSX_FG
    }
    BSX_HIP(hipGetLastError());
    return BSX_OK;
}

int bsx_search_fit_monte(bsx_search* h, const float* ft_dev, const double* mats_dev,
                         int32_t max_iter, double def_std, double shift_std,
                         const double* draws_dev, const double* defocus_start_dev,
                         double* origin_dev, double* defocus_dev, float* fom_dev, void* stream) {
    BSX_REQUIRE(h, BSX_ERR_ARG, "bsx_search_fit_monte: null handle");
    BSX_REQUIRE(h->nparts == 0 || (ft_dev && mats_dev && draws_dev && origin_dev && fom_dev),
                BSX_ERR_ARG, "bsx_search_fit_monte: null argument");
    BSX_REQUIRE(max_iter >= 1 && max_iter <= 1000000 && def_std >= 0 && std::isfinite(def_std) &&
                    shift_std >= 0 && std::isfinite(shift_std),
                BSX_ERR_ARG, "bsx_search_fit_monte: bad max_iter %d / def_std %g / shift_std %g",
                max_iter, def_std, shift_std);
    if (h->nparts == 0) return BSX_OK;
    BSX_REQUIRE(h->sp_cap >= h->nparts, BSX_ERR_STATE,
                "bsx_search_fit_monte: particle records missing (stage with the fused path)");
    hipStream_t s = (hipStream_t)stream;
    const int N = h->p.box;
    const double inv = 1.0 / (N * h->p.sampling), rs = N * h->p.sampling;
    const size_t shm = sizeof(float2) * h->Nd;
    switch (h->p.kernel_width) {
#define BSX_FM(W) case W: \
        BSX_HIP(hipFuncSetAttribute((const void*)k_fit_monte<W>, \
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm)); \
        hipLaunchKernelGGL((k_fit_monte<W>), dim3((unsigned)h->nparts), dim3(256), \
        shm, s, (const float2*)ft_dev, N, mats_dev, inv, rs, h->d_ktab, h->d_hk, h->d_ring, h->R, \
        h->Nd, h->d_fp, h->d_ctf, h->d_sp, max_iter, def_std, shift_std, draws_dev, \
        defocus_start_dev, origin_dev, defocus_dev, fom_dev); break;
        BSX_FM(1) BSX_FM(2) BSX_FM(3) BSX_FM(4) BSX_FM(5)
        BSX_FM(6) BSX_FM(7) BSX_FM(8) BSX_FM(9) BSX_FM(10)
#undef BSX_FM
    }
    BSX_HIP(hipGetLastError());
    return BSX_OK;
}

int bsx_search_particle_data(const bsx_search* h, float* values_dev, float* ctf_dev,
                             uint16_t* rows_dev, void* stream) {
    BSX_REQUIRE(h, BSX_ERR_ARG, "bsx_search_particle_data: null handle");
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = h->nparts;
    if (n == 0) return BSX_OK;
    if (values_dev)
        BSX_HIP(hipMemcpyAsync(values_dev, h->d_fp, sizeof(float2) * n * h->Nd, hipMemcpyDeviceToDevice, s));
    if (ctf_dev)
        BSX_HIP(hipMemcpyAsync(ctf_dev, h->d_ctf, sizeof(float) * n * h->Nd, hipMemcpyDeviceToDevice, s));
    if (rows_dev)
        BSX_HIP(hipMemcpyAsync(rows_dev, h->d_pbf, sizeof(uint16_t) * n * h->Kp, hipMemcpyDeviceToDevice, s));
    return BSX_OK;
}

int bsx_search_library_data(const bsx_search* h, float* values_dev, uint16_t* rows_dev,
                            void* stream) {
    BSX_REQUIRE(h, BSX_ERR_ARG, "bsx_search_library_data: null handle");
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = h->nviews;
    if (n == 0) return BSX_OK;
    if (values_dev)
        BSX_HIP(hipMemcpyAsync(values_dev, h->d_lib, sizeof(float2) * n * h->Nd, hipMemcpyDeviceToDevice, s));
    if (rows_dev)
        BSX_HIP(hipMemcpyAsync(rows_dev, h->d_lbf, sizeof(uint16_t) * n * h->Kp, hipMemcpyDeviceToDevice, s));
    return BSX_OK;
}

int bsx_search_view_mats(const bsx_search* h, const int32_t* views_dev, int64_t n,
                         double* mats_dev, void* stream) {
    BSX_REQUIRE(h && views_dev && mats_dev, BSX_ERR_ARG, "bsx_search_view_mats: null argument");
    BSX_REQUIRE(h->nviews > 0, BSX_ERR_STATE, "bsx_search_view_mats: no reference views");
    if (n <= 0) return BSX_OK;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_view_mats, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, h->d_libmats,
                       h->nviews, views_dev, n, mats_dev);
    BSX_HIP(hipGetLastError());
    return BSX_OK;
}

int bsx_search_refine(bsx_search* h, const float* ft_dev, double* mats_dev, float* frc_dev,
                      const double* steps, int32_t nsteps, void* stream) {
    BSX_REQUIRE(h && ft_dev && mats_dev && (steps || nsteps == 0), BSX_ERR_ARG,
                "bsx_search_refine: null argument");
    BSX_REQUIRE(nsteps >= 0 && nsteps <= 16, BSX_ERR_ARG, "bsx_search_refine: nsteps %d out of range",
                nsteps);
    for (int i = 0; i < nsteps; ++i)
        BSX_REQUIRE(std::isfinite(steps[i]) && steps[i] > 0 && steps[i] < 1, BSX_ERR_ARG,
                    "bsx_search_refine: step %d (%g rad) out of range (0, 1)", i, steps[i]);
    const int64_t n = h->nparts;
    if (n == 0 || nsteps == 0) return BSX_OK;
    BSX_REQUIRE(n * 27 < (1ll << 31), BSX_ERR_ARG, "bsx_search_refine: too many particles");
    int rc = grow(h, &h->d_lmats, &h->lmats_cap, n * 27 * 9, "refinement matrices");
    if (!rc) rc = grow(h, &h->d_lfrc, &h->lfrc_cap, n * 27, "refinement scores");
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    for (int i = 0; i < nsteps; ++i) {
        hipLaunchKernelGGL(k_local_mats, dim3((unsigned)((n * 27 + 255) / 256)), dim3(256), 0, s,
                           mats_dev, n, steps[i], h->d_lmats);
        BSX_HIP(hipGetLastError());
        rc = bsx_search_score_mats(h, ft_dev, h->d_lmats, 27, h->d_lfrc, stream);
        if (rc) return rc;
        hipLaunchKernelGGL(k_local_pick, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                           h->d_lmats, h->d_lfrc, n, mats_dev, frc_dev);
        BSX_HIP(hipGetLastError());
    }
    return BSX_OK;
}

int bsx_search_refine_views(bsx_search* h, const float* ft_dev, double* mats_dev,
                            float* frc_dev, double alpha_step, double accuracy, int32_t max_iter,
                            int32_t* iterations, void* stream) {
    BSX_REQUIRE(h && ft_dev && mats_dev, BSX_ERR_ARG, "bsx_search_refine_views: null argument");
    BSX_REQUIRE(std::isfinite(alpha_step) && std::isfinite(accuracy) && alpha_step > 0 &&
                    alpha_step < 1 && accuracy > 0 && alpha_step >= accuracy,
                BSX_ERR_ARG,
                "bsx_search_refine_views: step %g / accuracy %g rad invalid (0 < accuracy <= step < 1)",
                alpha_step, accuracy);
    BSX_REQUIRE(max_iter >= 1 && max_iter <= 10000, BSX_ERR_ARG,
                "bsx_search_refine_views: max_iter %d out of range 1..10000", max_iter);
    if (iterations) *iterations = 0;
    const int64_t n = h->nparts;
    if (n == 0) return BSX_OK;
    BSX_REQUIRE(n * 27 < (1ll << 31), BSX_ERR_ARG, "bsx_search_refine_views: too many particles");
    int rc = grow(h, &h->d_lmats, &h->lmats_cap, n * 27 * 9, "refinement matrices");
    if (!rc) rc = grow(h, &h->d_lfrc, &h->lfrc_cap, n * 27, "refinement scores");
    if (!rc) rc = grow(h, &h->d_rv, &h->rv_cap, n * 5, "refinement views and steps");
    if (!rc) rc = grow(h, &h->d_rvi, &h->rvi_cap, 2 * n + 1, "refinement state");
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    double* view = h->d_rv;
    double* st = h->d_rv + 4 * n;
    float* fom = (float*)(h->d_rvi + n);     // n floats beside the n flags
    int* active = h->d_rvi;
    int* left = h->d_rvi + 2 * n;
    const unsigned gp = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_rv_init, dim3(gp), dim3(256), 0, s, mats_dev, n, alpha_step, accuracy,
                       view, st, fom, active);
    BSX_HIP(hipGetLastError());
    int it = 0, remaining = 1;
    while (it < max_iter && remaining > 0) {
        hipLaunchKernelGGL(k_rv_mats, dim3(gp), dim3(256), 0, s, view, st, active, n, h->d_lmats);
        BSX_HIP(hipGetLastError());
        if (it == 0) {
            // particles that never iterate still need valid matrices to score
            hipLaunchKernelGGL(k_rv_fill, dim3(gp), dim3(256), 0, s, view, active, n, h->d_lmats);
            BSX_HIP(hipGetLastError());
        }
        rc = bsx_search_score_mats(h, ft_dev, h->d_lmats, 27, h->d_lfrc, stream);
        if (rc) return rc;
        BSX_HIP(hipMemsetAsync(left, 0, sizeof(int), s));
        hipLaunchKernelGGL(k_rv_pick, dim3(gp), dim3(256), 0, s, h->d_lfrc, n, accuracy, view, st,
                           fom, active, left);
        BSX_HIP(hipGetLastError());
        BSX_HIP(hipMemcpyAsync(&remaining, left, sizeof(int), hipMemcpyDeviceToHost, s));
        BSX_HIP(hipStreamSynchronize(s));
        ++it;
    }
    hipLaunchKernelGGL(k_rv_finish, dim3(gp), dim3(256), 0, s, view, fom, n, mats_dev, frc_dev);
    BSX_HIP(hipGetLastError());
    if (iterations) *iterations = it;
    return BSX_OK;
}

int bsx_views_for_refinement(const double view[4], double alpha_step, double* out) {
    BSX_REQUIRE(view && out, BSX_ERR_ARG, "bsx_views_for_refinement: null argument");
    BSX_REQUIRE(std::isfinite(alpha_step) && alpha_step > 0 && std::isfinite(view[0]) &&
                    std::isfinite(view[1]) && std::isfinite(view[2]) && std::isfinite(view[3]),
                BSX_ERR_ARG, "bsx_views_for_refinement: step %g or view not finite / positive",
                alpha_step);
    refine_view_grid(view, alpha_step, out);
    return BSX_OK;
}

int bsx_search_cnn_input(bsx_search* h, int32_t S, float* out_dev, void* stream) {
    BSX_REQUIRE(h && (out_dev || h->nparts == 0), BSX_ERR_ARG, "bsx_search_cnn_input: null argument");
    BSX_REQUIRE(S >= 8 && S <= 256 && S % 2 == 0 && fft_length_supported(S), BSX_ERR_ARG,
                "bsx_search_cnn_input: size %d unsupported", S);
    const int64_t n = h->nparts;
    if (n == 0) return BSX_OK;
    // grid point -> disk point table
    std::vector<int> gidx((size_t)S * S, INT_MIN);
    for (int i = 0; i < h->Nd; ++i) {
        const int hh = h->hk[i].x, kk = h->hk[i].y;
        if (2 * std::abs(hh) >= S || 2 * std::abs(kk) >= S) continue;
        gidx[(size_t)((kk + S) % S) * S + (hh + S) % S] = i;
        if (i > 0) gidx[(size_t)((-kk + S) % S) * S + (-hh + S) % S] = -i - 1;
    }
    hipStream_t s = (hipStream_t)stream;
    int* dg = nullptr;
    int64_t gcap = 0;
    int rc = grow(h, &dg, &gcap, (int64_t)S * S, "network input table");
    if (rc) return rc;
    const int64_t S2 = (int64_t)S * S;
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(n, (256ll << 20) / (8 * S2)));
    rc = grow(h, &h->d_spec, &h->spec_cap, chunk * S2, "network input grids");
    if (rc == BSX_OK &&
        hipMemcpyAsync(dg, gidx.data(), sizeof(int) * S2, hipMemcpyHostToDevice, s) != hipSuccess) {
        set_error("bsx_search_cnn_input: table upload failed");
        rc = BSX_ERR_DEVICE;
    }
    for (int64_t p0 = 0; rc == BSX_OK && p0 < n; p0 += chunk) {
        const int64_t nb = std::min(chunk, n - p0);
        hipLaunchKernelGGL(k_cnn_grid, dim3((unsigned)((nb * S2 + 255) / 256)), dim3(256), 0, s,
                           h->d_fp, h->d_ctf, h->Nd, dg, S, p0, nb, h->d_spec);
        if (hipGetLastError() != hipSuccess) { set_error("bsx_search_cnn_input: launch failed"); rc = BSX_ERR_DEVICE; break; }
        rc = fft_axis_ex(h->d_spec, nullptr, S, S, 1, nb, 0, +1, 1.0f, s);
        if (!rc) rc = fft_axis_ex(h->d_spec, nullptr, S, S, 1, nb, 1, +1, 1.0f, s);
        if (rc) break;
        hipLaunchKernelGGL(k_cnn_norm, dim3((unsigned)nb), dim3(256), 0, s, h->d_spec, S, p0, out_dev);
        if (hipGetLastError() != hipSuccess) { set_error("bsx_search_cnn_input: launch failed"); rc = BSX_ERR_DEVICE; break; }
    }
    (void)hipStreamSynchronize(s);   // the table and the grids are released / reused
    h->bytes -= sizeof(int) * gcap;
    (void)hipFree(dg);
    return rc;
}

int bsx_view_matrices(const double* views, int64_t n, double* mats) {
    BSX_REQUIRE((views && mats) || n == 0, BSX_ERR_ARG, "bsx_view_matrices: null argument");
    for (int64_t i = 0; i < n; ++i) {
        BSX_REQUIRE(std::isfinite(views[4 * i]) && std::isfinite(views[4 * i + 1]) &&
                        std::isfinite(views[4 * i + 2]) && std::isfinite(views[4 * i + 3]),
                    BSX_ERR_ARG, "bsx_view_matrices: view %lld is not finite", (long long)i);
        view_matrix(views + 4 * i, mats + 9 * i);
    }
    return BSX_OK;
}

// Matrix3::quaternion (include/Matrix3.h:129-166, the trace-based
// conversion) followed by View(Quaternion&) (include/View.h:52-63).
int bsx_matrix_views(const double* mats, int64_t n, double* views) {
    BSX_REQUIRE((mats && views) || n == 0, BSX_ERR_ARG, "bsx_matrix_views: null argument");
    for (int64_t i = 0; i < 9 * n; ++i)
        BSX_REQUIRE(std::isfinite(mats[i]), BSX_ERR_ARG, "bsx_matrix_views: matrix %lld is not finite",
                    (long long)(i / 9));
    host_parallel(n, 8192, [&](int64_t b, int64_t e, unsigned) {
        for (int64_t i = b; i < e; ++i) matrix_view(mats + 9 * i, views + 4 * i);
    });
    return BSX_OK;
}

// asymmetric_unit_views for point group C1 (src/util/symmetry.cpp:383-470)
// followed by view_list_expand_angles (src/util/View.cpp:519-556) over
// [0, 2 pi) in steps of angle_step (0 = no expansion).
int bsx_views_c1(double theta_step, double phi_step, int32_t flag, double angle_step,
                 double* views, int64_t cap, int64_t* count) {
    BSX_REQUIRE(count, BSX_ERR_ARG, "bsx_views_c1: null count");
    BSX_REQUIRE(angle_step >= 0 && std::isfinite(theta_step) && std::isfinite(phi_step), BSX_ERR_ARG,
                "bsx_views_c1: bad steps");
    if (theta_step < M_PI / 1800) theta_step = M_PI / 180;
    if (phi_step < M_PI / 1800) phi_step = M_PI / 180;
    const bool full = flag & 1, inplane = !(flag & 2);
    const double max_theta = full ? M_PI + 1e-6 : M_PI_2 + 1e-6;
    const int nphi = std::max(1, (int)(2 * M_PI / phi_step + 0.5));
    std::vector<double> base;
    auto add = [&](double x, double y, double z, double a) {
        const double r = sqrt(x * x + y * y + z * z);
        base.push_back(x / r);
        base.push_back(y / r);
        base.push_back(z / r);
        base.push_back(a);
    };
    add(0, 0, 1, 0);
    for (double th = theta_step; th <= max_theta; th += theta_step) {
        int nr = (int)(nphi * sin(th) / 2 + 0.5);
        double dphi = 0;
        int j0 = 0;
        if (nr) {
            dphi = M_PI / nr;
            j0 = -nr;
            nr -= 1;
        }
        for (int j = j0; j <= nr; ++j)
            add(sin(th) * cos(j * dphi), sin(th) * sin(j * dphi), cos(th), inplane ? j * dphi : 0);
    }
    if (full && base[base.size() - 2] > -0.999999) add(0, 0, -1, 0);
    const int64_t nb = (int64_t)base.size() / 4;
    int64_t na = 1;
    if (angle_step > 0) {
        na = 0;
        for (double a = 0; a <= 2 * M_PI - angle_step * 0.5; a += angle_step) ++na;
    }
    *count = nb * na;
    if (!views) return BSX_OK;
    int64_t o = 0;
    for (int64_t i = 0; i < nb; ++i) {
        double a = 0;
        for (int64_t q = 0; q < na; ++q, a += angle_step) {
            if (o < cap) {
                for (int c = 0; c < 3; ++c) views[4 * o + c] = base[4 * i + c];
                views[4 * o + 3] = base[4 * i + 3] + a;
            }
            ++o;
        }
    }
    return BSX_OK;
}

}  // extern "C"

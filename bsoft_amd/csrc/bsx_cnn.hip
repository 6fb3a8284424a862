// bsx_cnn.hip — the orientation-proposal network: a 5-layer CNN trained and
// run by hand-written HIP kernels (no library GEMMs).
//
// Architecture (input S x S, one channel; S a multiple of 8):
//   conv3x3 stride 1 (1 -> c1) + ReLU    S   x S
//   conv3x3 stride 2 (c1 -> c2) + ReLU   S/2 x S/2
//   conv3x3 stride 2 (c2 -> c3) + ReLU   S/4 x S/4
//   conv3x3 stride 2 (c3 -> c4) + ReLU   S/8 x S/8
//   fully connected (c4 (S/8)^2 -> nclasses), softmax cross-entropy
// Parameters are one flat FP32 array (W1 b1 W2 b2 W3 b3 W4 b4 Wfc bfc; conv
// weights [cout][cin][3][3], Wfc [class][feature]). Training: Adam on the
// FP32 master weights. Inference: BF16 copies of the weights with FP32
// accumulation and a fused per-row top-k of the softmax.
#include <cmath>
#include <cstring>
#include <algorithm>
#include <vector>

#include "bsx_common.h"

using namespace bsx;

namespace {

__device__ inline float bf16_to_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__host__ __device__ inline uint16_t f_to_bf16(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

template <typename T> __device__ inline float ldw(const T* p, long i);
template <> __device__ inline float ldw<float>(const float* p, long i) { return p[i]; }
template <> __device__ inline float ldw<uint16_t>(const uint16_t* p, long i) { return bf16_to_f(p[i]); }

// ---------------------------------------------------------------- conv fwd
// y[b][co][yo][xo] = relu(bias[co] + sum W[co][ci][ky][kx] x[b][ci][yo*s-1+ky][xo*s-1+kx]);
// one thread per (b, 4 output channels, yo, xo).
template <typename TW>
__global__ __launch_bounds__(256) void k_conv_fwd(const float* __restrict__ x, const TW* __restrict__ W,
                                                  const TW* __restrict__ bias, float* __restrict__ y,
                                                  int B, int Cin, int H, int Wd, int Cout, int Ho,
                                                  int Wo, int stride) {
    const long total = (long)B * (Cout / 4) * Ho * Wo;
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const int xo = (int)(g % Wo);
    const int yo = (int)((g / Wo) % Ho);
    const int c4 = (int)((g / ((long)Wo * Ho)) % (Cout / 4));
    const int b = (int)(g / ((long)Wo * Ho * (Cout / 4)));
    const int co = 4 * c4;
    float acc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = ldw(bias, co + q);
    const float* xb = x + (long)b * Cin * H * Wd;
    for (int ci = 0; ci < Cin; ++ci) {
        const float* xc = xb + (long)ci * H * Wd;
#pragma unroll
        for (int ky = 0; ky < 3; ++ky) {
            const int yi = yo * stride - 1 + ky;
            if (yi < 0 || yi >= H) continue;
#pragma unroll
            for (int kx = 0; kx < 3; ++kx) {
                const int xi = xo * stride - 1 + kx;
                if (xi < 0 || xi >= Wd) continue;
                const float v = xc[yi * Wd + xi];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    acc[q] = fmaf(ldw(W, (((long)(co + q) * Cin + ci) * 3 + ky) * 3 + kx), v, acc[q]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
        y[(((long)b * Cout + co + q) * Ho + yo) * Wo + xo] = fmaxf(acc[q], 0.f);
}

// g <- g * (y > 0): the ReLU backward, in place.
__global__ void k_relu_bwd(float* __restrict__ g, const float* __restrict__ y, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        if (!(y[i] > 0.f)) g[i] = 0.f;
}

// dx[b][ci][yi][xi] = sum_{co,ky,kx} W[co][ci][ky][kx] g[b][co][yo][xo],
// yo = (yi + 1 - ky) / s when divisible and in range.
__global__ __launch_bounds__(256) void k_conv_bwd_data(const float* __restrict__ g,
                                                       const float* __restrict__ W,
                                                       float* __restrict__ dx, int B, int Cin, int H,
                                                       int Wd, int Cout, int Ho, int Wo, int stride) {
    const long total = (long)B * Cin * H * Wd;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int xi = (int)(t % Wd);
    const int yi = (int)((t / Wd) % H);
    const int ci = (int)((t / ((long)Wd * H)) % Cin);
    const int b = (int)(t / ((long)Wd * H * Cin));
    float acc = 0.f;
    for (int co = 0; co < Cout; ++co) {
        const float* gc = g + ((long)b * Cout + co) * Ho * Wo;
        const float* wc = W + ((long)co * Cin + ci) * 9;
#pragma unroll
        for (int ky = 0; ky < 3; ++ky) {
            const int ny = yi + 1 - ky;
            if (ny < 0 || ny % stride) continue;
            const int yo = ny / stride;
            if (yo >= Ho) continue;
#pragma unroll
            for (int kx = 0; kx < 3; ++kx) {
                const int nx = xi + 1 - kx;
                if (nx < 0 || nx % stride) continue;
                const int xo = nx / stride;
                if (xo >= Wo) continue;
                acc = fmaf(wc[ky * 3 + kx], gc[yo * Wo + xo], acc);
            }
        }
    }
    dx[t] = acc;
}

// dW[co][ci][9] = sum_{b,yo,xo} g[b][co][yo][xo] x[b][ci][yo*s-1+ky][xo*s-1+kx];
// one workgroup per (co, ci); blockIdx.y == Cin gives db[co] = sum g.
__global__ __launch_bounds__(256) void k_conv_bwd_weight(const float* __restrict__ g,
                                                         const float* __restrict__ x,
                                                         float* __restrict__ dW, float* __restrict__ db,
                                                         int B, int Cin, int H, int Wd, int Cout,
                                                         int Ho, int Wo, int stride) {
    __shared__ float red[10][256];
    const int co = blockIdx.x, ci = blockIdx.y;
    float acc[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) acc[q] = 0.f;
    float sb = 0.f;
    const long per = (long)Ho * Wo;
    for (long e = threadIdx.x; e < (long)B * per; e += blockDim.x) {
        const int b = (int)(e / per);
        const int r = (int)(e % per);
        const int yo = r / Wo, xo = r % Wo;
        const float gv = g[((long)b * Cout + co) * per + r];
        if (ci == Cin) {
            sb += gv;
            continue;
        }
        const float* xc = x + ((long)b * Cin + ci) * H * Wd;
#pragma unroll
        for (int ky = 0; ky < 3; ++ky) {
            const int yi = yo * stride - 1 + ky;
            if (yi < 0 || yi >= H) continue;
#pragma unroll
            for (int kx = 0; kx < 3; ++kx) {
                const int xi = xo * stride - 1 + kx;
                if (xi < 0 || xi >= Wd) continue;
                acc[ky * 3 + kx] = fmaf(gv, xc[yi * Wd + xi], acc[ky * 3 + kx]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) red[q][threadIdx.x] = acc[q];
    red[9][threadIdx.x] = sb;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
#pragma unroll
            for (int q = 0; q < 10; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (ci == Cin) db[co] = red[9][0];
        else
            for (int q = 0; q < 9; ++q) dW[((long)co * Cin + ci) * 9 + q] = red[q][0];
    }
}

// ------------------------------------------------------------------ GEMM
// C[m][n] = sum_k A[m*am + k*ak] * B[k*bk + n*bn] (+ bias[n]); 64x64 tiles,
// K step 16 through LDS, 4x4 outputs per thread, edges guarded.
template <typename TB>
__global__ __launch_bounds__(256) void k_sgemm(int M, int N, int K, const float* __restrict__ A,
                                               long am, long ak, const TB* __restrict__ Bm, long bk,
                                               long bn, const TB* __restrict__ bias,
                                               float* __restrict__ Cm, long ldc) {
    __shared__ float As[16][64 + 4];
    __shared__ float Bs[16][64 + 4];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    float acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.f;
    for (int k0 = 0; k0 < K; k0 += 16) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = threadIdx.x + 256 * q;          // 0..1023
            // A tile: 64 rows x 16 k; consecutive threads walk k when ak == 1
            const int ar = ak == 1 ? e / 16 : e % 64, ac = ak == 1 ? e % 16 : e / 64;
            const int gm = m0 + ar, gk = k0 + ac;
            As[ac][ar] = (gm < M && gk < K) ? A[gm * am + gk * ak] : 0.f;
            const int br = bn == 1 ? e / 64 : e % 16, bc = bn == 1 ? e % 64 : e / 16;
            const int gk2 = k0 + br, gn = n0 + bc;
            Bs[br][bc] = (gk2 < K && gn < N) ? ldw(Bm, gk2 * bk + gn * bn) : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int gm = m0 + ty * 4 + i;
        if (gm >= M) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gn = n0 + tx * 4 + j;
            if (gn >= N) continue;
            Cm[(long)gm * ldc + gn] = acc[i][j] + (bias ? ldw(bias, gn) : 0.f);
        }
    }
}

// db[n] = sum_m g[m][n]
__global__ void k_colsum(const float* __restrict__ g, int M, int N, float* __restrict__ db) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.f;
    for (int m = 0; m < M; ++m) s += g[(long)m * N + n];
    db[n] = s;
}

// ------------------------------------------------------ softmax / top-k
// One workgroup per row: loss[b] = logsumexp - logit[label]; dl = (p - onehot) / B.
__global__ __launch_bounds__(256) void k_softmax_ce(const float* __restrict__ logits,
                                                    const int* __restrict__ labels, int B, int C,
                                                    float* __restrict__ dl, float* __restrict__ loss) {
    __shared__ float red[256];
    const int b = blockIdx.x;
    const float* l = logits + (long)b * C;
    float mx = -INFINITY;
    for (int c = threadIdx.x; c < C; c += blockDim.x) mx = fmaxf(mx, l[c]);
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    mx = red[0];
    __syncthreads();
    float se = 0.f;
    for (int c = threadIdx.x; c < C; c += blockDim.x) se += expf(l[c] - mx);
    red[threadIdx.x] = se;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    se = red[0];
    const int t = labels[b];
    const float inv = 1.f / (float)B;
    for (int c = threadIdx.x; c < C; c += blockDim.x)
        dl[(long)b * C + c] = (expf(l[c] - mx) / se - (c == t ? 1.f : 0.f)) * inv;
    if (threadIdx.x == 0 && loss) loss[b] = logf(se) + mx - l[t];
}

// One workgroup per row: the k largest logits (k <= 32) as classes and
// softmax probabilities, best first (ties: lower class).
template <int T>   // per-thread list length, >= k
__global__ __launch_bounds__(256) void k_row_topk(const float* __restrict__ logits, int C, int k,
                                                  int* __restrict__ cls, float* __restrict__ prob) {
    __shared__ float sv[256];
    __shared__ int si[256];
    __shared__ float red[256];
    const long b = blockIdx.x;
    const float* l = logits + b * C;
    float best[T];
    int bi[T];
#pragma unroll
    for (int q = 0; q < T; ++q) { best[q] = -INFINITY; bi[q] = 0x7fffffff; }
    float mx = -INFINITY;
    for (int c = threadIdx.x; c < C; c += blockDim.x) {
        const float v = l[c];
        mx = fmaxf(mx, v);
        if (v > best[T - 1]) {
            best[T - 1] = v;
            bi[T - 1] = c;
#pragma unroll
            for (int q = T - 1; q > 0; --q)
                if (best[q] > best[q - 1]) {
                    const float tv = best[q]; best[q] = best[q - 1]; best[q - 1] = tv;
                    const int ti = bi[q]; bi[q] = bi[q - 1]; bi[q - 1] = ti;
                }
        }
    }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    mx = red[0];
    __syncthreads();
    float se = 0.f;
    for (int c = threadIdx.x; c < C; c += blockDim.x) se += expf(l[c] - mx);
    red[threadIdx.x] = se;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    se = red[0];
    // k rounds of a block-wide arg-max over the heads of the per-thread lists
    int head = 0;
    for (int r = 0; r < k; ++r) {
        float hv = -INFINITY;
        int hi = 0x7fffffff;
#pragma unroll
        for (int q = 0; q < T; ++q)
            if (q == head) { hv = best[q]; hi = bi[q]; }
        sv[threadIdx.x] = hv;
        si[threadIdx.x] = hi;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (threadIdx.x < s) {
                const float ov = sv[threadIdx.x + s];
                const int oi = si[threadIdx.x + s];
                if (ov > sv[threadIdx.x] || (ov == sv[threadIdx.x] && oi < si[threadIdx.x])) {
                    sv[threadIdx.x] = ov;
                    si[threadIdx.x] = oi;
                }
            }
            __syncthreads();
        }
        const int wi = si[0];
        const float wv = sv[0];
        if (threadIdx.x == 0) {
            cls[b * k + r] = wi < C ? wi : -1;
            prob[b * k + r] = wi < C ? expf(wv - mx) / se : 0.f;
        }
        if (hi == wi && head < T) ++head;
        __syncthreads();
    }
}

// --------------------------------------------- contrastive (embedding) head
// Unit rows u = z / |z| of an n x d matrix (|z| floored at 1e-12); one
// thread per row.
__global__ void k_unit_rows(const float* __restrict__ z, long n, int d, float* __restrict__ u,
                            float* __restrict__ nrm) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* r = z + i * d;
    float ss = 0.f;
    for (int c = 0; c < d; ++c) ss = fmaf(r[c], r[c], ss);
    const float nz = fmaxf(sqrtf(ss), 1e-12f), inv = 1.f / nz;
    for (int c = 0; c < d; ++c) u[i * d + c] = r[c] * inv;
    if (nrm) nrm[i] = nz;
}

// Softmax of the pair scores S (B x B, cosines) over temperature tau, by row
// (dim 0: p_ij over j) or by column (dim 1: q_ij over i), one workgroup per
// row / column; loss[i] = -log of the matched pair's probability.
__global__ __launch_bounds__(256) void k_pair_softmax(const float* __restrict__ S, int B,
                                                      float inv_tau, int dim, float* __restrict__ P,
                                                      float* __restrict__ loss) {
    __shared__ float red[256];
    const int i = blockIdx.x;
    const long st = dim == 0 ? 1 : B, base = dim == 0 ? (long)i * B : i;
    float mx = -INFINITY;
    for (int j = threadIdx.x; j < B; j += blockDim.x) mx = fmaxf(mx, S[base + j * st] * inv_tau);
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    mx = red[0];
    __syncthreads();
    float se = 0.f;
    for (int j = threadIdx.x; j < B; j += blockDim.x) se += expf(S[base + j * st] * inv_tau - mx);
    red[threadIdx.x] = se;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    se = red[0];
    for (int j = threadIdx.x; j < B; j += blockDim.x)
        P[base + j * st] = expf(S[base + j * st] * inv_tau - mx) / se;
    if (threadIdx.x == 0) loss[i] = logf(se) + mx - S[base + (long)i * st] * inv_tau;
}

// dL/dS of the symmetric loss mean_i (row_i + col_i) / 2 over the B pairs:
// (p + q - 2 I) / (2 B tau); loss_out[i] = (row_i + col_i) / 2.
__global__ void k_pair_grad(const float* __restrict__ P, const float* __restrict__ Q, int B,
                            float inv_tau, const float* __restrict__ lrow,
                            const float* __restrict__ lcol, float* __restrict__ dS,
                            float* __restrict__ loss_out) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)B * B) return;
    const int i = (int)(t / B), j = (int)(t % B);
    dS[t] = (P[t] + Q[t] - (i == j ? 2.f : 0.f)) * inv_tau / (2.f * B);
    if (j == 0 && loss_out) loss_out[i] = 0.5f * (lrow[i] + lcol[i]);
}

// Backward of u = z / |z|: dz = (du - (du . u) u) / |z|, one thread per row.
__global__ void k_unit_rows_bwd(const float* __restrict__ du, const float* __restrict__ u,
                                const float* __restrict__ nrm, long n, int d,
                                float* __restrict__ dz) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float dot = 0.f;
    for (int c = 0; c < d; ++c) dot = fmaf(du[i * d + c], u[i * d + c], dot);
    const float inv = 1.f / nrm[i];
    for (int c = 0; c < d; ++c) dz[i * d + c] = (du[i * d + c] - dot * u[i * d + c]) * inv;
}

__global__ void k_scale(float* __restrict__ x, long n, float a) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        x[i] *= a;
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, long n, float lr, float b1, float b2, float eps,
                       float c1, float c2, float wd) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float gi = g[i] + wd * p[i];
        const float mi = b1 * m[i] + (1 - b1) * gi;
        const float vi = b2 * v[i] + (1 - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

__global__ void k_to_bf16(const float* __restrict__ p, uint16_t* __restrict__ q, long n) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        q[i] = f_to_bf16(p[i]);
}

// ------------------------------------------------------ MFMA inference
// Activations are NHWC BF16. Layer 1 (one input channel) runs on the VALU;
// layers 2-4 and the fully connected layer are implicit GEMMs on
// v_mfma_f32_32x32x16_bf16: rows = output positions (b, yo, xo), columns =
// output channels, k = (tap, input channel) with the channel fastest, so
// that 8 consecutive k are one 16-byte load from the NHWC input.
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;
constexpr int IM = 128, IN = 64, IK = 32, IROW = IK * 2 + 16;

// y[b][yo][xo][co] = bf16(relu(bias + sum_{ky,kx} W[co][0][ky][kx] x[b][yo-1+ky][xo-1+kx]))
__global__ __launch_bounds__(256) void k_conv1_nhwc(const float* __restrict__ x,
                                                    const uint16_t* __restrict__ W,
                                                    const uint16_t* __restrict__ bias, int B, int S,
                                                    int C1, uint16_t* __restrict__ y) {
    // thread = (output position, group of 8 channels): a wavefront's 16-byte
    // stores of 8 BF16 channels are contiguous in the NHWC output
    extern __shared__ float w[];   // C1 x 9 weights, then C1 biases
    for (int i = threadIdx.x; i < C1 * 9; i += blockDim.x) w[i] = bf16_to_f(W[i]);
    for (int i = threadIdx.x; i < C1; i += blockDim.x) w[C1 * 9 + i] = bf16_to_f(bias[i]);
    __syncthreads();
    const int ng = C1 / 8;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)B * S * S * ng) return;
    const int g = (int)(t % ng);
    const long pos = t / ng;
    const int xo = (int)(pos % S), yo = (int)((pos / S) % S);
    const long b = pos / ((long)S * S);
    float v[9];
#pragma unroll
    for (int ky = 0; ky < 3; ++ky)
#pragma unroll
        for (int kx = 0; kx < 3; ++kx) {
            const int yi = yo - 1 + ky, xi = xo - 1 + kx;
            v[ky * 3 + kx] = (yi >= 0 && yi < S && xi >= 0 && xi < S) ? x[(b * S + yi) * S + xi] : 0.f;
        }
    uint32_t packed[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int c = g * 8 + 2 * j;
        float a0 = w[C1 * 9 + c], a1 = w[C1 * 9 + c + 1];
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            a0 = fmaf(w[c * 9 + q], v[q], a0);
            a1 = fmaf(w[(c + 1) * 9 + q], v[q], a1);
        }
        packed[j] = (uint32_t)f_to_bf16(fmaxf(a0, 0.f)) |
                    ((uint32_t)f_to_bf16(fmaxf(a1, 0.f)) << 16);
    }
    *(uint4*)(y + pos * C1 + g * 8) = make_uint4(packed[0], packed[1], packed[2], packed[3]);
}

// Implicit-GEMM convolution (ks = 3: pad 1; ks = 1: a plain GEMM over
// H = W = 1) of NHWC BF16 input with BF16 weights [Cout][ks*ks][Cin] and
// BF16 biases. OUTF32: FP32 output rows of stride ldo without ReLU (logits);
// else BF16 NHWC output with ReLU. Cin must be a multiple of 32.
template <bool OUTF32>
__global__ __launch_bounds__(256) void k_conv_mfma(const uint16_t* __restrict__ x, int B, int H,
                                                   int Wd, int Cin, int ks, int stride, int Ho,
                                                   int Wo, const uint16_t* __restrict__ Wt,
                                                   const uint16_t* __restrict__ bias, int Cout,
                                                   long ldo, void* __restrict__ out) {
    __shared__ __attribute__((aligned(16))) char lds[2 * (IM + IN) * IROW];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const long M = (long)B * Ho * Wo;
    const int K = ks * ks * Cin;
    const int nk = K / IK;
    const long m0 = (long)blockIdx.x * IM;
    const int n0 = blockIdx.y * IN;
    const int pad = ks / 2;
    // A staging: chunks tid and tid + 256 (row = c >> 2, part = c & 3); the
    // positions live in named scalars (arrays captured by the loaders end up
    // in scratch memory)
    const long am0 = m0 + (tid >> 2), am1 = m0 + ((tid + 256) >> 2);
    const long mm0 = am0 < M ? am0 : 0, mm1 = am1 < M ? am1 : 0;
    const int ax0 = (int)(mm0 % Wo), ay0 = (int)((mm0 / Wo) % Ho), ab0 = (int)(mm0 / ((long)Wo * Ho));
    const int ax1 = (int)(mm1 % Wo), ay1 = (int)((mm1 / Wo) % Ho), ab1 = (int)(mm1 / ((long)Wo * Ho));
    const int apart0 = tid & 3;
    const int brow = tid >> 2, bpart = tid & 3;
    const int bco = n0 + brow;
    f32x16 acc0, acc1;
#pragma unroll
    for (int r = 0; r < 16; ++r) { acc0[r] = 0.f; acc1[r] = 0.f; }
    uint4 ra0, ra1, rb;
#define BSX_CGLOAD(KT)                                                                        \
    {                                                                                          \
        const int k0 = (KT) * IK;                                                              \
        const int tap = k0 / Cin;                                                              \
        const int ky = tap / ks, kx = tap % ks;                                                \
        const int ci0 = k0 % Cin + apart0 * 8;                                                 \
        const int yi0 = ay0 * stride - pad + ky, xi0 = ax0 * stride - pad + kx;                \
        const int yi1 = ay1 * stride - pad + ky, xi1 = ax1 * stride - pad + kx;                \
        ra0 = make_uint4(0, 0, 0, 0);                                                          \
        ra1 = make_uint4(0, 0, 0, 0);                                                          \
        rb = make_uint4(0, 0, 0, 0);                                                           \
        if (am0 < M && yi0 >= 0 && yi0 < H && xi0 >= 0 && xi0 < Wd)                            \
            ra0 = *(const uint4*)(x + (((long)ab0 * H + yi0) * Wd + xi0) * Cin + ci0);         \
        if (am1 < M && yi1 >= 0 && yi1 < H && xi1 >= 0 && xi1 < Wd)                            \
            ra1 = *(const uint4*)(x + (((long)ab1 * H + yi1) * Wd + xi1) * Cin + ci0);         \
        if (bco < Cout) rb = *(const uint4*)(Wt + (long)bco * K + k0 + bpart * 8);             \
    }
#define BSX_CLSTORE(BUF)                                                                      \
    {                                                                                          \
        *(uint4*)((BUF) + (tid >> 2) * IROW + apart0 * 16) = ra0;                              \
        *(uint4*)((BUF) + ((tid + 256) >> 2) * IROW + apart0 * 16) = ra1;                      \
        *(uint4*)((BUF) + (IM + brow) * IROW + bpart * 16) = rb;                               \
    }
    BSX_CGLOAD(0)
    BSX_CLSTORE(lds)
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const bool more = kt + 1 < nk;
        if (more) BSX_CGLOAD(kt + 1)
        const char* buf = lds + (kt & 1) * (IM + IN) * IROW;
#pragma unroll
        for (int sub = 0; sub < IK / 16; ++sub) {
            const int off = sub * 32 + (lane >> 5) * 16;
            const bf16x8 fa = *(const bf16x8*)(buf + (wave * 32 + (lane & 31)) * IROW + off);
            const bf16x8 fb0 = *(const bf16x8*)(buf + (IM + (lane & 31)) * IROW + off);
            const bf16x8 fb1 = *(const bf16x8*)(buf + (IM + 32 + (lane & 31)) * IROW + off);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa, fb0, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa, fb1, acc1, 0, 0, 0);
        }
        if (more) BSX_CLSTORE(lds + ((kt + 1) & 1) * (IM + IN) * IROW)
        __syncthreads();
    }
#undef BSX_CGLOAD
#undef BSX_CLSTORE
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const long m = m0 + wave * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            const int co = n0 + j * 32 + (lane & 31);
            if (m >= M || co >= Cout) continue;
            const float v = (j == 0 ? acc0[r] : acc1[r]) + bf16_to_f(bias[co]);
            if (OUTF32) ((float*)out)[m * ldo + co] = v;
            else ((uint16_t*)out)[m * Cout + co] = f_to_bf16(fmaxf(v, 0.f));
        }
}

// BF16 inference weights: conv [co][ci][ky][kx] -> [co][ky*3+kx][ci]; FC
// [class][c*hw + pos] -> [class][pos*C + c] (NHWC flattening).
__global__ void k_perm_conv(const float* __restrict__ W, int Cout, int Cin, uint16_t* __restrict__ o) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)Cout * Cin * 9) return;
    const int tap = (int)(t % 9);
    const int ci = (int)((t / 9) % Cin);
    const long co = t / (9l * Cin);
    o[(co * 9 + tap) * Cin + ci] = f_to_bf16(W[t]);
}

__global__ void k_perm_fc(const float* __restrict__ W, int ncls, int C, int hw, uint16_t* __restrict__ o) {
    const long F = (long)C * hw;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)ncls * F) return;
    const long cls = t / F;
    const int f = (int)(t % F);
    const int c = f / hw, pos = f % hw;
    o[cls * F + (long)pos * C + c] = f_to_bf16(W[t]);
}

unsigned nblk(long n, int t = 256) { return (unsigned)std::min<long>((n + t - 1) / t, 1l << 30); }

}  // namespace

// ================================================================ host
struct bsx_cnn {
    int S = 0, ch[5] = {1, 0, 0, 0, 0}, nclass = 0, maxb = 0;
    int hw[5];                 // spatial edge per activation level
    long woff[5], boff[5];     // offsets of W/b of layers 1..4 and the FC (index 4)
    long nparam = 0, feat = 0;
    float* d_p = nullptr;      // parameters
    float* d_g = nullptr;      // gradients
    float* d_m = nullptr;
    float* d_v = nullptr;
    uint16_t* d_pb = nullptr;  // BF16 copy for inference
    bool pb_valid = false;
    float* d_act[5] = {};      // activations 1..4 (index 1..4), logits at [0]
    float* d_gact[5] = {};
    float* d_loss = nullptr;
    long step = 0;
    // MFMA inference: permuted BF16 weights of layers 2-4 and the FC; NHWC
    // BF16 activations of layers 1-4 and FP32 logits, sized for chunks larger
    // than the training batch (a 512-particle chunk leaves most of the chip
    // idle in layers 3-4 and the FC)
    uint16_t* d_wi[5] = {};
    int icap = 0;
    uint16_t* d_ib[5] = {};
    float* d_ilog = nullptr;
    // contrastive training scratch (pairs of up to maxb / 2): unit rows,
    // norms, pair scores and their softmaxes, gradients
    float* d_nce = nullptr;
    long nce_cap = 0;
    float* d_match = nullptr;    // embedding top-k: scores of a particle chunk
    long match_cap = 0;
};

namespace {

void free_cnn(bsx_cnn* h) {
    void* ptrs[] = {h->d_p, h->d_g, h->d_m, h->d_v, h->d_pb, h->d_loss};
    for (void* q : ptrs)
        if (q) (void)hipFree(q);
    for (int i = 0; i < 5; ++i) {
        if (h->d_act[i]) (void)hipFree(h->d_act[i]);
        if (h->d_gact[i]) (void)hipFree(h->d_gact[i]);
        if (h->d_wi[i]) (void)hipFree(h->d_wi[i]);
        if (h->d_ib[i]) (void)hipFree(h->d_ib[i]);
    }
    if (h->d_ilog) (void)hipFree(h->d_ilog);
    if (h->d_nce) (void)hipFree(h->d_nce);
    if (h->d_match) (void)hipFree(h->d_match);
}

long act_size(const bsx_cnn* h, int l, int B) {
    if (l == 0) return (long)B * h->nclass;
    return (long)B * h->ch[l] * h->hw[l] * h->hw[l];
}

// splitmix64 + Box-Muller: reproducible He-normal initialisation
struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal() { return sqrt(-2 * log(uniform())) * cos(2 * M_PI * uniform()); }
};

template <typename TW>
int forward(bsx_cnn* h, const float* x, int B, const TW* P, hipStream_t s) {
    const float* in = x;
    for (int l = 1; l <= 4; ++l) {
        const int Cin = h->ch[l - 1], Cout = h->ch[l], H = h->hw[l - 1], Ho = h->hw[l];
        const int stride = l == 1 ? 1 : 2;
        const long tot = (long)B * (Cout / 4) * Ho * Ho;
        hipLaunchKernelGGL((k_conv_fwd<TW>), dim3(nblk(tot)), dim3(256), 0, s, in, P + h->woff[l - 1],
                           P + h->boff[l - 1], h->d_act[l], B, Cin, H, H, Cout, Ho, Ho, stride);
        BSX_HIP(hipGetLastError());
        in = h->d_act[l];
    }
    const int F = (int)h->feat, C = h->nclass;
    dim3 g((C + 63) / 64, (B + 63) / 64);
    hipLaunchKernelGGL((k_sgemm<TW>), g, dim3(256), 0, s, B, C, F, h->d_act[4], (long)F, 1l,
                       P + h->woff[4], 1l, (long)F, P + h->boff[4], h->d_act[0], (long)C);
    BSX_HIP(hipGetLastError());
    return BSX_OK;
}

// Backpropagation from dL/dlogits in d_gact[0] (B rows) to the parameter
// gradients in d_g.
int backward_body(bsx_cnn* h, const float* x, int B, hipStream_t s) {
    const int F = (int)h->feat, C = h->nclass;
    float* P = h->d_p;
    float* G = h->d_g;
    // FC: dW = dl^T x, db = colsum(dl), dx = dl W
    {
        dim3 g((F + 63) / 64, (C + 63) / 64);
        hipLaunchKernelGGL((k_sgemm<float>), g, dim3(256), 0, s, C, F, B, h->d_gact[0], 1l, (long)C,
                           h->d_act[4], (long)F, 1l, (const float*)nullptr, G + h->woff[4], (long)F);
        BSX_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_colsum, dim3((C + 255) / 256), dim3(256), 0, s, h->d_gact[0], B, C,
                           G + h->boff[4]);
        BSX_HIP(hipGetLastError());
        dim3 g2((F + 63) / 64, (B + 63) / 64);
        hipLaunchKernelGGL((k_sgemm<float>), g2, dim3(256), 0, s, B, F, C, h->d_gact[0], (long)C, 1l,
                           P + h->woff[4], (long)F, 1l, (const float*)nullptr, h->d_gact[4], (long)F);
        BSX_HIP(hipGetLastError());
    }
    for (int l = 4; l >= 1; --l) {
        const int Cin = h->ch[l - 1], Cout = h->ch[l], H = h->hw[l - 1], Ho = h->hw[l];
        const int stride = l == 1 ? 1 : 2;
        const long na = act_size(h, l, B);
        hipLaunchKernelGGL(k_relu_bwd, dim3(nblk(na)), dim3(256), 0, s, h->d_gact[l], h->d_act[l], na);
        BSX_HIP(hipGetLastError());
        const float* in = l == 1 ? x : h->d_act[l - 1];
        hipLaunchKernelGGL(k_conv_bwd_weight, dim3(Cout, Cin + 1), dim3(256), 0, s, h->d_gact[l], in,
                           G + h->woff[l - 1], G + h->boff[l - 1], B, Cin, H, H, Cout, Ho, Ho, stride);
        BSX_HIP(hipGetLastError());
        if (l > 1) {
            const long nx = (long)B * Cin * H * H;
            hipLaunchKernelGGL(k_conv_bwd_data, dim3(nblk(nx)), dim3(256), 0, s, h->d_gact[l],
                               P + h->woff[l - 1], h->d_gact[l - 1], B, Cin, H, H, Cout, Ho, Ho, stride);
            BSX_HIP(hipGetLastError());
        }
    }
    return BSX_OK;
}


int backward(bsx_cnn* h, const float* x, const int* labels, int B, hipStream_t s) {
    hipLaunchKernelGGL(k_softmax_ce, dim3(B), dim3(256), 0, s, h->d_act[0], labels, B, h->nclass,
                       h->d_gact[0], h->d_loss);
    BSX_HIP(hipGetLastError());
    return backward_body(h, x, B, s);
}

// Network outputs of n inputs in chunks (BF16 MFMA layers when the channel
// counts allow, else the FP32-accumulating BF16 path): calls
// use(logits, first, count) for each chunk, on stream s.
template <class Use>
int infer_chunks(bsx_cnn* h, const float* x_dev, int64_t n, hipStream_t s, const char* who, Use use) {
    // the MFMA layers need input channel counts (c1..c3 and the FC input) in multiples of 32
    const bool mfma = h->ch[1] % 32 == 0 && h->ch[2] % 32 == 0 && h->ch[3] % 32 == 0 &&
                      h->feat % 32 == 0;
    if (!h->pb_valid) {
        hipLaunchKernelGGL(k_to_bf16, dim3(nblk(h->nparam)), dim3(256), 0, s, h->d_p, h->d_pb, h->nparam);
        BSX_HIP(hipGetLastError());
        if (mfma) {
            for (int l = 1; l < 4; ++l) {
                const long nw = (long)h->ch[l + 1] * h->ch[l] * 9;
                hipLaunchKernelGGL(k_perm_conv, dim3(nblk(nw)), dim3(256), 0, s, h->d_p + h->woff[l],
                                   h->ch[l + 1], h->ch[l], h->d_wi[l]);
                BSX_HIP(hipGetLastError());
            }
            const long nf = h->feat * h->nclass;
            hipLaunchKernelGGL(k_perm_fc, dim3(nblk(nf)), dim3(256), 0, s, h->d_p + h->woff[4],
                               h->nclass, h->ch[4], h->hw[4] * h->hw[4], h->d_wi[4]);
            BSX_HIP(hipGetLastError());
        }
        h->pb_valid = true;
    }
    const long in = (long)h->S * h->S;
    // MFMA path: chunks of up to 4096 particles in the inference buffers
    int chunk = h->maxb;
    if (mfma) {
        chunk = (int)std::min<int64_t>(std::max<int64_t>(n, 1), 4096);
        chunk = std::max(chunk, std::min(h->maxb, 4096));
        if (chunk > h->icap) {
            for (int l = 1; l < 5; ++l) {
                if (h->d_ib[l]) (void)hipFree(h->d_ib[l]);
                h->d_ib[l] = nullptr;
            }
            if (h->d_ilog) (void)hipFree(h->d_ilog);
            h->d_ilog = nullptr;
            h->icap = 0;
            for (int l = 1; l < 5; ++l)
                if (hipMalloc(&h->d_ib[l], sizeof(uint16_t) * act_size(h, l, chunk)) != hipSuccess) {
                    set_error("%s: inference activations (%d particles)", who, chunk);
                    return BSX_ERR_NOMEM;
                }
            if (hipMalloc(&h->d_ilog, sizeof(float) * (size_t)chunk * h->nclass) != hipSuccess) {
                set_error("%s: logits (%d particles)", who, chunk);
                return BSX_ERR_NOMEM;
            }
            h->icap = chunk;
        }
        chunk = h->icap;
    }
    for (int64_t b0 = 0; b0 < n; b0 += chunk) {
        const int nb = (int)std::min<int64_t>(chunk, n - b0);
        if (!mfma) {
            int rc = forward<uint16_t>(h, x_dev + b0 * in, nb, h->d_pb, s);
            if (rc) return rc;
        } else {
            const uint16_t* P = h->d_pb;
            const long pos1 = (long)nb * in * (h->ch[1] / 8);
            hipLaunchKernelGGL(k_conv1_nhwc, dim3(nblk(pos1)), dim3(256),
                               sizeof(float) * 10 * h->ch[1], s, x_dev + b0 * in, P + h->woff[0],
                               P + h->boff[0], nb, h->S, h->ch[1], h->d_ib[1]);
            BSX_HIP(hipGetLastError());
            for (int l = 2; l <= 4; ++l) {
                const long M = (long)nb * h->hw[l] * h->hw[l];
                dim3 g((unsigned)((M + IM - 1) / IM), (unsigned)((h->ch[l] + IN - 1) / IN));
                hipLaunchKernelGGL((k_conv_mfma<false>), g, dim3(256), 0, s, h->d_ib[l - 1], nb,
                                   h->hw[l - 1], h->hw[l - 1], h->ch[l - 1], 3, 2, h->hw[l], h->hw[l],
                                   h->d_wi[l - 1], P + h->boff[l - 1], h->ch[l], (long)h->ch[l],
                                   (void*)h->d_ib[l]);
                BSX_HIP(hipGetLastError());
            }
            dim3 g((unsigned)((nb + IM - 1) / IM), (unsigned)((h->nclass + IN - 1) / IN));
            hipLaunchKernelGGL((k_conv_mfma<true>), g, dim3(256), 0, s, h->d_ib[4], nb, 1, 1,
                               (int)h->feat, 1, 1, 1, 1, h->d_wi[4], P + h->boff[4], h->nclass,
                               (long)h->nclass, (void*)h->d_ilog);
            BSX_HIP(hipGetLastError());
        }
        const int rc = use(mfma ? h->d_ilog : h->d_act[0], b0, nb);
        if (rc) return rc;
    }
    return BSX_OK;
}

}  // namespace

extern "C" {

int bsx_cnn_create(bsx_cnn** out, int32_t in_size, const int32_t* channels, int32_t nclasses,
                   int32_t max_batch, uint64_t seed) {
    BSX_REQUIRE(out && channels, BSX_ERR_ARG, "bsx_cnn_create: null argument");
    *out = nullptr;
    BSX_REQUIRE(in_size >= 8 && in_size <= 256 && in_size % 8 == 0, BSX_ERR_ARG,
                "bsx_cnn_create: input size %d must be a multiple of 8 in 8..256", in_size);
    for (int i = 0; i < 4; ++i)
        BSX_REQUIRE(channels[i] >= 4 && channels[i] <= 1024 && channels[i] % 4 == 0, BSX_ERR_ARG,
                    "bsx_cnn_create: channel count %d must be a multiple of 4 in 4..1024", channels[i]);
    BSX_REQUIRE(channels[0] <= 1024 && channels[0] % 2 == 0, BSX_ERR_ARG,
                "bsx_cnn_create: first-layer channel count must be even");
    BSX_REQUIRE(nclasses >= 2 && nclasses <= (1 << 20), BSX_ERR_ARG,
                "bsx_cnn_create: %d classes out of range", nclasses);
    BSX_REQUIRE(max_batch >= 1 && max_batch <= 65536, BSX_ERR_ARG,
                "bsx_cnn_create: max_batch %d out of range", max_batch);
    bsx_cnn* h = new bsx_cnn();
    h->S = in_size;
    for (int i = 0; i < 4; ++i) h->ch[i + 1] = channels[i];
    h->nclass = nclasses;
    h->maxb = max_batch;
    h->hw[0] = h->hw[1] = in_size;
    h->hw[2] = in_size / 2;
    h->hw[3] = in_size / 4;
    h->hw[4] = in_size / 8;
    h->feat = (long)h->ch[4] * h->hw[4] * h->hw[4];
    long o = 0;
    for (int l = 0; l < 4; ++l) {
        h->woff[l] = o;
        o += (long)h->ch[l + 1] * h->ch[l] * 9;
        h->boff[l] = o;
        o += h->ch[l + 1];
    }
    h->woff[4] = o;
    o += h->feat * nclasses;
    h->boff[4] = o;
    o += nclasses;
    h->nparam = o;
    // He-normal initialisation, zero biases
    std::vector<float> p((size_t)o, 0.f);
    Rng r{seed ? seed : 0x5eedull};
    for (int l = 0; l < 5; ++l) {
        const long fan_in = l < 4 ? (long)h->ch[l] * 9 : h->feat;
        const long nw = l < 4 ? (long)h->ch[l + 1] * h->ch[l] * 9 : h->feat * nclasses;
        const double sd = sqrt(2.0 / fan_in) * (l == 4 ? 0.5 : 1.0);
        for (long i = 0; i < nw; ++i) p[h->woff[l] + i] = (float)(sd * r.normal());
    }
    auto fail = [&](hipError_t e, const char* what) {
        set_error("bsx_cnn_create: %s: %s", what, hipGetErrorString(e));
        free_cnn(h);
        delete h;
        return BSX_ERR_NOMEM;
    };
    hipError_t e;
    if ((e = hipMalloc(&h->d_p, sizeof(float) * o)) != hipSuccess) return fail(e, "parameters");
    if ((e = hipMalloc(&h->d_g, sizeof(float) * o)) != hipSuccess) return fail(e, "gradients");
    if ((e = hipMalloc(&h->d_m, sizeof(float) * o)) != hipSuccess) return fail(e, "moments");
    if ((e = hipMalloc(&h->d_v, sizeof(float) * o)) != hipSuccess) return fail(e, "moments");
    if ((e = hipMalloc(&h->d_pb, sizeof(uint16_t) * o)) != hipSuccess) return fail(e, "bf16 copy");
    if ((e = hipMalloc(&h->d_loss, sizeof(float) * max_batch)) != hipSuccess) return fail(e, "loss");
    for (int l = 0; l < 5; ++l) {
        const long n = act_size(h, l, max_batch);
        if ((e = hipMalloc(&h->d_act[l], sizeof(float) * n)) != hipSuccess) return fail(e, "activations");
        if ((e = hipMalloc(&h->d_gact[l], sizeof(float) * n)) != hipSuccess) return fail(e, "activations");
    }
    for (int l = 1; l < 4; ++l)
        if ((e = hipMalloc(&h->d_wi[l], sizeof(uint16_t) * h->ch[l + 1] * h->ch[l] * 9)) != hipSuccess)
            return fail(e, "inference weights");
    if ((e = hipMalloc(&h->d_wi[4], sizeof(uint16_t) * h->feat * nclasses)) != hipSuccess)
        return fail(e, "inference weights");
    if ((e = hipMemcpy(h->d_p, p.data(), sizeof(float) * o, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemset(h->d_m, 0, sizeof(float) * o)) != hipSuccess ||
        (e = hipMemset(h->d_v, 0, sizeof(float) * o)) != hipSuccess ||
        (e = hipMemset(h->d_g, 0, sizeof(float) * o)) != hipSuccess)
        return fail(e, "initialisation");
    *out = h;
    return BSX_OK;
}

int bsx_cnn_destroy(bsx_cnn* h) {
    if (!h) return BSX_OK;
    free_cnn(h);
    delete h;
    return BSX_OK;
}

int bsx_cnn_nparams(const bsx_cnn* h, int64_t* n, int64_t* feature_size) {
    BSX_REQUIRE(h, BSX_ERR_ARG, "bsx_cnn_nparams: null handle");
    if (n) *n = h->nparam;
    if (feature_size) *feature_size = h->feat;
    return BSX_OK;
}

int bsx_cnn_params(bsx_cnn* h, float* get_dev, const float* set_dev, void* stream) {
    BSX_REQUIRE(h, BSX_ERR_ARG, "bsx_cnn_params: null handle");
    hipStream_t s = (hipStream_t)stream;
    if (get_dev)
        BSX_HIP(hipMemcpyAsync(get_dev, h->d_p, sizeof(float) * h->nparam, hipMemcpyDeviceToDevice, s));
    if (set_dev) {
        BSX_HIP(hipMemcpyAsync(h->d_p, set_dev, sizeof(float) * h->nparam, hipMemcpyDeviceToDevice, s));
        h->pb_valid = false;
    }
    return BSX_OK;
}

int bsx_cnn_forward(bsx_cnn* h, const float* x_dev, int32_t batch, float* logits_dev, void* stream) {
    BSX_REQUIRE(h && x_dev, BSX_ERR_ARG, "bsx_cnn_forward: null argument");
    BSX_REQUIRE(batch >= 1 && batch <= h->maxb, BSX_ERR_ARG, "bsx_cnn_forward: batch %d out of 1..%d",
                batch, h->maxb);
    hipStream_t s = (hipStream_t)stream;
    int rc = forward<float>(h, x_dev, batch, h->d_p, s);
    if (rc) return rc;
    if (logits_dev)
        BSX_HIP(hipMemcpyAsync(logits_dev, h->d_act[0], sizeof(float) * batch * h->nclass,
                               hipMemcpyDeviceToDevice, s));
    return BSX_OK;
}

int bsx_cnn_gradients(bsx_cnn* h, const float* x_dev, const int32_t* labels_dev, int32_t batch,
                      float* grad_dev, float* loss_dev, void* stream) {
    BSX_REQUIRE(h && x_dev && labels_dev, BSX_ERR_ARG, "bsx_cnn_gradients: null argument");
    BSX_REQUIRE(batch >= 1 && batch <= h->maxb, BSX_ERR_ARG,
                "bsx_cnn_gradients: batch %d out of 1..%d", batch, h->maxb);
    hipStream_t s = (hipStream_t)stream;
    int rc = forward<float>(h, x_dev, batch, h->d_p, s);
    if (!rc) rc = backward(h, x_dev, labels_dev, batch, s);
    if (rc) return rc;
    if (grad_dev)
        BSX_HIP(hipMemcpyAsync(grad_dev, h->d_g, sizeof(float) * h->nparam, hipMemcpyDeviceToDevice, s));
    if (loss_dev)
        BSX_HIP(hipMemcpyAsync(loss_dev, h->d_loss, sizeof(float) * batch, hipMemcpyDeviceToDevice, s));
    return BSX_OK;
}

int bsx_cnn_train_step(bsx_cnn* h, const float* x_dev, const int32_t* labels_dev, int32_t batch,
                       float lr, float weight_decay, float* loss_dev, void* stream) {
    BSX_REQUIRE(lr > 0 && std::isfinite(lr) && weight_decay >= 0, BSX_ERR_ARG,
                "bsx_cnn_train_step: bad learning rate / weight decay");
    int rc = bsx_cnn_gradients(h, x_dev, labels_dev, batch, nullptr, loss_dev, stream);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    h->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)h->step), c2 = 1.f - powf(b2, (float)h->step);
    hipLaunchKernelGGL(k_adam, dim3(nblk(h->nparam)), dim3(256), 0, s, h->d_p, h->d_g, h->d_m, h->d_v,
                       h->nparam, lr, b1, b2, 1e-8f, c1, c2, weight_decay);
    BSX_HIP(hipGetLastError());
    h->pb_valid = false;
    return BSX_OK;
}

int bsx_cnn_infer_topk(bsx_cnn* h, const float* x_dev, int64_t n, int32_t k, int32_t* cls_dev,
                       float* prob_dev, void* stream) {
    BSX_REQUIRE(h && (n == 0 || (x_dev && cls_dev && prob_dev)), BSX_ERR_ARG,
                "bsx_cnn_infer_topk: null argument");
    BSX_REQUIRE(k >= 1 && k <= 32 && k <= h->nclass, BSX_ERR_ARG, "bsx_cnn_infer_topk: k %d out of range", k);
    hipStream_t s = (hipStream_t)stream;
    return infer_chunks(h, x_dev, n, s, "bsx_cnn_infer_topk", [&](const float* lg, int64_t b0, int nb) {
        if (k <= 16)
            hipLaunchKernelGGL(k_row_topk<16>, dim3(nb), dim3(256), 0, s, lg, h->nclass, k,
                               cls_dev + b0 * k, prob_dev + b0 * k);
        else
            hipLaunchKernelGGL(k_row_topk<32>, dim3(nb), dim3(256), 0, s, lg, h->nclass, k,
                               cls_dev + b0 * k, prob_dev + b0 * k);
        BSX_HIP(hipGetLastError());
        return BSX_OK;
    });
}

int bsx_cnn_embed(bsx_cnn* h, const float* x_dev, int64_t n, float* emb_dev, void* stream) {
    BSX_REQUIRE(h && (n == 0 || (x_dev && emb_dev)), BSX_ERR_ARG, "bsx_cnn_embed: null argument");
    hipStream_t s = (hipStream_t)stream;
    const int d = h->nclass;
    return infer_chunks(h, x_dev, n, s, "bsx_cnn_embed", [&](const float* lg, int64_t b0, int nb) {
        hipLaunchKernelGGL(k_unit_rows, dim3(nblk(nb)), dim3(256), 0, s, lg, (long)nb, d,
                           emb_dev + b0 * d, (float*)nullptr);
        BSX_HIP(hipGetLastError());
        return BSX_OK;
    });
}

int bsx_cnn_match_topk(bsx_cnn* h, const float* p_dev, int64_t n, const float* r_dev, int32_t m,
                       int32_t k, float tau, int32_t* cls_dev, float* prob_dev, void* stream) {
    BSX_REQUIRE(h && (n == 0 || (p_dev && r_dev && cls_dev && prob_dev)), BSX_ERR_ARG,
                "bsx_cnn_match_topk: null argument");
    BSX_REQUIRE(m >= 1 && k >= 1 && k <= 32 && k <= m, BSX_ERR_ARG,
                "bsx_cnn_match_topk: k %d / references %d out of range", k, m);
    BSX_REQUIRE(tau > 0 && std::isfinite(tau), BSX_ERR_ARG, "bsx_cnn_match_topk: bad temperature");
    if (n == 0) return BSX_OK;
    hipStream_t s = (hipStream_t)stream;
    const int d = h->nclass;
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(n, (256ll << 20) / (4ll * m)));
    if (chunk * m > h->match_cap) {
        if (h->d_match) (void)hipFree(h->d_match);
        h->d_match = nullptr;
        h->match_cap = 0;
        BSX_HIP(hipMalloc(&h->d_match, sizeof(float) * chunk * m));
        h->match_cap = chunk * m;
    }
    for (int64_t b0 = 0; b0 < n; b0 += chunk) {
        const int nb = (int)std::min<int64_t>(chunk, n - b0);
        // scores = P R^T / tau: the rows' softmax over the references
        dim3 g((m + 63) / 64, (nb + 63) / 64);
        hipLaunchKernelGGL((k_sgemm<float>), g, dim3(256), 0, s, nb, m, d, p_dev + b0 * d, (long)d,
                           1l, r_dev, 1l, (long)d, (const float*)nullptr, h->d_match, (long)m);
        BSX_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_scale, dim3(nblk((long)nb * m)), dim3(256), 0, s, h->d_match,
                           (long)nb * m, 1.f / tau);
        BSX_HIP(hipGetLastError());
        if (k <= 16)
            hipLaunchKernelGGL(k_row_topk<16>, dim3(nb), dim3(256), 0, s, h->d_match, m, k,
                               cls_dev + b0 * k, prob_dev + b0 * k);
        else
            hipLaunchKernelGGL(k_row_topk<32>, dim3(nb), dim3(256), 0, s, h->d_match, m, k,
                               cls_dev + b0 * k, prob_dev + b0 * k);
        BSX_HIP(hipGetLastError());
    }
    return BSX_OK;
}

int bsx_cnn_pair_gradients(bsx_cnn* h, const float* x_dev, int32_t batch, float tau,
                           float* grad_dev, float* loss_dev, void* stream) {
    BSX_REQUIRE(h && x_dev, BSX_ERR_ARG, "bsx_cnn_pair_gradients: null argument");
    BSX_REQUIRE(batch >= 2 && 2 * batch <= h->maxb, BSX_ERR_ARG,
                "bsx_cnn_pair_gradients: batch %d out of 2..%d (pairs of the network's batch)",
                batch, h->maxb / 2);
    BSX_REQUIRE(tau > 0 && std::isfinite(tau), BSX_ERR_ARG,
                "bsx_cnn_pair_gradients: bad temperature");
    hipStream_t s = (hipStream_t)stream;
    const int B = batch, B2 = 2 * batch, d = h->nclass;
    // scratch: U (2B d), dU (2B d), norms (2B), S, P, Q, dS (B^2 each), losses (2B)
    const long need = 2l * B2 * d + 2l * B2 + 4l * B * B;
    if (need > h->nce_cap) {
        if (h->d_nce) (void)hipFree(h->d_nce);
        h->d_nce = nullptr;
        h->nce_cap = 0;
        BSX_HIP(hipMalloc(&h->d_nce, sizeof(float) * need));
        h->nce_cap = need;
    }
    float* U = h->d_nce;
    float* dU = U + (long)B2 * d;
    float* nrm = dU + (long)B2 * d;
    float* lrl = nrm + B2;                // row losses (B), column losses (B)
    float* S = lrl + B2;
    float* P = S + (long)B * B;
    float* Q = P + (long)B * B;
    float* dS = Q + (long)B * B;
    int rc = forward<float>(h, x_dev, B2, h->d_p, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_unit_rows, dim3(nblk(B2)), dim3(256), 0, s, h->d_act[0], (long)B2, d, U, nrm);
    BSX_HIP(hipGetLastError());
    const float* Up = U;
    const float* Ur = U + (long)B * d;
    // S = Up Ur^T (cosines)
    hipLaunchKernelGGL((k_sgemm<float>), dim3((B + 63) / 64, (B + 63) / 64), dim3(256), 0, s, B, B, d,
                       Up, (long)d, 1l, Ur, 1l, (long)d, (const float*)nullptr, S, (long)B);
    BSX_HIP(hipGetLastError());
    const float it = 1.f / tau;
    hipLaunchKernelGGL(k_pair_softmax, dim3(B), dim3(256), 0, s, S, B, it, 0, P, lrl);
    hipLaunchKernelGGL(k_pair_softmax, dim3(B), dim3(256), 0, s, S, B, it, 1, Q, lrl + B);
    hipLaunchKernelGGL(k_pair_grad, dim3(nblk((long)B * B)), dim3(256), 0, s, P, Q, B, it, lrl,
                       lrl + B, dS, h->d_loss);
    BSX_HIP(hipGetLastError());
    // dUp = dS Ur, dUr = dS^T Up
    dim3 gd((d + 63) / 64, (B + 63) / 64);
    hipLaunchKernelGGL((k_sgemm<float>), gd, dim3(256), 0, s, B, d, B, dS, (long)B, 1l, Ur, (long)d,
                       1l, (const float*)nullptr, dU, (long)d);
    hipLaunchKernelGGL((k_sgemm<float>), gd, dim3(256), 0, s, B, d, B, dS, 1l, (long)B, Up, (long)d,
                       1l, (const float*)nullptr, dU + (long)B * d, (long)d);
    BSX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_unit_rows_bwd, dim3(nblk(B2)), dim3(256), 0, s, dU, U, nrm, (long)B2, d,
                       h->d_gact[0]);
    BSX_HIP(hipGetLastError());
    rc = backward_body(h, x_dev, B2, s);
    if (rc) return rc;
    if (grad_dev)
        BSX_HIP(hipMemcpyAsync(grad_dev, h->d_g, sizeof(float) * h->nparam, hipMemcpyDeviceToDevice, s));
    if (loss_dev)
        BSX_HIP(hipMemcpyAsync(loss_dev, h->d_loss, sizeof(float) * B, hipMemcpyDeviceToDevice, s));
    return BSX_OK;
}

int bsx_cnn_train_pairs(bsx_cnn* h, const float* x_dev, int32_t batch, float tau, float lr,
                        float weight_decay, float* loss_dev, void* stream) {
    BSX_REQUIRE(lr > 0 && std::isfinite(lr) && weight_decay >= 0, BSX_ERR_ARG,
                "bsx_cnn_train_pairs: bad learning rate / weight decay");
    int rc = bsx_cnn_pair_gradients(h, x_dev, batch, tau, nullptr, loss_dev, stream);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    h->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)h->step), c2 = 1.f - powf(b2, (float)h->step);
    hipLaunchKernelGGL(k_adam, dim3(nblk(h->nparam)), dim3(256), 0, s, h->d_p, h->d_g, h->d_m, h->d_v,
                       h->nparam, lr, b1, b2, 1e-8f, c1, c2, weight_decay);
    BSX_HIP(hipGetLastError());
    h->pb_valid = false;
    return BSX_OK;
}

}  // extern "C"

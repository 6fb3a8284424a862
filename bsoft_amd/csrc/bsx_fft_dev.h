// bsx_fft_dev.h — in-LDS self-sorting (Stockham) FFT building blocks for gfx950.
//
// One line of L complex values lives in LDS; T = L/8 threads cooperate on it.
// Each pass loads all of a thread's butterfly inputs into registers, syncs,
// writes the outputs, syncs — so a single LDS buffer per line suffices.
// Radix sequences (compile time): 8/4/2/3 so that every supported length
// 16..512 and 3*2^k up to 384 is covered. Twiddles tw[a] = e^{+2 pi i a/L}
// come from a table (double-precision generated, stored FP32) that the caller
// has placed in LDS or global memory; SIGN selects the direction
// (-1 forward, +1 backward), matching FFTW's sign convention.
#pragma once

#include <hip/hip_runtime.h>

namespace bsx {

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by s*i (s = +-1)
template <int S>
__device__ __forceinline__ float2 mul_si(float2 a) {
    return S > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <int R, int S> struct Dft;

template <int S> struct Dft<2, S> {
    static __device__ __forceinline__ void run(float2* v) {
        float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <int S> struct Dft<3, S> {
    static __device__ __forceinline__ void run(float2* v) {
        const float c = -0.5f, sn = S * 0.86602540378443864676f;
        float2 a = v[0], b = v[1], d = v[2];
        float2 t = cadd(b, d), u = csub(b, d);
        float2 m = make_float2(a.x + c * t.x, a.y + c * t.y);
        float2 ju = make_float2(-sn * u.y, sn * u.x);   // i*sn*u
        v[0] = cadd(a, t);
        v[1] = cadd(m, ju);
        v[2] = csub(m, ju);
    }
};

template <int S> struct Dft<4, S> {
    static __device__ __forceinline__ void run(float2* v) {
        float2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
        float2 c = cadd(v[1], v[3]), d = mul_si<S>(csub(v[1], v[3]));
        v[0] = cadd(a, c);
        v[2] = csub(a, c);
        v[1] = cadd(b, d);
        v[3] = csub(b, d);
    }
};

template <int S> struct Dft<8, S> {
    static __device__ __forceinline__ void run(float2* v) {
        float2 e[4] = {v[0], v[2], v[4], v[6]};
        float2 o[4] = {v[1], v[3], v[5], v[7]};
        Dft<4, S>::run(e);
        Dft<4, S>::run(o);
        const float h = 0.70710678118654752440f;
        float2 o1 = cmul(o[1], make_float2(h, S * h));
        float2 o2 = mul_si<S>(o[2]);
        float2 o3 = cmul(o[3], make_float2(-h, S * h));
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    }
};

// One Stockham pass: Ns = product of radices already applied.
template <int L, int S, int T, int Ns, int R>
__device__ __forceinline__ void fft_pass(float2* buf, const float2* tw, int t) {
    constexpr int NB = L / R;                  // butterflies per line
    constexpr int NBT = (NB + T - 1) / T;      // butterflies per thread
    float2 v[NBT][R];
#pragma unroll
    for (int b = 0; b < NBT; ++b) {
        const int j = b * T + t;
        if (NBT * T == NB || j < NB) {
            const int k = j % Ns;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                float2 x = buf[j + r * NB];
                if (Ns > 1 && r > 0) {
                    float2 w = tw[k * r * (L / (Ns * R))];
                    w.y *= (float)S;
                    x = cmul(x, w);
                }
                v[b][r] = x;
            }
            Dft<R, S>::run(v[b]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < NBT; ++b) {
        const int j = b * T + t;
        if (NBT * T == NB || j < NB) {
            const int k = j % Ns;
            const int idxD = (j / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) buf[idxD + r * Ns] = v[b][r];
        }
    }
    __syncthreads();
}

template <int L, int S, int T, int Ns, int... Rs> struct Passes;
template <int L, int S, int T, int Ns> struct Passes<L, S, T, Ns> {
    static __device__ __forceinline__ void run(float2*, const float2*, int) {}
};
template <int L, int S, int T, int Ns, int R, int... Rs> struct Passes<L, S, T, Ns, R, Rs...> {
    static __device__ __forceinline__ void run(float2* buf, const float2* tw, int t) {
        fft_pass<L, S, T, Ns, R>(buf, tw, t);
        Passes<L, S, T, Ns * R, Rs...>::run(buf, tw, t);
    }
};

// Radix plans.
template <int L, int S, int T> struct FftLine;
#define BSX_FFT_PLAN(LEN, ...)                                                       \
    template <int S, int T> struct FftLine<LEN, S, T> {                             \
        static __device__ __forceinline__ void run(float2* buf, const float2* tw, int t) { \
            Passes<LEN, S, T, 1, __VA_ARGS__>::run(buf, tw, t);                     \
        }                                                                           \
    };
BSX_FFT_PLAN(16, 4, 4)
BSX_FFT_PLAN(32, 8, 4)
BSX_FFT_PLAN(64, 8, 8)
BSX_FFT_PLAN(128, 8, 4, 4)
BSX_FFT_PLAN(256, 8, 8, 4)
BSX_FFT_PLAN(512, 8, 8, 8)
BSX_FFT_PLAN(24, 8, 3)
BSX_FFT_PLAN(48, 4, 4, 3)
BSX_FFT_PLAN(96, 8, 4, 3)
BSX_FFT_PLAN(192, 8, 8, 3)
BSX_FFT_PLAN(384, 8, 4, 4, 3)
#undef BSX_FFT_PLAN

// Threads per line used by all kernels: one thread per 8 points.
template <int L> struct FftThreads { static constexpr int value = L / 8; };

}  // namespace bsx

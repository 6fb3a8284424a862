// bsx_fft.hip — batched C2C transforms along one axis (x, y or z) of complex
// arrays, built from the in-LDS Stockham line FFT in bsx_fft_dev.h.
//
// Replaces fft_setup_plan / fftw (src/util/fft.cpp:31-114) for the volume
// transforms of the finish (fft_back, Bimage.h:903-908) and of the map
// transform used for projection (Bimage_project.cpp:373).
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <map>
#include <mutex>
#include <vector>

#include "bsx_fft_dev.h"
#include "bsx_internal.h"

namespace bsx {

// ------------------------------------------------------------------ errors
static thread_local char g_err[1024] = "no error";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
const char* last_error() { return g_err; }

// ---------------------------------------------------------------- twiddles
const float2* twiddles(int n, hipStream_t s) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, float2*> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> g(mu);
    auto key = std::make_pair(dev, n);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    std::vector<float2> h(n);
    for (int a = 0; a < n; ++a) {
        double ang = 2.0 * M_PI * (double)a / (double)n;
        h[a] = make_float2((float)cos(ang), (float)sin(ang));
    }
    float2* d = nullptr;
    if (hipMalloc(&d, sizeof(float2) * n) != hipSuccess) return nullptr;
    if (hipMemcpyAsync(d, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice, s) != hipSuccess)
        return nullptr;
    if (hipStreamSynchronize(s) != hipSuccess) return nullptr;
    cache[key] = d;
    return d;
}

bool fft_length_supported(int n) {
    switch (n) {
        case 16: case 32: case 64: case 128: case 256: case 512:
        case 24: case 48: case 96: case 192: case 384:
            return true;
        default:
            return false;
    }
}

// ------------------------------------------------------------- axis kernel
// Lines: line index -> base = (line / inner) * outer + (line % inner);
// element i of a line at base + i * stride. B lines per workgroup, processed
// G = 256/T at a time. If real_out is set, the real part (x scale) is written
// to real_out at the same element index instead of updating `data`.
template <int L, int S, int B>
__global__ __launch_bounds__(256) void k_fft_axis(float2* __restrict__ data,
                                                  float* __restrict__ real_out, long long nlines,
                                                  long long inner, long long outer,
                                                  long long stride, float scale,
                                                  const float2* __restrict__ twg) {
    constexpr int T = FftThreads<L>::value;
    constexpr int G = 256 / T;
    constexpr int NT = G * T;
    static_assert(B % G == 0, "bundle must be a multiple of the line groups");
    __shared__ float2 lds[B * L + L];
    float2* tw = lds + B * L;
    const int tid = threadIdx.x;
    if (tid >= NT) return;   // uniform per wavefront only when NT % 64 == 0; see launch
    for (int i = tid; i < L; i += NT) tw[i] = twg[i];
    const long long line0 = (long long)blockIdx.x * B;
    if (stride == 1) {
        for (int idx = tid; idx < B * L; idx += NT) {
            const int l = idx / L, i = idx % L;
            const long long line = line0 + l;
            if (line < nlines) {
                const long long base = (line / inner) * outer + (line % inner);
                lds[l * L + i] = data[base + i];
            }
        }
    } else {
        for (int idx = tid; idx < B * L; idx += NT) {
            const int l = idx % B, i = idx / B;
            const long long line = line0 + l;
            if (line < nlines) {
                const long long base = (line / inner) * outer + (line % inner);
                lds[l * L + i] = data[base + (long long)i * stride];
            }
        }
    }
    __syncthreads();
    const int g = tid / T, t = tid % T;
#pragma unroll 1
    for (int r0 = 0; r0 < B; r0 += G) FftLine<L, S, T>::run(lds + (r0 + g) * L, tw, t);
    if (stride == 1) {
        for (int idx = tid; idx < B * L; idx += NT) {
            const int l = idx / L, i = idx % L;
            const long long line = line0 + l;
            if (line < nlines) {
                const long long base = (line / inner) * outer + (line % inner);
                float2 v = lds[l * L + i];
                if (real_out) real_out[base + i] = v.x * scale;
                else data[base + i] = make_float2(v.x * scale, v.y * scale);
            }
        }
    } else {
        for (int idx = tid; idx < B * L; idx += NT) {
            const int l = idx % B, i = idx / B;
            const long long line = line0 + l;
            if (line < nlines) {
                const long long base = (line / inner) * outer + (line % inner);
                float2 v = lds[l * L + i];
                const long long e = base + (long long)i * stride;
                if (real_out) real_out[e] = v.x * scale;
                else data[e] = make_float2(v.x * scale, v.y * scale);
            }
        }
    }
}

// Workgroup sizes: NT = (256/T)*T threads. For lengths where NT is not a
// multiple of 64 the early return above is never taken because the launch
// uses exactly NT threads.
template <int L, int S>
static int launch_axis(float2* data, float* real_out, long long nlines, long long inner,
                       long long outer, long long stride, float scale, hipStream_t s) {
    constexpr int T = FftThreads<L>::value;
    constexpr int G = 256 / T;
    constexpr int NT = G * T;
    const float2* tw = twiddles(L, s);
    BSX_REQUIRE(tw != nullptr, BSX_ERR_DEVICE, "twiddle table allocation failed for L=%d", L);
    if (stride == 1) {
        constexpr int B = G;
        long long nb = (nlines + B - 1) / B;
        hipLaunchKernelGGL((k_fft_axis<L, S, B>), dim3((unsigned)nb), dim3(NT), 0, s, data,
                           real_out, nlines, inner, outer, stride, scale, tw);
    } else {
        // bundle of >= 16 adjacent lines so each load touches >= 128 B runs
        constexpr int B0 = G >= 16 ? G : ((16 + G - 1) / G) * G;
        constexpr int B = (B0 * L * 8 + L * 8 > 64 * 1024) ? G * ((64 * 1024 / (L * 8) - 1) / G) : B0;
        static_assert(B >= G, "bundle");
        long long nb = (nlines + B - 1) / B;
        hipLaunchKernelGGL((k_fft_axis<L, S, B>), dim3((unsigned)nb), dim3(NT), 0, s, data,
                           real_out, nlines, inner, outer, stride, scale, tw);
    }
    BSX_HIP(hipGetLastError());
    return BSX_OK;
}

template <int S>
static int dispatch_axis(int L, float2* data, float* real_out, long long nlines, long long inner,
                         long long outer, long long stride, float scale, hipStream_t s) {
    switch (L) {
        case 16: return launch_axis<16, S>(data, real_out, nlines, inner, outer, stride, scale, s);
        case 32: return launch_axis<32, S>(data, real_out, nlines, inner, outer, stride, scale, s);
        case 64: return launch_axis<64, S>(data, real_out, nlines, inner, outer, stride, scale, s);
        case 128: return launch_axis<128, S>(data, real_out, nlines, inner, outer, stride, scale, s);
        case 256: return launch_axis<256, S>(data, real_out, nlines, inner, outer, stride, scale, s);
        case 512: return launch_axis<512, S>(data, real_out, nlines, inner, outer, stride, scale, s);
        case 24: return launch_axis<24, S>(data, real_out, nlines, inner, outer, stride, scale, s);
        case 48: return launch_axis<48, S>(data, real_out, nlines, inner, outer, stride, scale, s);
        case 96: return launch_axis<96, S>(data, real_out, nlines, inner, outer, stride, scale, s);
        case 192: return launch_axis<192, S>(data, real_out, nlines, inner, outer, stride, scale, s);
        case 384: return launch_axis<384, S>(data, real_out, nlines, inner, outer, stride, scale, s);
        default:
            set_error("FFT length %d not supported (16..512 powers of 2, 3*2^k to 384)", L);
            return BSX_ERR_ARG;
    }
}

// Public launcher used by the finish and map-transform paths.
int fft_axis_ex(float2* data, float* real_out, int nx, int ny, int nz, int64_t batch, int axis,
                int sign, float scale, hipStream_t s) {
    long long nxy = (long long)nx * ny, n3 = nxy * nz;
    long long L, nlines, inner, outer, stride;
    if (axis == 0) { L = nx; nlines = n3 / nx * batch; inner = 1; outer = nx; stride = 1; }
    else if (axis == 1) { L = ny; nlines = (long long)nx * nz * batch; inner = nx; outer = nxy; stride = nx; }
    else { L = nz; nlines = nxy * batch; inner = nxy; outer = n3; stride = nxy; }
    if (L == 1) return BSX_OK;
    return sign < 0 ? dispatch_axis<-1>((int)L, data, real_out, nlines, inner, outer, stride, scale, s)
                    : dispatch_axis<+1>((int)L, data, real_out, nlines, inner, outer, stride, scale, s);
}

int fft_axis(float2* data, int nx, int ny, int nz, int64_t batch, int axis, int sign, float scale,
             hipStream_t s) {
    return fft_axis_ex(data, nullptr, nx, ny, nz, batch, axis, sign, scale, s);
}

}  // namespace bsx

using namespace bsx;

extern "C" int bsx_fft(float* data_dev, int nx, int ny, int nz, int64_t batch, int dir, int norm,
                       void* stream) {
    BSX_REQUIRE(data_dev != nullptr, BSX_ERR_ARG, "bsx_fft: null data");
    BSX_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1 && batch >= 1, BSX_ERR_ARG, "bsx_fft: bad sizes");
    BSX_REQUIRE((nx == 1 || fft_length_supported(nx)) && (ny == 1 || fft_length_supported(ny)) &&
                    (nz == 1 || fft_length_supported(nz)),
                BSX_ERR_ARG, "bsx_fft: unsupported length %dx%dx%d", nx, ny, nz);
    BSX_REQUIRE(dir == -1 || dir == 1, BSX_ERR_ARG, "bsx_fft: dir must be -1 or +1");
    hipStream_t s = (hipStream_t)stream;
    double n = (double)nx * ny * nz;
    double sc = norm == 1 ? 1.0 / sqrt(n) : norm == 2 ? 1.0 / n : 1.0;
    float2* d = (float2*)data_dev;
    int axes[3] = {0, 1, 2};
    int last = nz > 1 ? 2 : ny > 1 ? 1 : 0;
    for (int a : axes) {
        int L = a == 0 ? nx : a == 1 ? ny : nz;
        if (L == 1) continue;
        int r = fft_axis(d, nx, ny, nz, batch, a, dir, a == last ? (float)sc : 1.0f, s);
        if (r != BSX_OK) return r;
    }
    return BSX_OK;
}

extern "C" const char* bsx_last_error(void) { return bsx::last_error(); }

// bsx_extract.hip — the mirror of insertion and the quality measures:
//   k_real_to_complex + 3 axis FFTs + k_phase_origin : map transform
//       (Bimage::project prologue, src/img/Bimage_project.cpp:373-375)
//   k_extract        : central sections by FSI-kernel gather interpolation
//       (Bimage::central_section, Bimage_project.cpp:288-338;
//        Bimage::fspace_interpolate, Bimage_fspace.cpp:85-147;
//        FSI_Kernel, include/FSI_Kernel.h:28-68), centre shift as project()
//   k_fsc            : FP64 shell sums of Bimage::fsc (Bimage_resolution.cpp:330-406)
//   k_frc_pairs      : img_recip_space_fom, fom_type 0 (src/mg/mg_refine.cpp:660-781)
#include <cmath>
#include <mutex>
#include <vector>

#include "bsx_common.h"

namespace bsx {
int fft_axis_ex(float2* data, float* real_out, int nx, int ny, int nz, int64_t batch, int axis,
                int sign, float scale, hipStream_t s);
}
using namespace bsx;

// ------------------------------------------------------------ map transform
__global__ void k_real_to_complex(const float* __restrict__ in, float2* __restrict__ out,
                                  long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        out[i] = make_float2(in[i], 0.f);
}

// phase_shift_to_origin (Bimage_complex.cpp:490-543): x e^{i phi},
// phi = h*sx + (l*sz + k*sy), s = origin/size * 2 pi.
__global__ void k_phase_origin(float2* __restrict__ v, int N, double sx, double sy, double sz) {
    const long long n3 = (long long)N * N * N;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
         i += (long long)gridDim.x * blockDim.x) {
        const long long xx = i % N, yy = (i / N) % N, zz = i / ((long long)N * N);
        const long h = xx > (N - 1) / 2 ? xx - N : xx;
        const long k = yy > (N - 1) / 2 ? yy - N : yy;
        const long l = zz > (N - 1) / 2 ? zz - N : zz;
        const double phi = h * sx + (l * sz + k * sy);
        double sn, cs;
        sincos(phi, &sn, &cs);
        const float2 a = v[i];
        v[i] = make_float2((float)(a.x * cs - a.y * sn), (float)(a.x * sn + a.y * cs));
    }
}

// ------------------------------------------------------------- extraction
// One thread per section pixel; blockIdx.y = view. Kernel table in LDS.
template <int WD>
__global__ __launch_bounds__(256) void k_extract(const float2* __restrict__ ft, int N,
                                                 const double* __restrict__ mats, double inv,
                                                 double rs, double maxs2, int center,
                                                 const float* __restrict__ ktab,
                                                 float2* __restrict__ out) {
    __shared__ float K[(FSI_NDIV + 1) * WD];
    for (int i = threadIdx.x; i < (FSI_NDIV + 1) * WD; i += blockDim.x) K[i] = ktab[i];
    __syncthreads();
    const int v = blockIdx.y;
    const long long npix = (long long)N * N;
    const long long pix = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= npix) return;
    const double* M = mats + 9 * v;
    const int xx = (int)(pix % N), yy = (int)(pix / N);
    const int hx = (N - 1) / 2;
    const double ivy = (double)(yy > hx ? yy - N : yy) * inv;
    const double ivx = (double)(xx > hx ? xx - N : xx) * inv;
    const double s2 = ivx * ivx + ivy * ivy;
    double vr = 0, vi = 0;
    if (s2 <= maxs2) {
        double m[3];
        for (int k = 0; k < 3; ++k) m[k] = (ivx * M[3 * k] + ivy * M[3 * k + 1] + 0.0 * M[3 * k + 2]) * rs;
        fsi_sample<WD>(ft, N, m, K, vr, vi);
    }
    if (center) {
        // phase_shift_to_center of a 2D section: x (-1)^(h+k) for even N
        const long h = xx > hx ? xx - N : xx, k = yy > hx ? yy - N : yy;
        if ((N & 1) == 0) {
            if ((h + k) & 1) { vr = -vr; vi = -vi; }
        } else {
            const double t = -2 * M_PI * (double)(N / 2) / N;
            double sn, cs;
            sincos(h * t + k * t, &sn, &cs);
            const double a = vr, b = vi;
            vr = a * cs - b * sn;
            vi = a * sn + b * cs;
        }
    }
    out[(long long)v * npix + pix] = make_float2((float)vr, (float)vi);
}

// -------------------------------------------------------------------- FSC
__global__ __launch_bounds__(256) void k_fsc(const float2* __restrict__ a,
                                             const float2* __restrict__ b, int N, long maxrad,
                                             double* __restrict__ sums) {
    extern __shared__ double sh[];   // 3 * maxrad
    for (long i = threadIdx.x; i < 3 * maxrad; i += blockDim.x) sh[i] = 0;
    __syncthreads();
    const long long n3 = (long long)N * N * N;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n3;
         i += (long long)gridDim.x * blockDim.x) {
        const long long xx = i % N, yy = (i / N) % N, zz = i / ((long long)N * N);
        const double h = xx > (N - 1) / 2 ? (double)(xx - N) : (double)xx;
        const double k = yy > (N - 1) / 2 ? (double)(yy - N) : (double)yy;
        const double l = zz > (N - 1) / 2 ? (double)(zz - N) : (double)zz;
        // radius = rad_scale*sqrt(rx^2+ry^2+rz^2), rx = h/(N pix): voxel units
        const double radius = sqrt(h * h + k * k + l * l);
        const long ir = (long)radius;
        if (ir >= maxrad) continue;
        const double f = radius - ir, f1 = 1 - f;
        const float2 u = a[i], v = b[i];
        const double p1 = (double)u.x * u.x + (double)u.y * u.y;
        const double p2 = (double)v.x * v.x + (double)v.y * v.y;
        const double re = (double)u.x * v.x + (double)u.y * v.y;
        atomicAdd(&sh[ir], f1 * p1);
        atomicAdd(&sh[maxrad + ir], f1 * p2);
        atomicAdd(&sh[2 * maxrad + ir], f1 * re);
        if (ir + 1 < maxrad) {
            atomicAdd(&sh[ir + 1], f * p1);
            atomicAdd(&sh[maxrad + ir + 1], f * p2);
            atomicAdd(&sh[2 * maxrad + ir + 1], f * re);
        }
    }
    __syncthreads();
    for (long i = threadIdx.x; i < 3 * maxrad; i += blockDim.x)
        if (sh[i] != 0) atomicAdd(&sums[i], sh[i]);
}

// FP64 shell sums of Bimage::fsc_dpr (Bimage_resolution.cpp:93-181) for two
// real maps transformed separately (the same F1, F2 as unpacking the packed
// transform): over x < N/2, linear interpolation between shells i and i+1
// (both must exist), sums[c * maxrad + i] for c = |F1|^2, |F2|^2,
// Re F1 conj F2, |F1|^2 + |F2|^2, (|F1|^2 + |F2|^2) dphi^2.
__global__ __launch_bounds__(256) void k_fsc_dpr(const float2* __restrict__ a,
                                                 const float2* __restrict__ b, int N, int nz,
                                                 double inv_ratio, long maxrad,
                                                 double* __restrict__ sums) {
    extern __shared__ double sh[];   // 5 * maxrad
    for (long i = threadIdx.x; i < 5 * maxrad; i += blockDim.x) sh[i] = 0;
    __syncthreads();
    const int hx = N / 2;
    const long long nv = (long long)hx * N * nz;   // nz = 1: a 2D image
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nv;
         t += (long long)gridDim.x * blockDim.x) {
        const long long xx = t % hx, yz = t / hx, yy = yz % N, zz = yz / N;
        const double h = xx > (N - 1) / 2 ? (double)(xx - N) : (double)xx;
        const double k = yy > (N - 1) / 2 ? (double)(yy - N) : (double)yy;
        const double l = zz > (nz - 1) / 2 ? (double)(zz - nz) : (double)zz;
        const double radius = inv_ratio * sqrt(h * h + k * k + l * l);
        const long ir = (long)radius;
        if (ir + 1 >= maxrad) continue;
        const double f = radius - ir, f1 = 1 - f;
        const long long i = (zz * N + yy) * N + xx;
        const float2 u = a[i], v = b[i];
        const double I1 = (double)u.x * u.x + (double)u.y * u.y;
        const double I2 = (double)v.x * v.x + (double)v.y * v.y;
        double dphi = atan2((double)u.y, (double)u.x) - atan2((double)v.y, (double)v.x);
        dphi = remainder(dphi, 2 * M_PI);
        dphi *= dphi;
        const double re = (double)u.x * v.x + (double)u.y * v.y;
        const double is = I1 + I2;
        const double val[5] = {I1, I2, re, is, is * dphi};
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            atomicAdd(&sh[c * maxrad + ir], f1 * val[c]);
            atomicAdd(&sh[c * maxrad + ir + 1], f * val[c]);
        }
    }
    __syncthreads();
    for (long i = threadIdx.x; i < 5 * maxrad; i += blockDim.x)
        if (sh[i] != 0) atomicAdd(&sums[i], sh[i]);
}

// ------------------------------------------------------------ FRC scoring
// One workgroup per (particle, reference) pair. Ring sums in LDS (FP64).
__global__ __launch_bounds__(256) void k_frc_pairs(const float2* __restrict__ part,
                                                   const float2* __restrict__ ref, int N,
                                                   const int* __restrict__ pairs,
                                                   double min_rad_sq, double max_rad_cv_sq,
                                                   double max_rad, double max_rad_cv,
                                                   double* __restrict__ out,
                                                   double* __restrict__ out_cv) {
    extern __shared__ double rs[];   // 3 * N
    double* fom = rs;
    double* dsum = rs + N;
    double* msum = rs + 2 * N;
    for (int i = threadIdx.x; i < 3 * N; i += blockDim.x) rs[i] = 0;
    __syncthreads();
    const int ip = pairs[2 * blockIdx.x], ir = pairs[2 * blockIdx.x + 1];
    const long long np = (long long)N * N;
    const float2* P = part + (long long)ip * np;
    const float2* R = ref + (long long)ir * np;
    const int hx = (N - 1) / 2;
    for (long long i = threadIdx.x; i < np; i += blockDim.x) {
        const int xx = (int)(i % N), yy = (int)(i / N);
        const double ix = xx > hx ? (double)xx - N : (double)xx;
        const double iy = yy > hx ? (double)yy - N : (double)yy;
        const double d2 = ix * ix + iy * iy;
        if (d2 >= min_rad_sq && d2 <= max_rad_cv_sq) {
            const int j = (int)sqrt(d2);
            if (j >= N) continue;
            const float2 d = P[i], r = R[i];
            atomicAdd(&fom[j], (double)d.x * r.x + (double)d.y * r.y);
            atomicAdd(&msum[j], (double)r.x * r.x + (double)r.y * r.y);
            atomicAdd(&dsum[j], (double)d.x * d.x + (double)d.y * d.y);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double f = 0, wsum = 0, cv = 0;
        for (int j = 0; j < max_rad && j < N; ++j) {
            wsum += 1;
            const double den = dsum[j] * msum[j];
            if (den > 0) f += fom[j] / sqrt(den);
        }
        f = wsum > 0 ? f / wsum : 0;
        wsum = 0;
        for (int j = (int)max_rad; j < max_rad_cv && j < N; ++j) {
            wsum += 1;
            const double den = dsum[j] * msum[j];
            if (den > 0) cv += fom[j] / sqrt(den);
        }
        out[blockIdx.x] = f;
        if (out_cv) out_cv[blockIdx.x] = wsum > 0 ? cv / wsum : 0;
    }
}

// ================================================================ C-ABI
static double check_res(double res, int n, double sampling, int ndim) {
    if (res > n * sampling) return n * sampling;
    double rm = 2 / sqrt(ndim / (sampling * sampling));
    return res < rm ? rm : res;
}

extern "C" {

int bsx_map_transform(const float* map_dev, int n, const double origin[3], float* ft_dev,
                      void* stream) {
    BSX_REQUIRE(map_dev && ft_dev && origin, BSX_ERR_ARG, "bsx_map_transform: null argument");
    BSX_REQUIRE(fft_length_supported(n), BSX_ERR_ARG, "bsx_map_transform: size %d unsupported", n);
    hipStream_t s = (hipStream_t)stream;
    const long long n3 = (long long)n * n * n;
    float2* v = (float2*)ft_dev;
    hipLaunchKernelGGL(k_real_to_complex, dim3(2048), dim3(256), 0, s, map_dev, v, n3);
    BSX_HIP(hipGetLastError());
    const float sc = (float)(1.0 / sqrt((double)n3));
    int r = fft_axis_ex(v, nullptr, n, n, n, 1, 0, -1, 1.0f, s);
    if (r == BSX_OK) r = fft_axis_ex(v, nullptr, n, n, n, 1, 1, -1, 1.0f, s);
    if (r == BSX_OK) r = fft_axis_ex(v, nullptr, n, n, n, 1, 2, -1, sc, s);
    if (r != BSX_OK) return r;
    const double sx = origin[0] / n * 2 * M_PI, sy = origin[1] / n * 2 * M_PI,
                 sz = origin[2] / n * 2 * M_PI;
    hipLaunchKernelGGL(k_phase_origin, dim3(2048), dim3(256), 0, s, v, n, sx, sy, sz);
    BSX_HIP(hipGetLastError());
    return BSX_OK;
}

// Argument checks shared by bsx_extract and bsx_extract_noalloc; normalises
// width, sampling and resolution as the reference's defaults do.
static int extract_args(const float* ft_dev, int n, double& sampling, const double* mats,
                        int64_t nviews, double& resolution, int& kernel_width, int kernel_power,
                        const float* out_dev, const char* who) {
    BSX_REQUIRE(ft_dev && mats && out_dev, BSX_ERR_ARG, "%s: null argument", who);
    BSX_REQUIRE(n >= 2 && nviews >= 0 && nviews < 65536, BSX_ERR_ARG,
                "%s: n %d / nviews %lld out of range", who, n, (long long)nviews);
    if (kernel_power < 0) kernel_width = 1;
    BSX_REQUIRE(kernel_width >= 1 && kernel_width <= 10, BSX_ERR_ARG,
                "%s: kernel width %d out of range 1..10", who, kernel_width);
    if (sampling < 1e-10) sampling = 1;
    if (resolution < 1e-10) resolution = sampling;
    for (int64_t i = 0; i < 9 * nviews; ++i)
        BSX_REQUIRE(std::isfinite(mats[i]) && fabs(mats[i]) <= 1.0 + 1e-9, BSX_ERR_ARG,
                    "%s: matrix %lld is not a rotation", who, (long long)(i / 9));
    return BSX_OK;
}

// Launches k_extract<width> over nviews sections; dtab / dmats are device
// copies of the kernel table and the view matrices.
static int extract_launch(const float* ft_dev, int n, double sampling, const double* dmats,
                          int64_t nviews, double resolution, int kernel_width, int center,
                          const float* dtab, float* out_dev, hipStream_t s) {
    const double inv = 1.0 / (n * sampling), rs = n * sampling;
    const double maxs2 = 1.0 / (resolution * resolution);
    dim3 grid((unsigned)(((long long)n * n + 255) / 256), (unsigned)nviews);
    const float2* ft = (const float2*)ft_dev;
    float2* out = (float2*)out_dev;
    switch (kernel_width) {
#define BSX_EXT(W) case W: hipLaunchKernelGGL((k_extract<W>), grid, dim3(256), 0, s, ft, n, dmats, inv, rs, maxs2, center, dtab, out); break;
        BSX_EXT(1) BSX_EXT(2) BSX_EXT(3) BSX_EXT(4) BSX_EXT(5)
        BSX_EXT(6) BSX_EXT(7) BSX_EXT(8) BSX_EXT(9) BSX_EXT(10)
#undef BSX_EXT
    }
    BSX_HIP(hipGetLastError());
    return BSX_OK;
}

int bsx_extract(const float* ft_dev, int n, double sampling, const double* mats, int64_t nviews,
                double resolution, int kernel_width, int kernel_power, int center, float* out_dev,
                void* stream) {
    int r = extract_args(ft_dev, n, sampling, mats, nviews, resolution, kernel_width,
                         kernel_power, out_dev, "bsx_extract");
    if (r != BSX_OK || nviews == 0) return r;
    hipStream_t s = (hipStream_t)stream;
    std::vector<float> tab = fsi_table(kernel_width, kernel_power);
    float* dtab = nullptr;
    double* dmats = nullptr;
    BSX_HIP(hipMalloc(&dtab, sizeof(float) * tab.size()));
    BSX_HIP(hipMalloc(&dmats, sizeof(double) * 9 * nviews));
    BSX_HIP(hipMemcpyAsync(dtab, tab.data(), sizeof(float) * tab.size(), hipMemcpyHostToDevice, s));
    BSX_HIP(hipMemcpyAsync(dmats, mats, sizeof(double) * 9 * nviews, hipMemcpyHostToDevice, s));
    r = extract_launch(ft_dev, n, sampling, dmats, nviews, resolution, kernel_width, center, dtab,
                       out_dev, s);
    hipError_t se = hipStreamSynchronize(s);
    (void)hipFree(dtab);
    (void)hipFree(dmats);
    if (r != BSX_OK) return r;
    BSX_HIP(se);
    return BSX_OK;
}

// ---------------------------------------------- allocation-free extraction
// Scratch of bsx_extract_noalloc: the view matrices (FP64, at most
// BSX_NOALLOC_MAX_VIEWS) followed by the FSI kernel table (at most width 10).
// One copy per GPU, held until the process ends. In host code `scratch` is
// only a symbol: its device address comes from hipGetSymbolAddress, once per
// device, and only that address is handed to copies and kernels.
constexpr size_t kScratchMats = sizeof(double) * 9 * BSX_NOALLOC_MAX_VIEWS;
constexpr size_t kScratchTab = sizeof(float) * (FSI_NDIV + 1) * 10;
static __device__ __align__(256) unsigned char scratch[kScratchMats + kScratchTab];

constexpr int kMaxDevices = 64;
static std::mutex scratch_mu;
static void* scratch_ptr[kMaxDevices];

// Device address of the current device's scratch (looked up on the first
// call on that device).
static int scratch_address(void** out) {
    int dev = 0;
    BSX_HIP(hipGetDevice(&dev));
    BSX_REQUIRE(dev >= 0 && dev < kMaxDevices, BSX_ERR_DEVICE, "device %d out of range", dev);
    std::lock_guard<std::mutex> g(scratch_mu);
    if (!scratch_ptr[dev]) {
        void* p = nullptr;
        BSX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        scratch_ptr[dev] = p;
    }
    *out = scratch_ptr[dev];
    return BSX_OK;
}

int bsx_extract_noalloc(const float* ft_dev, int n, double sampling, const double* mats,
                        int64_t nviews, double resolution, int kernel_width, int kernel_power,
                        int center, float* out_dev, void* stream) {
    int r = extract_args(ft_dev, n, sampling, mats, nviews, resolution, kernel_width,
                         kernel_power, out_dev, "bsx_extract_noalloc");
    if (r != BSX_OK) return r;
    BSX_REQUIRE(nviews <= BSX_NOALLOC_MAX_VIEWS, BSX_ERR_ARG,
                "bsx_extract_noalloc: %lld views exceed the scratch capacity of %d",
                (long long)nviews, BSX_NOALLOC_MAX_VIEWS);
    if (nviews == 0) return BSX_OK;
    void* base = nullptr;
    r = scratch_address(&base);
    if (r != BSX_OK) return r;
    double* dmats = (double*)base;
    float* dtab = (float*)((unsigned char*)base + kScratchMats);
    // the kernel table on the host stack (no heap), then both into scratch
    float tab[(FSI_NDIV + 1) * 10];
    fsi_table_into(tab, kernel_width, kernel_power);
    const size_t ntab = (size_t)(FSI_NDIV + 1) * (kernel_power < 0 ? 1 : kernel_width);
    hipStream_t s = (hipStream_t)stream;
    BSX_HIP(hipMemcpyAsync(dtab, tab, sizeof(float) * ntab, hipMemcpyHostToDevice, s));
    BSX_HIP(hipMemcpyAsync(dmats, mats, sizeof(double) * 9 * nviews, hipMemcpyHostToDevice, s));
    r = extract_launch(ft_dev, n, sampling, dmats, nviews, resolution, kernel_width, center, dtab,
                       out_dev, s);
    if (r != BSX_OK) return r;
    // the host copies above read `tab` and `mats`: wait before returning
    BSX_HIP(hipStreamSynchronize(s));
    return BSX_OK;
}

int bsx_fsc(const float* a_dev, const float* b_dev, int n, double sampling, double hi_res,
            double* fsc, int* nshell, void* stream) {
    BSX_REQUIRE(a_dev && b_dev && fsc && nshell, BSX_ERR_ARG, "bsx_fsc: null argument");
    BSX_REQUIRE(n >= 2 && sampling > 0, BSX_ERR_ARG, "bsx_fsc: bad size/sampling");
    hipStream_t s = (hipStream_t)stream;
    hi_res = check_res(hi_res, n, sampling, 3);
    long maxrad = (long)(n * sampling / hi_res);
    BSX_REQUIRE(maxrad <= *nshell, BSX_ERR_ARG, "bsx_fsc: %ld shells exceed capacity %d", maxrad,
                *nshell);
    double* d = nullptr;
    BSX_HIP(hipMalloc(&d, sizeof(double) * 3 * (maxrad + 1)));
    BSX_HIP(hipMemsetAsync(d, 0, sizeof(double) * 3 * (maxrad + 1), s));
    hipLaunchKernelGGL(k_fsc, dim3(1024), dim3(256), sizeof(double) * 3 * maxrad, s,
                       (const float2*)a_dev, (const float2*)b_dev, n, maxrad, d);
    hipError_t le = hipGetLastError();
    std::vector<double> h(3 * maxrad);
    if (maxrad > 0)
        BSX_HIP(hipMemcpyAsync(h.data(), d, sizeof(double) * 3 * maxrad, hipMemcpyDeviceToHost, s));
    BSX_HIP(hipStreamSynchronize(s));
    (void)hipFree(d);
    BSX_HIP(le);
    for (long i = 0; i < maxrad; ++i) {
        double den = h[i] * h[maxrad + i];
        fsc[i] = den > 0 ? h[2 * maxrad + i] / sqrt(den) : 0;
    }
    if (maxrad > 0) fsc[0] = 1;
    *nshell = (int)maxrad;
    return BSX_OK;
}

// n x n x nz maps, nz = n (3D) or 1 (2D images).
static int fsc_dpr_impl(const float* map1_dev, const float* map2_dev, int n, int nz,
                        double sampling, double hi_res, double sampling_ratio, double* curves,
                        int* nrows, void* stream) {
    BSX_REQUIRE(map1_dev && map2_dev && curves && nrows, BSX_ERR_ARG, "bsx_fsc_dpr: null argument");
    BSX_REQUIRE(n >= 2 && n % 2 == 0 && sampling > 0, BSX_ERR_ARG, "bsx_fsc_dpr: bad size/sampling");
    if (sampling_ratio <= 0) sampling_ratio = 1;
    hipStream_t s = (hipStream_t)stream;
    hi_res = check_res(hi_res, n, sampling, nz > 1 ? 3 : 2);
    const double rad_scale = n * sampling / sampling_ratio;
    const long maxrad = (long)(2 + rad_scale / hi_res);
    BSX_REQUIRE(maxrad <= *nrows, BSX_ERR_ARG, "bsx_fsc_dpr: %ld rows exceed capacity %d", maxrad,
                *nrows);
    const long long n3 = (long long)n * n * nz;
    float2* v = nullptr;
    double* d = nullptr;
    BSX_HIP(hipMalloc(&v, sizeof(float2) * 2 * n3));
    if (hipMalloc(&d, sizeof(double) * 5 * maxrad) != hipSuccess) {
        (void)hipFree(v);
        set_error("bsx_fsc_dpr: allocation failed");
        return BSX_ERR_NOMEM;
    }
    int r = BSX_OK;
    for (int m = 0; m < 2 && r == BSX_OK; ++m) {
        hipLaunchKernelGGL(k_real_to_complex, dim3(2048), dim3(256), 0, s, m ? map2_dev : map1_dev,
                           v + m * n3, n3);
        for (int axis = 0; axis < (nz > 1 ? 3 : 2) && r == BSX_OK; ++axis)
            r = fft_axis_ex(v + m * n3, nullptr, n, n, nz, 1, axis, -1, 1.0f, s);
    }
    if (r == BSX_OK) {
        (void)hipMemsetAsync(d, 0, sizeof(double) * 5 * maxrad, s);
        hipLaunchKernelGGL(k_fsc_dpr, dim3(1024), dim3(256), sizeof(double) * 5 * maxrad, s, v,
                           v + n3, n, nz, 1.0 / sampling_ratio, maxrad, d);
        if (hipGetLastError() != hipSuccess) r = BSX_ERR_DEVICE;
    }
    std::vector<double> h(5 * maxrad);
    if (r == BSX_OK &&
        (hipMemcpyAsync(h.data(), d, sizeof(double) * 5 * maxrad, hipMemcpyDeviceToHost, s) !=
             hipSuccess ||
         hipStreamSynchronize(s) != hipSuccess))
        r = BSX_ERR_DEVICE;
    (void)hipFree(v);
    (void)hipFree(d);
    if (r) {
        set_error("bsx_fsc_dpr: device error");
        return r;
    }
    // curves: (s, FSC, DPR in degrees) per row (Bimage_resolution.cpp:225-239)
    for (long i = 0; i < maxrad; ++i) {
        double fsc = h[2 * maxrad + i], dpr = 0;
        if (i == 0) {
            fsc = 1;
        } else {
            const double f12 = h[i] * h[maxrad + i];
            if (f12 > 1e-30) fsc /= sqrt(f12);
            fsc = std::min(1.0, std::max(-1.0, fsc));
            const double fs = h[3 * maxrad + i];
            dpr = fs > 1e-30 ? sqrt(h[4 * maxrad + i] / fs) * 180.0 / M_PI : 0;
        }
        curves[3 * i] = i / rad_scale;
        curves[3 * i + 1] = fsc;
        curves[3 * i + 2] = dpr;
    }
    *nrows = (int)maxrad;
    return BSX_OK;
}

int bsx_fsc_dpr(const float* map1_dev, const float* map2_dev, int n, double sampling,
                double hi_res, double sampling_ratio, double* curves, int* nrows, void* stream) {
    return fsc_dpr_impl(map1_dev, map2_dev, n, n, sampling, hi_res, sampling_ratio, curves, nrows,
                        stream);
}

int bsx_fsc_dpr_2d(const float* img1_dev, const float* img2_dev, int n, double sampling,
                   double hi_res, double sampling_ratio, double* curves, int* nrows,
                   void* stream) {
    return fsc_dpr_impl(img1_dev, img2_dev, n, 1, sampling, hi_res, sampling_ratio, curves, nrows,
                        stream);
}

int bsx_score_frc(const float* part_dev, int64_t nparts, const float* ref_dev, int64_t nrefs,
                  int n, double sampling, double hi_res, double lo_res, const int32_t* pairs,
                  int64_t npairs, double* out, void* stream) {
    BSX_REQUIRE(part_dev && ref_dev && pairs && out, BSX_ERR_ARG, "bsx_score_frc: null argument");
    BSX_REQUIRE(n >= 2 && npairs >= 0, BSX_ERR_ARG, "bsx_score_frc: bad sizes");
    if (npairs == 0) return BSX_OK;
    for (int64_t i = 0; i < npairs; ++i)
        BSX_REQUIRE(pairs[2 * i] >= 0 && pairs[2 * i] < nparts && pairs[2 * i + 1] >= 0 &&
                        pairs[2 * i + 1] < nrefs,
                    BSX_ERR_ARG, "bsx_score_frc: pair %lld (%d, %d) out of range", (long long)i,
                    pairs[2 * i], pairs[2 * i + 1]);
    hipStream_t s = (hipStream_t)stream;
    if (lo_res > 0 && hi_res > lo_res) std::swap(hi_res, lo_res);
    hi_res = check_res(hi_res, n, sampling, 2);
    double max_rad = n * sampling / hi_res;
    double max_rad_cv = 1.1 * max_rad;
    if (max_rad_cv <= max_rad) max_rad_cv = max_rad + 1;
    double min_rad_sq = 0;
    if (lo_res > 0) { min_rad_sq = n * sampling / lo_res; min_rad_sq *= min_rad_sq; }
    int* dp = nullptr;
    double* dout = nullptr;
    BSX_HIP(hipMalloc(&dp, sizeof(int) * 2 * npairs));
    BSX_HIP(hipMalloc(&dout, sizeof(double) * npairs));
    BSX_HIP(hipMemcpyAsync(dp, pairs, sizeof(int) * 2 * npairs, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_frc_pairs, dim3((unsigned)npairs), dim3(256), sizeof(double) * 3 * n, s,
                       (const float2*)part_dev, (const float2*)ref_dev, n, dp, min_rad_sq,
                       max_rad_cv * max_rad_cv, max_rad, max_rad_cv, dout, (double*)nullptr);
    hipError_t le = hipGetLastError();
    BSX_HIP(hipMemcpyAsync(out, dout, sizeof(double) * npairs, hipMemcpyDeviceToHost, s));
    BSX_HIP(hipStreamSynchronize(s));
    (void)hipFree(dp);
    (void)hipFree(dout);
    BSX_HIP(le);
    return BSX_OK;
}

}  // extern "C"

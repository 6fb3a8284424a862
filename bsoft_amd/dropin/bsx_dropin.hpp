// bsx_dropin.hpp — the Bsoft-free core of the drop-in particle_reconstruct /
// img_reconstruction_sum_weigh (bsoft/include/mg_reconstruct.h:16-20): plain
// host arrays in and out, libbsx (include/bsx.h) underneath.
//
//   PartialMap  one partial map of particle_reconstruct
//               (src/mg/mg_reconstruct.cpp:100-264): particle images staged
//               from host memory to the GPU in chunks and inserted, then the
//               accumulator chain the reference returns, as host planes:
//               complex sum(w F) (interleaved re, im), sum(w |F|^2), sum(w),
//               sum(w^2), each N^3 FP32 (N^2 for a 2D map), voxel order
//               (z*N + y)*N + x.
//   sum_weigh   img_reconstruction_sum_weigh (mg_reconstruct.cpp:266-361) of
//               such planes: the selected partial maps summed in the
//               reference's order (FP32 adds), then weighed, SNR figure of
//               merit, centre shift on the GPU; out: the weighed transform
//               (interleaved complex) and the FOM map.
//   partial_index / half_of  the reference's round-robin assignment of the
//               k-th selected particle of a class to a partial map
//               (mg_reconstruct.cpp:741-744) and of partial maps to half sets
//               (mg_reconstruct.cpp:287-295).
//
// Header-only C++17; needs the HIP runtime (hipMalloc / hipMemcpy) and
// libbsx. Every function returns a bsx status (BSX_OK or BSX_ERR_*), with
// bsx_last_error() or the message passed back for text.
#ifndef BSX_DROPIN_HPP
#define BSX_DROPIN_HPP

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bsx.h"

namespace bsx_dropin {

// Partial map (0-based) of the k-th (1-based) selected particle of class c
// (1-based) with T partial maps per class: sel = (c-1) T + ((k-1) mod T) + 1
// is the reference's selection number; the map index is sel - 1.
inline int partial_index(long k, int c, int T) { return (c - 1) * T + (int)((k - 1) % T); }

// Half set of a partial map when nmaps is 2 or 3: sum_weigh takes every
// second partial map from st = imap % 2 (nmaps 2) or imap % 3 - 1 (nmaps 3),
// and the configuration makes T even whenever half sets are asked for
// (mg_reconstruct.cpp:676), so partial map i belongs to half i % 2.
inline int half_of(int partial, int T) { return (partial % T) % 2; }

namespace detail {
// Voxels of the accumulator planes (always N^3: a 2D map writes its z = 0
// plane) and of a map, transform or partial-map image (N^2 for a 2D map).
inline size_t acc_voxels(const bsx_recon_params& p) { return (size_t)p.size * p.size * p.size; }
inline size_t map_voxels(const bsx_recon_params& p) {
    return (size_t)p.size * p.size * ((p.flags & BSX_FLAG_2D) ? 1 : p.size);
}

inline int hip_status(hipError_t e, std::string* msg, const char* what) {
    if (e == hipSuccess) return BSX_OK;
    if (msg) *msg = std::string(what) + ": " + hipGetErrorString(e);
    return BSX_ERR_DEVICE;
}
}  // namespace detail

// One partial map on the GPU (one libbsx handle with the FOM sums, as the
// reference always accumulates sum(w|F|^2) and sum(w^2)).
class PartialMap {
public:
    // `chunk`: particles staged per host-to-device copy.
    explicit PartialMap(const bsx_recon_params& params, long chunk = 1024)
        : p_(params), chunk_(chunk < 1 ? 1 : chunk) {
        p_.flags |= BSX_FLAG_FOM;
        p_.nhalves = 1;
        status_ = bsx_recon_create(&h_, &p_);
        if (status_ != BSX_OK) msg_ = bsx_last_error();
    }
    ~PartialMap() {
        if (h_) bsx_recon_destroy(h_);
        if (dimg_) (void)hipFree(dimg_);
    }
    PartialMap(const PartialMap&) = delete;
    PartialMap& operator=(const PartialMap&) = delete;

    int status() const { return status_; }
    const std::string& message() const { return msg_; }
    long count() const { return count_; }

    // Inserts n particles: images n x box x box FP32 (host), records as
    // particle_reconstruct fills them (view, origin, CTF of the micrograph
    // with the particle's defocus, weight).
    int insert(const float* imgs, const bsx_particle* parts, long n) {
        if (status_ != BSX_OK) return status_;
        const size_t per = (size_t)p_.box * p_.box;
        if (!dimg_) {
            int r = detail::hip_status(hipMalloc((void**)&dimg_, sizeof(float) * per * chunk_),
                                       &msg_, "hipMalloc");
            if (r) return status_ = r;
        }
        for (long i = 0; i < n; i += chunk_) {
            const long m = n - i < chunk_ ? n - i : chunk_;
            int r = detail::hip_status(hipMemcpy(dimg_, imgs + (size_t)i * per,
                                                 sizeof(float) * per * m, hipMemcpyHostToDevice),
                                       &msg_, "hipMemcpy");
            if (r) return status_ = r;
            r = bsx_recon_insert(h_, dimg_, parts + i, m, nullptr);
            if (r) {
                msg_ = bsx_last_error();
                return status_ = r;
            }
        }
        count_ += n;
        return BSX_OK;
    }

    // Point-group symmetry of the insertion (symmetry mode 0): every later
    // particle is added at all its related views. sym as Bsymmetry's label.
    int set_symmetry(const char* sym) {
        if (status_ != BSX_OK) return status_;
        int r = bsx_recon_set_symmetry(h_, sym);
        if (r) msg_ = bsx_last_error();
        return r;
    }

    // The accumulator chain as host planes (each N^3 floats, N^2 for a 2D
    // map; cplx twice that).
    int planes(float* cplx, float* power, float* weight, float* weight2) {
        if (status_ != BSX_OK) return status_;
        const size_t n3 = detail::acc_voxels(p_), nv = detail::map_voxels(p_);
        float* d = nullptr;
        int r = detail::hip_status(hipMalloc((void**)&d, sizeof(float) * 5 * n3), &msg_,
                                   "hipMalloc");
        if (r) return r;
        std::vector<float> h(5 * n3);
        r = bsx_recon_accumulators(h_, 0, d, nullptr);
        if (r) msg_ = bsx_last_error();
        if (!r) r = detail::hip_status(hipMemcpy(h.data(), d, sizeof(float) * 5 * n3,
                                                 hipMemcpyDeviceToHost), &msg_, "hipMemcpy");
        (void)hipFree(d);
        if (r) return r;
        for (size_t i = 0; i < nv; ++i) {
            cplx[2 * i] = h[i];
            cplx[2 * i + 1] = h[n3 + i];
        }
        std::memcpy(weight, h.data() + 2 * n3, sizeof(float) * nv);
        std::memcpy(power, h.data() + 3 * n3, sizeof(float) * nv);
        std::memcpy(weight2, h.data() + 4 * n3, sizeof(float) * nv);
        return BSX_OK;
    }

private:
    bsx_recon_params p_;
    long chunk_;
    bsx_recon* h_ = nullptr;
    float* dimg_ = nullptr;
    long count_ = 0;
    int status_ = BSX_OK;
    std::string msg_;
};

// Host planes of one partial map (borrowed pointers).
struct PartialPlanes {
    const float* cplx;      // 2 N^3 (2 N^2 in 2D): sum(w F), interleaved
    const float* power;     // N^3 (N^2): sum(w |F|^2)
    const float* weight;    // N^3 (N^2): sum(w)
    const float* weight2;   // N^3 (N^2): sum(w^2)
    long count;             // particles in it
};

// img_reconstruction_sum_weigh(pacc, imap, nmaps, maps_per_class, hi_res):
// partial maps of class imap / nmaps, every one (nmaps 1, or imap 0 of 3) or
// every second from the half's first, summed in that order with FP32 adds as
// Bimage::add does; then the finish on the GPU. params: as for the partial
// maps (hi_res: the limit of the shell table). out_cplx (2 N^3, 2 N^2 in 2D)
// receives the weighed, centre-shifted transform, out_fom (N^3 / N^2) the SNR
// figure of merit; stats may be null. rows, if not null, receives the
// fspace_reconstruction_stats table (9 values per shell, bsx_recon_stats).
inline int sum_weigh(const bsx_recon_params& params, const PartialPlanes* pacc, int imap,
                     int nmaps, int maps_per_class, float* out_cplx, float* out_fom,
                     bsx_stats* stats, std::string* msg, std::vector<double>* rows = nullptr) {
    if (nmaps < 1 || nmaps > 3 || imap < 0 || maps_per_class < 1) {
        if (msg) *msg = "sum_weigh: bad imap / nmaps / maps_per_class";
        return BSX_ERR_ARG;
    }
    const int iclass = imap / nmaps;
    int st = 0, inc = 1;
    if (nmaps == 2) {
        st = imap % 2;
        inc = 2;
    }
    if (nmaps == 3 && imap % 3) {
        st = imap % 3 - 1;
        inc = 2;
    }
    const size_t n3 = detail::acc_voxels(params), nv = detail::map_voxels(params);
    std::vector<float> planes(5 * n3, 0.0f);
    long count = 0;
    for (int i = iclass * maps_per_class + st, j = 0; j < maps_per_class; i += inc, j += inc) {
        const PartialPlanes& a = pacc[i];
        for (size_t k = 0; k < nv; ++k) {
            planes[k] += a.cplx[2 * k];
            planes[n3 + k] += a.cplx[2 * k + 1];
            planes[2 * n3 + k] += a.weight[k];
            planes[3 * n3 + k] += a.power[k];
            planes[4 * n3 + k] += a.weight2[k];
        }
        count += a.count;
    }
    bsx_recon_params p = params;
    p.flags |= BSX_FLAG_FOM;
    p.nhalves = 1;
    bsx_recon* h = nullptr;
    int r = bsx_recon_create(&h, &p);
    if (r) {
        if (msg) *msg = bsx_last_error();
        return r;
    }
    float *dplanes = nullptr, *dft = nullptr, *dfom = nullptr;
    r = detail::hip_status(hipMalloc((void**)&dplanes, sizeof(float) * 5 * n3), msg, "hipMalloc");
    if (!r) r = detail::hip_status(hipMalloc((void**)&dft, sizeof(float) * 2 * nv), msg, "hipMalloc");
    if (!r) r = detail::hip_status(hipMalloc((void**)&dfom, sizeof(float) * nv), msg, "hipMalloc");
    if (!r) r = detail::hip_status(hipMemcpy(dplanes, planes.data(), sizeof(float) * 5 * n3,
                                             hipMemcpyHostToDevice), msg, "hipMemcpy");
    if (!r) {
        r = bsx_recon_set_accumulators(h, 0, dplanes, count, nullptr);
        if (!r && rows) {
            std::vector<double> tab((size_t)9 * params.size);
            int32_t nrows = 0;
            r = bsx_recon_stats(h, 0, 1, params.hi_res, tab.data(), params.size, &nrows, nullptr,
                                nullptr);
            if (!r) rows->assign(tab.begin(), tab.begin() + (size_t)9 * nrows);
        }
        if (!r) r = bsx_recon_finish(h, 0, 1, nullptr, dfom, dft, stats, nullptr);
        if (r && msg) *msg = bsx_last_error();
    }
    if (!r) r = detail::hip_status(hipMemcpy(out_cplx, dft, sizeof(float) * 2 * nv,
                                             hipMemcpyDeviceToHost), msg, "hipMemcpy");
    if (!r) r = detail::hip_status(hipMemcpy(out_fom, dfom, sizeof(float) * nv,
                                             hipMemcpyDeviceToHost), msg, "hipMemcpy");
    if (dplanes) (void)hipFree(dplanes);
    if (dft) (void)hipFree(dft);
    if (dfom) (void)hipFree(dfom);
    bsx_recon_destroy(h);
    return r;
}

}  // namespace bsx_dropin

#endif

// bsx_common.h — helpers shared by the reconstruction, extraction and search
// sources: electron wavelength, view -> rotation matrix, CTF aberration
// weights, the FSI interpolation kernel table and its device-side gather.
#pragma once

#include <cmath>
#include <vector>
#include <thread>
#include <algorithm>

#include "bsx_internal.h"

namespace bsx {

// CTFparam::lambda (include/ctf.h:309-317) with the constants of
// include/utilities.h:55-58.
inline double lambda_of(double av) {
    const double PLANCK = 6.626070E-34, ECHARGE = 1.602177E-19;
    const double EMASS = 9.109384E-31, LIGHTSPEED = 299792458.0;
    if (av <= 0) return 0;
    return PLANCK * 1e10 /
           sqrt(2.0 * EMASS * ECHARGE * av *
                (1.0 + ECHARGE * av / (2.0 * EMASS * LIGHTSPEED * LIGHTSPEED)));
}

// Host loop over [0, n) split into contiguous ranges on up to 16 threads
// (f(begin, end, thread)); runs inline below `grain` items. Used for the
// per-particle record preparation, which is independent per item.
template <class F>
inline void host_parallel(int64_t n, int64_t grain, F f) {
    unsigned nt = std::thread::hardware_concurrency();
    nt = std::max(1u, std::min(nt, 16u));
    if (n < grain || nt == 1) {
        f((int64_t)0, n, 0u);
        return;
    }
    nt = (unsigned)std::min<int64_t>(nt, (n + grain - 1) / grain);
    std::vector<std::thread> pool;
    const int64_t per = (n + nt - 1) / nt;
    for (unsigned t = 1; t < nt; ++t) {
        const int64_t b = t * per, e = std::min(n, b + per);
        if (b < e) pool.emplace_back(f, b, e, t);
    }
    f((int64_t)0, std::min(n, per), 0u);
    for (auto& th : pool) th.join();
}

// View::quaternion + Matrix3::from_quaternion (include/View.h:131-157,
// include/Matrix3.h:30-42): row-major d[3*i+j].
__host__ __device__ inline void view_matrix(const double view[4], double d[9]) {
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

// View (x, y, z, angle) after View::normalize (include/View.h:110-116): a
// unit view vector ((0, 0, 1) when it vanishes), components below 1e-30
// set to 0, the angle brought to (-pi, pi] by subtracting an even multiple
// of pi (angle_set_negPI_to_PI, src/util/utilities.cpp:606-615) and zeroed
// below 1e-30.
__host__ __device__ inline void view_normalize4(double v[4]) {
    const double s = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (s < 1e-30) {
        v[0] = v[1] = 0;
        v[2] = 1;
    } else {
        v[0] /= s;
        v[1] /= s;
        v[2] /= s;
    }
    for (int i = 0; i < 3; ++i)
        if (fabs(v[i]) < 1e-30) v[i] = 0;
    double a = v[3];
    if (!(fabs(a) <= 1.7976931348623157e308)) {   // not finite
        a = 0;
    } else {
        const long half_turns = (long)(a / M_PI);            // truncated
        a -= (double)(half_turns + half_turns % 2) * M_PI;   // an even multiple of pi
    }
    v[3] = fabs(a) < 1e-30 ? 0 : a;
}

// Quaternions (s, x, y, z) as the reference's Quaternion class multiplies
// and rotates them (include/Quaternion.h:34-41, 117-124, 161-177).
struct Quat {
    double s, x, y, z;
};

// Rotation by `angle` about `axis` (normalised here); parts below 1e-30 set
// to 0.
__host__ __device__ inline Quat quat_axis_angle(const double axis[3], double angle) {
    const double n = sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
    const double half = angle / 2.0, sh = sin(half);
    Quat q{cos(half), axis[0] / n * sh, axis[1] / n * sh, axis[2] / n * sh};
    if (fabs(q.s) < 1e-30) q.s = 0;
    if (fabs(q.x) < 1e-30) q.x = 0;
    if (fabs(q.y) < 1e-30) q.y = 0;
    if (fabs(q.z) < 1e-30) q.z = 0;
    return q;
}

__host__ __device__ inline Quat quat_mul(const Quat& a, const Quat& b) {
    Quat r;
    r.s = a.s * b.s - (a.x * b.x + a.y * b.y + a.z * b.z);
    r.x = a.s * b.x + a.x * b.s + a.y * b.z - a.z * b.y;
    r.y = a.s * b.y + a.y * b.s + a.z * b.x - a.x * b.z;
    r.z = a.s * b.z + a.z * b.s + a.x * b.y - a.y * b.x;
    return r;
}

// q p q^-1, with q^-1 = conj(q) / |q|^2.
__host__ __device__ inline Quat quat_rotate(const Quat& q, const Quat& p) {
    const double n2 = q.s * q.s + q.x * q.x + q.y * q.y + q.z * q.z, d = 1 / n2;
    const Quat qi{q.s * d, -q.x * d, -q.y * d, -q.z * d};
    return quat_mul(quat_mul(q, p), qi);
}

// views_for_refinement (src/util/View.cpp:371-436): the 27 views around
// `view` for an angular step > 0. With v the normalised view vector,
// axis1 = v x (0, 0, 1) and axis2 = v x axis1 (each set to (1, ., .) /
// (., 1, .) when shorter than 0.001, then normalised); for a1, a2 in
// {-step, 0, step} (the reference's floating loop) the vector is turned by
// R(axis2, a2) R(axis1, a1), and for a3 in the same set the view carries the
// in-plane angle angle + a3. out: 27 x (x, y, z, angle) in that loop order,
// each normalised as a View; the centre (0, 0, 0) is entry 13.
__host__ __device__ inline void refine_view_grid(const double view[4], double step,
                                                 double* out) {
    double v[4] = {view[0], view[1], view[2], view[3]};
    view_normalize4(v);
    double ax1[3] = {v[1], -v[0], 0.0};                 // v x (0, 0, 1)
    if (sqrt(ax1[0] * ax1[0] + ax1[1] * ax1[1] + ax1[2] * ax1[2]) < 0.001) ax1[0] = 1;
    double n1 = sqrt(ax1[0] * ax1[0] + ax1[1] * ax1[1] + ax1[2] * ax1[2]);
    for (double& c : ax1) c /= n1;
    double ax2[3] = {v[1] * ax1[2] - v[2] * ax1[1], v[2] * ax1[0] - v[0] * ax1[2],
                     v[0] * ax1[1] - v[1] * ax1[0]};   // v x axis1
    if (sqrt(ax2[0] * ax2[0] + ax2[1] * ax2[1] + ax2[2] * ax2[2]) < 0.001) ax2[1] = 1;
    double n2 = sqrt(ax2[0] * ax2[0] + ax2[1] * ax2[1] + ax2[2] * ax2[2]);
    for (double& c : ax2) c /= n2;
    const Quat qv{0.0, v[0], v[1], v[2]};
    int i = 0;
    for (double a1 = -step; a1 <= step; a1 += step) {
        const Quat q1 = quat_axis_angle(ax1, a1);
        for (double a2 = -step; a2 <= step; a2 += step) {
            const Quat r = quat_rotate(quat_mul(quat_axis_angle(ax2, a2), q1), qv);
            for (double a3 = -step; a3 <= step; a3 += step) {
                if (i < 27) {
                    double* o = out + 4 * i;
                    o[0] = r.x;
                    o[1] = r.y;
                    o[2] = r.z;
                    o[3] = v[3] + a3;
                    view_normalize4(o);
                }
                ++i;
            }
        }
    }
}

// Matrix3::quaternion (include/Matrix3.h:129-166) and View(Quaternion&)
// (include/View.h:52-63): the view (x, y, z, angle) of a rotation matrix
// (row-major), angle in [-pi, pi].
__host__ __device__ inline void matrix_view(const double* d, double* view) {
    double q[4];
    const double tr = 1 + d[0] + d[4] + d[8];
    if (tr > 1) {
        const double t = 0.5 * sqrt(tr), u = 0.25 / t;
        q[0] = t; q[1] = (d[7] - d[5]) * u; q[2] = (d[2] - d[6]) * u; q[3] = (d[3] - d[1]) * u;
    } else if (d[0] > d[4] && d[0] > d[8]) {
        const double t = 0.5 * sqrt(1 + d[0] - d[4] - d[8]), u = 0.25 / t;
        q[1] = t; q[0] = (d[7] - d[5]) * u; q[2] = (d[1] + d[3]) * u; q[3] = (d[2] + d[6]) * u;
    } else if (d[4] > d[8]) {
        const double t = 0.5 * sqrt(1 - d[0] + d[4] - d[8]), u = 0.25 / t;
        q[2] = t; q[0] = (d[2] - d[6]) * u; q[1] = (d[1] + d[3]) * u; q[3] = (d[5] + d[7]) * u;
    } else {
        const double t = 0.5 * sqrt(1 - d[0] - d[4] + d[8]), u = 0.25 / t;
        q[3] = t; q[0] = (d[3] - d[1]) * u; q[1] = (d[2] + d[6]) * u; q[2] = (d[5] + d[7]) * u;
    }
    const double qn = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    for (double& x : q) x /= qn;
    double ang = atan2(q[3], q[0]);
    const double ca = cos(ang), sa = sin(ang);
    const double q2 = q[0] * q[0] + q[3] * q[3];
    const double f2 = 2 * sqrt(q2);
    if (q2 < 1e-30) ang = atan2(q[1], q[2]);
    double v0 = f2 * (q[2] * ca + q[1] * sa), v1 = f2 * (q[2] * sa - q[1] * ca), v2 = 2 * q2 - 1;
    const double r = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
    if (r > 0) { v0 /= r; v1 /= r; v2 /= r; }
    view[0] = v0;
    view[1] = v1;
    view[2] = v2;
    view[3] = remainder(2 * ang, 2 * M_PI);
}

// Even aberration weights of CTFparam (include/ctf.h:79-83, 321-335;
// src/util/ctf.cpp:21-57): chi = c0 + c20 s^2 + c22 s^2 cos2a + c2m2 s^2 sin2a
// + c40 s^4, with the particle's defocus, astigmatism and Cs.
struct CtfCoef {
    double c0, c20, c22, c2m2, c40;
};

inline CtfCoef ctf_coef(const bsx_particle& q) {
    double wl = lambda_of(q.volt);
    double t1 = M_PI_2 * wl * wl * wl, t2 = -M_PI * wl, dev = fabs(q.def_dev);
    CtfCoef c;
    c.c0 = -q.amp_shift;
    c.c20 = t2 * q.def_avg;
    c.c22 = t2 * dev * cos(2 * q.ast_angle);
    c.c2m2 = t2 * dev * sin(2 * q.ast_angle);
    c.c40 = t1 * q.cs;
    return c;
}

// chi at spatial frequency (fx, fy) in 1/A, FP64.
__host__ __device__ inline double ctf_chi(const CtfCoef& c, double fx, double fy) {
    const double s2 = fx * fx + fy * fy;
    double c2a = 1, s2a = 0;
    if (s2 > 0) {
        c2a = (fx * fx - fy * fy) / s2;
        s2a = 2 * fx * fy / s2;
    }
    double dphi = c.c0;
    dphi += c.c2m2 * s2 * s2a;
    dphi += c.c20 * s2;
    dphi += c.c22 * s2 * c2a;
    dphi += c.c40 * s2 * s2;
    return dphi;
}

// FSI interpolation kernel table. Its values are those of FSI_Kernel
// (include/FSI_Kernel.h:28-68), evaluated here from the closed form
//   K(x) = sin(pi x) / (wd sin(pi x / wd)) * cos(pi x / wd)^power
// at the tap distance x = frac + (wd - 1)/2 - j of tap j, for the (ndiv + 1)
// fractions frac = row / 512 (row = round(frac * 512) at lookup). At integer
// distances (rows 0 and 512) K takes its limit: 1 at x = 0, else 0.
// power < 0: the nearest-neighbour table, one tap of weight 1.
// Computed in double, stored FP32, row-major by fraction.
constexpr int FSI_NDIV = 512;

inline double fsi_weight(double x, int wd, int power) {
    const double r = std::nearbyint(x);
    if (std::fabs(x - r) < 1e-12) return r == 0 ? 1.0 : 0.0;
    const double a = M_PI * x / wd;
    return std::sin(M_PI * x) / (wd * std::sin(a)) * std::pow(std::cos(a), power);
}

// Writes the table into `t` ((FSI_NDIV + 1) x wd floats, or FSI_NDIV + 1
// when power < 0).
inline void fsi_table_into(float* t, int wd, int power) {
    if (power < 0) {
        for (int row = 0; row <= FSI_NDIV; ++row) t[row] = 1.0f;
        return;
    }
    const int centre = (wd - 1) / 2;
    for (int row = 0; row <= FSI_NDIV; ++row) {
        const double frac = (double)row / FSI_NDIV;
        for (int j = 0; j < wd; ++j)
            t[(size_t)row * wd + j] = (float)fsi_weight(frac + centre - j, wd, power);
    }
}

inline std::vector<float> fsi_table(int wd, int power) {
    std::vector<float> t((size_t)(FSI_NDIV + 1) * (power < 0 ? 1 : wd));
    fsi_table_into(t.data(), wd, power);
    return t;
}

// Periodic index in [0, N): one compare in the usual in-range case, a
// modulo otherwise.
__device__ __forceinline__ int wrap_index(int i, int N) {
    if ((unsigned)i < (unsigned)N) return i;
    const int r = i % N;
    return r < 0 ? r + N : r;
}

// Bimage::fspace_interpolate (src/img/Bimage_fspace.cpp:85-147): the value of
// an N^3 complex transform at voxel coordinate m by WD^3 kernel taps around
// floor(m) - (WD-1)/2, indices wrapped periodically (the reference wraps one
// step; the total wrap here keeps every index inside the volume for any m).
// K: the kernel table (in LDS); the x and y weights of a sample are read once.
// FP32 taps, per-plane sums added in double into (vr, vi).
template <int WD>
__device__ inline void fsi_sample(const float2* __restrict__ ft, int N, const double m[3],
                                  const float* K, double& vr, double& vi) {
    constexpr int HW = (WD - 1) / 2;
    const double fx = floor(m[0]), fy = floor(m[1]), fz = floor(m[2]);
    const int ikx = ((int)((m[0] - fx) * FSI_NDIV + 0.5)) * WD;
    const int iky = ((int)((m[1] - fy) * FSI_NDIV + 0.5)) * WD;
    const int ikz = ((int)((m[2] - fz) * FSI_NDIV + 0.5)) * WD;
    const int ix = (int)fx - HW, iy = (int)fy - HW, iz = (int)fz - HW;
    float kx[WD], ky[WD];
    int jxs[WD];
#pragma unroll
    for (int q = 0; q < WD; ++q) {
        kx[q] = K[ikx + q];
        ky[q] = K[iky + q];
        jxs[q] = wrap_index(ix + q, N);
    }
    vr = 0;
    vi = 0;
#pragma unroll 1
    for (int zz = 0; zz < WD; ++zz) {
        const long jz = wrap_index(iz + zz, N);
        const float kzz = K[ikz + zz];
        float pr = 0.f, pi = 0.f;
#pragma unroll
        for (int yy2 = 0; yy2 < WD; ++yy2) {
            const long jy = wrap_index(iy + yy2, N);
            const float w = ky[yy2] * kzz;
            const float2* row = ft + (jz * N + jy) * (long)N;
#pragma unroll
            for (int q = 0; q < WD; ++q) {
                const float kw = kx[q] * w;
                const float2 a = row[jxs[q]];
                pr = fmaf(a.x, kw, pr);
                pi = fmaf(a.y, kw, pi);
            }
        }
        vr += pr;
        vi += pi;
    }
}

}  // namespace bsx

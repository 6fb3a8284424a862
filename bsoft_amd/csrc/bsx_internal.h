// bsx_internal.h — shared declarations of libbsx (HIP, gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/bsx.h"

namespace bsx {

// ---- error handling: thread-local last error, status codes, no throws ----
void set_error(const char* fmt, ...);
const char* last_error();

#define BSX_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t _e = (call);                                                         \
        if (_e != hipSuccess) {                                                         \
            ::bsx::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,                 \
                             hipGetErrorString(_e));                                    \
            return BSX_ERR_DEVICE;                                                      \
        }                                                                               \
    } while (0)

#define BSX_REQUIRE(cond, code, ...)                                                    \
    do {                                                                                \
        if (!(cond)) {                                                                  \
            ::bsx::set_error(__VA_ARGS__);                                              \
            return (code);                                                              \
        }                                                                               \
    } while (0)

// ---- FFT launchers (bsx_fft.hip) ----
// Batched in-place C2C transform along one axis of a (nz, ny, nx) complex
// array, `batch` arrays of nx*ny*nz each. axis 0 = x, 1 = y, 2 = z.
// sign -1 forward, +1 backward. scale multiplies every output.
int fft_axis(float2* data, int nx, int ny, int nz, int64_t batch, int axis, int sign,
             float scale, hipStream_t s);
bool fft_length_supported(int n);

// Per-particle transform stages (bsx_recon.hip).
struct PartRec;   // device record of one particle (see bsx_recon.hip)

// Twiddle table for length n, sign-independent: tw[a] = (cos 2pi a/n, sin 2pi a/n).
const float2* twiddles(int n, hipStream_t s);

}  // namespace bsx

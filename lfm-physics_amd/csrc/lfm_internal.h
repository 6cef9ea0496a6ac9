// Internal helpers shared by the HIP translation units of liblfm_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <cstdio>

#include "lfm_hip.h"

namespace lfm {

// Per-thread last-error message (lfm_last_error).
void set_error(const char* fmt, ...);

#define LFM_HIP_TRY(expr)                                                        \
    do {                                                                         \
        hipError_t e_ = (expr);                                                  \
        if (e_ != hipSuccess) {                                                  \
            ::lfm::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                             __FILE__, __LINE__);                                \
            return LFM_E_HIP;                                                    \
        }                                                                        \
    } while (0)

// Launch-error check after a kernel launch.
#define LFM_LAUNCH_CHECK(name)                                                   \
    do {                                                                         \
        hipError_t e_ = hipGetLastError();                                       \
        if (e_ != hipSuccess) {                                                  \
            ::lfm::set_error("launch of %s failed: %s", name, hipGetErrorString(e_)); \
            return LFM_E_HIP;                                                    \
        }                                                                        \
    } while (0)

// True when the byte ranges [a, a+na) and [b, b+nb) overlap.
inline bool overlaps(const void* a, size_t na, const void* b, size_t nb) {
    if (!a || !b) return false;
    auto pa = reinterpret_cast<uintptr_t>(a), pb = reinterpret_cast<uintptr_t>(b);
    return pa < pb + nb && pb < pa + na;
}

// View of an S_a plan for kernels in other translation units (lfm_scv.hip).
struct ScvView {
    float2* spec;        // half spectra [a][i][j][kk], row pitch Hp
    const float2* tw;    // exp(-2 pi i m / N), m < N (direct method only)
    int N, nc, Hp, method;
};
// False when `plan` is NULL.
bool scv_view(void* plan, ScvView* v);
// Direct method: spectrum := the S_a filter applied along j and i to the
// k-axis half spectra of |Psi_a|^2 (pr, pi of nc colours), i.e. the k-axis
// spectra of S_a.  Runs |Psi|^2, the k pass and the three column passes.
int scv_prepare(void* plan, const float* pr, const float* pi, float gamma, float D,
                hipStream_t st);
// Direct method: the three column passes (j forward, i forward / filter /
// inverse, j inverse) on the spectrum in place.
int scv_columns(void* plan, float gamma, float D, hipStream_t st);

}  // namespace lfm

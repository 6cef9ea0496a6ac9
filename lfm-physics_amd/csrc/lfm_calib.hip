// Bandwidth calibration: a pure streaming kernel with the same number of
// planar float32 read and write streams as a step kernel, so a step kernel's
// achieved HBM rate can be set against what that access pattern reaches on
// this device with no stencil, halo or arithmetic in the way.
#include "lfm_internal.h"

namespace lfm {
namespace {

constexpr int kMaxStreams = 32;

struct StreamArgs {
    const float* in[kMaxStreams];
    float* out[kMaxStreams];
    int nin, nout;
    size_t n;
};

// R, W > 0: stream counts fixed at compile time (loads fully unrolled, as in
// a step kernel); R = W = 0: counts taken from A at run time.  V floats per
// lane per stream (1 or 2, as 4- or 8-byte accesses).
template <int R, int W, int V, bool NT>
__global__ __launch_bounds__(256) void stream_kernel(StreamArgs A) {
    typedef float T __attribute__((ext_vector_type(V)));
    const int nin = R > 0 ? R : A.nin, nout = W > 0 ? W : A.nout;
    const size_t nv = A.n / V;
    for (size_t t = size_t(blockIdx.x) * 256 + threadIdx.x; t < nv; t += size_t(gridDim.x) * 256) {
        T s = T(0.0f);
#pragma unroll
        for (int r = 0; r < (R > 0 ? R : kMaxStreams); ++r) {
            if (R == 0 && r >= nin) break;
            const T* p = reinterpret_cast<const T*>(A.in[r]) + t;
            s += NT ? __builtin_nontemporal_load(p) : *p;
        }
#pragma unroll
        for (int w = 0; w < (W > 0 ? W : kMaxStreams); ++w) {
            if (W == 0 && w >= nout) break;
            const T v = s * float(w + 1);
            T* p = reinterpret_cast<T*>(A.out[w]) + t;
            if (NT) __builtin_nontemporal_store(v, p);
            else *p = v;
        }
    }
}

template <int V, bool NT>
void launch_stream(const StreamArgs& A, int grid, hipStream_t st) {
    const dim3 g(grid), b(256);
    const int key = A.nin * 100 + A.nout;
    switch (key) {
        case 1810: hipLaunchKernelGGL((stream_kernel<18, 10, V, NT>), g, b, 0, st, A); break;
        case 1610: hipLaunchKernelGGL((stream_kernel<16, 10, V, NT>), g, b, 0, st, A); break;
        case 603: hipLaunchKernelGGL((stream_kernel<6, 3, V, NT>), g, b, 0, st, A); break;
        case 703: hipLaunchKernelGGL((stream_kernel<7, 3, V, NT>), g, b, 0, st, A); break;
        case 502: hipLaunchKernelGGL((stream_kernel<5, 2, V, NT>), g, b, 0, st, A); break;
        case 1003: hipLaunchKernelGGL((stream_kernel<10, 3, V, NT>), g, b, 0, st, A); break;
        case 101: hipLaunchKernelGGL((stream_kernel<1, 1, V, NT>), g, b, 0, st, A); break;
        case 606: hipLaunchKernelGGL((stream_kernel<6, 6, V, NT>), g, b, 0, st, A); break;
        case 706: hipLaunchKernelGGL((stream_kernel<7, 6, V, NT>), g, b, 0, st, A); break;
        case 504: hipLaunchKernelGGL((stream_kernel<5, 4, V, NT>), g, b, 0, st, A); break;
        default: hipLaunchKernelGGL((stream_kernel<0, 0, V, NT>), g, b, 0, st, A); break;
    }
}

}  // namespace
}  // namespace lfm

using namespace lfm;

extern "C" int lfm_calib_stream_f32(const float* const* in, int32_t nin, float* const* out,
                                    int32_t nout, int64_t n, int32_t flags, int32_t grid,
                                    void* stream) {
    if (!in || !out || nin < 0 || nout < 1 || nin > kMaxStreams || nout > kMaxStreams || n < 1 ||
        grid < 0 || (flags & ~3)) {
        set_error("lfm_calib_stream_f32: bad argument (1 <= nout, nin <= %d, flags in 0..3, grid >= 0)",
                  kMaxStreams);
        return LFM_E_ARG;
    }
    const bool vec2 = (flags & LFM_CALIB_VEC2) != 0;
    if (vec2 && (n & 1)) {
        set_error("lfm_calib_stream_f32: two floats per lane need an even n");
        return LFM_E_ARG;
    }
    StreamArgs A{};
    for (int r = 0; r < nin; ++r) {
        if (!in[r]) { set_error("lfm_calib_stream_f32: input %d is NULL", r); return LFM_E_ARG; }
        if (vec2 && (reinterpret_cast<uintptr_t>(in[r]) & 7)) {
            set_error("lfm_calib_stream_f32: input %d is not 8-byte aligned", r);
            return LFM_E_ARG;
        }
        A.in[r] = in[r];
    }
    for (int w = 0; w < nout; ++w) {
        if (!out[w]) { set_error("lfm_calib_stream_f32: output %d is NULL", w); return LFM_E_ARG; }
        if (vec2 && (reinterpret_cast<uintptr_t>(out[w]) & 7)) {
            set_error("lfm_calib_stream_f32: output %d is not 8-byte aligned", w);
            return LFM_E_ARG;
        }
        A.out[w] = out[w];
    }
    A.nin = nin;
    A.nout = nout;
    A.n = size_t(n);
    const int g = grid > 0 ? grid : 16384;
    const hipStream_t st = (hipStream_t)stream;
    const bool nt = (flags & LFM_CALIB_STREAMING) != 0;
    if (vec2) nt ? launch_stream<2, true>(A, g, st) : launch_stream<2, false>(A, g, st);
    else nt ? launch_stream<1, true>(A, g, st) : launch_stream<1, false>(A, g, st);
    LFM_LAUNCH_CHECK("stream_kernel");
    return LFM_OK;
}

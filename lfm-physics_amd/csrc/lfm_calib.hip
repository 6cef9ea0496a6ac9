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
// a step kernel); R = W = 0: counts taken from A at run time.
template <int R, int W, bool NT>
__global__ __launch_bounds__(256) void stream_kernel(StreamArgs A) {
    const int nin = R > 0 ? R : A.nin, nout = W > 0 ? W : A.nout;
    for (size_t t = size_t(blockIdx.x) * 256 + threadIdx.x; t < A.n; t += size_t(gridDim.x) * 256) {
        float s = 0.0f;
#pragma unroll
        for (int r = 0; r < (R > 0 ? R : kMaxStreams); ++r) {
            if (R == 0 && r >= nin) break;
            s += NT ? __builtin_nontemporal_load(A.in[r] + t) : A.in[r][t];
        }
#pragma unroll
        for (int w = 0; w < (W > 0 ? W : kMaxStreams); ++w) {
            if (W == 0 && w >= nout) break;
            const float v = s * float(w + 1);
            if (NT) __builtin_nontemporal_store(v, A.out[w] + t);
            else A.out[w][t] = v;
        }
    }
}

template <bool NT>
void launch_stream(const StreamArgs& A, hipStream_t st) {
    const dim3 g(16384), b(256);
    const int key = A.nin * 100 + A.nout;
    switch (key) {
        case 1810: hipLaunchKernelGGL((stream_kernel<18, 10, NT>), g, b, 0, st, A); break;
        case 603: hipLaunchKernelGGL((stream_kernel<6, 3, NT>), g, b, 0, st, A); break;
        case 703: hipLaunchKernelGGL((stream_kernel<7, 3, NT>), g, b, 0, st, A); break;
        case 502: hipLaunchKernelGGL((stream_kernel<5, 2, NT>), g, b, 0, st, A); break;
        case 1003: hipLaunchKernelGGL((stream_kernel<10, 3, NT>), g, b, 0, st, A); break;
        case 101: hipLaunchKernelGGL((stream_kernel<1, 1, NT>), g, b, 0, st, A); break;
        case 606: hipLaunchKernelGGL((stream_kernel<6, 6, NT>), g, b, 0, st, A); break;
        case 706: hipLaunchKernelGGL((stream_kernel<7, 6, NT>), g, b, 0, st, A); break;
        case 504: hipLaunchKernelGGL((stream_kernel<5, 4, NT>), g, b, 0, st, A); break;
        default: hipLaunchKernelGGL((stream_kernel<0, 0, NT>), g, b, 0, st, A); break;
    }
}

}  // namespace
}  // namespace lfm

using namespace lfm;

extern "C" int lfm_calib_stream_f32(const float* const* in, int32_t nin, float* const* out,
                                    int32_t nout, int64_t n, int32_t streaming, void* stream) {
    if (!in || !out || nin < 0 || nout < 1 || nin > kMaxStreams || nout > kMaxStreams || n < 1) {
        set_error("lfm_calib_stream_f32: bad argument (1 <= nout, nin <= %d)", kMaxStreams);
        return LFM_E_ARG;
    }
    StreamArgs A{};
    for (int r = 0; r < nin; ++r) {
        if (!in[r]) { set_error("lfm_calib_stream_f32: input %d is NULL", r); return LFM_E_ARG; }
        A.in[r] = in[r];
    }
    for (int w = 0; w < nout; ++w) {
        if (!out[w]) { set_error("lfm_calib_stream_f32: output %d is NULL", w); return LFM_E_ARG; }
        A.out[w] = out[w];
    }
    A.nin = nin;
    A.nout = nout;
    A.n = size_t(n);
    if (streaming) launch_stream<true>(A, (hipStream_t)stream);
    else launch_stream<false>(A, (hipStream_t)stream);
    LFM_LAUNCH_CHECK("stream_kernel");
    return LFM_OK;
}

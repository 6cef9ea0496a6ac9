// Imaginary-time eigenmode relaxation (lfm/particles/solver.py:620-826):
//   E_new = E + dt (lap E - chi^2 E),   E_new *= amplitude / max|E_new|,
// in frozen chi, between Poisson updates of chi.  The stencil update folds
// max|E_new| into a device word; the rescale either runs as its own kernel
// (which, at convergence checks, also folds max|E_new - E| into another
// word) or is folded into the next update, which reads the unscaled values
// and multiplies each by amplitude / max as it loads it: the same float32
// products, one launch per step.  Relaxed lattices are small (the reference
// runs them at N = 64) and stay in the last-level cache, so the update
// gathers its 19 neighbours directly instead of staging tiles.
#include "lfm_internal.h"
#include "lfm_stencil.h"

namespace lfm {
namespace {

constexpr int kThreads = 256;
// A max "slot" is LFM_RELAX_SLOT_WORDS words: kLines words, one per 128-byte
// line, each taking the atomic max of every kLines-th workgroup, so the
// atomics of one launch spread over kLines addresses instead of queueing on
// one.  The reader takes the max of the kLines words.
constexpr int kLines = 64, kStride = 32;
static_assert(kLines * kStride == LFM_RELAX_SLOT_WORDS, "slot layout");

// Max of a slot's lines, in every thread (wave-wide reduction of the first
// 64 lanes' loads, broadcast through LDS).
__device__ __forceinline__ float slot_read(const uint32_t* slot) {
    __shared__ float bcast;
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    if (tid < 64) {
        float v = __uint_as_float(slot[tid * kStride]);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
        if (tid == 0) bcast = v;
    }
    __syncthreads();
    return bcast;
}

// Block-wide max of non-negative values, published with one atomic max on
// the float's bits (non-negative floats order as their unsigned bits) into
// the block's line of the slot.
__device__ __forceinline__ void block_max_publish(float v, uint32_t* slot) {
    __shared__ float wmax[kThreads / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;   // 64-wide rows: one wave per row
    const int w = tid >> 6;
    if ((tid & 63) == 0) wmax[w] = v;
    __syncthreads();
    if (tid == 0) {
        float m = wmax[0];
#pragma unroll
        for (int i = 1; i < kThreads / 64; ++i) m = fmaxf(m, wmax[i]);
        const unsigned b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
        atomicMax(slot + (b % kLines) * kStride, __float_as_uint(m));
    }
}

// amplitude / max formed in double and rounded once, as the reference's
// `E_new *= amplitude / E_max` applies it; 1 (no rescale) when max <= 1e-8
__device__ __forceinline__ float rescale_factor(float mx, float amplitude, bool& scale) {
    scale = double(mx) > 1e-8;
    return scale ? float(double(amplitude) / double(mx)) : 1.0f;
}

// One thread per site on a (k, j, i) grid of 64 x 4 x 1 workgroups, so the
// site coordinates come from the launch, not from integer division.

__global__ __launch_bounds__(kThreads) void relax_step_kernel(const float* __restrict__ E_raw,
                                                              const float* __restrict__ chi2,
                                                              float* __restrict__ out,
                                                              uint32_t* max_slot,
                                                              const uint32_t* scale_slot,
                                                              uint32_t* clear_slot,
                                                              float amplitude, int N, float dt) {
    const int k = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, i = blockIdx.z;
    bool scale = false;
    const float s = scale_slot ? rescale_factor(slot_read(scale_slot), amplitude, scale) : 1.0f;
    // E values as the previous step's rescale leaves them
    struct Rd {
        const float* p;
        float s;
        bool on;
        __device__ __forceinline__ float operator[](size_t t) const {
            const float v = p[t];
            return on ? mul_rn(v, s) : v;
        }
    } const E{E_raw, s, scale};
    float m = 0.0f;
    if (k < N && j < N) {
        const size_t nn = size_t(N) * N;
        const int jm = j == 0 ? N - 1 : j - 1, jp = j == N - 1 ? 0 : j + 1;
        const int km = k == 0 ? N - 1 : k - 1, kp = k == N - 1 ? 0 : k + 1;
        const size_t pl = size_t(i) * nn, pm = size_t(i == 0 ? N - 1 : i - 1) * nn,
                     pp = size_t(i == N - 1 ? 0 : i + 1) * nn;
        const size_t rm = size_t(jm) * N, rc = size_t(j) * N, rp = size_t(jp) * N;
        // in-plane partial of plane i, off-plane partials of planes i -+ 1
        const float c = E[pl + rc + k];
        const float faces = (E[pl + rm + k] + E[pl + rp + k]) + (E[pl + rc + km] + E[pl + rc + kp]);
        const float diag = (E[pl + rm + km] + E[pl + rm + kp]) + (E[pl + rp + km] + E[pl + rp + kp]);
        const float bm = lap_part_b(E[pm + rc + k], (E[pm + rm + k] + E[pm + rp + k]) +
                                                        (E[pm + rc + km] + E[pm + rc + kp]));
        const float bp = lap_part_b(E[pp + rc + k], (E[pp + rm + k] + E[pp + rp + k]) +
                                                        (E[pp + rc + km] + E[pp + rc + kp]));
        const float lap = (lap_part_a(faces, diag, c) + bm) + bp;
        // E + dt * (lap - chi2 * E), one rounding per operation as the
        // reference's elementwise array arithmetic
        const size_t t = pl + rc + k;
        const float v = add_rn(c, mul_rn(dt, sub_rn(lap, mul_rn(chi2[t], c))));
        out[t] = v;
        m = fabsf(v);
    }
    if (clear_slot && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.y == 0)
        clear_slot[threadIdx.x * kStride] = 0u;   // 64 lanes, one word per line
    block_max_publish(m, max_slot);
}

__global__ __launch_bounds__(kThreads) void relax_scale_kernel(float* __restrict__ E,
                                                               const float* __restrict__ E_prev,
                                                               const uint32_t* max_slot,
                                                               uint32_t* clear_slot,
                                                               uint32_t* diff_slot, float amplitude,
                                                               size_t n) {
    bool scale = false;
    const float s = rescale_factor(slot_read(max_slot), amplitude, scale);
    float d = 0.0f;
    for (size_t t = size_t(blockIdx.x) * kThreads + threadIdx.x; t < n;
         t += size_t(gridDim.x) * kThreads) {
        float v = E[t];
        if (scale) {
            v = mul_rn(v, s);
            E[t] = v;
        }
        if (E_prev) d = fmaxf(d, fabsf(sub_rn(v, E_prev[t])));
    }
    if (clear_slot && blockIdx.x == 0 && threadIdx.x < kLines) clear_slot[threadIdx.x * kStride] = 0u;
    if (diff_slot) block_max_publish(d, diff_slot);
}

unsigned grid_of(size_t n) {
    const size_t g = (n + kThreads - 1) / kThreads;
    return unsigned(g > 16384 ? 16384 : (g ? g : 1));
}

}  // namespace
}  // namespace lfm

using namespace lfm;

extern "C" int lfm_relax_step_f32(const float* E, const float* chi2, float* out, uint32_t* max_slot,
                                  const uint32_t* scale_slot, uint32_t* clear_slot, float amplitude,
                                  int32_t N, float dt, void* stream) {
    if (!E || !chi2 || !out || !max_slot || N < 3 || N > 2048) {
        set_error("lfm_relax_step_f32: bad argument (pointers non-NULL, 3 <= N <= 2048)");
        return LFM_E_ARG;
    }
    const size_t n3 = size_t(N) * N * N;
    if (overlaps(out, n3 * 4, E, n3 * 4) || overlaps(out, n3 * 4, chi2, n3 * 4)) {
        set_error("lfm_relax_step_f32: out overlaps an input");
        return LFM_E_ARG;
    }
    constexpr size_t SB = LFM_RELAX_SLOT_WORDS * 4;
    if ((scale_slot && overlaps(scale_slot, SB, max_slot, SB)) ||
        (clear_slot && (overlaps(clear_slot, SB, max_slot, SB) ||
                        (scale_slot && overlaps(clear_slot, SB, scale_slot, SB))))) {
        set_error("lfm_relax_step_f32: max_slot, scale_slot and clear_slot must be distinct");
        return LFM_E_ARG;
    }
    const dim3 grid(unsigned((N + 63) / 64), unsigned((N + 3) / 4), unsigned(N));
    hipLaunchKernelGGL(relax_step_kernel, grid, dim3(64, 4), 0, (hipStream_t)stream, E, chi2, out,
                       max_slot, scale_slot, clear_slot, amplitude, N, dt);
    LFM_LAUNCH_CHECK("relax_step_kernel");
    return LFM_OK;
}

extern "C" int lfm_relax_scale_f32(float* E, const float* E_prev, const uint32_t* max_slot,
                                   uint32_t* clear_slot, uint32_t* diff_slot, float amplitude,
                                   int64_t n, void* stream) {
    if (!E || !max_slot || n < 1 || (E_prev && !diff_slot) || (diff_slot && !E_prev)) {
        set_error("lfm_relax_scale_f32: bad argument (E_prev and diff_slot go together)");
        return LFM_E_ARG;
    }
    const auto apart = [](const void* a, const void* b) {
        return !overlaps(a, LFM_RELAX_SLOT_WORDS * 4, b, LFM_RELAX_SLOT_WORDS * 4);
    };
    if (clear_slot && (!apart(clear_slot, max_slot) || (diff_slot && !apart(clear_slot, diff_slot)))) {
        set_error("lfm_relax_scale_f32: clear_slot must differ from max_slot and diff_slot");
        return LFM_E_ARG;
    }
    hipLaunchKernelGGL(relax_scale_kernel, dim3(grid_of(size_t(n))), dim3(kThreads), 0,
                       (hipStream_t)stream, E, E_prev, max_slot, clear_slot, diff_slot, amplitude,
                       size_t(n));
    LFM_LAUNCH_CHECK("relax_scale_kernel");
    return LFM_OK;
}

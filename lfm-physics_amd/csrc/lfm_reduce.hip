// FP64 reductions for energy and chi metrics.
//
// energy: lfm/analysis/energy.py:26-135 --
//   T = 1/2 ((Psi - Psi_prev)/dt)^2, G = 1/2 c^2 sum_ax np.gradient(Psi, ax)^2
//   (second-order central interior, first-order one-sided edges, NOT periodic),
//   V = 1/2 chi^2 |Psi|^2, each summed over components and sites in float64.
// chi statistics: lfm/analysis/structure.py:26-108,148-171 over the interior
//   ball dist((i,j,k), N/2) <= (1-bf) N/2 (float64 distance, as the reference).
//
// Each grid-stride kernel accumulates per-thread float64 sums, reduces a
// wavefront with lane shuffles (64 lanes), then the workgroup through LDS,
// and writes one partial per workgroup.  A one-workgroup kernel sums the
// partials in a fixed order, so results are deterministic run to run.
#include <cfloat>
#include <mutex>

#include "lfm_internal.h"

namespace lfm {
namespace {

constexpr int RB = 256;          // threads per reduction workgroup
constexpr int RMAXB = 2048;      // maximum reduction workgroups
constexpr int RSLOTS = 8;        // doubles per partial record

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// op: 0 sum, 1 min, 2 max.  Returns the block result in thread 0.
template <int NV>
__device__ void block_reduce(double (&v)[NV], const int (&op)[NV]) {
    __shared__ double sh[NV][RB / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        v[q] = op[q] == 0 ? wave_sum(v[q]) : op[q] == 1 ? wave_min(v[q]) : wave_max(v[q]);
        if (lane == 0) sh[q][w] = v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double r = sh[q][0];
            for (int t = 1; t < RB / 64; ++t)
                r = op[q] == 0 ? r + sh[q][t] : op[q] == 1 ? fmin(r, sh[q][t]) : fmax(r, sh[q][t]);
            v[q] = r;
        }
    }
}

// np.gradient along one axis at position p of n, values at p-1, p, p+1.
__device__ __forceinline__ double grad1(int p, int n, double fm, double f0, double fp) {
    if (p == 0) return fp - f0;
    if (p == n - 1) return f0 - fm;
    return (fp - fm) / 2.0;
}

struct EArgs {
    const float* pr; const float* prp; const float* pi; const float* pip;
    const float* chi; double chi_off; const float* mask;
    int damp, ncomp, N;
    double dt, c;
};

__device__ double grad_sq_at(const float* F, int i, int j, int k, int N, size_t idx) {
    const size_t NN = size_t(N) * N;
    const double f0 = F[idx];
    const double xm = i > 0 ? F[idx - NN] : 0.0, xp = i < N - 1 ? F[idx + NN] : 0.0;
    const double ym = j > 0 ? F[idx - N] : 0.0, yp = j < N - 1 ? F[idx + N] : 0.0;
    const double zm = k > 0 ? F[idx - 1] : 0.0, zp = k < N - 1 ? F[idx + 1] : 0.0;
    const double gx = grad1(i, N, xm, f0, xp);
    const double gy = grad1(j, N, ym, f0, yp);
    const double gz = grad1(k, N, zm, f0, zp);
    return gx * gx + gy * gy + gz * gz;
}

__global__ __launch_bounds__(RB) void energy_partial(EArgs a, double* partials) {
    const int N = a.N;
    const size_t n3 = size_t(N) * N * N;
    double T = 0.0, G = 0.0, V = 0.0, S = 0.0;
    for (size_t idx = size_t(blockIdx.x) * RB + threadIdx.x; idx < n3;
         idx += size_t(gridDim.x) * RB) {
        const int i = int(idx / (size_t(N) * N)), j = int((idx / N) % N), k = int(idx % N);
        // prev as the reference stores it: float32 (1-m)*Psi_prev
        const float af = (a.mask && a.damp) ? 1.0f - a.mask[idx] : 1.0f;
        double psq = 0.0, t = 0.0, g = 0.0;
        for (int cc = 0; cc < a.ncomp; ++cc) {
            const size_t o = size_t(cc) * n3 + idx;
            {
                const double f = a.pr[o];
                const double fp = a.prp[o] * af;
                const double d = (f - fp) / a.dt;
                t += 0.5 * (d * d);
                g += grad_sq_at(a.pr + size_t(cc) * n3, i, j, k, N, idx);
                psq += f * f;
            }
            if (a.pi) {
                const double f = a.pi[o];
                const double fp = a.pip[o] * af;
                const double d = (f - fp) / a.dt;
                t += 0.5 * (d * d);
                g += grad_sq_at(a.pi + size_t(cc) * n3, i, j, k, N, idx);
                psq += f * f;
            }
        }
        const double x = a.chi_off + double(a.chi[idx]);
        T += t;
        G += 0.5 * a.c * a.c * g;
        V += 0.5 * (x * x) * psq;
        S += psq;
    }
    double v[4] = {T, G, V, S};
    const int op[4] = {0, 0, 0, 0};
    block_reduce<4>(v, op);
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; ++q) partials[size_t(blockIdx.x) * RSLOTS + q] = v[q];
}

__global__ __launch_bounds__(RB) void sum_partials(const double* partials, int nblocks,
                                                   int nvals, double* out) {
    // One workgroup: thread t sums partials t, t+RB, ... then a tree reduce.
    double v[RSLOTS];
    int op[RSLOTS];
    for (int q = 0; q < RSLOTS; ++q) { v[q] = 0.0; op[q] = 0; }
    for (int b = threadIdx.x; b < nblocks; b += RB)
        for (int q = 0; q < nvals; ++q) v[q] += partials[size_t(b) * RSLOTS + q];
    block_reduce<RSLOTS>(v, op);
    if (threadIdx.x == 0)
        for (int q = 0; q < nvals; ++q) out[q] = v[q];
}

__device__ __forceinline__ bool interior(int i, int j, int k, int N, double bf) {
    const double c = N / 2.0;
    const double dx = double(i) - c, dy = double(j) - c, dz = double(k) - c;
    return sqrt(dx * dx + dy * dy + dz * dz) <= c * (1.0 - bf);
}

// pass 1: count, sum, min, max, wells, voids
__global__ __launch_bounds__(RB) void chi_pass1(const float* chi, double off, int N, double bf,
                                                double wt, double vt, double* partials) {
    const size_t n3 = size_t(N) * N * N;
    double cnt = 0, sum = 0, mn = DBL_MAX, mx = -DBL_MAX, wl = 0, vd = 0;
    for (size_t idx = size_t(blockIdx.x) * RB + threadIdx.x; idx < n3;
         idx += size_t(gridDim.x) * RB) {
        const int i = int(idx / (size_t(N) * N)), j = int((idx / N) % N), k = int(idx % N);
        if (!interior(i, j, k, N, bf)) continue;
        const double x = off + double(chi[idx]);
        cnt += 1.0; sum += x; mn = fmin(mn, x); mx = fmax(mx, x);
        wl += x < wt ? 1.0 : 0.0;
        vd += x > vt ? 1.0 : 0.0;
    }
    double v[6] = {cnt, sum, mn, mx, wl, vd};
    const int op[6] = {0, 0, 1, 2, 0, 0};
    block_reduce<6>(v, op);
    if (threadIdx.x == 0)
        for (int q = 0; q < 6; ++q) partials[size_t(blockIdx.x) * RSLOTS + q] = v[q];
}

// finalize pass 1 -> out[0..6] (std filled later); one workgroup.
__global__ __launch_bounds__(RB) void chi_final1(const double* partials, int nblocks, double* out) {
    double v[6] = {0, 0, DBL_MAX, -DBL_MAX, 0, 0};
    const int op[6] = {0, 0, 1, 2, 0, 0};
    for (int b = threadIdx.x; b < nblocks; b += RB) {
        const double* p = partials + size_t(b) * RSLOTS;
        v[0] += p[0]; v[1] += p[1]; v[2] = fmin(v[2], p[2]); v[3] = fmax(v[3], p[3]);
        v[4] += p[4]; v[5] += p[5];
    }
    block_reduce<6>(v, op);
    if (threadIdx.x == 0) {
        const double cnt = v[0];
        out[0] = v[2];
        out[1] = v[3];
        out[2] = cnt > 0 ? v[1] / cnt : 0.0;
        out[3] = 0.0;
        out[4] = cnt > 0 ? v[4] / cnt : 0.0;
        out[5] = cnt > 0 ? v[5] / cnt : 0.0;
        out[6] = cnt;
    }
}

// pass 2: sum (x - mean)^2
__global__ __launch_bounds__(RB) void chi_pass2(const float* chi, double off, int N, double bf,
                                                const double* out, double* partials) {
    const size_t n3 = size_t(N) * N * N;
    const double mean = out[2];
    double s = 0.0;
    for (size_t idx = size_t(blockIdx.x) * RB + threadIdx.x; idx < n3;
         idx += size_t(gridDim.x) * RB) {
        const int i = int(idx / (size_t(N) * N)), j = int((idx / N) % N), k = int(idx % N);
        if (!interior(i, j, k, N, bf)) continue;
        const double d = (off + double(chi[idx])) - mean;
        s += d * d;
    }
    double v[1] = {s};
    const int op[1] = {0};
    block_reduce<1>(v, op);
    if (threadIdx.x == 0) partials[size_t(blockIdx.x) * RSLOTS] = v[0];
}

__global__ __launch_bounds__(RB) void chi_final2(const double* partials, int nblocks, double* out) {
    double v[1] = {0.0};
    const int op[1] = {0};
    for (int b = threadIdx.x; b < nblocks; b += RB) v[0] += partials[size_t(b) * RSLOTS];
    block_reduce<1>(v, op);
    if (threadIdx.x == 0) out[3] = out[6] > 0 ? sqrt(v[0] / out[6]) : 0.0;
}

int nblocks_for(int N) {
    const size_t n3 = size_t(N) * N * N;
    size_t b = (n3 + RB - 1) / RB;
    return int(b > RMAXB ? RMAXB : b);
}

// Partials of lfm_energy_f64_noalloc: one file-scope array per GPU, held
// until the process ends (RMAXB records of RSLOTS doubles = 128 KiB).  Host
// code never passes `scratch` itself (on the host it names a stand-in, not
// the device copy): each device's address comes from hipGetSymbolAddress,
// taken once per device with that device current, and only that pointer is
// handed to kernels.
constexpr int SCRATCH_DOUBLES = RMAXB * RSLOTS;
__device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_DEVICES] = {};

// The current device's copy of `scratch`; NULL (error set) on failure.
double* device_scratch() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
        set_error("lfm_energy_f64_noalloc: no current HIP device (or device index >= %d)",
                  MAX_DEVICES);
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            set_error("lfm_energy_f64_noalloc: hipGetSymbolAddress(scratch) on device %d failed: %s",
                      dev, hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = static_cast<double*>(ptr);
    }
    return g_scratch_ptr[dev];
}

}  // namespace
}  // namespace lfm

using namespace lfm;

extern "C" int64_t lfm_reduce_partials_len(int32_t N) {
    (void)N;
    return int64_t(RMAXB) * RSLOTS;
}

extern "C" int lfm_energy_f64(const float* pr, const float* pr_prev, const float* pi,
                              const float* pi_prev, const float* chi, double chi_offset,
                              const float* mask, int32_t damp_prev, int32_t ncomp, int32_t N,
                              double dt, double c, double* partials, double* out, void* stream) {
    if (!pr || !pr_prev || !chi || !partials || !out || ncomp < 1 || N < 2 || !(dt > 0.0)) {
        set_error("lfm_energy_f64: bad argument");
        return LFM_E_ARG;
    }
    if ((pi == nullptr) != (pi_prev == nullptr)) {
        set_error("lfm_energy_f64: pi and pi_prev must both be given or both NULL");
        return LFM_E_ARG;
    }
    EArgs a{pr, pr_prev, pi, pi_prev, chi, chi_offset, mask, damp_prev, ncomp, N, dt, c};
    hipStream_t st = (hipStream_t)stream;
    const int nb = nblocks_for(N);
    hipLaunchKernelGGL(energy_partial, dim3(nb), dim3(RB), 0, st, a, partials);
    LFM_LAUNCH_CHECK("energy_partial");
    hipLaunchKernelGGL(sum_partials, dim3(1), dim3(RB), 0, st, (const double*)partials, nb, 4, out);
    LFM_LAUNCH_CHECK("sum_partials");
    return LFM_OK;
}

extern "C" int lfm_energy_f64_noalloc(const float* pr, const float* pr_prev, const float* pi,
                                      const float* pi_prev, const float* chi, double chi_offset,
                                      const float* mask, int32_t damp_prev, int32_t ncomp,
                                      int32_t N, double dt, double c, double* out, void* stream) {
    // Same kernels and summation order as lfm_energy_f64, so the same bits;
    // the per-workgroup partials live in this device's `scratch` array.  Every
    // partial record is written by energy_partial before sum_partials reads
    // it, so nothing needs clearing between calls.
    if (N >= 2 && int64_t(nblocks_for(N)) * RSLOTS > SCRATCH_DOUBLES) {
        set_error("lfm_energy_f64_noalloc: N=%d needs %lld partial doubles, the static scratch "
                  "holds %d", N, (long long)(int64_t(nblocks_for(N)) * RSLOTS), SCRATCH_DOUBLES);
        return LFM_E_SIZE;
    }
    double* partials = device_scratch();
    if (!partials) return LFM_E_HIP;
    return lfm_energy_f64(pr, pr_prev, pi, pi_prev, chi, chi_offset, mask, damp_prev, ncomp, N, dt,
                          c, partials, out, stream);
}

extern "C" int lfm_chi_stats_f64(const float* chi, double chi_offset, int32_t N,
                                 double boundary_fraction,
                                 double well_threshold, double void_threshold, double* partials,
                                 double* out, void* stream) {
    if (!chi || !partials || !out || N < 2) {
        set_error("lfm_chi_stats_f64: bad argument");
        return LFM_E_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    const int nb = nblocks_for(N);
    hipLaunchKernelGGL(chi_pass1, dim3(nb), dim3(RB), 0, st, chi, chi_offset, N, boundary_fraction,
                       well_threshold, void_threshold, partials);
    LFM_LAUNCH_CHECK("chi_pass1");
    hipLaunchKernelGGL(chi_final1, dim3(1), dim3(RB), 0, st, (const double*)partials, nb, out);
    LFM_LAUNCH_CHECK("chi_final1");
    hipLaunchKernelGGL(chi_pass2, dim3(nb), dim3(RB), 0, st, chi, chi_offset, N, boundary_fraction,
                       (const double*)out, partials);
    LFM_LAUNCH_CHECK("chi_pass2");
    hipLaunchKernelGGL(chi_final2, dim3(1), dim3(RB), 0, st, (const double*)partials, nb, out);
    LFM_LAUNCH_CHECK("chi_final2");
    return LFM_OK;
}

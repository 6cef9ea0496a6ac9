// S_a Helmholtz transform for the SCV term (COLOR level).
//
//   S_a = irfftn( gamma / (gamma + D |k|^2) * rfftn(|Psi_a|^2) ),
//   k_d = 2 pi fftfreq(N)   (continuous |k|^2, not the stencil symbol)
//
// Reference: lfm/core/backends/numpy_backend.py:338-358 and
// cupy_backend.py:260-279, which rebuild the k-grid and the float64 filter on
// every step and run one transform pair per colour.  Here one batched
// (ncolors) single-precision rocFFT R2C / C2R plan pair is built once per
// (N, ncolors); the filter is evaluated on the fly by one k-space kernel that
// also folds in the 1/N^3 normalisation of the inverse transform.  The
// clip S_a >= 0 happens where S_a is consumed (the step kernel).
#include <rocfft/rocfft.h>

#include <mutex>

#include "lfm_internal.h"

namespace lfm {
namespace {

struct ScvPlan {
    int N = 0, nc = 0;
    rocfft_plan fwd = nullptr, inv = nullptr;
    rocfft_execution_info info = nullptr;
    void* work = nullptr;
    size_t work_bytes = 0;
    float2* spec = nullptr;  // nc * N * N * (N/2+1)
    float* e = nullptr;      // nc * N^3 (|Psi_a|^2 when computed here)
};

std::once_flag g_fft_once;
rocfft_status g_fft_setup = rocfft_status_success;

#define LFM_FFT_TRY(expr)                                                   \
    do {                                                                    \
        rocfft_status s_ = (expr);                                          \
        if (s_ != rocfft_status_success) {                                  \
            set_error("%s failed: rocfft status %d", #expr, (int)s_);       \
            return LFM_E_FFT;                                               \
        }                                                                   \
    } while (0)

__global__ __launch_bounds__(256) void psisq_kernel(const float* __restrict__ pr,
                                                    const float* __restrict__ pi,
                                                    float* __restrict__ e, size_t n) {
    for (size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
         t += size_t(gridDim.x) * blockDim.x) {
        const float r = pr[t], m = pi[t];
        e[t] = r * r + m * m;
    }
}

// Multiply the half spectrum (layout [a][i][j][kk], kk < N/2+1) by
// gamma/(gamma + D k^2) / N^3.
__global__ __launch_bounds__(256) void helmholtz_filter(float2* __restrict__ spec, int N,
                                                        int nc, float gamma, float D,
                                                        float inv_n3) {
    const int H = N / 2 + 1;
    const size_t per = size_t(N) * N * H;
    const size_t n = per * nc;
    const float two_pi_over_n = 6.283185307179586f / float(N);
    for (size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
         t += size_t(gridDim.x) * blockDim.x) {
        const size_t r = t % per;
        const int kk = int(r % H);
        const int j = int((r / H) % N);
        const int i = int(r / (size_t(H) * N));
        const int fi = i < (N + 1) / 2 ? i : i - N;
        const int fj = j < (N + 1) / 2 ? j : j - N;
        const float kx = two_pi_over_n * float(fi);
        const float ky = two_pi_over_n * float(fj);
        const float kz = two_pi_over_n * float(kk);
        const float k2 = kx * kx + ky * ky + kz * kz;
        const float h = gamma / (gamma + D * k2) * inv_n3;
        float2 v = spec[t];
        v.x *= h;
        v.y *= h;
        spec[t] = v;
    }
}

// rhs = kappa * (sum_c (pr_c^2 + pi_c^2) - e0_sq)   (GOV-04 source)
__global__ __launch_bounds__(256) void poisson_rhs(const float* __restrict__ pr,
                                                   const float* __restrict__ pi, int ncomp,
                                                   float kappa, float e0_sq,
                                                   float* __restrict__ rhs, size_t n3) {
    for (size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n3;
         t += size_t(gridDim.x) * blockDim.x) {
        float s = 0.0f;
        for (int c = 0; c < ncomp; ++c) {
            const float r = pr[c * n3 + t];
            s += r * r;
            if (pi) { const float m = pi[c * n3 + t]; s += m * m; }
        }
        rhs[t] = kappa * (s - e0_sq);
    }
}

// Multiply the half spectrum by -1/(N^3 k^2) with the DC term set to 0.
__global__ __launch_bounds__(256) void poisson_filter(float2* __restrict__ spec, int N,
                                                      float inv_n3) {
    const int H = N / 2 + 1;
    const size_t n = size_t(N) * N * H;
    const float two_pi_over_n = 6.283185307179586f / float(N);
    for (size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
         t += size_t(gridDim.x) * blockDim.x) {
        const int kk = int(t % H);
        const int j = int((t / H) % N);
        const int i = int(t / (size_t(H) * N));
        const int fi = i < (N + 1) / 2 ? i : i - N;
        const int fj = j < (N + 1) / 2 ? j : j - N;
        const float kx = two_pi_over_n * float(fi), ky = two_pi_over_n * float(fj);
        const float kz = two_pi_over_n * float(kk);
        const float k2 = kx * kx + ky * ky + kz * kz;
        const float h = t == 0 ? 0.0f : -inv_n3 / k2;
        float2 v = spec[t];
        v.x *= h;
        v.y *= h;
        spec[t] = v;
    }
}

// chi = chi0 + delta; outside the interior ball (dist > (1-bf) N/2, float64
// distance as lfm/analysis/structure.py:148-171) chi = chi0 when bf > 0.
__global__ __launch_bounds__(256) void poisson_finish(float* __restrict__ chi, int N, float chi0,
                                                      double bf) {
    const size_t n3 = size_t(N) * N * N;
    const double c = N / 2.0;
    for (size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n3;
         t += size_t(gridDim.x) * blockDim.x) {
        float v = chi0 + chi[t];
        if (bf > 0.0) {
            const int i = int(t / (size_t(N) * N)), j = int((t / N) % N), k = int(t % N);
            const double dx = double(i) - c, dy = double(j) - c, dz = double(k) - c;
            if (!(sqrt(dx * dx + dy * dy + dz * dz) <= c * (1.0 - bf))) v = chi0;
        }
        chi[t] = v;
    }
}

__global__ __launch_bounds__(256) void clip_kernel(const float* __restrict__ x,
                                                   float* __restrict__ y, size_t n) {
    for (size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
         t += size_t(gridDim.x) * blockDim.x) {
        const float v = x[t];
        y[t] = v > 0.0f ? v : 0.0f;
    }
}

unsigned grid_for(size_t n) {
    size_t g = (n + 255) / 256;
    return unsigned(g > 65536 ? 65536 : (g ? g : 1));
}

void destroy(ScvPlan* p) {
    if (!p) return;
    if (p->fwd) rocfft_plan_destroy(p->fwd);
    if (p->inv) rocfft_plan_destroy(p->inv);
    if (p->info) rocfft_execution_info_destroy(p->info);
    if (p->work) (void)hipFree(p->work);
    if (p->spec) (void)hipFree(p->spec);
    if (p->e) (void)hipFree(p->e);
    delete p;
}

int run(ScvPlan* p, const float* e, float gamma, float D, float* sa_out, hipStream_t st) {
    LFM_FFT_TRY(rocfft_execution_info_set_stream(p->info, st));
    void* in_f[1] = {const_cast<float*>(e)};
    void* out_f[1] = {p->spec};
    LFM_FFT_TRY(rocfft_execute(p->fwd, in_f, out_f, p->info));
    const size_t ns = size_t(p->nc) * p->N * p->N * (p->N / 2 + 1);
    const double n3 = double(p->N) * p->N * p->N;
    hipLaunchKernelGGL(helmholtz_filter, dim3(grid_for(ns)), dim3(256), 0, st, p->spec, p->N,
                       p->nc, gamma, D, float(1.0 / n3));
    LFM_LAUNCH_CHECK("helmholtz_filter");
    void* in_i[1] = {p->spec};
    void* out_i[1] = {sa_out};
    LFM_FFT_TRY(rocfft_execute(p->inv, in_i, out_i, p->info));
    return LFM_OK;
}

}  // namespace
}  // namespace lfm

using namespace lfm;

extern "C" int lfm_scv_plan_create(int32_t N, int32_t ncolors, void** plan) {
    if (!plan) { set_error("lfm_scv_plan_create: plan is NULL"); return LFM_E_ARG; }
    *plan = nullptr;
    if (N < 8 || N > 2048 || ncolors < 1 || ncolors > 8) {
        set_error("lfm_scv_plan_create: N=%d ncolors=%d out of range", N, ncolors);
        return LFM_E_ARG;
    }
    std::call_once(g_fft_once, [] { g_fft_setup = rocfft_setup(); });
    if (g_fft_setup != rocfft_status_success) {
        set_error("rocfft_setup failed: status %d", (int)g_fft_setup);
        return LFM_E_FFT;
    }
    auto* p = new ScvPlan();
    p->N = N;
    p->nc = ncolors;
    // rocFFT lengths are fastest-first: (k, j, i) for the C-order (i, j, k) array.
    const size_t len[3] = {size_t(N), size_t(N), size_t(N)};
    rocfft_status s = rocfft_plan_create(&p->fwd, rocfft_placement_notinplace,
                                         rocfft_transform_type_real_forward,
                                         rocfft_precision_single, 3, len, size_t(ncolors), nullptr);
    if (s == rocfft_status_success)
        s = rocfft_plan_create(&p->inv, rocfft_placement_notinplace,
                               rocfft_transform_type_real_inverse, rocfft_precision_single, 3, len,
                               size_t(ncolors), nullptr);
    size_t wf = 0, wi = 0;
    if (s == rocfft_status_success) s = rocfft_plan_get_work_buffer_size(p->fwd, &wf);
    if (s == rocfft_status_success) s = rocfft_plan_get_work_buffer_size(p->inv, &wi);
    if (s == rocfft_status_success) s = rocfft_execution_info_create(&p->info);
    if (s != rocfft_status_success) {
        set_error("lfm_scv_plan_create: rocFFT plan setup failed (status %d)", (int)s);
        destroy(p);
        return LFM_E_FFT;
    }
    p->work_bytes = wf > wi ? wf : wi;
    hipError_t he = hipSuccess;
    if (p->work_bytes) he = hipMalloc(&p->work, p->work_bytes);
    const size_t ns = size_t(ncolors) * N * N * (N / 2 + 1);
    if (he == hipSuccess) he = hipMalloc((void**)&p->spec, ns * sizeof(float2));
    if (he == hipSuccess) he = hipMalloc((void**)&p->e, size_t(ncolors) * N * N * N * sizeof(float));
    if (he != hipSuccess) {
        set_error("lfm_scv_plan_create: hipMalloc failed: %s", hipGetErrorString(he));
        destroy(p);
        return LFM_E_HIP;
    }
    if (p->work_bytes) {
        s = rocfft_execution_info_set_work_buffer(p->info, p->work, p->work_bytes);
        if (s != rocfft_status_success) {
            set_error("rocfft_execution_info_set_work_buffer failed: status %d", (int)s);
            destroy(p);
            return LFM_E_FFT;
        }
    }
    *plan = p;
    return LFM_OK;
}

extern "C" int lfm_scv_plan_destroy(void* plan) {
    destroy(static_cast<ScvPlan*>(plan));
    return LFM_OK;
}

extern "C" int lfm_scv_apply_psisq(void* plan, const float* e, float gamma, float D,
                                   float* sa_out, void* stream) {
    auto* p = static_cast<ScvPlan*>(plan);
    if (!p || !e || !sa_out) { set_error("lfm_scv_apply_psisq: NULL argument"); return LFM_E_ARG; }
    if (!(gamma > 0.0f) || D < 0.0f) {
        set_error("lfm_scv_apply_psisq: need gamma > 0 and D >= 0 (got %g, %g)", gamma, D);
        return LFM_E_ARG;
    }
    return run(p, e, gamma, D, sa_out, (hipStream_t)stream);
}

extern "C" int lfm_scv_apply(void* plan, const float* pr, const float* pi, float gamma, float D,
                             float* sa_out, void* stream) {
    auto* p = static_cast<ScvPlan*>(plan);
    if (!p || !pr || !pi || !sa_out) { set_error("lfm_scv_apply: NULL argument"); return LFM_E_ARG; }
    if (!(gamma > 0.0f) || D < 0.0f) {
        set_error("lfm_scv_apply: need gamma > 0 and D >= 0 (got %g, %g)", gamma, D);
        return LFM_E_ARG;
    }
    const size_t n = size_t(p->nc) * p->N * p->N * p->N;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(psisq_kernel, dim3(grid_for(n)), dim3(256), 0, st, pr, pi, p->e, n);
    LFM_LAUNCH_CHECK("psisq_kernel");
    return run(p, p->e, gamma, D, sa_out, st);
}

extern "C" int lfm_clip_nonneg_f32(const float* x, float* y, int64_t n, void* stream) {
    if (!x || !y || n < 0) { set_error("lfm_clip_nonneg_f32: bad argument"); return LFM_E_ARG; }
    if (n == 0) return LFM_OK;
    hipLaunchKernelGGL(clip_kernel, dim3(grid_for(size_t(n))), dim3(256), 0, (hipStream_t)stream,
                       x, y, size_t(n));
    LFM_LAUNCH_CHECK("clip_kernel");
    return LFM_OK;
}

extern "C" int lfm_poisson_chi_f32(void* plan, const float* pr, const float* pi, int32_t ncomp,
                                   float kappa, float e0_sq, float chi0,
                                   double boundary_fraction, float* chi_out, void* stream) {
    auto* p = static_cast<ScvPlan*>(plan);
    if (!p || !pr || !chi_out || ncomp < 1) {
        set_error("lfm_poisson_chi_f32: bad argument");
        return LFM_E_ARG;
    }
    if (p->nc != 1) {
        set_error("lfm_poisson_chi_f32: needs a plan created with ncolors = 1 (got %d)", p->nc);
        return LFM_E_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    const size_t n3 = size_t(p->N) * p->N * p->N;
    hipLaunchKernelGGL(poisson_rhs, dim3(grid_for(n3)), dim3(256), 0, st, pr, pi, ncomp, kappa,
                       e0_sq, p->e, n3);
    LFM_LAUNCH_CHECK("poisson_rhs");
    LFM_FFT_TRY(rocfft_execution_info_set_stream(p->info, st));
    void* in_f[1] = {p->e};
    void* out_f[1] = {p->spec};
    LFM_FFT_TRY(rocfft_execute(p->fwd, in_f, out_f, p->info));
    const size_t ns = size_t(p->N) * p->N * (p->N / 2 + 1);
    hipLaunchKernelGGL(poisson_filter, dim3(grid_for(ns)), dim3(256), 0, st, p->spec, p->N,
                       float(1.0 / double(n3)));
    LFM_LAUNCH_CHECK("poisson_filter");
    void* in_i[1] = {p->spec};
    void* out_i[1] = {chi_out};
    LFM_FFT_TRY(rocfft_execute(p->inv, in_i, out_i, p->info));
    hipLaunchKernelGGL(poisson_finish, dim3(grid_for(n3)), dim3(256), 0, st, chi_out, p->N, chi0,
                       boundary_fraction);
    LFM_LAUNCH_CHECK("poisson_finish");
    return LFM_OK;
}

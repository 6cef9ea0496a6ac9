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
    int method = 0;          // 1: rocFFT 3-D plans; 2: 1-D rocFFT along k + column passes
    rocfft_plan fwd = nullptr, inv = nullptr;
    rocfft_execution_info info = nullptr;
    void* work = nullptr;
    size_t work_bytes = 0;
    float2* spec = nullptr;  // nc * N * N * (N/2+1)
    float* e = nullptr;      // nc * N^3 (|Psi_a|^2 when computed here)
    float2* tw = nullptr;    // exp(-2 pi i m / N), m < N (column passes)
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

// ---------------------------------------------------------------------------
// Column passes: 1-D FFTs of length N along the j or the i axis of the half
// spectrum [a][i][j][kk] (kk < H = N/2+1 contiguous).  A workgroup loads
// NCOL adjacent kk columns of one (a, i) or (a, j) slab (NCOL contiguous
// complex values per row, rows `rs` apart) into LDS, runs a Stockham
// autosort FFT per column (radix-8 stages, one radix-2/4 stage when log2 N is
// not a multiple of 3; every thread owns 8 values per stage) and writes the
// columns back.  The fused mode transforms along i, multiplies by the
// spectral filter and transforms back, so the i axis is read and written
// once for the forward transform, the filter and the inverse transform.
// ---------------------------------------------------------------------------
constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by -i (forward, sign = -1) or +i (inverse, sign = +1)
__device__ __forceinline__ float2 rot(float2 a, int sign) {
    return sign < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

template <int R>
__device__ __forceinline__ void dft(float2 (&x)[R], int sign) {
    if constexpr (R == 2) {
        const float2 a = x[0], b = x[1];
        x[0] = cadd(a, b);
        x[1] = csub(a, b);
    } else if constexpr (R == 4) {
        const float2 t0 = cadd(x[0], x[2]), t1 = csub(x[0], x[2]);
        const float2 t2 = cadd(x[1], x[3]), t3 = rot(csub(x[1], x[3]), sign);
        x[0] = cadd(t0, t2);
        x[2] = csub(t0, t2);
        x[1] = cadd(t1, t3);
        x[3] = csub(t1, t3);
    } else {
        const float c = 0.70710678118654752f;
        const float2 t0 = cadd(x[0], x[4]), t1 = csub(x[0], x[4]);
        const float2 t2 = cadd(x[2], x[6]), t3 = rot(csub(x[2], x[6]), sign);
        const float2 t4 = cadd(x[1], x[5]), t5 = csub(x[1], x[5]);
        const float2 t6 = cadd(x[3], x[7]), t7 = rot(csub(x[3], x[7]), sign);
        const float2 e0 = cadd(t0, t2), e2 = csub(t0, t2), e1 = cadd(t1, t3), e3 = csub(t1, t3);
        const float2 o0 = cadd(t4, t6), o2 = csub(t4, t6), o1 = cadd(t5, t7), o3 = csub(t5, t7);
        // W8^1 = (c, sign c), W8^2 = sign i, W8^3 = (-c, sign c)
        const float2 w1o1 = make_float2(c * (o1.x - sign * o1.y), c * (o1.y + sign * o1.x));
        const float2 w2o2 = rot(o2, sign);
        const float2 w3o3 = make_float2(-c * (o3.x + sign * o3.y), c * (sign * o3.x - o3.y));
        x[0] = cadd(e0, o0);
        x[4] = csub(e0, o0);
        x[1] = cadd(e1, w1o1);
        x[5] = csub(e1, w1o1);
        x[2] = cadd(e2, w2o2);
        x[6] = csub(e2, w2o2);
        x[3] = cadd(e3, w3o3);
        x[7] = csub(e3, w3o3);
    }
}

// One Stockham stage of radix R on one column `col` (length N, LDS stride
// CS between columns) by thread t of TC = N/8 column threads.
template <int N, int R, int CS>
__device__ __forceinline__ void stage(float2* lds, int col, int t, int Ns, int sign,
                                      const float2* __restrict__ tw) {
    constexpr int B = N / R;           // butterflies in the stage
    constexpr int TC = N / 8;          // threads per column
    constexpr int NB = B / TC;         // butterflies per thread (8 / R)
    float2 x[NB][R];
    float2* c = lds + col * CS;
#pragma unroll
    for (int u = 0; u < NB; ++u) {
        const int j = t + u * TC;
        const int k = j % Ns;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            float2 v = c[j + r * B];
            if (r > 0 && Ns > 1) {
                float2 w = tw[(r * k * (N / (Ns * R))) & (N - 1)];
                if (sign > 0) w.y = -w.y;
                v = cmul(v, w);
            }
            x[u][r] = v;
        }
        dft<R>(x[u], sign);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < NB; ++u) {
        const int j = t + u * TC;
        const int k = j % Ns;
        const int base = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) c[base + r * Ns] = x[u][r];
    }
    __syncthreads();
}

template <int N, int CS>
__device__ __forceinline__ void fft_col(float2* lds, int col, int t, int sign,
                                        const float2* __restrict__ tw) {
    constexpr int L = ilog2(N);
    constexpr int n8 = L / 3;
    constexpr int rem = L % 3;
    int Ns = 1;
#pragma unroll
    for (int st = 0; st < n8; ++st) {
        stage<N, 8, CS>(lds, col, t, Ns, sign, tw);
        Ns *= 8;
    }
    if constexpr (rem == 1) stage<N, 2, CS>(lds, col, t, Ns, sign, tw);
    if constexpr (rem == 2) stage<N, 4, CS>(lds, col, t, Ns, sign, tw);
}

struct ColArgs {
    float2* spec;
    const float2* tw;
    int nc;
    long long slab_stride;   // complex elements between consecutive slab indices x
    long long row_stride;    // complex elements between consecutive rows (FFT axis)
    int mode;                // 0 forward, 1 inverse, 2 forward + filter + inverse (i axis)
    int filter;              // 0 Helmholtz gamma/(gamma + D k^2), 1 Poisson -1/k^2
    float gamma, D, scale;
};

template <int N>
constexpr int ncol_for() { return N >= 1024 ? 8 : 16; }

template <int N>
__global__ __launch_bounds__(ncol_for<N>() * (N / 8)) void fft_columns(ColArgs A) {
    constexpr int NCOL = ncol_for<N>();
    constexpr int TC = N / 8;
    constexpr int H = N / 2 + 1;
    constexpr int CS = N + 1;                 // LDS column stride (odd: spreads banks)
    __shared__ float2 lds[NCOL * CS];
    constexpr int NCB = (H + NCOL - 1) / NCOL;
    const int cb = blockIdx.x % NCB;
    const int slab = blockIdx.x / NCB;         // a * N + x
    const int a = slab / N, x = slab % N;
    const int kk0 = cb * NCOL;
    const int nv = min(NCOL, H - kk0);
    float2* base = A.spec + (long long)a * N * N * H + (long long)x * A.slab_stride + kk0;
    const int tid = threadIdx.x;
    constexpr int NT = NCOL * TC;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int e = tid + q * NT;
        const int row = e / NCOL, col = e % NCOL;
        lds[col * CS + row] = col < nv ? base[(long long)row * A.row_stride + col] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    const int col = tid / TC, t = tid % TC;
    const int sign = A.mode == 1 ? 1 : -1;
    fft_col<N, CS>(lds, col, t, sign, A.tw);
    if (A.mode == 2) {
        // multiply by the filter at (k_i = row, k_j = x, k_kk = kk0 + col)
        const float two_pi_over_n = 6.283185307179586f / float(N);
        const int fj = x < (N + 1) / 2 ? x : x - N;
        const float kx2 = (two_pi_over_n * fj) * (two_pi_over_n * fj);
        const float kz = two_pi_over_n * float(kk0 + col);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = t + u * TC;
            const int fi = i < (N + 1) / 2 ? i : i - N;
            const float ki = two_pi_over_n * float(fi);
            const float k2 = ki * ki + kx2 + kz * kz;
            float h;
            if (A.filter == 0) h = A.gamma / (A.gamma + A.D * k2) * A.scale;
            else h = (k2 > 0.0f) ? -A.scale / k2 : 0.0f;
            float2 v = lds[col * CS + i];
            v.x *= h;
            v.y *= h;
            lds[col * CS + i] = v;
        }
        __syncthreads();
        fft_col<N, CS>(lds, col, t, 1, A.tw);
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int e = tid + q * NT;
        const int row = e / NCOL, cl = e % NCOL;
        if (cl < nv) base[(long long)row * A.row_stride + cl] = lds[cl * CS + row];
    }
}

template <int N>
int launch_columns(const ColArgs& A, hipStream_t st) {
    constexpr int NCOL = ncol_for<N>();
    constexpr int H = N / 2 + 1;
    constexpr int NCB = (H + NCOL - 1) / NCOL;
    const long long nb = (long long)A.nc * N * NCB;
    hipLaunchKernelGGL(fft_columns<N>, dim3(unsigned(nb)), dim3(NCOL * (N / 8)), 0, st, A);
    LFM_LAUNCH_CHECK("fft_columns");
    return LFM_OK;
}

int columns(int N, const ColArgs& A, hipStream_t st) {
    switch (N) {
        case 16: return launch_columns<16>(A, st);
        case 32: return launch_columns<32>(A, st);
        case 64: return launch_columns<64>(A, st);
        case 128: return launch_columns<128>(A, st);
        case 256: return launch_columns<256>(A, st);
        case 512: return launch_columns<512>(A, st);
        case 1024: return launch_columns<1024>(A, st);
        default: set_error("column FFT: N=%d unsupported", N); return LFM_E_ARG;
    }
}

bool columns_supported(int N) { return N >= 16 && N <= 1024 && (N & (N - 1)) == 0; }

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
    if (p->tw) (void)hipFree(p->tw);
    delete p;
}

// Transform e -> spectrum -> filter -> out.  filter 0: Helmholtz; 1: Poisson.
int run_filter(ScvPlan* p, const float* e, int filter, float gamma, float D, float* out,
               hipStream_t st, int ncomp_used) {
    LFM_FFT_TRY(rocfft_execution_info_set_stream(p->info, st));
    const int N = p->N;
    const double n3 = double(N) * N * N;
    void* in_f[1] = {const_cast<float*>(e)};
    void* out_f[1] = {p->spec};
    LFM_FFT_TRY(rocfft_execute(p->fwd, in_f, out_f, p->info));
    if (p->method == 2) {
        const long long H = N / 2 + 1;
        ColArgs A{p->spec, p->tw, ncomp_used, 0, 0, 0, filter, gamma, D, float(1.0 / n3)};
        // forward along j: slab (a, i), rows j
        A.slab_stride = (long long)N * H; A.row_stride = H; A.mode = 0;
        int rc = columns(N, A, st);
        if (rc) return rc;
        // along i: forward, filter, inverse in one pass; slab (a, j), rows i
        A.slab_stride = H; A.row_stride = (long long)N * H; A.mode = 2;
        rc = columns(N, A, st);
        if (rc) return rc;
        // inverse along j
        A.slab_stride = (long long)N * H; A.row_stride = H; A.mode = 1;
        rc = columns(N, A, st);
        if (rc) return rc;
    } else {
        const size_t ns = size_t(p->nc) * N * N * (N / 2 + 1);
        if (filter == 0) {
            hipLaunchKernelGGL(helmholtz_filter, dim3(grid_for(ns)), dim3(256), 0, st, p->spec, N,
                               p->nc, gamma, D, float(1.0 / n3));
            LFM_LAUNCH_CHECK("helmholtz_filter");
        } else {
            hipLaunchKernelGGL(poisson_filter, dim3(grid_for(ns)), dim3(256), 0, st, p->spec, N,
                               float(1.0 / n3));
            LFM_LAUNCH_CHECK("poisson_filter");
        }
    }
    void* in_i[1] = {p->spec};
    void* out_i[1] = {out};
    LFM_FFT_TRY(rocfft_execute(p->inv, in_i, out_i, p->info));
    return LFM_OK;
}

int run(ScvPlan* p, const float* e, float gamma, float D, float* sa_out, hipStream_t st) {
    return run_filter(p, e, 0, gamma, D, sa_out, st, p->nc);
}

}  // namespace
}  // namespace lfm

using namespace lfm;

extern "C" int lfm_scv_plan_create_ex(int32_t N, int32_t ncolors, int32_t method, void** plan) {
    if (!plan) { set_error("lfm_scv_plan_create: plan is NULL"); return LFM_E_ARG; }
    *plan = nullptr;
    if (N < 8 || N > 2048 || ncolors < 1 || ncolors > 8) {
        set_error("lfm_scv_plan_create: N=%d ncolors=%d out of range", N, ncolors);
        return LFM_E_ARG;
    }
    if (method == 0) method = columns_supported(N) ? 2 : 1;
    if (method == 2 && !columns_supported(N)) {
        set_error("lfm_scv_plan_create: column passes need a power-of-two N in [16, 1024] (N=%d)", N);
        return LFM_E_ARG;
    }
    if (method != 1 && method != 2) {
        set_error("lfm_scv_plan_create: method %d unknown", method);
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
    p->method = method;
    rocfft_status s;
    if (method == 1) {
        // rocFFT lengths are fastest-first: (k, j, i) for the C-order (i, j, k) array.
        const size_t len[3] = {size_t(N), size_t(N), size_t(N)};
        s = rocfft_plan_create(&p->fwd, rocfft_placement_notinplace, rocfft_transform_type_real_forward,
                               rocfft_precision_single, 3, len, size_t(ncolors), nullptr);
        if (s == rocfft_status_success)
            s = rocfft_plan_create(&p->inv, rocfft_placement_notinplace,
                                   rocfft_transform_type_real_inverse, rocfft_precision_single, 3,
                                   len, size_t(ncolors), nullptr);
    } else {
        // 1-D real transforms along the contiguous k axis, batch ncolors * N * N
        const size_t len[1] = {size_t(N)};
        const size_t batch = size_t(ncolors) * N * N;
        s = rocfft_plan_create(&p->fwd, rocfft_placement_notinplace, rocfft_transform_type_real_forward,
                               rocfft_precision_single, 1, len, batch, nullptr);
        if (s == rocfft_status_success)
            s = rocfft_plan_create(&p->inv, rocfft_placement_notinplace,
                                   rocfft_transform_type_real_inverse, rocfft_precision_single, 1,
                                   len, batch, nullptr);
    }
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
    if (he == hipSuccess && method == 2) {
        // twiddles exp(-2 pi i m / N), computed in double on the host
        float2* h = new float2[N];
        for (int m = 0; m < N; ++m) {
            const double ang = -2.0 * 3.14159265358979323846 * double(m) / double(N);
            h[m] = make_float2(float(cos(ang)), float(sin(ang)));
        }
        he = hipMalloc((void**)&p->tw, N * sizeof(float2));
        if (he == hipSuccess) he = hipMemcpy(p->tw, h, N * sizeof(float2), hipMemcpyHostToDevice);
        delete[] h;
    }
    if (he != hipSuccess) {
        set_error("lfm_scv_plan_create: %s", hipGetErrorString(he));
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

extern "C" int lfm_scv_plan_create(int32_t N, int32_t ncolors, void** plan) {
    return lfm_scv_plan_create_ex(N, ncolors, 0, plan);
}

extern "C" int32_t lfm_scv_plan_method(void* plan) {
    return plan ? static_cast<ScvPlan*>(plan)->method : -1;
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
    const int rc = run_filter(p, p->e, 1, 0.0f, 0.0f, chi_out, st, 1);
    if (rc) return rc;
    hipLaunchKernelGGL(poisson_finish, dim3(grid_for(n3)), dim3(256), 0, st, chi_out, p->N, chi0,
                       boundary_fraction);
    LFM_LAUNCH_CHECK("poisson_finish");
    return LFM_OK;
}

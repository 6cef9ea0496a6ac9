// S_a Helmholtz transform for the SCV term (COLOR level).
//
//   S_a = irfftn( gamma / (gamma + D |k|^2) * rfftn(|Psi_a|^2) ),
//   k_d = 2 pi fftfreq(N)   (continuous |k|^2, not the stencil symbol)
//
// Reference: lfm/core/backends/numpy_backend.py:338-358 and
// cupy_backend.py:260-279, which rebuild the k-grid and the float64 filter on
// every step and run one transform pair per colour.  Here a plan is built
// once per (N, ncolors) and all colours go through one batched transform;
// the filter (with the 1/N^3 normalisation of the inverse) is evaluated on
// the fly.  Two methods:
//   2 (default for power-of-two N in [16, 1024]): five passes written here --
//     real lines along k to half spectra, complex lines along j, one fused
//     forward / filter / inverse pass along i, lines along j, back to real;
//   1 (any other N): rocFFT 3-D R2C / C2R plans and one filter kernel.
// The clip S_a >= 0 happens where S_a is consumed (the step kernel).
#include <rocfft/rocfft.h>

#include <mutex>

#include "lfm_internal.h"
#include "lfm_fft.h"
#include "lfm_stencil.h"

namespace lfm {
namespace {

struct ScvPlan {
    int N = 0, nc = 0;
    int method = 0;          // 1: rocFFT 3-D plans; 2: row and column passes written here
    int Hp = 0;              // row pitch of the half spectrum along kk
    rocfft_plan fwd = nullptr, inv = nullptr;
    rocfft_execution_info info = nullptr;
    void* work = nullptr;
    size_t work_bytes = 0;
    float2* spec = nullptr;  // nc * N * N * Hp
    float* e = nullptr;      // nc * N^3 (|Psi_a|^2 when computed here)
    float2* tw = nullptr;    // exp(-2 pi i m / N), m < N (direct path)
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
        e[t] = psi_sq(pr[t], pi[t]);
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
// Transform passes of the direct path (method 2): a length-N complex FFT per
// line, every thread holding V = 8 values of its line (16 in the column
// passes at N = 1024).
//
// Stockham autosort with radix-8 stages and one radix-2/4 stage when log2 N
// is not a multiple of 3.  Thread t of the TC = N/V threads of a line holds
// the values at positions t + e*TC, e = 0..V-1.  With that assignment every
// stage reads exactly those positions, and the last stage (Ns * R = N) writes
// exactly those positions, so the first stage starts from the values the
// thread loaded from HBM, the last stage's results go straight back to HBM
// (or into the filter and the first inverse stage), and only the stage
// boundaries in between pass through LDS.  Twiddles come from an N-entry
// LDS table filled from a host-computed (double precision) table.
//
// Passes over the half spectrum [a][i][j][kk] (kk < H = N/2+1, row pitch
// Hp = H rounded up to 16 complex values so every row starts on a 128-byte
// line):
//   rows   : real lines along k, two per complex FFT (z = x_2p + i x_2p+1),
//            split / merged with Hermitian symmetry (forward: e -> spectrum,
//            inverse: spectrum -> S_a);
//   columns: NCOL adjacent kk columns of one (a, i) or (a, j) slab, lines
//            along j or along i; the i pass transforms forward, multiplies
//            by the filter and transforms back in one read and one write.
// ---------------------------------------------------------------------------
struct ColArgs {
    float2* spec;
    const float2* tw;
    int nc;
    int Hp;                  // row pitch of the half spectrum along kk
    long long slab_stride;   // complex elements between consecutive slab indices x
    long long row_stride;    // complex elements between consecutive rows (FFT axis)
    int mode;                // 0 forward, 1 inverse, 2 forward + filter + inverse (i axis)
    int filter;              // 0 Helmholtz gamma/(gamma + D k^2), 1 Poisson -1/k^2
    float gamma, D, scale;
    int rounds;              // column groups of NCOL one workgroup takes in turn
};

// Column passes: 16 adjacent kk columns per workgroup (128-byte row segments)
// and V values per thread, so a workgroup has 16 N / V <= 1024 lanes: V = 8
// up to N = 512, 16 at N = 1024.  With V = 8 the stage boundaries alternate
// between two LDS buffers; at N = 1024 one buffer (128 KiB) is all that fits.
template <int N>
constexpr int ncol_for() { return 16; }
template <int N>
constexpr int col_vals() { return N >= 1024 ? 16 : 8; }
template <int N>
constexpr Sync col_sync() { return N >= 1024 ? Sync::Block : Sync::Alt; }

// The i-axis filter on a line held as positions t + e*N/V: multiply by
// gamma/(gamma + D k^2) (FILTER 0) or -1/k^2 (FILTER 1, zero at k = 0), times
// the 1/N^3 normalisation, at (k_i = position, k_j = ky, k_kk = kk).  The
// quotient is the hardware reciprocal refined by one Newton step (within
// 1 ulp of the correctly rounded one).
template <int N, int FILTER, int V>
__device__ __forceinline__ void filter_line(float2 (&v)[V], int t, int kk, float ky,
                                            const ColArgs& A) {
    constexpr int TC = N / V;
    const float two_pi_over_n = 6.283185307179586f / float(N);
    const float kz = two_pi_over_n * float(kk);
    const float k2jk = ky * ky + kz * kz;
    const float num = FILTER == 0 ? A.gamma * A.scale : -A.scale;
#pragma unroll
    for (int e = 0; e < V; ++e) {
        const int i = t + e * TC;
        const int fi = i < (N + 1) / 2 ? i : i - N;
        const float kx = two_pi_over_n * float(fi);
        const float k2 = kx * kx + k2jk;
        const float d = FILTER == 0 ? A.gamma + A.D * k2 : k2;
        float r = __builtin_amdgcn_rcpf(d);
        r = fmaf(fmaf(-d, r, 1.0f), r, r);
        float h = num * r;
        if (FILTER != 0 && !(k2 > 0.0f)) h = 0.0f;
        v[e].x *= h;
        v[e].y *= h;
    }
}

// Workgroup = NCOL adjacent kk columns x TC threads per column; lane order
// is column-fastest so each load/store instruction moves whole 128-byte rows
// and LDS rows ([pos][col]) are read and written contiguously.  A workgroup
// takes `rounds` consecutive column groups of one slab in turn, so the N rows
// of a slab, which for the i axis lie 8 N Hp bytes apart, are opened once per
// `rounds` groups.  With V = 8 the next group's loads are issued before the
// current group is transformed; with V = 16 the next group's 32 registers do
// not fit beside the transform's in 128 VGPRs (1024 lanes per CU), so the
// groups are taken one after the other.
// MODE (0 forward, 1 inverse, 2 forward + filter + inverse) and FILTER are
// template parameters so the butterflies' rotations and twiddle conjugates
// fold at compile time.
template <int N, int MODE, int FILTER>
__global__ __launch_bounds__(ncol_for<N>() * (N / col_vals<N>())) void fft_columns(ColArgs A) {
    constexpr int NCOL = ncol_for<N>();
    constexpr int V = col_vals<N>();
    constexpr int TC = N / V;
    constexpr int H = N / 2 + 1;
    constexpr Sync S = col_sync<N>();
    constexpr int NBUF = S == Sync::Alt ? 2 : 1;
    __shared__ float2 lds[NBUF][N * NCOL];          // Alt: stage boundaries alternate
    __shared__ float2 twl[N];
    int par = 0;                                    // boundaries passed (Alt)
    constexpr int NCB = (H + NCOL - 1) / NCOL;      // column groups per slab
    const int nbs = (NCB + A.rounds - 1) / A.rounds;  // workgroups per slab
    const int cb0 = (blockIdx.x % nbs) * A.rounds;
    const int slab = blockIdx.x / nbs;         // a * N + x
    const int a = slab / N, x = slab % N;
    const int col = threadIdx.x % NCOL, t = threadIdx.x / NCOL;
    float2* sbase = A.spec + (long long)a * N * N * A.Hp + (long long)x * A.slab_stride;
    const int ng = min(A.rounds, NCB - cb0);   // column groups this workgroup takes
    load_twiddles<N>(twl, A.tw);
    auto at = [col](int pos) { return pos * NCOL + col; };
    const float two_pi_over_n = 6.283185307179586f / float(N);
    const int fj = x < (N + 1) / 2 ? x : x - N;
    const float ky = two_pi_over_n * float(fj);
    float2 v[V], vn[V];
    // Loads carry no predicate: lanes past the last live column read column
    // H - 1, a line the live lanes fetch anyway.  Branch-free loads let the
    // compiler wait only for the current group's values, so the next
    // group's loads stay in flight while this one is transformed.
    auto load = [&](float2 (&dst)[V], int cb) {
        const float2* b = sbase + min(cb * NCOL + col, H - 1);
#pragma unroll
        for (int e = 0; e < V; ++e) dst[e] = ld2nt<N>(b + (long long)(t + e * TC) * A.row_stride);
    };
    // `pending` (the prefetched group, or none) is waited for before this
    // group's stores issue: the wait counter covers loads and stores
    // together, and a wait issued while stores are pending drains them too.
    auto transform_store = [&](int cb, float2 (*pending)[V]) {
        const int kk = cb * NCOL + col;
        // the first stage needs no twiddles; the first boundary's barrier
        // publishes the table before any stage reads it
        fft_line<N, S, V>(v, lds[0], t, MODE == 1 ? 1 : -1, twl, at, lds[NBUF - 1], &par);
        if constexpr (MODE == 2) {
            filter_line<N, FILTER, V>(v, t, kk, ky, A);
            fft_line<N, S, V>(v, lds[0], t, 1, twl, at, lds[NBUF - 1], &par);
        }
        if (pending) {
#pragma unroll
            for (int e = 0; e < V; ++e) asm volatile("" ::"v"((*pending)[e].x), "v"((*pending)[e].y));
        }
        if (kk < H) {
            float2* b = sbase + kk;
#pragma unroll
            for (int e = 0; e < V; ++e) st2nt<N>(b + (long long)(t + e * TC) * A.row_stride, v[e]);
        }
    };
    if constexpr (V == 8) {
        load(v, cb0);
        // wait for the first group here, so no load is pending where the loop
        // is entered and the loop's waits need not cover earlier stores
#pragma unroll
        for (int e = 0; e < V; ++e) asm volatile("" ::"v"(v[e].x), "v"(v[e].y));
        for (int r = 0; r + 1 < ng; ++r) {
            load(vn, cb0 + r + 1);
            transform_store(cb0 + r, &vn);
#pragma unroll
            for (int e = 0; e < V; ++e) v[e] = vn[e];
        }
        transform_store(cb0 + ng - 1, nullptr);
    } else {
        for (int r = 0; r < ng; ++r) {
            load(v, cb0 + r);
            transform_store(cb0 + r, nullptr);
        }
    }
}

// Lines along the contiguous k axis.  Two real lines p0 = 2q, p1 = 2q+1 form
// z = x_p0 + i x_p1; after Z = FFT(z):
//   X_p0[k] = (Z[k] + conj Z[N-k]) / 2,  X_p1[k] = (Z[k] - conj Z[N-k]) / (2i).
// The inverse builds Z[k] = X_p0[k] + i X_p1[k] over all k from the half
// spectra (X[N-k] = conj X[k]; the imaginary parts of X[0] and X[N/2] are
// dropped, as a C2R transform does) and takes x_p0 = Re z, x_p1 = Im z.
// LDS slot of position pos: pos + pos/8, which spreads the radix-8
// boundaries' stride-8 writes over all banks.
// Values per thread: 8, and 16 at N = 1024, so a line of N/V <= 64 threads
// lies in one wavefront at every size and its stage boundaries need no
// workgroup barrier.
template <int N>
constexpr int row_vals() { return N >= 1024 ? 16 : 8; }
template <int N>
constexpr int rows_threads() { return 256; }
template <int N>
constexpr int rows_pairs() { return rows_threads<N>() / (N / row_vals<N>()); }
template <int N>
constexpr int rows_pitch() { return N + N / 8; }
template <int N>
constexpr Sync rows_sync() { return N / row_vals<N>() <= 64 ? Sync::Wave : Sync::Block; }

// PSI: the real lines are |Psi_a|^2, formed here from the Pr lines of `in`
// and the Pi lines of `in_i` (same layout), so no |Psi|^2 field is stored.
template <int N, bool PSI>
__global__ __launch_bounds__(rows_threads<N>()) void fft_rows_r2c(const float* __restrict__ in,
                                                                 const float* __restrict__ in_i,
                                                                 float2* __restrict__ spec,
                                                                 const float2* __restrict__ tw,
                                                                 int Hp) {
    constexpr int V = row_vals<N>();
    constexpr int TC = N / V;
    constexpr int PB = rows_pairs<N>();
    constexpr int PS = rows_pitch<N>();
    __shared__ float2 lds[PB * PS];
    __shared__ float2 twl[N];
    const int q = threadIdx.x / TC, t = threadIdx.x % TC;
    const long long pair = (long long)blockIdx.x * PB + q;
    const float* r0 = in + pair * 2 * N;
    const float* r1 = r0 + N;
    load_twiddles<N>(twl, tw);
    float2 v[V];
    if constexpr (PSI) {
        const float* m0 = in_i + pair * 2 * N;
        const float* m1 = m0 + N;
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const int k = t + e * TC;
            v[e] = make_float2(psi_sq(ld1nt<N>(r0 + k), ld1nt<N>(m0 + k)),
                               psi_sq(ld1nt<N>(r1 + k), ld1nt<N>(m1 + k)));
        }
    } else {
#pragma unroll
        for (int e = 0; e < V; ++e)
            v[e] = make_float2(ld1nt<N>(r0 + t + e * TC), ld1nt<N>(r1 + t + e * TC));
    }
    float2* L = lds + q * PS;
    auto at = [](int pos) { return pos + pos / 8; };
    constexpr Sync S = rows_sync<N>();
    // with wavefront-level boundaries the twiddle table, which the whole
    // workgroup fills, is published here
    if constexpr (S == Sync::Wave) __syncthreads();
    fft_line<N, S, V>(v, L, t, -1, twl, at);
    line_sync<S>();
#pragma unroll
    for (int e = 0; e < V; ++e) L[at(t + e * TC)] = v[e];
    line_sync<S>();
    float2* o0 = spec + pair * 2 * Hp;
    float2* o1 = o0 + Hp;
#pragma unroll
    for (int e = 0; e <= V / 2; ++e) {
        const int k = t + e * TC;
        if (e == V / 2 && t != 0) break;  // k = N/2 once
        const float2 zk = L[at(k)];
        const float2 zm = L[at((N - k) & (N - 1))];
        st2nt<N>(o0 + k, make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y)));
        st2nt<N>(o1 + k, make_float2(0.5f * (zk.y + zm.y), 0.5f * (zm.x - zk.x)));
    }
}

template <int N>
__global__ __launch_bounds__(rows_threads<N>()) void fft_rows_c2r(const float2* __restrict__ spec,
                                                                 float* __restrict__ out,
                                                                 const float2* __restrict__ tw,
                                                                 int Hp) {
    constexpr int V = row_vals<N>();
    constexpr int TC = N / V;
    constexpr int PB = rows_pairs<N>();
    constexpr int PS = rows_pitch<N>();
    __shared__ float2 lds[PB * PS];
    __shared__ float2 twl[N];
    const int q = threadIdx.x / TC, t = threadIdx.x % TC;
    const long long pair = (long long)blockIdx.x * PB + q;
    const float2* s0 = spec + pair * 2 * Hp;
    const float2* s1 = s0 + Hp;
    load_twiddles<N>(twl, tw);
    // The two half spectra go through LDS once (coalesced, each value read
    // once from HBM); the mirrored reads X[N-k] = conj X[k] then come from LDS.
    constexpr int H = N / 2 + 1;
    static_assert(2 * H <= PS, "both half spectra fit the line's LDS slot");
    float2* L = lds + q * PS;
#pragma unroll
    for (int e = 0; e < (H + TC - 1) / TC; ++e) {
        const int k = t + e * TC;
        if (k < H) {
            L[k] = ld2nt<N>(s0 + k);
            L[H + k] = ld2nt<N>(s1 + k);
        }
    }
    __syncthreads();
    float2 v[V];
#pragma unroll
    for (int e = 0; e < V; ++e) {
        const int k = t + e * TC;
        const bool lo = k <= N / 2;
        const int m = lo ? k : N - k;
        float2 a = L[m], b = L[H + m];
        if (m == 0 || m == N / 2) { a.y = 0.0f; b.y = 0.0f; }
        if (!lo) { a.y = -a.y; b.y = -b.y; }
        v[e] = make_float2(a.x - b.y, a.y + b.x);
    }
    // the transform's first boundary orders these reads before its writes
    fft_line<N, rows_sync<N>(), V>(v, L, t, 1, twl, [](int pos) { return pos + pos / 8; });
    float* r0 = out + pair * 2 * N;
    float* r1 = r0 + N;
#pragma unroll
    for (int e = 0; e < V; ++e) {
        st1nt<N>(r0 + t + e * TC, v[e].x);
        st1nt<N>(r1 + t + e * TC, v[e].y);
    }
}

// Half spectra -> the SCV term of the step kernel, for three colours:
// fft_rows_c2r's inverse transform on the same row pair (i, j), (i, j+1) of
// every colour in one workgroup (PJ row pairs per colour), then
// scv[site] = scv_term(S_0, S_1, S_2) -- the clipped-S_a combination the
// step kernel otherwise forms from the three S_a fields it reads, here
// formed where S_a is produced, so one float per site is written instead of
// three (and read by the step instead of three).  Colours 1 and 2 pass their
// values to colour 0's threads through LDS (same positions t + e*TC).
template <int N>
constexpr int scv_pairs() { return N / row_vals<N>() >= 64 ? 1 : 64 / (N / row_vals<N>()); }
template <int N>
constexpr int scv_threads() { return 3 * scv_pairs<N>() * (N / row_vals<N>()); }

template <int N>
__global__ __launch_bounds__(scv_threads<N>()) void fft_rows_c2r_scv(
        const float2* __restrict__ spec, float* __restrict__ scv,
        const float2* __restrict__ tw, int Hp) {
    constexpr int V = row_vals<N>();
    constexpr int TC = N / V;
    constexpr int PJ = scv_pairs<N>();
    constexpr int PS = rows_pitch<N>();
    constexpr long long NP = (long long)N * N / 2;   // row pairs per colour
    __shared__ float2 lds[3 * PJ * PS];
    __shared__ float2 twl[N];
    const int q = threadIdx.x / TC, t = threadIdx.x % TC;
    const int a = q / PJ, lp = q % PJ;               // colour, row pair in the workgroup
    const long long pair = (long long)blockIdx.x * PJ + lp;
    const float2* s0 = spec + (a * NP + pair) * 2 * Hp;
    const float2* s1 = s0 + Hp;
    load_twiddles<N>(twl, tw);
    constexpr int H = N / 2 + 1;
    static_assert(2 * H <= PS, "both half spectra fit the line's LDS slot");
    float2* L = lds + q * PS;
#pragma unroll
    for (int e = 0; e < (H + TC - 1) / TC; ++e) {
        const int k = t + e * TC;
        if (k < H) {
            L[k] = ld2nt<N>(s0 + k);
            L[H + k] = ld2nt<N>(s1 + k);
        }
    }
    __syncthreads();
    float2 v[V];
#pragma unroll
    for (int e = 0; e < V; ++e) {
        const int k = t + e * TC;
        const bool lo = k <= N / 2;
        const int m = lo ? k : N - k;
        float2 x = L[m], y = L[H + m];
        if (m == 0 || m == N / 2) { x.y = 0.0f; y.y = 0.0f; }
        if (!lo) { x.y = -x.y; y.y = -y.y; }
        v[e] = make_float2(x.x - y.y, x.y + y.x);
    }
    fft_line<N, rows_sync<N>(), V>(v, L, t, 1, twl, [](int pos) { return pos + pos / 8; });
    // v[e] = (S_a at row 2 pair, S_a at row 2 pair + 1), position t + e TC.
    // The transform's last LDS reads precede this barrier; colours 1 and 2
    // then park their results in their own slots for colour 0.
    __syncthreads();
    if (a > 0) {
#pragma unroll
        for (int e = 0; e < V; ++e) L[t + e * TC] = v[e];
    }
    __syncthreads();
    if (a == 0) {
        const float2* L1 = lds + (PJ + lp) * PS;
        const float2* L2 = lds + (2 * PJ + lp) * PS;
        float* r0 = scv + pair * 2 * N;
        float* r1 = r0 + N;
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const float2 w1 = L1[t + e * TC], w2 = L2[t + e * TC];
            float clip[3];
            st1nt<N>(r0 + t + e * TC, scv_term(v[e].x, w1.x, w2.x, clip));
            st1nt<N>(r1 + t + e * TC, scv_term(v[e].y, w1.y, w2.y, clip));
        }
    }
}

template <int N>
int launch_c2r_scv(const float2* spec, float* scv, const float2* tw, int Hp, hipStream_t st) {
    const long long nb = (long long)N * N / 2 / scv_pairs<N>();
    hipLaunchKernelGGL(fft_rows_c2r_scv<N>, dim3(unsigned(nb)), dim3(scv_threads<N>()), 0, st,
                       spec, scv, tw, Hp);
    LFM_LAUNCH_CHECK("fft_rows_c2r_scv");
    return LFM_OK;
}

template <int N>
int launch_columns(const ColArgs& A, hipStream_t st) {
    constexpr int NCOL = ncol_for<N>();
    constexpr int H = N / 2 + 1;
    constexpr int NCB = (H + NCOL - 1) / NCOL;
    const long long nb = (long long)A.nc * N * ((NCB + A.rounds - 1) / A.rounds);
    const dim3 g{unsigned(nb)}, b{unsigned(NCOL * (N / col_vals<N>()))};
    if (A.mode == 0) hipLaunchKernelGGL((fft_columns<N, 0, 0>), g, b, 0, st, A);
    else if (A.mode == 1) hipLaunchKernelGGL((fft_columns<N, 1, 0>), g, b, 0, st, A);
    else if (A.filter == 0) hipLaunchKernelGGL((fft_columns<N, 2, 0>), g, b, 0, st, A);
    else hipLaunchKernelGGL((fft_columns<N, 2, 1>), g, b, 0, st, A);
    LFM_LAUNCH_CHECK("fft_columns");
    return LFM_OK;
}

// dir 0: real lines -> half spectra (of |Psi|^2 when in_i, the Pi lines, is
// given); 1: half spectra -> real lines
template <int N>
int launch_rows(int dir, const float* in, const float* in_i, float2* spec, float* out,
                const float2* tw, int Hp, int nc, hipStream_t st) {
    const long long pairs = (long long)nc * N * N / 2;
    const long long nb = pairs / rows_pairs<N>();
    if (dir == 0 && in_i)
        hipLaunchKernelGGL((fft_rows_r2c<N, true>), dim3(unsigned(nb)), dim3(rows_threads<N>()), 0,
                           st, in, in_i, spec, tw, Hp);
    else if (dir == 0)
        hipLaunchKernelGGL((fft_rows_r2c<N, false>), dim3(unsigned(nb)), dim3(rows_threads<N>()), 0,
                           st, in, nullptr, spec, tw, Hp);
    else
        hipLaunchKernelGGL(fft_rows_c2r<N>, dim3(unsigned(nb)), dim3(rows_threads<N>()), 0, st,
                           spec, out, tw, Hp);
    LFM_LAUNCH_CHECK(dir == 0 ? "fft_rows_r2c" : "fft_rows_c2r");
    return LFM_OK;
}

#define LFM_FFT_SIZES(X) X(16) X(32) X(64) X(128) X(256) X(512) X(1024)

int columns(int N, const ColArgs& A, hipStream_t st) {
    switch (N) {
#define LFM_CASE(n) case n: return launch_columns<n>(A, st);
        LFM_FFT_SIZES(LFM_CASE)
#undef LFM_CASE
        default: set_error("column FFT: N=%d unsupported", N); return LFM_E_ARG;
    }
}

int rows(int N, int dir, const float* in, const float* in_i, float2* spec, float* out,
         const float2* tw, int Hp, int nc, hipStream_t st) {
    switch (N) {
#define LFM_CASE(n) case n: return launch_rows<n>(dir, in, in_i, spec, out, tw, Hp, nc, st);
        LFM_FFT_SIZES(LFM_CASE)
#undef LFM_CASE
        default: set_error("row FFT: N=%d unsupported", N); return LFM_E_ARG;
    }
}

int rows_scv(int N, const float2* spec, float* scv, const float2* tw, int Hp, hipStream_t st) {
    switch (N) {
#define LFM_CASE(n) case n: return launch_c2r_scv<n>(spec, scv, tw, Hp, st);
        LFM_FFT_SIZES(LFM_CASE)
#undef LFM_CASE
        default: set_error("row FFT: N=%d unsupported", N); return LFM_E_ARG;
    }
}

bool columns_supported(int N) { return N >= 16 && N <= 1024 && (N & (N - 1)) == 0; }

// half-spectrum row pitch of the direct path
int spectrum_pitch(int N) { return ((N / 2 + 1) + 15) / 16 * 16; }

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

// Column groups per workgroup: along j 1 at N = 1024, 4 at N = 512, 2 below;
// along i a workgroup takes its whole slab (N rows of N/2+1 columns), which
// opens each row of the slab once.  Measured (three passes), j/i rounds:
// 1024^3 (16 values per thread): 4/whole 20.5 ms, 1/1 20.3, 1/4 20.1, 1/8 19.9,
// 1/whole 19.4;
// 512^3: 2/4 2.10 ms, 4/4 2.08, 2/whole 2.05, 4/whole 2.04, 6/whole 2.09;
// 256^3: 4/whole 0.218, 2/whole 0.212, 2/4 0.221;
// 128^3: 4/whole 0.045, 2/whole 0.037, 2/4 0.039.
int rounds_j(int N) { return N >= 1024 ? 1 : N >= 512 ? 4 : 2; }
int whole_slab_rounds(int N) {
    const int ncol = 16;   // ncol_for<N>()
    return (N / 2 + 1 + ncol - 1) / ncol;
}

// The three column passes of the direct method on the spectrum in place.
int run_columns(ScvPlan* p, int filter, float gamma, float D, hipStream_t st, int ncomp_used) {
    const int N = p->N;
    const double n3 = double(N) * N * N;
    ColArgs A{p->spec, p->tw, ncomp_used, p->Hp, 0, 0, 0, filter, gamma, D, float(1.0 / n3), 1};
    // forward along j: slab (a, i), rows j
    A.slab_stride = (long long)N * p->Hp; A.row_stride = p->Hp; A.mode = 0;
    A.rounds = rounds_j(N);
    int rc = columns(N, A, st);
    if (rc) return rc;
    // along i: forward, filter, inverse in one pass; slab (a, j), rows i
    A.slab_stride = p->Hp; A.row_stride = (long long)N * p->Hp; A.mode = 2;
    A.rounds = whole_slab_rounds(N);
    rc = columns(N, A, st);
    if (rc) return rc;
    // inverse along j
    A.slab_stride = (long long)N * p->Hp; A.row_stride = p->Hp; A.mode = 1;
    A.rounds = rounds_j(N);
    return columns(N, A, st);
}

// Transform e -> spectrum -> filter -> out.  filter 0: Helmholtz; 1: Poisson.
// With e_i given, e and e_i are the Pr and Pi fields and the transformed lines
// are |Psi|^2: the direct method forms it in its first pass, the rocFFT
// method in the plan's real-space buffer first.
int run_filter(ScvPlan* p, const float* e, const float* e_i, int filter, float gamma, float D,
               float* out, hipStream_t st, int ncomp_used) {
    const int N = p->N;
    const double n3 = double(N) * N * N;
    if (p->method == 2) {
        int rc = rows(N, 0, e, e_i, p->spec, nullptr, p->tw, p->Hp, ncomp_used, st);
        if (rc) return rc;
        rc = run_columns(p, filter, gamma, D, st, ncomp_used);
        if (rc) return rc;
        return rows(N, 1, nullptr, nullptr, p->spec, out, p->tw, p->Hp, ncomp_used, st);
    }
    if (e_i) {
        const size_t n = size_t(ncomp_used) * N * N * N;
        hipLaunchKernelGGL(psisq_kernel, dim3(grid_for(n)), dim3(256), 0, st, e, e_i, p->e, n);
        LFM_LAUNCH_CHECK("psisq_kernel");
        e = p->e;
    }
    LFM_FFT_TRY(rocfft_execution_info_set_stream(p->info, st));
    void* in_f[1] = {const_cast<float*>(e)};
    void* out_f[1] = {p->spec};
    LFM_FFT_TRY(rocfft_execute(p->fwd, in_f, out_f, p->info));
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
    void* in_i[1] = {p->spec};
    void* out_i[1] = {out};
    LFM_FFT_TRY(rocfft_execute(p->inv, in_i, out_i, p->info));
    return LFM_OK;
}

int run(ScvPlan* p, const float* e, const float* e_i, float gamma, float D, float* sa_out,
        hipStream_t st) {
    return run_filter(p, e, e_i, 0, gamma, D, sa_out, st, p->nc);
}

// Direct method, three colours: e (or |Psi|^2 of Pr = e, Pi = e_i) ->
// spectrum -> Helmholtz filter -> the SCV term (one field) instead of the
// three S_a fields.
int run_scv(ScvPlan* p, const float* e, const float* e_i, float gamma, float D, float* scv_out,
            hipStream_t st) {
    const int N = p->N;
    int rc = rows(N, 0, e, e_i, p->spec, nullptr, p->tw, p->Hp, 3, st);
    if (rc) return rc;
    rc = run_columns(p, 0, gamma, D, st, 3);
    if (rc) return rc;
    return rows_scv(N, p->spec, scv_out, p->tw, p->Hp, st);
}

}  // namespace

bool scv_view(void* plan, ScvView* v) {
    auto* p = static_cast<ScvPlan*>(plan);
    if (!p) return false;
    *v = ScvView{p->spec, p->tw, p->N, p->nc, p->Hp, p->method};
    return true;
}

int scv_prepare(void* plan, const float* pr, const float* pi, float gamma, float D,
                hipStream_t st) {
    auto* p = static_cast<ScvPlan*>(plan);
    if (!p || p->method != 2) { set_error("scv_prepare: needs a direct-method plan"); return LFM_E_ARG; }
    int rc = rows(p->N, 0, pr, pi, p->spec, nullptr, p->tw, p->Hp, p->nc, st);
    if (rc) return rc;
    return run_columns(p, 0, gamma, D, st, p->nc);
}

int scv_columns(void* plan, float gamma, float D, hipStream_t st) {
    auto* p = static_cast<ScvPlan*>(plan);
    if (!p || p->method != 2) { set_error("scv_columns: needs a direct-method plan"); return LFM_E_ARG; }
    return run_columns(p, 0, gamma, D, st, p->nc);
}

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
    auto* p = new ScvPlan();
    p->N = N;
    p->nc = ncolors;
    p->method = method;
    p->Hp = method == 2 ? spectrum_pitch(N) : N / 2 + 1;
    if (method == 1) {
        std::call_once(g_fft_once, [] { g_fft_setup = rocfft_setup(); });
        rocfft_status s = g_fft_setup;
        // rocFFT lengths are fastest-first: (k, j, i) for the C-order (i, j, k) array.
        const size_t len[3] = {size_t(N), size_t(N), size_t(N)};
        if (s == rocfft_status_success)
            s = rocfft_plan_create(&p->fwd, rocfft_placement_notinplace,
                                   rocfft_transform_type_real_forward, rocfft_precision_single, 3,
                                   len, size_t(ncolors), nullptr);
        if (s == rocfft_status_success)
            s = rocfft_plan_create(&p->inv, rocfft_placement_notinplace,
                                   rocfft_transform_type_real_inverse, rocfft_precision_single, 3,
                                   len, size_t(ncolors), nullptr);
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
    }
    hipError_t he = hipSuccess;
    if (p->work_bytes) he = hipMalloc(&p->work, p->work_bytes);
    const size_t ns = size_t(ncolors) * N * N * p->Hp;
    if (he == hipSuccess) he = hipMalloc((void**)&p->spec, ns * sizeof(float2));
    // the real-space buffer: the Poisson solve's right-hand side (one-colour
    // plans) and the rocFFT method's |Psi|^2 input; the direct method forms
    // |Psi|^2 in its first pass and needs none for several colours
    if (he == hipSuccess && (method == 1 || ncolors == 1))
        he = hipMalloc((void**)&p->e, size_t(ncolors) * N * N * N * sizeof(float));
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
        const rocfft_status s = rocfft_execution_info_set_work_buffer(p->info, p->work, p->work_bytes);
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
    return run(p, e, nullptr, gamma, D, sa_out, (hipStream_t)stream);
}

extern "C" int lfm_scv_apply(void* plan, const float* pr, const float* pi, float gamma, float D,
                             float* sa_out, void* stream) {
    auto* p = static_cast<ScvPlan*>(plan);
    if (!p || !pr || !pi || !sa_out) { set_error("lfm_scv_apply: NULL argument"); return LFM_E_ARG; }
    if (!(gamma > 0.0f) || D < 0.0f) {
        set_error("lfm_scv_apply: need gamma > 0 and D >= 0 (got %g, %g)", gamma, D);
        return LFM_E_ARG;
    }
    return run(p, pr, pi, gamma, D, sa_out, (hipStream_t)stream);
}

extern "C" int lfm_scv_prepare_f32(void* plan, const float* pr, const float* pi, float gamma,
                                   float D, void* stream) {
    auto* p = static_cast<ScvPlan*>(plan);
    if (!p || !pr || !pi) { set_error("lfm_scv_prepare_f32: NULL argument"); return LFM_E_ARG; }
    if (!(gamma > 0.0f) || D < 0.0f) {
        set_error("lfm_scv_prepare_f32: need gamma > 0 and D >= 0 (got %g, %g)", gamma, D);
        return LFM_E_ARG;
    }
    return scv_prepare(plan, pr, pi, gamma, D, (hipStream_t)stream);
}

extern "C" int lfm_scv_columns_f32(void* plan, float gamma, float D, void* stream) {
    auto* p = static_cast<ScvPlan*>(plan);
    if (!p) { set_error("lfm_scv_columns_f32: NULL plan"); return LFM_E_ARG; }
    if (!(gamma > 0.0f) || D < 0.0f) {
        set_error("lfm_scv_columns_f32: need gamma > 0 and D >= 0 (got %g, %g)", gamma, D);
        return LFM_E_ARG;
    }
    return scv_columns(plan, gamma, D, (hipStream_t)stream);
}

extern "C" int lfm_scv_apply_scv_f32(void* plan, const float* pr, const float* pi,
                                     const float* e, float gamma, float D, float* scv_out,
                                     void* stream) {
    auto* p = static_cast<ScvPlan*>(plan);
    if (!p || !scv_out || (!e && (!pr || !pi))) {
        set_error("lfm_scv_apply_scv_f32: bad argument");
        return LFM_E_ARG;
    }
    if (p->method != 2 || p->nc != 3) {
        set_error("lfm_scv_apply_scv_f32: needs a direct-method plan for 3 colours "
                  "(method %d, %d colours)", p->method, p->nc);
        return LFM_E_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    return e ? run_scv(p, e, nullptr, gamma, D, scv_out, st) : run_scv(p, pr, pi, gamma, D, scv_out, st);
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
    const int rc = run_filter(p, p->e, nullptr, 1, 0.0f, 0.0f, chi_out, st, 1);
    if (rc) return rc;
    hipLaunchKernelGGL(poisson_finish, dim3(grid_for(n3)), dim3(256), 0, st, chi_out, p->N, chi0,
                       boundary_fraction);
    LFM_LAUNCH_CHECK("poisson_finish");
    return LFM_OK;
}

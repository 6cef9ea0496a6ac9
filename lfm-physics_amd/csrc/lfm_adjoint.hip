// Adjoint (reverse-mode) step of the COMPLEX leapfrog.
//
// Forward map (reference semantics, lfm/core/backends/numpy_backend.py:132-205
// with the Evolver's buffers, lfm/core/evolver.py:182-281), a = 1 - m:
//   cur'  = a (2 cur - prev + dt2 (L cur - chi^2 cur))
//   prev' = a cur
//   chi'  = m chi0 + a max(H, -chi0),  H = 2 chi - chi_p + dt2 acc
//   chi_p' = chi
//   acc = L chi - (kappa/chi0) chi S - 4 lam chi (chi^2 - chi0^2),
//   S = |Psi|^2 + eps_w j - E0^2,  j = 1/2 sum_d [Pr dPi_d - Pi dPr_d].
// Adjoint state of step n+1: U = dL/dcur, X = dL/dchi; dL/dprev and
// dL/dchi_p are -Ut and -Xt of the following step (held by the caller), with
//   Ut = a U,  Xt = a 1[H > -chi0] X.
// The caller keeps Ut, Xt of the two following steps.  One reverse step
// n <- n+1 is one kernel (adj_fused):
//   U_n = 2Ut + dt2 (L Ut - chi^2 Ut) - a Ut_{n+2} + (dH/dPsi)^T Xt
//   X_n = Xt (2 + dt2 dacc/dchi) + dt2 L Xt - 2 dt2 chi (Pr Utr + Pi Uti) - Xt_{n+2}
// with (dH/dPr)^T Xt = 2 Pr q + w/2 sum_d dPi_d + 1/2 sum_d [Pi w](+d) - [Pi w](-d),
//      (dH/dPi)^T Xt = 2 Pi q - w/2 sum_d dPr_d + 1/2 sum_d [Pr w](-d) - [Pr w](+d),
// q = -dt2 (kappa/chi0) chi Xt, w = eps_w q (L is symmetric, so L^T = L),
// and writes either U_n, X_n (the last reverse step) or directly the damped
// Ut_n = a U_n, Xt_n = a 1[no clamp in step n-1 -> n] X_n for the next one.
// The clamp indicator comes from the stored chi_n: with deviation storage
// the forward writes a * max(H, -2 chi0), so the clamp was taken exactly when
// the stored value equals a * (-2 chi0) (for a = 1 this is H <= floor; in the
// taper shell a value of H within rounding of the floor also reads as
// clamped).  The same kernel adds the kappa-gradient term
// sum(Xt dt2 (-chi S / chi0)) of the step into a per-workgroup partial.
// All chi buffers hold chi - chi0 (the step kernels' deviation storage).
#include "lfm_internal.h"

namespace lfm {
namespace {

constexpr float WF = 1.0f / 3.0f, WE = 1.0f / 6.0f;
// Workgroup tile: BK (k) x BJ (j) columns marching along i over a chunk of
// planes; one thread per column.
constexpr int BK = 64, BJ = 4, AB = BK * BJ;
constexpr int LWK = BK + 2, LWJ = BJ + 2, LPL = LWK * LWJ;   // staged plane with halo
constexpr int NLF = 7;                                         // staged fields

struct AdjArgs {
    const float* pr; const float* pi; const float* chi; const float* mask;   // state n
    const float* Ur; const float* Ui; const float* X;       // Ut, Xt of step n+1
    const float* U2r; const float* U2i; const float* X2;    // Ut, Xt of step n+2 (or NULL)
    float* Or; float* Oi; float* Ox;                        // step n outputs
    double* kpart;
    int N, flags, damp_out, CH;
    float dt2, kappa, lam, chi0, e0sq, epsw;
};

__device__ __forceinline__ double block_sum(double v) {
    __shared__ double sh[AB / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int t = threadIdx.x;
    if ((t & 63) == 0) sh[t >> 6] = v;
    __syncthreads();
    double r = 0.0;
    if (t == 0)
        for (int q = 0; q < AB / 64; ++q) r += sh[q];
    return r;
}

// Ut = a U, Xt = a 1[no clamp] X for the state whose stored chi is given.
__device__ __forceinline__ float clamp_free(float cs, float a, float chi0) {
    return cs != a * (-2.0f * chi0) ? 1.0f : 0.0f;
}

__global__ __launch_bounds__(AB) void adj_seed(const float* __restrict__ chi,
                                               const float* __restrict__ mask,
                                               const float* __restrict__ Ur,
                                               const float* __restrict__ Ui,
                                               const float* __restrict__ X, float* __restrict__ Or,
                                               float* __restrict__ Oi, float* __restrict__ Ox,
                                               uint32_t n3, float chi0) {
    const uint32_t idx = blockIdx.x * AB + threadIdx.x;
    if (idx >= n3) return;
    const float a = mask ? 1.0f - mask[idx] : 1.0f;
    Or[idx] = a * Ur[idx];
    Oi[idx] = a * Ui[idx];
    Ox[idx] = a * clamp_free(chi[idx], a, chi0) * X[idx];
}

// Loads and stores at a 32-bit byte offset from a kernel-argument base, so
// the address is one uniform base plus one VGPR (N^3 * 4 < 2^32, N <= 1023).
__device__ __forceinline__ float ldb(const float* p, uint32_t b) {
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(p) + b);
}
// NT: streaming (non-temporal) form for data touched once, used when the
// step's working set is far beyond the last-level cache.
template <bool NT>
__device__ __forceinline__ float ldbp(const float* p, uint32_t b) {
    const float* q = reinterpret_cast<const float*>(reinterpret_cast<const char*>(p) + b);
    if constexpr (NT) return __builtin_nontemporal_load(q);
    else return *q;
}
template <bool NT>
__device__ __forceinline__ void stb(float* p, uint32_t b, float v) {
    float* q = reinterpret_cast<float*>(reinterpret_cast<char*>(p) + b);
    if constexpr (NT) __builtin_nontemporal_store(v, q);
    else *q = v;
}

// 2.5-D form: each plane q of the chunk (plus one halo plane on each side) is
// staged once into LDS as Ut_r, Ut_i, Xt, Pr, Pi, Pi w, Pr w (w = eps_w q at
// that site); the 19-point Laplacians split as L(i) = A(i) + B(i-1) + B(i+1)
// and the face sums split into in-plane parts and the +-i differences, so a
// site i is finished when plane i+1 has been staged.
template <bool NT>
__global__ __launch_bounds__(AB) void adj_fused(AdjArgs A) {
    __shared__ float lds[NLF * LPL];
    const int N = A.N;
    const int nkt = (N + BK - 1) / BK, njt = (N + BJ - 1) / BJ;
    // Blocks are dispatched round-robin over the 8 XCDs (one L2 each): give
    // each XCD a contiguous run of tiles so neighbouring tiles' halo lines
    // are fetched into the same L2.
    const int nb = gridDim.x, b = blockIdx.x;
    const int lb = (nb & 7) == 0 ? (b & 7) * (nb >> 3) + (b >> 3) : b;
    const int kt = lb % nkt;
    const int jt = (lb / nkt) % njt;
    const int ch = lb / (nkt * njt);
    const int k0 = kt * BK, j0 = jt * BJ;
    const int i0 = ch * A.CH;
    const int i1 = min(i0 + A.CH, N);
    const int tx = threadIdx.x % BK, ty = threadIdx.x / BK;
    const int k = k0 + tx, j = j0 + ty;
    const bool live = k < N && j < N;
    const float kc = A.kappa / A.chi0;
    const float wk = -A.epsw * A.dt2 * kc;
    const uint32_t NN4 = uint32_t(N) * N * 4u;
    // staging assignment: positions p = threadIdx.x + r * AB of the LPL plane
    uint32_t soff[2];
    bool sact[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int p = threadIdx.x + r * AB;
        sact[r] = p < LPL;
        const int row = sact[r] ? p / LWK : 0, col = sact[r] ? p % LWK : 0;
        int gj = j0 + row - 1, gk = k0 + col - 1;
        gj = gj < 0 ? gj + N : (gj >= N ? gj - N : gj);
        gk = gk < 0 ? gk + N : (gk >= N ? gk - N : gk);
        gj = gj >= N ? gj % N : gj;   // tiles wider than N (N < BK)
        gk = gk >= N ? gk % N : gk;
        soff[r] = (uint32_t(gj) * N + gk) * 4u;
    }
    const int c0 = (ty + 1) * LWK + (tx + 1);    // thread's column in a staged plane
    const uint32_t cb = (uint32_t(j < N ? j : 0) * N + uint32_t(k < N ? k : 0)) * 4u;
    // Running sums per column.  For the site of the previous plane (p) and
    // of the current plane (c): Laplacian parts of Ut_r, Ut_i, Xt and the face
    // sums sum_d dPr_d, sum_d dPi_d, sum_d [Pi w](+d) - [Pi w](-d),
    // sum_d [Pr w](-d) - [Pr w](+d).  Plane q contributes A(q) and the in-plane
    // differences to its own site, B(q) and the +-i terms to sites q -+ 1.
    float accp[7] = {}, accc[7] = {};
    float cpr = 0.f, cpi = 0.f, cur = 0.f, cui = 0.f, cxt = 0.f;   // centres of site p
    double kp = 0.0;
    const int nplanes = i1 - i0 + 2;
    // Staged values are loaded two planes ahead (sa: next plane to stage, sb:
    // the one after) and the pointwise data of a site one iteration ahead, so
    // global-load latency stays off each iteration's barrier-bound path.
    float sa[2][6], sb[2][6];
    // Loads carry no per-lane predicate (idle staging positions read a valid
    // in-lattice address and discard it; absent optional fields read one
    // shared word): branch-free loads let the compiler place its waits where
    // the values are used instead of draining everything at each join.
    auto load_plane = [&](float (&v)[2][6], int s) {
        int q = i0 - 1 + s;
        q = q < 0 ? q + N : (q >= N ? q - N : q);
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const uint32_t o = uint32_t(q) * NN4 + soff[r];
            v[r][0] = ldb(A.Ur, o);
            v[r][1] = ldb(A.Ui, o);
            v[r][2] = ldb(A.X, o);
            v[r][3] = ldb(A.pr, o);
            v[r][4] = ldb(A.pi, o);
            v[r][5] = ldb(A.chi, o);
        }
    };
    // pointwise data of the site finished at iteration s (plane i0 + s - 2)
    float pw[5] = {}, pn[5] = {};
    const float* mask_src = A.mask ? A.mask : A.chi;
    const bool has_u2 = A.U2r != nullptr;
    const float* u2r = has_u2 ? A.U2r : A.chi;
    const float* u2i = has_u2 ? A.U2i : A.chi;
    const float* x2 = has_u2 ? A.X2 : A.chi;
    auto load_point = [&](float (&v)[5], int s) {
        // planes outside the chunk's sites load plane i0 and are never used
        const int im = (s >= 2 && s < nplanes) ? i0 + s - 2 : i0;
        const uint32_t ob = uint32_t(im) * NN4 + cb;
        const uint32_t om = A.mask ? ob : 0u, o2 = has_u2 ? ob : 0u;
        v[0] = ldb(A.chi, ob);
        const float mv = ldbp<NT>(mask_src, om);
        v[1] = A.mask ? mv : 0.0f;
        const float a2 = ldbp<NT>(u2r, o2), b2 = ldbp<NT>(u2i, o2), c2 = ldbp<NT>(x2, o2);
        v[2] = has_u2 ? a2 : 0.0f;
        v[3] = has_u2 ? b2 : 0.0f;
        v[4] = has_u2 ? c2 : 0.0f;
    };
    load_plane(sa, 0);
    if (nplanes > 1) load_plane(sb, 1);
    for (int s = 0; s < nplanes; ++s) {
        load_point(pn, s + 1);
        // stage plane s from registers, then prefetch plane s + 2
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            if (!sact[r]) continue;
            const float xt = sa[r][2];
            const float w = wk * (A.chi0 + sa[r][5]) * xt;
            const int p = threadIdx.x + r * AB;
            lds[0 * LPL + p] = sa[r][0];
            lds[1 * LPL + p] = sa[r][1];
            lds[2 * LPL + p] = xt;
            lds[3 * LPL + p] = sa[r][3];
            lds[4 * LPL + p] = sa[r][4];
            lds[5 * LPL + p] = sa[r][4] * w;
            lds[6 * LPL + p] = sa[r][3] * w;
        }
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int f = 0; f < 6; ++f) sa[r][f] = sb[r][f];
        if (s + 2 < nplanes) load_plane(sb, s + 2);
        __syncthreads();
        // fold plane s into the running sums: finished sums of site p in fin[]
        float fin[7];
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            const float* L = lds + f * LPL + c0;
            const float c = L[0];
            const float fs = L[-1] + L[1] + L[-LWK] + L[LWK];
            const float es = L[-LWK - 1] + L[-LWK + 1] + L[LWK - 1] + L[LWK + 1];
            const float B = WF * c + WE * fs;
            fin[f] = accp[f] + B;
            accp[f] = accc[f] + (WF * fs + WE * es - 4.0f * c);
            accc[f] = B;
        }
        const float* P3 = lds + 3 * LPL + c0;
        const float* P4 = lds + 4 * LPL + c0;
        const float* P5 = lds + 5 * LPL + c0;
        const float* P6 = lds + 6 * LPL + c0;
        const float npr = P3[0], npi = P4[0], npw = P5[0], nrw = P6[0];
        fin[3] = accp[3] + npr;
        fin[4] = accp[4] + npi;
        fin[5] = accp[5] + npw;
        fin[6] = accp[6] - nrw;
        accp[3] = accc[3] + (P3[1] - P3[-1]) + (P3[LWK] - P3[-LWK]);
        accp[4] = accc[4] + (P4[1] - P4[-1]) + (P4[LWK] - P4[-LWK]);
        accp[5] = accc[5] + (P5[1] - P5[-1]) + (P5[LWK] - P5[-LWK]);
        accp[6] = accc[6] + (P6[-1] - P6[1]) + (P6[-LWK] - P6[LWK]);
        accc[3] = -npr;
        accc[4] = -npi;
        accc[5] = -npw;
        accc[6] = nrw;
        const int im = i0 + s - 2;
        {
            // site (im, j, k); formed on every lane, stored where it exists
            const uint32_t ob = uint32_t(im < 0 ? 0 : im) * NN4 + cb;
            const float cs = pw[0], m = pw[1], v2r = pw[2], v2i = pw[3], y2 = pw[4];
            const float chi = A.chi0 + cs;
            const float chi_sq = chi * chi;
            const float jsum = cpr * fin[4] - cpi * fin[3];
            const float S = cpr * cpr + cpi * cpi + A.epsw * (0.5f * jsum) - A.e0sq;
            const float qv = -A.dt2 * kc * chi * cxt;
            const float w = A.epsw * qv;
            const float a = 1.0f - m;
            const float ur = 2.0f * cur + A.dt2 * (fin[0] - chi_sq * cur) - a * v2r +
                             2.0f * cpr * qv + 0.5f * w * fin[4] + 0.5f * fin[5];
            const float ui = 2.0f * cui + A.dt2 * (fin[1] - chi_sq * cui) - a * v2i +
                             2.0f * cpi * qv - 0.5f * w * fin[3] + 0.5f * fin[6];
            const float dacc = -kc * S - 4.0f * A.lam * (3.0f * chi_sq - A.chi0 * A.chi0);
            const float xn = cxt * (2.0f + A.dt2 * dacc) + A.dt2 * fin[2] -
                             2.0f * A.dt2 * chi * (cpr * cur + cpi * cui) - y2;
            const bool site = s >= 2 && live;
            if (site) kp += double(cxt) * (double(chi) * double(S));
            float o_r = ur, o_i = ui, o_x = xn;
            if (A.damp_out) {
                o_r = a * ur;
                o_i = a * ui;
                o_x = a * clamp_free(cs, a, A.chi0) * xn;
            }
            // The prefetched point and plane data are waited for here, after
            // the outputs are formed and before they are stored: loads and
            // stores share one wait counter, and a wait behind these stores
            // (at the next iteration's use of the prefetch) would drain them.
#pragma unroll
            for (int f = 0; f < 5; ++f) asm volatile("" : "+v"(o_r) : "v"(pn[f]));
#pragma unroll
            for (int r = 0; r < 2; ++r)
#pragma unroll
                for (int f = 0; f < 6; ++f) asm volatile("" : "+v"(o_r) : "v"(sb[r][f]));
            if (site) {
                stb<NT>(A.Or, ob, o_r);
                stb<NT>(A.Oi, ob, o_i);
                stb<NT>(A.Ox, ob, o_x);
            }
        }
        // plane s's centre values belong to the site finished next iteration
        cpr = npr;
        cpi = npi;
        cur = lds[0 * LPL + c0];
        cui = lds[1 * LPL + c0];
        cxt = lds[2 * LPL + c0];
#pragma unroll
        for (int f = 0; f < 5; ++f) pw[f] = pn[f];
    }
    // kappa-gradient term: sum Xt dt2 (-chi S / chi0)
    const double bs = block_sum(kp) * (-double(A.dt2) / double(A.chi0));
    if (threadIdx.x == 0) A.kpart[blockIdx.x] = bs;
}

__global__ __launch_bounds__(256) void sum_into(const double* part, int n, double* acc) {
    double v = 0.0;
    for (int t = threadIdx.x; t < n; t += 256) v += part[t];
    __shared__ double sh[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) *acc += sh[0] + sh[1] + sh[2] + sh[3];
}

// Workgroups of adj_fused the current device holds at once (CUs x
// resident blocks per CU); 1024 when it cannot be queried.
int adj_capacity() {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, adj_fused<true>, AB, 0) != hipSuccess ||
        cus <= 0 || per <= 0) {
        (void)hipGetLastError();
        return 1024;
    }
    return cus * per;
}

// Planes per chunk: as many chunks as keep every tile's workgroup resident in
// one round (a partial second round would double the time), at least 4
// planes per chunk so the two halo planes stay a modest share of the reads.
int adj_chunk(int N) {
    static const int cap = adj_capacity();
    const int tiles = ((N + BK - 1) / BK) * ((N + BJ - 1) / BJ);
    int nch = cap / tiles;
    if (nch < 1) nch = 1;
    int ch = (N + nch - 1) / nch;
    if (ch < 4) ch = 4;
    return ch < N ? ch : N;
}

unsigned adj_blocks(int N) {
    const int ch = adj_chunk(N);
    return unsigned(((N + BK - 1) / BK) * ((N + BJ - 1) / BJ) * ((N + ch - 1) / ch));
}

}  // namespace
}  // namespace lfm

using namespace lfm;

extern "C" int64_t lfm_adjoint_partials_len(int32_t N) {
    if (N < 1) return 0;
    return int64_t(adj_blocks(N));
}

static int adj_check_common(const char* fn, const lfm_params_t* p, const float* mask) {
    if (!p) { set_error("%s: params is NULL", fn); return LFM_E_ARG; }
    if (p->N < 8 || p->N > 1023) { set_error("%s: N=%d out of range (8..1023)", fn, p->N); return LFM_E_ARG; }
    if (!(p->flags & LFM_F_CHI_DELTA)) {
        set_error("%s: needs chi in deviation storage (LFM_F_CHI_DELTA)", fn);
        return LFM_E_ARG;
    }
    if ((p->flags & LFM_F_MASK) && !mask) {
        set_error("%s: LFM_F_MASK set but mask is NULL", fn);
        return LFM_E_ARG;
    }
    if (p->flags & (LFM_F_KEEP_PSI | LFM_F_SOURCE | LFM_F_DETECTOR)) {
        set_error("%s: LFM_F_KEEP_PSI / LFM_F_SOURCE / LFM_F_DETECTOR are not supported", fn);
        return LFM_E_ARG;
    }
    return LFM_OK;
}

extern "C" int lfm_adjoint_complex_seed_f32(const float* chi, const float* mask, const float* U_r,
                                            const float* U_i, const float* X, float* Ut_r,
                                            float* Ut_i, float* Xt, const lfm_params_t* p,
                                            void* stream) {
    if (int rc = adj_check_common("lfm_adjoint_complex_seed_f32", p, mask)) return rc;
    if (!chi || !U_r || !U_i || !X || !Ut_r || !Ut_i || !Xt) {
        set_error("lfm_adjoint_complex_seed_f32: required pointer is NULL");
        return LFM_E_ARG;
    }
    const uint32_t n3 = uint32_t(p->N) * p->N * p->N;
    hipLaunchKernelGGL(adj_seed, dim3((n3 + AB - 1) / AB), dim3(AB), 0, (hipStream_t)stream, chi,
                       (p->flags & LFM_F_MASK) ? mask : nullptr, U_r, U_i, X, Ut_r, Ut_i, Xt, n3,
                       p->chi0);
    LFM_LAUNCH_CHECK("adj_seed");
    return LFM_OK;
}

extern "C" int lfm_adjoint_complex_step_f32(const float* pr, const float* pi, const float* chi,
                                            const float* mask, const float* Ut_r,
                                            const float* Ut_i, const float* Xt,
                                            const float* Ut2_r, const float* Ut2_i,
                                            const float* Xt2, float* out_r, float* out_i,
                                            float* out_x, int32_t damp_out, double* partials,
                                            double* kappa_grad, const lfm_params_t* p,
                                            void* stream) {
    if (int rc = adj_check_common("lfm_adjoint_complex_step_f32", p, mask)) return rc;
    if (!pr || !pi || !chi || !Ut_r || !Ut_i || !Xt || !out_r || !out_i || !out_x || !partials) {
        set_error("lfm_adjoint_complex_step_f32: required pointer is NULL");
        return LFM_E_ARG;
    }
    if ((Ut2_r == nullptr) != (Ut2_i == nullptr) || (Ut2_r == nullptr) != (Xt2 == nullptr)) {
        set_error("lfm_adjoint_complex_step_f32: Ut2_r, Ut2_i, Xt2 must all be given or all NULL");
        return LFM_E_ARG;
    }
    const float* ins[] = {Ut_r, Ut_i, Xt};
    const float* outs[] = {out_r, out_i, out_x};
    for (const float* o : outs)
        for (const float* in : ins)
            if (o == in) {
                set_error("lfm_adjoint_complex_step_f32: outputs must not alias Ut / Xt");
                return LFM_E_ARG;
            }
    AdjArgs A{pr, pi, chi, (p->flags & LFM_F_MASK) ? mask : nullptr, Ut_r, Ut_i, Xt, Ut2_r,
              Ut2_i, Xt2, out_r, out_i, out_x, partials, p->N, p->flags, damp_out ? 1 : 0,
              adj_chunk(p->N), p->dt2, p->kappa, p->lambda_self, p->chi0, p->e0_sq, p->epsilon_w};
    hipStream_t st = (hipStream_t)stream;
    const unsigned nb = adj_blocks(p->N);
    // 13 float fields per site; streaming once far beyond the 256 MB last-level cache
    if (size_t(p->N) * p->N * p->N * 52 > (size_t(512) << 20))
        hipLaunchKernelGGL(adj_fused<true>, dim3(nb), dim3(AB), 0, st, A);
    else
        hipLaunchKernelGGL(adj_fused<false>, dim3(nb), dim3(AB), 0, st, A);
    LFM_LAUNCH_CHECK("adj_fused");
    if (kappa_grad) {
        hipLaunchKernelGGL(sum_into, dim3(1), dim3(256), 0, st, (const double*)partials,
                           int(nb), kappa_grad);
        LFM_LAUNCH_CHECK("sum_into");
    }
    return LFM_OK;
}

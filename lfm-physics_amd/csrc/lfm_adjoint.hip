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
// One reverse step n <- n+1 is two kernels:
//   adj_prep : Ut, Xt of step n+1 from the forward state n (recomputes H) and
//              the kappa-gradient partial sum(Xt dt2 (-chi S / chi0)).
//   adj_step : U_n = 2Ut + dt2 (L Ut - chi^2 Ut) - a Ut_{n+2} + (dH/dPsi)^T Xt
//              X_n = Xt (2 + dt2 dacc/dchi) + dt2 L Xt - 2 dt2 chi (Pr Utr + Pi Uti)
//                    - Xt_{n+2}
// with (dH/dPr)^T Xt = 2 Pr q + w/2 sum_d dPi_d + 1/2 sum_d [Pi w](+d) - [Pi w](-d),
//      (dH/dPi)^T Xt = 2 Pi q - w/2 sum_d dPr_d + 1/2 sum_d [Pr w](-d) - [Pr w](+d),
// q = -dt2 (kappa/chi0) chi Xt, w = eps_w q.  L is symmetric, so L^T = L.
// All chi buffers use the step kernels' storage (chi or chi - chi0).
#include "lfm_internal.h"

namespace lfm {
namespace {

constexpr float WF = 1.0f / 3.0f, WE = 1.0f / 6.0f;
constexpr int AB = 256;

struct AdjArgs {
    const float* pr; const float* pi; const float* chi; const float* chip; const float* mask;
    const float* U_r; const float* U_i; const float* X;          // step n+1
    float* Ut_r; float* Ut_i; float* Xt;                         // prep outputs
    const float* Ut2_r; const float* Ut2_i; const float* Xt2;    // step n+2 (NULL at the end)
    float* Un_r; float* Un_i; float* Xn;                         // step n outputs
    double* kpart;
    int N, flags;
    float dt2, kappa, lam, chi0, e0sq, epsw;
};

struct Nb {
    size_t o[18];   // 6 faces (+i,-i,+j,-j,+k,-k) then 12 edges
};

__device__ __forceinline__ Nb neighbours(size_t idx, int N) {
    const size_t NN = size_t(N) * N;
    const int i = int(idx / NN), j = int((idx / N) % N), k = int(idx % N);
    const size_t rc = size_t(i) * NN, rp = size_t((i + 1) % N) * NN, rm = size_t((i + N - 1) % N) * NN;
    const size_t cc = size_t(j) * N, cp = size_t((j + 1) % N) * N, cm = size_t((j + N - 1) % N) * N;
    const size_t dc = k, dp = (k + 1) % N, dm = (k + N - 1) % N;
    Nb n;
    const size_t v[18] = {rp + cc + dc, rm + cc + dc, rc + cp + dc, rc + cm + dc, rc + cc + dp,
                          rc + cc + dm, rp + cp + dc, rp + cm + dc, rm + cp + dc, rm + cm + dc,
                          rp + cc + dp, rp + cc + dm, rm + cc + dp, rm + cc + dm, rc + cp + dp,
                          rc + cp + dm, rc + cm + dp, rc + cm + dm};
#pragma unroll
    for (int t = 0; t < 18; ++t) n.o[t] = v[t];
    return n;
}

__device__ __forceinline__ float lap(const float* F, const Nb& n, float c) {
    float fs = 0.0f, es = 0.0f;
#pragma unroll
    for (int t = 0; t < 6; ++t) fs += F[n.o[t]];
#pragma unroll
    for (int t = 6; t < 18; ++t) es += F[n.o[t]];
    return WF * fs + WE * es - 4.0f * c;
}

__device__ __forceinline__ double block_sum(double v) {
    __shared__ double sh[AB / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int t = 0; t < AB / 64; ++t) r += sh[t];
    return r;
}

__global__ __launch_bounds__(AB) void adj_prep(AdjArgs A) {
    const size_t n3 = size_t(A.N) * A.N * A.N;
    const size_t idx = size_t(blockIdx.x) * AB + threadIdx.x;
    double kp = 0.0;
    if (idx < n3) {
        const Nb nb = neighbours(idx, A.N);
        const bool delta = (A.flags & LFM_F_CHI_DELTA) != 0;
        const float cs = A.chi[idx];
        const float chi = delta ? A.chi0 + cs : cs;
        const float pr = A.pr[idx], pi = A.pi[idx];
        float jsum = 0.0f;
#pragma unroll
        for (int d = 0; d < 3; ++d)
            jsum += pr * (A.pi[nb.o[2 * d]] - A.pi[nb.o[2 * d + 1]]) -
                    pi * (A.pr[nb.o[2 * d]] - A.pr[nb.o[2 * d + 1]]);
        const float S = pr * pr + pi * pi + A.epsw * (0.5f * jsum) - A.e0sq;
        const float hat = delta ? cs * (2.0f * A.chi0 + cs) : chi * chi - A.chi0 * A.chi0;
        const float acc = lap(A.chi, nb, cs) - (A.kappa / A.chi0) * chi * S - 4.0f * A.lam * chi * hat;
        const float H = 2.0f * cs - A.chip[idx] + A.dt2 * acc;
        const float floor_v = delta ? -2.0f * A.chi0 : -A.chi0;
        const float a = (A.flags & LFM_F_MASK) ? 1.0f - A.mask[idx] : 1.0f;
        const float xt = H > floor_v ? a * A.X[idx] : 0.0f;
        A.Xt[idx] = xt;
        A.Ut_r[idx] = a * A.U_r[idx];
        A.Ut_i[idx] = a * A.U_i[idx];
        kp = double(xt) * double(A.dt2) * (-double(chi) * double(S) / double(A.chi0));
    }
    const double bs = block_sum(kp);
    if (threadIdx.x == 0) A.kpart[blockIdx.x] = bs;
}

__global__ __launch_bounds__(AB) void adj_step(AdjArgs A) {
    const size_t n3 = size_t(A.N) * A.N * A.N;
    const size_t idx = size_t(blockIdx.x) * AB + threadIdx.x;
    if (idx >= n3) return;
    const Nb nb = neighbours(idx, A.N);
    const bool delta = (A.flags & LFM_F_CHI_DELTA) != 0;
    const float cs = A.chi[idx];
    const float chi = delta ? A.chi0 + cs : cs;
    const float chi_sq = chi * chi;
    const float pr = A.pr[idx], pi = A.pi[idx];
    const float utr = A.Ut_r[idx], uti = A.Ut_i[idx], xt = A.Xt[idx];
    const float kc = A.kappa / A.chi0;
    // forward source S and its colour current
    float dpr[3], dpi[3], jsum = 0.0f;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        dpr[d] = A.pr[nb.o[2 * d]] - A.pr[nb.o[2 * d + 1]];
        dpi[d] = A.pi[nb.o[2 * d]] - A.pi[nb.o[2 * d + 1]];
        jsum += pr * dpi[d] - pi * dpr[d];
    }
    const float S = pr * pr + pi * pi + A.epsw * (0.5f * jsum) - A.e0sq;
    // transposed current: gathers of Pi w and Pr w from the face neighbours
    const float q = -A.dt2 * kc * chi * xt;
    const float w = A.epsw * q;
    float gr = 0.0f, gi = 0.0f, sdpi = 0.0f, sdpr = 0.0f;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const size_t up = nb.o[2 * d], dn = nb.o[2 * d + 1];
        const float chu = delta ? A.chi0 + A.chi[up] : A.chi[up];
        const float chd = delta ? A.chi0 + A.chi[dn] : A.chi[dn];
        const float wu = A.epsw * (-A.dt2 * kc * chu * A.Xt[up]);
        const float wd = A.epsw * (-A.dt2 * kc * chd * A.Xt[dn]);
        gr += A.pi[up] * wu - A.pi[dn] * wd;
        gi += A.pr[dn] * wd - A.pr[up] * wu;
        sdpi += dpi[d];
        sdpr += dpr[d];
    }
    const float a = (A.flags & LFM_F_MASK) ? 1.0f - A.mask[idx] : 1.0f;
    const float v2r = A.Ut2_r ? A.Ut2_r[idx] : 0.0f;
    const float v2i = A.Ut2_i ? A.Ut2_i[idx] : 0.0f;
    const float y2 = A.Xt2 ? A.Xt2[idx] : 0.0f;
    const float ur = 2.0f * utr + A.dt2 * (lap(A.Ut_r, nb, utr) - chi_sq * utr) - a * v2r +
                     2.0f * pr * q + 0.5f * w * sdpi + 0.5f * gr;
    const float ui = 2.0f * uti + A.dt2 * (lap(A.Ut_i, nb, uti) - chi_sq * uti) - a * v2i +
                     2.0f * pi * q - 0.5f * w * sdpr + 0.5f * gi;
    const float dacc = -kc * S - 4.0f * A.lam * (3.0f * chi_sq - A.chi0 * A.chi0);
    const float xn = xt * (2.0f + A.dt2 * dacc) + A.dt2 * lap(A.Xt, nb, xt) -
                     2.0f * A.dt2 * chi * (pr * utr + pi * uti) - y2;
    A.Un_r[idx] = ur;
    A.Un_i[idx] = ui;
    A.Xn[idx] = xn;
}

__global__ __launch_bounds__(AB) void sum_into(const double* part, int n, double* acc) {
    double v = 0.0;
    for (int t = threadIdx.x; t < n; t += AB) v += part[t];
    const double s = block_sum(v);
    if (threadIdx.x == 0) *acc += s;
}

}  // namespace
}  // namespace lfm

using namespace lfm;

extern "C" int64_t lfm_adjoint_partials_len(int32_t N) {
    const size_t n3 = size_t(N) * N * N;
    return int64_t((n3 + AB - 1) / AB);
}

extern "C" int lfm_adjoint_complex_step_f32(
    const float* pr, const float* pi, const float* chi, const float* chi_prev, const float* mask,
    const float* U_r, const float* U_i, const float* X, float* Ut_r, float* Ut_i, float* Xt,
    const float* Ut2_r, const float* Ut2_i, const float* Xt2, float* Un_r, float* Un_i,
    float* Xn, double* partials, double* kappa_grad, const lfm_params_t* p, void* stream) {
    if (!p || !pr || !pi || !chi || !chi_prev || !U_r || !U_i || !X || !Ut_r || !Ut_i || !Xt ||
        !Un_r || !Un_i || !Xn || !partials) {
        set_error("lfm_adjoint_complex_step_f32: required pointer is NULL");
        return LFM_E_ARG;
    }
    if ((p->flags & LFM_F_MASK) && !mask) {
        set_error("lfm_adjoint_complex_step_f32: LFM_F_MASK set but mask is NULL");
        return LFM_E_ARG;
    }
    if (p->N < 8 || p->N > 1290) {
        set_error("lfm_adjoint_complex_step_f32: N=%d out of range", p->N);
        return LFM_E_ARG;
    }
    if ((Ut2_r == nullptr) != (Ut2_i == nullptr)) {
        set_error("lfm_adjoint_complex_step_f32: Ut2_r and Ut2_i must both be given or both NULL");
        return LFM_E_ARG;
    }
    AdjArgs A{pr, pi, chi, chi_prev, mask, U_r, U_i, X, Ut_r, Ut_i, Xt, Ut2_r, Ut2_i, Xt2,
              Un_r, Un_i, Xn, partials, p->N, p->flags, p->dt2, p->kappa, p->lambda_self,
              p->chi0, p->e0_sq, p->epsilon_w};
    const size_t n3 = size_t(p->N) * p->N * p->N;
    const unsigned nb = unsigned((n3 + AB - 1) / AB);
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(adj_prep, dim3(nb), dim3(AB), 0, st, A);
    LFM_LAUNCH_CHECK("adj_prep");
    hipLaunchKernelGGL(adj_step, dim3(nb), dim3(AB), 0, st, A);
    LFM_LAUNCH_CHECK("adj_step");
    if (kappa_grad) {
        hipLaunchKernelGGL(sum_into, dim3(1), dim3(AB), 0, st, (const double*)partials, int(nb),
                           kappa_grad);
        LFM_LAUNCH_CHECK("sum_into");
    }
    return LFM_OK;
}

// GOV-01/GOV-02 leapfrog step kernels for gfx950 (MI355X).
//
// Two kernels compute the same step:
//  * step_tiled  -- the production kernel.  A workgroup owns a 64 (k) x TJ (j)
//                   column tile and marches along i (the slowest axis) over a
//                   chunk of planes.  Each plane of every stencilled field is
//                   read from HBM once (plus a one-cell j/k halo), staged in
//                   LDS, and split into an in-plane partial A(q) and an
//                   off-plane partial B(q) of the 19-point Laplacian:
//                       A(q) = 1/3 (4 in-plane faces) + 1/6 (4 in-plane edges) - 4 f
//                       B(q) = 1/3 f + 1/6 (4 in-plane faces)
//                       lap(i) = A(i) + B(i-1) + B(i+1)
//                   A register queue carries A and B from plane to plane, so
//                   the i-direction never re-reads HBM.
//  * step_direct -- one thread per site, 19 scattered loads per field: an
//                   independent second implementation for testing, and the
//                   faster one on small cache-resident lattices (32^3),
//                   where the tiled march has too few planes to hide latency.
//
// Physics per site (reference: lfm/core/backends/kernel_source.py:86-233 for
// COLOR, :377-417 REAL, :483-538 COMPLEX; numpy_backend.py:80-382):
//   Psi' = 2 Psi - Psi_p + dt2 (lap Psi - chi^2 Psi) [- dt2 eps_cc chi^2 (Psi_a - Psi_bar)]
//   chi' = 2 chi - chi_p + dt2 (lap chi - (kappa/chi0) chi (|Psi|^2 + eps_w J - E0^2)
//                               - CV + chi_self - kappa_s CCV - kappa_t SCV)
//   chi' = max(chi', -chi0); chi' = m chi0 + (1-m) chi'; Psi' *= (1-m).
#include "lfm_internal.h"
#include "lfm_fft.h"
#include "lfm_stencil.h"

#include <algorithm>
#include <type_traits>
#include <vector>

namespace lfm {
namespace {

// Current arithmetic with every rounding spelled out, as the Laplacian
// partials in lfm_stencil.h: a b - c d (one current component)
__device__ __forceinline__ float cross2(float a, float b, float c, float d) {
    return fmaf(a, b, -mul_rn(c, d));
}

enum Level { REAL = 0, COMPLEX = 1, COLOR = 2 };

// Number of real Psi component fields (Pr_0..Pr_{C-1}, Pi_0..Pi_{C-1}) and of
// complex colours per level.
template <int L> struct Lv;
template <> struct Lv<REAL> { static constexpr int NPF = 1, NC = 0; };
template <> struct Lv<COMPLEX> { static constexpr int NPF = 2, NC = 1; };
template <> struct Lv<COLOR> { static constexpr int NPF = 6, NC = 3; };

struct KParams {
    int N, flags;
    float dt2, kappa, lam, chi0, e0sq, epsw, kc, ecc, ks, kt;
    int tile, chunk;
    // kappa / chi0 and kappa_c / chi0, each one correctly rounded float32
    // division done once on the host (the same value a device division with
    // round-to-nearest gives), so no kernel divides per site
    float kc0, kcc0;
    // planes [ib, ie) along i are computed (0, N for a whole lattice; a slab
    // share computes its interior planes and reads the two halo planes)
    int ib, ie;
    // step index of this launch = *counter + step_off (LFM_F_SOURCE, LFM_F_DETECTOR)
    int step_off;
    // plane source (LFM_F_SOURCE): axis, plane, step indexes in the amount table
    int src_axis, src_plane;
    long long src_len;
    // detector plane (LFM_F_DETECTOR): axis, plane, recording interval, field
    // (0 |Psi|^2, 1 Pr), first recorded step index, frames in the buffer
    int det_axis, det_plane, det_every, det_field;
    long long det_first, det_len;
};

struct Ptrs {
    const float* pr;     // Psi real parts (or E), NC*N^3 (N^3 for REAL)
    const float* pi;     // Psi imaginary parts (NULL for REAL)
    const float* prp;    // prev
    const float* pip;
    const float* chi;
    const float* chip;
    const float* mask;   // NULL when !LFM_F_MASK
    const float* sa;     // NULL unless COLOR with kappa_tube > 0
    float* pr_out;
    float* pi_out;
    float* prp_out;      // NULL in rotation mode
    float* pip_out;
    float* chi_out;
    float* chip_out;     // NULL in rotation mode
    float* sa_clip;      // NULL or max(sa, 0)
    float* psisq;        // NULL or |Psi'_a|^2 of the new fields
    const float* keep;   // NULL unless LFM_F_KEEP_PSI: per-site factor on Psi'
    // NULL unless LFM_F_SOURCE: N*N envelope, 2 amounts per step index
    const float* src_env;
    const float* src_amp;
    // NULL unless LFM_F_DETECTOR: det_len frames of N*N values
    float* det_frames;
    // NULL unless LFM_F_SOURCE or LFM_F_DETECTOR: the device step counter
    const long long* counter;
    size_t n3;           // N^3
};

// Component field f of a level: f < NC -> real part of colour f, else the
// imaginary part of colour f-NC (REAL: f = 0 is E).
template <int L>
__device__ __forceinline__ const float* field(const float* pr, const float* pi, int f,
                                              size_t n3) {
    constexpr int C = Lv<L>::NC;
    if constexpr (L == REAL) return pr;
    return f < C ? pr + f * n3 : pi + (f - C) * n3;
}
template <int L>
__device__ __forceinline__ float* field_out(float* pr, float* pi, int f, size_t n3) {
    constexpr int C = Lv<L>::NC;
    if constexpr (L == REAL) return pr;
    return f < C ? pr + f * n3 : pi + (f - C) * n3;
}

// Site addressing.  Off32 is a byte offset shared by every field (all fields
// have the same layout), so each access is one uniform 64-bit base held in
// SGPRs plus one 32-bit lane offset (the saddr form of global loads); it
// covers N <= 1024 (the largest offset, 4 N^3 - 4, is below 2^32).  Off64 is
// a plain element index for any N.
// With NT the pointwise inputs (ldp) and every output (st) are streaming
// (non-temporal) accesses: on lattices far larger than the last-level cache
// they stop data touched once per step from pushing the stencil fields' halo
// lines, which neighbouring tiles re-read, out of L2.  On lattices whose
// state fits in the last-level cache the plain accesses keep that reuse
// across steps, so the launcher picks NT by working-set size.
template <bool NT>
struct Off32T {
    uint32_t b;
    __device__ __forceinline__ const float* at(const float* p) const {
        return reinterpret_cast<const float*>(reinterpret_cast<const char*>(p) + b);
    }
    __device__ __forceinline__ float ld(const float* p) const { return *at(p); }
    __device__ __forceinline__ float ldp(const float* p) const {
        if constexpr (NT) return __builtin_nontemporal_load(at(p));
        else return *at(p);
    }
    __device__ __forceinline__ void st(float* p, float v) const {
        float* q = reinterpret_cast<float*>(reinterpret_cast<char*>(p) + b);
        if constexpr (NT) __builtin_nontemporal_store(v, q);
        else *q = v;
    }
};
using Off32 = Off32T<false>;
struct Off64 {
    size_t i;
    __device__ __forceinline__ float ld(const float* p) const { return p[i]; }
    __device__ __forceinline__ float ldp(const float* p) const { return p[i]; }
    __device__ __forceinline__ void st(float* p, float v) const { p[i] = v; }
};

// Pointwise inputs of one site update (read once, not stencilled): the prev
// Psi components, chi_prev, the mask and the three S_a values.  They are
// loaded separately from the update so the tiled kernel can issue them one
// plane ahead and hide their latency behind the current plane's work.
template <int L>
struct PointData {
    float prev[Lv<L>::NPF];
    float chip, m;
    float sa[3];
};

template <int L, class O, bool SA = true>
__device__ __forceinline__ PointData<L> load_point(const KParams& kp, const Ptrs& P, O idx) {
    constexpr int NPF = Lv<L>::NPF;
    PointData<L> d;
#pragma unroll
    for (int f = 0; f < NPF; ++f) d.prev[f] = idx.ldp(field<L>(P.prp, P.pip, f, P.n3));
    d.m = (kp.flags & LFM_F_MASK) ? idx.ldp(P.mask) : 0.0f;
    d.chip = (kp.flags & LFM_F_FREEZE_CHI) ? 0.0f : idx.ldp(P.chip);
    if constexpr (L == COLOR && SA) {
        if (P.sa) {
            if (kp.flags & LFM_F_SA_SCV) {
                d.sa[0] = idx.ldp(P.sa);          // the SCV term itself
            } else {
#pragma unroll
                for (int q = 0; q < 3; ++q) d.sa[q] = idx.ldp(P.sa + q * P.n3);
            }
        }
    }
    return d;
}

// Per-site arithmetic of one leapfrog step, shared by every kernel.
//  c[f]     Psi component values at the site      lap[f]  their Laplacians
//  jsum     sum over colours and directions of j_{a,d}   (COMPLEX/COLOR)
//  ccv      sum_d (sum_a j_{a,d}^2 - 1/3 (sum_a j_{a,d})^2)  (COLOR)
//  chi_st, lap_chi  chi as stored and the Laplacian of the stored field
//  scv_on   the S_a term is on (COLOR); pd.sa holds the raw S_a, or with
//           LFM_F_SA_SCV the SCV term in pd.sa[0]
// Out: nw (new Psi), cn (new chi as stored; chi_st under LFM_F_FREEZE_CHI),
// sclip (S_a clipped at 0, COLOR with scv_on).
template <int L>
__device__ __forceinline__ void site_compute(const KParams& kp, const PointData<L>& pd,
                                             bool scv_on, const float (&c)[Lv<L>::NPF],
                                             const float (&lap)[Lv<L>::NPF], float chi_st,
                                             float lap_chi, float jsum, float ccv,
                                             float (&nw)[Lv<L>::NPF], float& cn,
                                             float (&sclip)[3]) {
    // Every rounding is explicit (fmaf, mul_rn, add_rn, sub_rn: none is
    // fused into a multiply-add), so the result does not depend on the code
    // around the call in a given kernel: all kernels produce bit-identical
    // updates for identical inputs.
    constexpr int NPF = Lv<L>::NPF;
    constexpr int C = Lv<L>::NC;
    // chi_st is chi as stored: chi itself, or chi - chi0 with LFM_F_CHI_DELTA.
    const bool delta = (kp.flags & LFM_F_CHI_DELTA) != 0;
    const float chi_c = delta ? add_rn(kp.chi0, chi_st) : chi_st;
    const bool has_mask = (kp.flags & LFM_F_MASK) != 0;
    const float m = has_mask ? pd.m : 0.0f;
    const float a = sub_rn(1.0f, m);
    const float chi_sq = mul_rn(chi_c, chi_c);

    float avg[2] = {0.0f, 0.0f};
    if constexpr (L == COLOR) {
        if (kp.ecc > 0.0f) {
            avg[0] = mul_rn(add_rn(add_rn(c[0], c[1]), c[2]), 1.0f / 3.0f);
            avg[1] = mul_rn(add_rn(add_rn(c[3], c[4]), c[5]), 1.0f / 3.0f);
        }
    }
    const bool damp = has_mask && (kp.flags & LFM_F_DAMP_PREV);
#pragma unroll
    for (int f = 0; f < NPF; ++f) {
        float pv = pd.prev[f];
        if (damp) pv = mul_rn(pv, a);
        // 2 c - pv + dt2 (lap - chi^2 c); 2 c is exact, so fmaf(2, c, -pv)
        // rounds 2 c - pv once, as the subtraction does
        float v = fmaf(kp.dt2, fmaf(-chi_sq, c[f], lap[f]), fmaf(2.0f, c[f], -pv));
        if constexpr (L == COLOR) {
            if (kp.ecc > 0.0f)
                v = fmaf(-mul_rn(mul_rn(kp.dt2, kp.ecc), chi_sq), sub_rn(c[f], avg[f / C]), v);
        }
        nw[f] = mul_rn(v, a);
    }
    cn = chi_st;
    if (kp.flags & LFM_F_FREEZE_CHI) return;

    const float kc0 = kp.kc0;
    float psi_sq;
    // the coupling source (kappa/chi0) chi (|Psi|^2 + eps_w j - E0^2) as
    // factor src_f = (kappa/chi0) chi times src_e, and the colour-variance
    // term CV as cv_f psi_sq
    float src_f, src_e;
    float cv_f = 0.0f, scv = 0.0f;
    if constexpr (L == REAL) {
        psi_sq = mul_rn(c[0], c[0]);
        src_f = mul_rn(kc0, chi_c);
        src_e = sub_rn(psi_sq, kp.e0sq);
    } else {
        float e[C > 0 ? C : 1];
        psi_sq = 0.0f;
#pragma unroll
        for (int q = 0; q < C; ++q) {
            e[q] = fmaf(c[q], c[q], mul_rn(c[q + C], c[q + C]));
            psi_sq = add_rn(psi_sq, e[q]);
        }
        const float j_total = mul_rn(0.5f, jsum);
        src_f = mul_rn(kc0, chi_c);
        src_e = sub_rn(fmaf(kp.epsw, j_total, psi_sq), kp.e0sq);
        if constexpr (L == COLOR) {
            if (kp.kc > 0.0f) {
                const float total_sq = mul_rn(psi_sq, psi_sq);
                if (total_sq > 1e-30f) {
                    const float sum_sq = fmaf(e[0], e[0], fmaf(e[1], e[1], mul_rn(e[2], e[2])));
                    const float f_c = sub_rn(__fdiv_rn(sum_sq, total_sq), 1.0f / 3.0f);
                    cv_f = mul_rn(mul_rn(kp.kcc0, chi_c), f_c);
                }
            }
            if (scv_on) {
                // LFM_F_SA_SCV: pd.sa[0] already holds scv_term of the S_a
                scv = (kp.flags & LFM_F_SA_SCV) ? pd.sa[0]
                                                : scv_term(pd.sa[0], pd.sa[1], pd.sa[2], sclip);
            }
        }
    }
    // chi_self = -4 lam chi (chi^2 - chi0^2); with deviation storage the
    // factor chi^2 - chi0^2 is formed as d (2 chi0 + d), free of cancellation.
    const float hat = delta ? mul_rn(chi_st, fmaf(2.0f, kp.chi0, chi_st))
                            : fmaf(-kp.chi0, kp.chi0, chi_sq);
    // lap chi - src - CV + chi_self, each product folded into its sum (fmaf)
    float accel = fmaf(-src_f, src_e, lap_chi);
    if constexpr (L == COLOR) accel = fmaf(-cv_f, psi_sq, accel);
    accel = fmaf(mul_rn(-4.0f * kp.lam, chi_c), hat, accel);
    if constexpr (L == COLOR) accel = fmaf(-kp.kt, scv, fmaf(-kp.ks, ccv, accel));
    // leapfrog in storage units (a constant offset drops out of 2c - c_prev and
    // of lap); 2 chi_st is exact, so the fmaf rounds 2 chi_st - chi_prev once
    float cnv = fmaf(kp.dt2, accel, fmaf(2.0f, chi_st, -pd.chip));
    const float floor_v = delta ? -2.0f * kp.chi0 : -kp.chi0;
    if (cnv < floor_v) cnv = floor_v;   // BH excision clamp chi >= -chi0 (NaN passes through)
    cn = delta ? mul_rn(a, cnv) : fmaf(m, kp.chi0, mul_rn(a, cnv));   // frozen boundary: chi -> chi0
}

// Plane source (LFM_F_SOURCE) at site (i, j, k): a continuous source's
// `psi[plane] += amount * envelope` after the step, folded in.  The amounts of
// step index *counter + step_off are read only on the source plane.
template <int L>
__device__ __forceinline__ void plane_source(const KParams& kp, const Ptrs& P, int i, int j,
                                             int k, float (&nw)[Lv<L>::NPF]) {
    const int ax = kp.src_axis;
    if ((ax == 0 ? i : ax == 1 ? j : k) != kp.src_plane) return;
    const long long s = *P.counter + kp.step_off;
    if (s < 0 || s >= kp.src_len) return;
    // transverse coordinates in ascending axis order
    const int t0 = ax == 0 ? j : i, t1 = ax == 2 ? j : k;
    const float e = P.src_env[t0 * kp.N + t1];
    // product and sum rounded separately, as the callback's two array operations
    const float ar = P.src_amp[2 * s];
    if (ar != 0.0f) nw[0] = add_rn(nw[0], mul_rn(ar, e));
    if constexpr (L == COMPLEX) {
        const float ai = P.src_amp[2 * s + 1];
        if (ai != 0.0f) nw[1] = add_rn(nw[1], mul_rn(ai, e));
    }
}

// Detector plane (LFM_F_DETECTOR) at site (i, j, k): a detector screen's
// record of the plane after the step, |Psi'|^2 (Pr'^2 + Pi'^2, each square
// and the sum rounded once) or Pr', into the frame of step index
// *counter + step_off when that step is recorded.
template <int L>
__device__ __forceinline__ void plane_detector(const KParams& kp, const Ptrs& P, int i, int j,
                                               int k, const float (&nw)[Lv<L>::NPF]) {
    const int ax = kp.det_axis;
    if ((ax == 0 ? i : ax == 1 ? j : k) != kp.det_plane) return;
    const long long r = *P.counter + kp.step_off - kp.det_first;
    if (r < 0 || r % kp.det_every != 0) return;
    const long long f = r / kp.det_every;
    if (f >= kp.det_len) return;
    const int t0 = ax == 0 ? j : i, t1 = ax == 2 ? j : k;
    float v = nw[0];
    if (kp.det_field == 0) {
        v = mul_rn(nw[0], nw[0]);
        if constexpr (L == COMPLEX) v = add_rn(v, mul_rn(nw[1], nw[1]));
    }
    P.det_frames[(size_t)f * kp.N * kp.N + size_t(t0) * kp.N + t1] = v;
}

// Per-site update with its stores (site_compute, then the outputs of P).
// EXTRA: the kernels for LFM_F_SOURCE, LFM_F_KEEP_PSI and LFM_F_DETECTOR (a
// separate instantiation, so the others carry none of their code); (i, j, k)
// is the site, used by the source and detector planes only.
template <int L, class O, bool EXTRA = false>
__device__ __forceinline__ void site_update(const KParams& kp, const Ptrs& P, O idx,
                                            const PointData<L>& pd,
                                            const float (&c)[Lv<L>::NPF],
                                            const float (&lap)[Lv<L>::NPF], float chi_st,
                                            float lap_chi, float jsum, float ccv, int si, int sj,
                                            int sk, float* ret3 = nullptr) {
    constexpr int NPF = Lv<L>::NPF;
    constexpr int C = Lv<L>::NC;
    const size_t n3 = P.n3;
    float nw[NPF], cn, sclip[3];
    site_compute<L>(kp, pd, P.sa != nullptr, c, lap, chi_st, lap_chi, jsum, ccv, nw, cn, sclip);
    if constexpr (EXTRA && L != COLOR) {
        if (P.src_env) plane_source<L>(kp, P, si, sj, sk, nw);
    }
    if constexpr (EXTRA) {
        if (P.keep) {
            // a driver's `psi *= factor` after the step, folded in
            const float k = idx.ldp(P.keep);
#pragma unroll
            for (int f = 0; f < NPF; ++f) nw[f] = mul_rn(nw[f], k);
        }
    }
    if constexpr (EXTRA && L != COLOR) {
        if (P.det_frames) plane_detector<L>(kp, P, si, sj, sk, nw);
    }
    const float a = 1.0f - ((kp.flags & LFM_F_MASK) ? pd.m : 0.0f);
#pragma unroll
    for (int f = 0; f < NPF; ++f) {
        idx.st(field_out<L>(P.pr_out, P.pi_out, f, n3), nw[f]);
        if (P.prp_out) idx.st(field_out<L>(P.prp_out, P.pip_out, f, n3), c[f] * a);
    }
    if constexpr (L == COLOR) {
        if (P.psisq) {
#pragma unroll
            for (int q = 0; q < C; ++q) idx.st(P.psisq + q * n3, psi_sq(nw[q], nw[q + C]));
        }
        if (ret3) {
#pragma unroll
            for (int q = 0; q < C; ++q) ret3[q] = psi_sq(nw[q], nw[q + C]);
        }
    }
    if (kp.flags & LFM_F_FREEZE_CHI) return;
    if constexpr (L == COLOR) {
        if (P.sa && P.sa_clip) {
#pragma unroll
            for (int q = 0; q < C; ++q) idx.st(P.sa_clip + q * n3, sclip[q]);
        }
    }
    idx.st(P.chi_out, cn);
    if (P.chip_out) idx.st(P.chip_out, chi_st);
}

// ---------------------------------------------------------------------------
// Direct kernel: one thread per site, on a (k, j, i) grid of 64 x 4
// workgroups, so the coordinates come from the launch.
// ---------------------------------------------------------------------------
template <int L, bool EXTRA = false>
__global__ __launch_bounds__(256) void step_direct(Ptrs P, KParams kp) {
    constexpr int NPF = Lv<L>::NPF;
    constexpr int C = Lv<L>::NC;
    const int N = kp.N;
    const size_t n3 = P.n3;
    const int k = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, i = blockIdx.z;
    if (k >= N || j >= N) return;
    const size_t idx = (size_t(i) * N + j) * N + k;
    const size_t NN = size_t(N) * N;
    const int ip = i + 1 == N ? 0 : i + 1, im = i == 0 ? N - 1 : i - 1;
    const int jp = j + 1 == N ? 0 : j + 1, jm = j == 0 ? N - 1 : j - 1;
    const size_t rc = size_t(i) * NN, rp = size_t(ip) * NN, rm = size_t(im) * NN;
    const size_t cc = size_t(j) * N, cp = size_t(jp) * N, cm = size_t(jm) * N;
    const size_t dc = k, dp = k + 1 == N ? 0 : k + 1, dm = k == 0 ? N - 1 : k - 1;
    // neighbour offsets: 6 faces then 12 edges
    const size_t nb[18] = {rp + cc + dc, rm + cc + dc, rc + cp + dc, rc + cm + dc, rc + cc + dp,
                           rc + cc + dm, rp + cp + dc, rp + cm + dc, rm + cp + dc, rm + cm + dc,
                           rp + cc + dp, rp + cc + dm, rm + cc + dp, rm + cc + dm, rc + cp + dp,
                           rc + cp + dm, rc + cm + dp, rc + cm + dm};
    auto lap19 = [&](const float* F, float ctr) {
        float fs = F[nb[0]] + F[nb[1]] + F[nb[2]] + F[nb[3]] + F[nb[4]] + F[nb[5]];
        float es = 0.0f;
#pragma unroll
        for (int t = 6; t < 18; ++t) es += F[nb[t]];
        return WF * fs + WE * es - 4.0f * ctr;
    };
    float c[NPF], lap[NPF];
#pragma unroll
    for (int f = 0; f < NPF; ++f) {
        const float* F = field<L>(P.pr, P.pi, f, n3);
        c[f] = F[idx];
        lap[f] = lap19(F, c[f]);
    }
    const float chi_c = P.chi[idx];
    const float lap_chi = lap19(P.chi, chi_c);
    float jsum = 0.0f, ccv = 0.0f;
    if constexpr (L != REAL) {
        float jd[3][C];
#pragma unroll
        for (int q = 0; q < C; ++q) {
            const float* R = field<L>(P.pr, P.pi, q, n3);
            const float* I = field<L>(P.pr, P.pi, q + C, n3);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const size_t up = nb[2 * d], dn = nb[2 * d + 1];
                jd[d][q] = c[q] * (I[up] - I[dn]) - c[q + C] * (R[up] - R[dn]);
                jsum += jd[d][q];
            }
        }
        if constexpr (L == COLOR) {
            if (kp.ks > 0.0f) {
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const float s = jd[d][0] + jd[d][1] + jd[d][2];
                    ccv += (jd[d][0] * jd[d][0] + jd[d][1] * jd[d][1] + jd[d][2] * jd[d][2]) -
                           (1.0f / 3.0f) * s * s;
                }
            }
        }
    }
    site_update<L, Off64, EXTRA>(kp, P, Off64{idx}, load_point<L>(kp, P, Off64{idx}), c, lap,
                                chi_c, lap_chi, jsum, ccv, i, j, k);
}

// ---------------------------------------------------------------------------
// Tiled 2.5-D kernel.
//   block = 64 x BJ threads (one wavefront per j-row group), each thread owns
//   RJ consecutive j rows at one k column: tile TK=64 (k) x TJ=BJ*RJ (j).
//   The grid is (k tiles) x (j tiles) x (i chunks of CH planes); the block id
//   is remapped so that consecutive tiles of one i chunk land on the same XCD
//   (blocks are dispatched round-robin over the 8 XCDs), which keeps the j/k
//   halo rows shared between neighbouring tiles inside one L2.
// ---------------------------------------------------------------------------
constexpr int TK = 64;

// One workgroup's work: the column tile at (k0, j0) over planes [i0, i1) of
// one lattice (P, kp).  Shared by the single-lattice and batched kernels.
template <int L, int BJ, int RJ, bool NT, bool EXTRA = false>
__device__ __forceinline__ void tile_body(const Ptrs& P, const KParams& kp, int k0, int j0,
                                          int i0, int i1) {
    constexpr int NPF = Lv<L>::NPF;
    constexpr int C = Lv<L>::NC;
    constexpr int NF = NPF + 1;              // + chi
    constexpr int TJ = BJ * RJ;
    constexpr int LW = TK + 2;
    constexpr int LH = TJ + 2;
    static_assert(BJ >= 4, "halo loading uses four wavefronts");
    // Field-fastest LDS layout: the NF values of one (row, column) position are
    // adjacent (padded to NFP), so all 9 stencil positions of all fields are
    // reached from one per-lane base address with immediate offsets, and the
    // 64 lanes of a wavefront touch consecutive positions (stride NFP floats).
    constexpr int NFP = NF == 7 ? 8 : NF == 3 ? 4 : NF;
    __shared__ float lds[LH][LW][NFP];

    const int N = kp.N;
    const size_t n3 = P.n3;
    const uint32_t rowB = uint32_t(N) * 4u;         // bytes per j row
    const uint32_t planeB = rowB * uint32_t(N);     // bytes per i plane
    const int nk = min(TK, N - k0), nj = min(TJ, N - j0);
    const int tx = threadIdx.x & 63;
    const int ty = threadIdx.x >> 6;

    const float* src[NF];
#pragma unroll
    for (int f = 0; f < NPF; ++f) src[f] = field<L>(P.pr, P.pi, f, n3);
    src[NPF] = P.chi;

    // In-plane byte offsets of this thread's loads.
    //   interior: rows ty*RJ + rr, column tx
    //   halo (wave 0: row -1, wave 1: row nj, wave 2: column -1, wave 3: column nk)
    const bool kin = tx < nk;
    bool hin = false;
    // lanes without a halo job load a valid in-plane address and discard it,
    // so every global load is unconditional (no per-load branches)
    uint32_t hoff = uint32_t(min(j0, N - 1)) * rowB + uint32_t(min(k0 + tx, N - 1)) * 4u;
    int hrow = 0, hcol = 0;  // LDS destination of the halo value
    if (ty == 0 && kin) {
        hin = true; hoff = ((j0 - 1 + N) % N) * rowB + (k0 + tx) * 4u; hrow = 0; hcol = tx + 1;
    } else if (ty == 1 && kin) {
        hin = true; hoff = ((j0 + nj) % N) * rowB + (k0 + tx) * 4u; hrow = nj + 1; hcol = tx + 1;
    } else if (ty == 2 && tx < nj + 2) {
        hin = true; hoff = ((j0 - 1 + tx + N) % N) * rowB + ((k0 - 1 + N) % N) * 4u;
        hrow = tx; hcol = 0;
    } else if (ty == 3 && tx < nj + 2) {
        hin = true; hoff = ((j0 - 1 + tx + N) % N) * rowB + ((k0 + nk) % N) * 4u;
        hrow = tx; hcol = nk + 1;
    }
    bool rin[RJ];
    uint32_t loff[RJ];   // load offsets, clamped into the lattice for idle lanes
    const uint32_t ioff0 = uint32_t(j0 + ty * RJ) * rowB + uint32_t(k0 + tx) * 4u;
#pragma unroll
    for (int rr = 0; rr < RJ; ++rr) {
        rin[rr] = kin && (ty * RJ + rr) < nj;
        loff[rr] = uint32_t(min(j0 + ty * RJ + rr, N - 1)) * rowB + uint32_t(min(k0 + tx, N - 1)) * 4u;
    }

    float pre[NF][RJ + 1];
    auto prefetch = [&](int q) {
        const uint32_t pb = uint32_t((q + N) % N) * planeB;
#pragma unroll
        for (int f = 0; f < NF; ++f) {
#pragma unroll
            for (int rr = 0; rr < RJ; ++rr)
                pre[f][rr] = Off32{pb + loff[rr]}.ld(src[f]);
            pre[f][RJ] = Off32{pb + hoff}.ld(src[f]);
        }
    };
    auto stage = [&]() {
#pragma unroll
        for (int rr = 0; rr < RJ; ++rr)
            if (rin[rr]) {
#pragma unroll
                for (int f = 0; f < NF; ++f) lds[ty * RJ + rr + 1][tx + 1][f] = pre[f][rr];
            }
        if (hin) {
#pragma unroll
            for (int f = 0; f < NF; ++f) lds[hrow][hcol][f] = pre[f][RJ];
        }
    };

    // Register queue (per owned row).
    float acc[RJ][NF];     // A(q-1) + B(q-2): Laplacian of the pending site minus B(q)
    float bprev[RJ][NF];   // B(q-1)
    float cm2[RJ][NPF];    // Psi at plane q-2 (for the i-direction current)
    float cm1[RJ][NPF];    // Psi at plane q-1 (the pending site)
    float chim1[RJ];       // chi at the pending site
    float jyz[RJ];         // sum over colours of j_y + j_z at the pending site
    float ccvyz[RJ];       // y,z part of CCV at the pending site
#pragma unroll
    for (int rr = 0; rr < RJ; ++rr) {
#pragma unroll
        for (int f = 0; f < NF; ++f) { acc[rr][f] = 0.0f; bprev[rr][f] = 0.0f; }
#pragma unroll
        for (int f = 0; f < NPF; ++f) { cm2[rr][f] = 0.0f; cm1[rr][f] = 0.0f; }
        chim1[rr] = 0.0f; jyz[rr] = 0.0f; ccvyz[rr] = 0.0f;
    }

    // Pointwise inputs of the pending sites, loaded one plane ahead.
    PointData<L> ptc[RJ], ptn[RJ];
    const int cnt = (i1 - i0) + 2;  // planes i0-1 .. i1
    prefetch(i0 - 1);
    for (int s = 0; s < cnt; ++s) {
        const int q = i0 - 1 + s;
        __syncthreads();
        stage();
        __syncthreads();
        if (s + 1 < cnt) prefetch(q + 1);
        if (s >= 1 && q < i1) {
            // site q is finalised in the next iteration: fetch its pointwise data now
            const uint32_t qb = uint32_t(q) * planeB;
#pragma unroll
            for (int rr = 0; rr < RJ; ++rr) ptn[rr] = load_point<L>(kp, P, Off32T<NT>{qb + loff[rr]});
        }
        const uint32_t pend_pb = uint32_t((q - 1 + N) % N) * planeB;
#pragma unroll
        for (int rr = 0; rr < RJ; ++rr) {
            const int r = ty * RJ + rr + 1;  // LDS row of this site
            const int x = tx + 1;
            float A[NF], B[NF], cq[NF], dy[NPF], dz[NPF];
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                const float ctr = lds[r][x][f];
                const float nn = lds[r - 1][x][f], ss = lds[r + 1][x][f];
                const float ww = lds[r][x - 1][f], ee = lds[r][x + 1][f];
                const float faces = (nn + ss) + (ww + ee);
                const float diag = (lds[r - 1][x - 1][f] + lds[r - 1][x + 1][f]) +
                                   (lds[r + 1][x - 1][f] + lds[r + 1][x + 1][f]);
                A[f] = lap_part_a(faces, diag, ctr);
                B[f] = lap_part_b(ctr, faces);
                cq[f] = ctr;
                if (f < NPF) { dy[f] = ss - nn; dz[f] = ee - ww; }
            }
            // In-plane currents (y, z) of this plane's sites.
            float jyz_new = 0.0f, ccv_new = 0.0f;
            if constexpr (L != REAL) {
                float sy = 0.0f, sz = 0.0f, qy = 0.0f, qz = 0.0f;
#pragma unroll
                for (int a2 = 0; a2 < C; ++a2) {
                    const float jy = cross2(cq[a2], dy[a2 + C], cq[a2 + C], dy[a2]);
                    const float jz = cross2(cq[a2], dz[a2 + C], cq[a2 + C], dz[a2]);
                    sy += jy; sz += jz; qy += jy * jy; qz += jz * jz;
                }
                jyz_new = sy + sz;
                if constexpr (L == COLOR)
                    ccv_new = (qy - (1.0f / 3.0f) * sy * sy) + (qz - (1.0f / 3.0f) * sz * sz);
            }
            if (s >= 2 && rin[rr]) {
                // Finalise the pending site at plane q-1.
                float lap[NPF], cpend[NPF];
#pragma unroll
                for (int f = 0; f < NPF; ++f) { lap[f] = acc[rr][f] + B[f]; cpend[f] = cm1[rr][f]; }
                const float lap_chi = acc[rr][NPF] + B[NPF];
                float jsum = 0.0f, ccv = 0.0f;
                if constexpr (L != REAL) {
                    float sx = 0.0f, qx = 0.0f;
#pragma unroll
                    for (int a2 = 0; a2 < C; ++a2) {
                        const float jx = cross2(cm1[rr][a2], cq[a2 + C] - cm2[rr][a2 + C],
                                                cm1[rr][a2 + C], cq[a2] - cm2[rr][a2]);
                        sx += jx; qx += jx * jx;
                    }
                    jsum = sx + jyz[rr];
                    if constexpr (L == COLOR) {
                        if (kp.ks > 0.0f) ccv = ccvyz[rr] + (qx - (1.0f / 3.0f) * sx * sx);
                    }
                }
                site_update<L, Off32T<NT>, EXTRA>(kp, P, Off32T<NT>{pend_pb + ioff0 + rr * rowB},
                                                 ptc[rr], cpend, lap, chim1[rr], lap_chi, jsum,
                                                 ccv, q - 1, j0 + ty * RJ + rr, k0 + tx);
            }
            // Advance the queue.
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                acc[rr][f] = A[f] + bprev[rr][f];
                bprev[rr][f] = B[f];
            }
#pragma unroll
            for (int f = 0; f < NPF; ++f) { cm2[rr][f] = cm1[rr][f]; cm1[rr][f] = cq[f]; }
            chim1[rr] = cq[NPF];
            jyz[rr] = jyz_new;
            ccvyz[rr] = ccv_new;
            ptc[rr] = ptn[rr];
        }
    }
}

template <int L, int BJ, int RJ, bool NT, bool EXTRA = false>
__global__ __launch_bounds__(64 * BJ) void step_tiled(Ptrs P, KParams kp, int nkt, int njt, int CH) {
    // XCD-aware block remap: blocks are dispatched round-robin over the 8
    // XCDs; this makes consecutive tiles of one i chunk share an XCD (and its
    // L2) so the halo rows neighbouring tiles both read hit in L2.
    const int nb = gridDim.x;
    const int b = blockIdx.x;
    const int lb = (nb & 7) == 0 ? (b & 7) * (nb >> 3) + (b >> 3) : b;
    const int kt = lb % nkt;
    const int rest = lb / nkt;
    const int jt = rest % njt;
    const int ch = rest / njt;
    const int i0 = kp.ib + ch * CH;
    tile_body<L, BJ, RJ, NT, EXTRA>(P, kp, kt * TK, jt * BJ * RJ, i0, min(i0 + CH, kp.ie));
}

// Batched sweep: one launch advances many independent lattices.  Each
// workgroup reads its tile descriptor {lattice, k0, j0, i0} and the
// lattice's record (arena offset, N, chunk, scalar parameters).
struct BatchLattice {
    long long off;   // element offset of the lattice in every arena
    int N, CH, ncomp, pad;
    KParams kp;
};

struct BatchArenas {
    const float* pr; const float* prp; const float* pi; const float* pip;
    const float* chi; const float* chip; const float* mask;
    float* pr_out; float* pi_out; float* chi_out;
    const float* sa;   // COLOR: raw S_a of every lattice (3 fields each), or NULL
    float* psisq;      // COLOR: |Psi'_a|^2 of the new fields, or NULL
};

template <int L, int BJ, int RJ, bool NT>
__global__ __launch_bounds__(64 * BJ) void step_tiled_batch(BatchArenas A,
                                                             const BatchLattice* __restrict__ lat,
                                                             const int4* __restrict__ desc,
                                                             int flags_or) {
    const int4 d = desc[blockIdx.x];
    if (d.x < 0) return;   // padding descriptor (whole workgroup exits together)
    const BatchLattice& Lb = lat[d.x];
    KParams kp = Lb.kp;
    kp.flags |= flags_or;
    const long long o = Lb.off;
    const size_t n3 = size_t(Lb.N) * Lb.N * Lb.N;
    // colour fields of one lattice are colour-major within its own span
    const long long oc = o * Lb.ncomp;
    Ptrs P{};
    P.pr = A.pr + oc; P.prp = A.prp + oc; P.chi = A.chi + o; P.chip = A.chip + o;
    if (A.pi) { P.pi = A.pi + oc; P.pip = A.pip + oc; P.pi_out = A.pi_out + oc; }
    P.mask = (kp.flags & LFM_F_MASK) ? A.mask + o : nullptr;
    P.pr_out = A.pr_out + oc; P.chi_out = A.chi_out + o;
    if (A.sa && kp.kt > 0.0f) {
        // the S_a term of this lattice (three S_a fields, or under
        // LFM_F_SA_SCV the SCV term, one field at the lattice's scalar
        // offset), and |Psi'|^2 for its next transform
        P.sa = A.sa + ((kp.flags & LFM_F_SA_SCV) ? o : oc);
        P.psisq = A.psisq ? A.psisq + oc : nullptr;
    }
    P.n3 = n3;
    tile_body<L, BJ, RJ, NT>(P, kp, d.y, d.z, d.w, min(d.w + Lb.CH, Lb.N));
}

int choose_chunk(int N, int tiles, int min_ch = 8, int target = 8192) {
    // Aim for ~target workgroups (8192: 32 per CU on 256 CUs) so every CU keeps
    // several tiles in flight, with at least min_ch planes per chunk: 8 when
    // the step streams from HBM (the two halo planes a chunk re-reads cost at
    // most a quarter of its traffic), fewer when the lattice's fields stay in
    // the last-level cache and a small lattice needs the extra workgroups
    // (the caller picks min_ch).
    const int nch = (target + tiles - 1) / tiles;
    int ch = (N + nch - 1) / nch;
    if (ch < min_ch) ch = min_ch;
    return ch < N ? ch : N;
}

template <int L, int BJ, int RJ, bool NT, bool EXTRA = false>
int launch_tiled(const Ptrs& P, const KParams& kp, hipStream_t st) {
    constexpr int TJ = BJ * RJ;
    const int N = kp.N;
    const int nkt = (N + TK - 1) / TK;
    const int njt = (N + TJ - 1) / TJ;
    const int span = kp.ie - kp.ib;       // planes computed
    // cache-resident lattices: 2-plane chunks up to 64^3 (the few tiles need
    // the workgroups), 4 from 96^3 up (COLOR 128^3: 44.9 against 49.9 us,
    // COMPLEX 160^3: 30.2 against 32.1 us at tile 1)
    const int min_ch = NT ? 8 : (N >= 96 ? 4 : 2);
    const int CH = kp.chunk > 0 ? min(kp.chunk, span) : choose_chunk(span, nkt * njt, min_ch);
    const int nch = (span + CH - 1) / CH;
    const long long nblocks = (long long)nkt * njt * nch;
    hipLaunchKernelGGL((step_tiled<L, BJ, RJ, NT, EXTRA>), dim3((unsigned)nblocks), dim3(64 * BJ), 0,
                       st, P, kp, nkt, njt, CH);
    LFM_LAUNCH_CHECK("step_tiled");
    return LFM_OK;
}

// Streaming accesses pay off once one step's traffic (algorithmic bytes per
// site x sites) is well beyond the 256 MB last-level cache.
template <int L>
bool streaming(size_t sites) {
    constexpr size_t bytes = L == REAL ? 20 : L == COMPLEX ? 40 : 112;
    return sites * bytes > (size_t(512) << 20);
}

template <int L>
int launch(const Ptrs& P, const KParams& kp, hipStream_t st) {
    const bool extra = (kp.flags & (LFM_F_KEEP_PSI | LFM_F_SOURCE | LFM_F_DETECTOR)) != 0;
    if ((kp.flags & LFM_F_DIRECT) || kp.N > 1024) {
        const dim3 grid(unsigned((kp.N + 63) / 64), unsigned((kp.N + 3) / 4), unsigned(kp.N));
        if (extra) hipLaunchKernelGGL((step_direct<L, true>), grid, dim3(64, 4), 0, st, P, kp);
        else hipLaunchKernelGGL((step_direct<L>), grid, dim3(64, 4), 0, st, P, kp);
        LFM_LAUNCH_CHECK("step_direct");
        return LFM_OK;
    }
    const bool nt = streaming<L>(P.n3);
    if (extra) {
        // the LFM_F_KEEP_PSI / _SOURCE / _DETECTOR kernels come in one tile shape, {4 x 2}
        return nt ? launch_tiled<L, 4, 2, true, true>(P, kp, st)
                  : launch_tiled<L, 4, 2, false, true>(P, kp, st);
    }
    // Tile shapes (wavefronts BJ x rows per thread RJ); 0 picks the level's default.
    const int tile = kp.tile > 0 ? kp.tile : 1;
#define LFM_TILES(NT)                                                  \
    switch (tile) {                                                    \
        case 1: return launch_tiled<L, 4, 1, NT>(P, kp, st);           \
        case 2: return launch_tiled<L, 4, 2, NT>(P, kp, st);           \
        case 3: return launch_tiled<L, 4, 4, NT>(P, kp, st);           \
        case 4: return launch_tiled<L, 8, 1, NT>(P, kp, st);           \
        case 5: return launch_tiled<L, 8, 2, NT>(P, kp, st);           \
        case 6: return launch_tiled<L, 16, 1, NT>(P, kp, st);          \
        default:                                                       \
            set_error("tile variant %d unknown (1..%d)", tile, LFM_TILE_VARIANTS); \
            return LFM_E_ARG;                                          \
    }
    if (nt) { LFM_TILES(true) } else { LFM_TILES(false) }
#undef LFM_TILES
}

// ---------------------------------------------------------------------------
// COLOR step with the k-axis passes of the S_a transform folded in.
//
// The direct S_a transform is five passes: real k-lines to half spectra,
// lines along j, the fused i pass (forward, filter, inverse), lines along j,
// half spectra back to real k-lines.  Here a workgroup owns one full k-row j
// of every plane of a chunk, so the first and last of those passes run inside
// the step: the workgroup reads the k-axis spectra of S_a for its row (what
// the three column passes left in the plan's spectrum), turns them into S_a
// with two inverse line FFTs (colours 0 and 1 packed as z = x_0 + i x_1,
// colour 2 alone), updates the sites, and turns |Psi'_a|^2 of the new row
// into its k-axis half spectra with two forward line FFTs, written in place
// over the rows it has read.  Each spectrum row belongs to one workgroup and
// is read two planes before it is rewritten.  Per site this replaces the S_a
// read and the |Psi'|^2 write of the plain step (12 B each) by spectrum reads
// and writes of the same size and removes two whole transform passes.
//
// The stencil part is the plain kernel's plane march with a full-row tile:
// thread k owns column k; the staged plane is the rows j-1, j, j+1
// (periodic); the k halo is the row's own wrap-around.  Four of the
// N/8-thread line groups run the two inverse and the two forward FFTs of a
// plane side by side; every thread takes part in the barriers.
// ---------------------------------------------------------------------------
template <int N>
struct KfftLds {
    static constexpr int LW = N + 2;
    static constexpr int PS = N + N / 8;      // line FFT buffer (slot p + p/8)
    static constexpr int H = N / 2 + 1;
    float st[3][LW][8];                       // stencil fields of rows j-1, j, j+1
    float fz[3][N];                           // |Psi'_a|^2 of the finished row
    float2 fx[4][PS];                         // line FFT exchange buffers
    float sab[2][3][N];                       // S_a of the pending plane (parity of q)
    float2 specin[3][H];                      // k-axis spectra of S_a, next plane
    float2 twl[N];
};

template <int N, bool NT>
__global__ __launch_bounds__(N) void step_color_kfft(Ptrs P, KParams kp, float2* __restrict__ spec,
                                                     const float2* __restrict__ tw, int Hp,
                                                     int CH) {
    constexpr int L = COLOR;
    constexpr int NPF = Lv<L>::NPF;            // 6
    constexpr int C = Lv<L>::NC;               // 3
    constexpr int NF = NPF + 1;                // + chi
    constexpr int TC = N / 8;
    constexpr int H = N / 2 + 1;
    static_assert(4 * TC <= N, "four line groups per workgroup");
    __shared__ KfftLds<N> S;

    const uint32_t rowB = uint32_t(N) * 4u;
    const uint32_t planeB = rowB * uint32_t(N);
    const size_t n3 = P.n3;
    // XCD grouping as in step_tiled: consecutive rows share an L2
    const int nb = gridDim.x, b = blockIdx.x;
    const int lb = (nb & 7) == 0 ? (b & 7) * (nb >> 3) + (b >> 3) : b;
    const int j = lb % N;
    const int i0 = (lb / N) * CH;
    const int i1 = min(i0 + CH, N);
    const int k = threadIdx.x;

    S.twl[k] = tw[k];

    const float* src[NF];
#pragma unroll
    for (int f = 0; f < NPF; ++f) src[f] = field<L>(P.pr, P.pi, f, n3);
    src[NPF] = P.chi;
    uint32_t roff[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) roff[r] = uint32_t((j - 1 + r + N) % N) * rowB + uint32_t(k) * 4u;

    float pre[NF][3];
    auto prefetch = [&](int q) {
        const uint32_t pb = uint32_t((q + N) % N) * planeB;
#pragma unroll
        for (int f = 0; f < NF; ++f)
#pragma unroll
            for (int r = 0; r < 3; ++r) pre[f][r] = Off32{pb + roff[r]}.ld(src[f]);
    };
    auto stage = [&]() {
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int f = 0; f < NF; ++f) S.st[r][k + 1][f] = pre[f][r];
        if (k == N - 1) {   // column -1 is column N-1 of the same row
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int f = 0; f < NF; ++f) S.st[r][0][f] = pre[f][r];
        }
        if (k == 0) {       // column N is column 0
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int f = 0; f < NF; ++f) S.st[r][N + 1][f] = pre[f][r];
        }
    };
    // k-axis spectra of S_a of plane q, all colours
    constexpr int NSPEC = 3 * H;
    constexpr int NSU = (NSPEC + N - 1) / N;
    float2 spre[NSU];
    auto spec_row = [&](int c, int q) {
        return spec + ((size_t(c) * N + q) * N + j) * size_t(Hp);
    };
    auto spec_prefetch = [&](int q) {
#pragma unroll
        for (int u = 0; u < NSU; ++u) {
            const int id = k + u * N;
            if (id < NSPEC) spre[u] = ld2nt<N>(spec_row(id / H, q) + id % H);
        }
    };
    auto spec_stage = [&]() {
#pragma unroll
        for (int u = 0; u < NSU; ++u) {
            const int id = k + u * N;
            if (id < NSPEC) S.specin[id / H][id % H] = spre[u];
        }
    };

    // line groups: 0 inverse of colours 0+1, 1 inverse of colour 2,
    // 2 forward of colours 0+1, 3 forward of colour 2
    const int grp = k / TC, tq = k % TC;
    const bool fjob = grp < 4;
    const bool inv = grp < 2;
    const bool pair = (grp & 1) == 0;
    float2* xb = S.fx[fjob ? grp : 0];
    auto slot = [](int pos) { return pos + pos / 8; };
    // One FFT phase: S_a of plane qi (when inv_ok) into sab; the k-axis half
    // spectra of the |Psi'|^2 row of plane qf (when fwd_ok) to the spectrum.
    auto fft_phase = [&](int qi, bool inv_ok, int qf, bool fwd_ok) {
        float2 v[8];
        if (inv) {
            const int c0 = pair ? 0 : 2;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int kk = tq + e * TC;
                const bool lo = kk <= N / 2;
                const int m = lo ? kk : N - kk;
                float2 a = S.specin[c0][m];
                float2 bb = pair ? S.specin[1][m] : make_float2(0.0f, 0.0f);
                if (m == 0 || m == N / 2) { a.y = 0.0f; bb.y = 0.0f; }
                if (!lo) { a.y = -a.y; bb.y = -bb.y; }
                v[e] = make_float2(a.x - bb.y, a.y + bb.x);
            }
        } else if (fjob) {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int kk = tq + e * TC;
                v[e] = pair ? make_float2(S.fz[0][kk], S.fz[1][kk]) : make_float2(S.fz[2][kk], 0.0f);
            }
        } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = make_float2(0.0f, 0.0f);
        }
        fft_line_m<N>(v, xb, tq, inv ? 1 : -1, S.twl, slot, fjob);
        if (inv && inv_ok) {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int kk = tq + e * TC;
                if (pair) {
                    S.sab[qi & 1][0][kk] = v[e].x;
                    S.sab[qi & 1][1][kk] = v[e].y;
                } else {
                    S.sab[qi & 1][2][kk] = v[e].x;
                }
            }
        }
        // forward: split the packed pair with Hermitian symmetry
        __syncthreads();
        if (fjob && !inv) {
#pragma unroll
            for (int e = 0; e < 8; ++e) xb[slot(tq + e * TC)] = v[e];
        }
        __syncthreads();
        if (fjob && !inv && fwd_ok) {
            float2* o0 = spec_row(pair ? 0 : 2, qf);
            float2* o1 = spec_row(1, qf);
#pragma unroll
            for (int e = 0; e < 5; ++e) {
                const int kk = tq + e * TC;
                if (e == 4 && tq != 0) break;   // kk = N/2 once
                const float2 zk = xb[slot(kk)];
                const float2 zm = xb[slot((N - kk) & (N - 1))];
                st2nt<N>(o0 + kk, make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y)));
                if (pair) st2nt<N>(o1 + kk, make_float2(0.5f * (zk.y + zm.y), 0.5f * (zm.x - zk.x)));
            }
        }
    };

    // register queue (as in tile_body, one row)
    float acc[NF], bprev[NF], cm2[NPF], cm1[NPF], chim1 = 0.0f, jyz = 0.0f, ccvyz = 0.0f;
#pragma unroll
    for (int f = 0; f < NF; ++f) { acc[f] = 0.0f; bprev[f] = 0.0f; }
#pragma unroll
    for (int f = 0; f < NPF; ++f) { cm2[f] = 0.0f; cm1[f] = 0.0f; }
    const uint32_t ioff = uint32_t(j) * rowB + uint32_t(k) * 4u;
    const int cnt = (i1 - i0) + 2;   // planes i0-1 .. i1
    prefetch(i0 - 1);
    spec_prefetch(i0);
    for (int s = 0; s < cnt; ++s) {
        const int q = i0 - 1 + s;
        const bool site_q = q >= i0 && q < i1;
        PointData<L> pd;
        if (s >= 2) {
            // pointwise inputs of the site finished this iteration (plane q-1);
            // the staging and the FFT phase below cover their latency
            pd = load_point<L, Off32T<NT>, false>(kp, P, Off32T<NT>{uint32_t(q - 1) * planeB + ioff});
        }
        __syncthreads();
        stage();
        if (site_q) spec_stage();
        __syncthreads();
        if (s + 1 < cnt) prefetch(q + 1);
        if (q + 1 >= i0 && q + 1 < i1) spec_prefetch(q + 1);
        // S_a of plane q; k-axis spectra of |Psi'|^2 of plane q-2 (finished last iteration)
        fft_phase(q, site_q, q - 2, s >= 3);
        const int x = k + 1;
        float A[NF], B[NF], cq[NF], dy[NPF], dz[NPF];
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            const float ctr = S.st[1][x][f];
            const float nn = S.st[0][x][f], ss = S.st[2][x][f];
            const float ww = S.st[1][x - 1][f], ee = S.st[1][x + 1][f];
            const float faces = (nn + ss) + (ww + ee);
            const float diag = (S.st[0][x - 1][f] + S.st[0][x + 1][f]) +
                               (S.st[2][x - 1][f] + S.st[2][x + 1][f]);
            A[f] = lap_part_a(faces, diag, ctr);
            B[f] = lap_part_b(ctr, faces);
            cq[f] = ctr;
            if (f < NPF) { dy[f] = ss - nn; dz[f] = ee - ww; }
        }
        float sy = 0.0f, sz = 0.0f, qy = 0.0f, qz = 0.0f;
#pragma unroll
        for (int a2 = 0; a2 < C; ++a2) {
            const float jy = cross2(cq[a2], dy[a2 + C], cq[a2 + C], dy[a2]);
            const float jz = cross2(cq[a2], dz[a2 + C], cq[a2 + C], dz[a2]);
            sy += jy; sz += jz; qy += jy * jy; qz += jz * jz;
        }
        const float jyz_new = sy + sz;
        const float ccv_new = (qy - (1.0f / 3.0f) * sy * sy) + (qz - (1.0f / 3.0f) * sz * sz);
        if (s >= 2) {
            // finish the site at plane q-1 (always a site plane of this chunk)
            float lap[NPF], cpend[NPF];
#pragma unroll
            for (int f = 0; f < NPF; ++f) { lap[f] = acc[f] + B[f]; cpend[f] = cm1[f]; }
            const float lap_chi = acc[NPF] + B[NPF];
            float sx = 0.0f, qx = 0.0f;
#pragma unroll
            for (int a2 = 0; a2 < C; ++a2) {
                const float jx = cross2(cm1[a2], cq[a2 + C] - cm2[a2 + C], cm1[a2 + C],
                                        cq[a2] - cm2[a2]);
                sx += jx; qx += jx * jx;
            }
            const float jsum = sx + jyz;
            const float ccv = kp.ks > 0.0f ? ccvyz + (qx - (1.0f / 3.0f) * sx * sx) : 0.0f;
#pragma unroll
            for (int c = 0; c < C; ++c) pd.sa[c] = S.sab[(q - 1) & 1][c][k];
            float e2[C];
            site_update<L>(kp, P, Off32T<NT>{uint32_t(q - 1) * planeB + ioff}, pd, cpend, lap,
                           chim1, lap_chi, jsum, ccv, 0, 0, 0, e2);
#pragma unroll
            for (int c = 0; c < C; ++c) S.fz[c][k] = e2[c];
        }
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            acc[f] = A[f] + bprev[f];
            bprev[f] = B[f];
        }
#pragma unroll
        for (int f = 0; f < NPF; ++f) { cm2[f] = cm1[f]; cm1[f] = cq[f]; }
        chim1 = cq[NPF];
        jyz = jyz_new;
        ccvyz = ccv_new;
    }
    // the last finished plane's |Psi'|^2 row
    __syncthreads();
    fft_phase(0, false, i1 - 1, true);
}

// Workgroups of step_color_kfft<N> one device holds at once; 0 if unknown.
template <int N, bool NT>
int kfft_capacity() {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, step_color_kfft<N, NT>, N, 0) !=
            hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return cus * per;
}

template <int N, bool NT>
int launch_kfft(const Ptrs& P, const KParams& kp, const ScvView& v, hipStream_t st) {
    // one round of resident workgroups: as many i chunks as fit beside the
    // N rows (at least one)
    static const int cap = kfft_capacity<N, NT>();
    int CH = kp.chunk;
    if (CH <= 0) {
        int nch = cap > 0 ? cap / N : 1;
        if (nch < 1) nch = 1;
        CH = (N + nch - 1) / nch;
    }
    CH = min(CH, N);
    const int nch = (N + CH - 1) / CH;
    const unsigned nblocks = unsigned(N * nch);
    hipLaunchKernelGGL((step_color_kfft<N, NT>), dim3(nblocks), dim3(N), 0, st, P, kp, v.spec, v.tw,
                       v.Hp, CH);
    LFM_LAUNCH_CHECK("step_color_kfft");
    return LFM_OK;
}

bool kfft_supported(int N) { return N == 64 || N == 128 || N == 256 || N == 512; }

int launch_kfft_any(const Ptrs& P, const KParams& kp, const ScvView& v, hipStream_t st) {
    const bool nt = streaming<COLOR>(P.n3);
    switch (kp.N) {
        case 64: return nt ? launch_kfft<64, true>(P, kp, v, st) : launch_kfft<64, false>(P, kp, v, st);
        case 128: return nt ? launch_kfft<128, true>(P, kp, v, st) : launch_kfft<128, false>(P, kp, v, st);
        case 256: return nt ? launch_kfft<256, true>(P, kp, v, st) : launch_kfft<256, false>(P, kp, v, st);
        case 512: return nt ? launch_kfft<512, true>(P, kp, v, st) : launch_kfft<512, false>(P, kp, v, st);
        default: set_error("fused COLOR step: N=%d unsupported", kp.N); return LFM_E_ARG;
    }
}

// ---------------------------------------------------------------------------
// Two leapfrog steps per launch (REAL and COMPLEX).
//
// A workgroup owns a 64 (k) x 8 (j) tile and marches along i.  Its
// wavefronts take one of two roles, which run side by side between one
// barrier per plane:
//  * 11 first-step wavefronts compute step n -> n+1 on the tile plus a
//    one-cell ring (660 positions, one per lane, each with its own
//    plane-march queue).  Fed plane q of Psi_n / chi_n from LDS window 0,
//    they finish plane q-1: the new Psi and chi go to LDS window 1 and, on
//    the tile, to global memory; Psi_n, chi_n and the mask of the tile's
//    sites go to a three-plane buffer as the second step's pointwise inputs.
//  * 4 second-step wavefronts compute step n+1 -> n+2 on the tile, two rows
//    per lane.  Fed plane q-2 of window 1 (written one plane earlier), they
//    finish plane q-3.
// Every lane also stages plane q+1 of Psi_n and chi_n on the tile plus a
// two-cell ring (816 positions) into the other half of window 0.  Both
// windows are double-buffered, so the one barrier per plane orders every
// write before its reads.  Each role is its own loop and keeps only its own
// queues in registers.
//
// Per site the two steps read Psi_n, Psi_{n-1}, chi_n, chi_{n-1} and the
// mask once and write Psi and chi of both steps (the rotation keeps step
// n+1 as the next prev): 52 B for two COMPLEX steps with mask instead of
// 2 x 40 B.  Both steps use site_compute, and the Laplacian and current are
// summed in the same order as in tile_body, so the results equal two single
// steps bit for bit.
// ---------------------------------------------------------------------------
constexpr int T2K = 64, T2J = 8;                 // tile
constexpr int T2W1 = 11, T2W2 = 4;               // wavefronts of each role
constexpr int T2T = 64 * (T2W1 + T2W2);          // 960 lanes per workgroup

struct Ptrs2 {
    Ptrs P;            // inputs, and step n+1 outputs in pr_out / pi_out / chi_out
    float* pr_out2;    // step n+2
    float* pi_out2;
    float* chi_out2;
};

template <int L>
struct Win2 {
    static constexpr int NPF = Lv<L>::NPF;
    static constexpr int NF = NPF + 1;                 // + chi
    static constexpr int R0 = T2J + 4, C0 = T2K + 4;  // two-cell ring
    static constexpr int R1 = T2J + 2, C1 = T2K + 2;  // one-cell ring
    static_assert(R0 * C0 <= T2T, "one staging position per lane");
    static_assert(R1 * C1 <= 64 * T2W1, "one first-step position per lane");
    static_assert(T2J == 2 * T2W2, "two second-step rows per lane");
    // Psi (float2 for COMPLEX, float for REAL) and chi in separate arrays:
    // lanes read consecutive 8- or 4-byte slots, free of bank conflicts.
    using VP = std::conditional_t<NPF == 2, float2, float>;
    VP p0[2][R0][C0];         // window 0: Psi_n
    float c0[2][R0][C0];      //           chi_n
    VP p1[2][R1][C1];         // window 1: Psi_{n+1}
    float c1[2][R1][C1];      //           chi_{n+1}
    VP pp[3][T2J][T2K];       // second step's pointwise inputs on the tile:
    float cp[3][T2J][T2K];    //   Psi_n, chi_n and the mask
    float mp[3][T2J][T2K];
};

__device__ __forceinline__ float fld(const float2& v, int f) { return f == 0 ? v.x : v.y; }
__device__ __forceinline__ float fld(float v, int) { return v; }
__device__ __forceinline__ void put(float2& d, const float (&v)[2]) { d = make_float2(v[0], v[1]); }
__device__ __forceinline__ void put(float& d, const float (&v)[1]) { d = v[0]; }

// In-plane parts of one plane at (r, c) of a window slot (Psi array P, chi
// array X): the partials A and B (lap_part_a / lap_part_b), the centre and
// the j / k differences of every field (Psi components, then chi), summed
// exactly as tile_body sums them.
template <int NF, class VP, int RR, int CC>
__device__ __forceinline__ void plane_parts(const VP (&P)[RR][CC], const float (&X)[RR][CC], int r,
                                            int c, float (&A)[NF], float (&B)[NF], float (&ctr)[NF],
                                            float (&dj)[NF], float (&dk)[NF]) {
    const VP pnw = P[r - 1][c - 1], pnc = P[r - 1][c], pne = P[r - 1][c + 1];
    const VP pcw = P[r][c - 1], pcc = P[r][c], pce = P[r][c + 1];
    const VP psw = P[r + 1][c - 1], psc = P[r + 1][c], pse = P[r + 1][c + 1];
    float v[NF][9];   // nw nc ne cw cc ce sw sc se
#pragma unroll
    for (int f = 0; f < NF - 1; ++f) {
        v[f][0] = fld(pnw, f); v[f][1] = fld(pnc, f); v[f][2] = fld(pne, f);
        v[f][3] = fld(pcw, f); v[f][4] = fld(pcc, f); v[f][5] = fld(pce, f);
        v[f][6] = fld(psw, f); v[f][7] = fld(psc, f); v[f][8] = fld(pse, f);
    }
    v[NF - 1][0] = X[r - 1][c - 1]; v[NF - 1][1] = X[r - 1][c]; v[NF - 1][2] = X[r - 1][c + 1];
    v[NF - 1][3] = X[r][c - 1];     v[NF - 1][4] = X[r][c];     v[NF - 1][5] = X[r][c + 1];
    v[NF - 1][6] = X[r + 1][c - 1]; v[NF - 1][7] = X[r + 1][c]; v[NF - 1][8] = X[r + 1][c + 1];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        const float x = v[f][4];
        const float nn = v[f][1], ss = v[f][7], ww = v[f][3], ee = v[f][5];
        const float faces = (nn + ss) + (ww + ee);
        const float diag = (v[f][0] + v[f][2]) + (v[f][6] + v[f][8]);
        A[f] = lap_part_a(faces, diag, x);
        B[f] = lap_part_b(x, faces);
        ctr[f] = x;
        dj[f] = ss - nn;
        dk[f] = ee - ww;
    }
}

// The plane-march queue of one site column (as in tile_body): fed one plane
// at a time, it finishes the site one plane behind.
template <int L>
struct Queue2 {
    static constexpr int NPF = Lv<L>::NPF, NF = NPF + 1;
    float acc[NF], bprev[NF], cm1[NPF], cm2[NPF], chim1, jyz;
    __device__ __forceinline__ void clear() {
#pragma unroll
        for (int f = 0; f < NF; ++f) { acc[f] = 0.0f; bprev[f] = 0.0f; }
#pragma unroll
        for (int f = 0; f < NPF; ++f) { cm1[f] = 0.0f; cm2[f] = 0.0f; }
        chim1 = 0.0f;
        jyz = 0.0f;
    }
    // Feed plane q's parts; when `finish`, complete the site at plane q-1
    // with the pointwise inputs pd into nw, cn.
    __device__ __forceinline__ void feed(const KParams& kp, const float (&A)[NF], const float (&B)[NF],
                                         const float (&ctr)[NF], const float (&dj)[NF],
                                         const float (&dk)[NF], bool finish, const PointData<L>& pd,
                                         float (&nw)[NPF], float& cn) {
        if (finish) {
            float lap[NPF];
#pragma unroll
            for (int f = 0; f < NPF; ++f) lap[f] = acc[f] + B[f];
            const float lap_chi = acc[NPF] + B[NPF];
            float jsum = 0.0f;
            if constexpr (L == COMPLEX) {
                const float sx = 0.0f + cross2(cm1[0], ctr[1] - cm2[1], cm1[1], ctr[0] - cm2[0]);
                jsum = sx + jyz;
            }
            float sclip[3];
            site_compute<L>(kp, pd, false, cm1, lap, chim1, lap_chi, jsum, 0.0f, nw, cn, sclip);
        }
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            acc[f] = A[f] + bprev[f];
            bprev[f] = B[f];
        }
#pragma unroll
        for (int f = 0; f < NPF; ++f) { cm2[f] = cm1[f]; cm1[f] = ctr[f]; }
        chim1 = ctr[NPF];
        if constexpr (L == COMPLEX) {
            const float sy = 0.0f + cross2(ctr[0], dj[1], ctr[1], dj[0]);
            const float sz = 0.0f + cross2(ctr[0], dk[1], ctr[1], dk[0]);
            jyz = sy + sz;
        }
    }
};

template <int L, bool NT>
__device__ __forceinline__ void tile2_body(const Ptrs2& P2, const KParams& kp, int k0, int j0,
                                           int i0, int i1) {
    using Wn = Win2<L>;
    using VP = typename Wn::VP;
    constexpr int NPF = Lv<L>::NPF;
    constexpr int NF = Wn::NF;
    constexpr int C0 = Wn::C0, C1 = Wn::C1;
    constexpr int P0 = Wn::R0 * C0;           // staged positions of window 0
    constexpr int P1 = Wn::R1 * C1;           // first-step positions
    __shared__ Wn S;
    const Ptrs& P = P2.P;
    const int N = kp.N;
    const size_t n3 = P.n3;
    const uint32_t rowB = uint32_t(N) * 4u;
    const uint32_t planeB = rowB * uint32_t(N);
    const int t = threadIdx.x;
    const int nk = min(T2K, N - k0), nj = min(T2J, N - j0);
    auto wrap = [N](int x) { x %= N; return x < 0 ? x + N : x; };

    const float* src[NF];
#pragma unroll
    for (int f = 0; f < NPF; ++f) src[f] = field<L>(P.pr, P.pi, f, n3);
    src[NPF] = P.chi;
    // this lane's staging position (two-cell ring around the tile); lanes
    // past the last position load a valid address and discard it
    const bool stg = t < P0;
    const int ps = min(t, P0 - 1);
    const int sr = ps / C0, sc = ps % C0;
    const uint32_t o0 = uint32_t(wrap(j0 - 2 + sr)) * rowB + uint32_t(wrap(k0 - 2 + sc)) * 4u;
    float pre[NF];
    auto prefetch = [&](int q) {
        const uint32_t pb = uint32_t(wrap(q)) * planeB;
#pragma unroll
        for (int f = 0; f < NF; ++f) pre[f] = Off32{pb + o0}.ld(src[f]);
    };
    auto stage = [&](int slot) {
        if (stg) {
            float pv[NPF];
#pragma unroll
            for (int f = 0; f < NPF; ++f) pv[f] = pre[f];
            put(S.p0[slot][sr][sc], pv);
            S.c0[slot][sr][sc] = pre[NPF];
        }
    };
    // iterations: the first step is fed input planes i0-2 .. i1+1 (iterations
    // 0 .. cnt-2), the second step window-1 planes i0-1 .. i1 (3 .. cnt-1)
    const int cnt = (i1 - i0) + 5;
    // at iteration s the lanes stage plane q+1 and load plane q+2 (q = i0-2+s)
    auto stage_next = [&](int s) {
        if (s <= cnt - 3) {
            stage((s + 1) & 1);
            if (s <= cnt - 4) prefetch(i0 + s);
        }
    };
    prefetch(i0 - 2);
    stage(0);
    prefetch(i0 - 1);
    __syncthreads();

    if ((t >> 6) < T2W1) {
        // ---- first step: n -> n+1 on the tile plus a one-cell ring ----
        const bool act = t < P1;
        const int p1 = min(t, P1 - 1);
        const int r1 = p1 / C1, c1 = p1 % C1;
        const uint32_t o1 = uint32_t(wrap(j0 - 1 + r1)) * rowB + uint32_t(wrap(k0 - 1 + c1)) * 4u;
        const int rt = r1 - 1, ct = c1 - 1;   // tile coordinates
        const bool on_tile = act && rt >= 0 && rt < T2J && ct >= 0 && ct < T2K;
        const bool in_lat = on_tile && rt < nj && ct < nk;
        const bool has_mask = (kp.flags & LFM_F_MASK) != 0;
        PointData<L> pa{}, pn{};
        Queue2<L> qa;
        qa.clear();
        // one plane; cur = s & 1, passed as a constant by the loop below
        auto iter = [&](const int s, const int cur) {
            const int q = i0 - 2 + s;
            stage_next(s);
            if (s <= cnt - 2) {
                if (s >= 2) pa = pn;
                // site q is finished next iteration: its pointwise inputs now
                if (s >= 1 && q <= i1) pn = load_point<L>(kp, P, Off32T<NT>{uint32_t(wrap(q)) * planeB + o1});
                if (act) {
                    float A[NF], B[NF], ctr[NF], dj[NF], dk[NF], nw[NPF], cn = 0.0f;
                    plane_parts<NF>(S.p0[cur], S.c0[cur], r1 + 1, c1 + 1, A, B, ctr, dj, dk);
                    // Psi_n and chi_n of the pending site (plane q-1)
                    float cpend[NPF];
#pragma unroll
                    for (int f = 0; f < NPF; ++f) cpend[f] = qa.cm1[f];
                    const float chpend = qa.chim1;
                    qa.feed(kp, A, B, ctr, dj, dk, s >= 2, pa, nw, cn);
                    if (s >= 2) {
                        put(S.p1[cur][r1][c1], nw);
                        S.c1[cur][r1][c1] = cn;
                        if (on_tile) {
                            const int sl = s % 3;
                            put(S.pp[sl][rt][ct], cpend);
                            S.cp[sl][rt][ct] = chpend;
                            S.mp[sl][rt][ct] = has_mask ? pa.m : 0.0f;
                            if (in_lat && q - 1 >= i0 && q - 1 < i1) {
                                const Off32T<NT> o{uint32_t(q - 1) * planeB + o1};
#pragma unroll
                                for (int f = 0; f < NPF; ++f)
                                    o.st(field_out<L>(P.pr_out, P.pi_out, f, n3), nw[f]);
                                o.st(P.chi_out, cn);
                            }
                        }
                    }
                }
            }
            __syncthreads();
        };
        // two planes per trip, so the window halves are fixed in each copy
        for (int s = 0; s < cnt; s += 2) {
            iter(s, 0);
            if (s + 1 < cnt) iter(s + 1, 1);
        }
    } else {
        // ---- second step: n+1 -> n+2 on the tile, rows r0, r0+1 ----
        const int tx = t & 63, r0 = 2 * ((t >> 6) - T2W1);
        KParams kpB = kp;                      // its prev is Psi_n, read damped
        if (kp.flags & LFM_F_MASK) kpB.flags |= LFM_F_DAMP_PREV;
        Queue2<L> qb[2];
        qb[0].clear();
        qb[1].clear();
        auto iter = [&](const int s, const int cur) {
            const int q = i0 - 2 + s;
            stage_next(s);
            if (s >= 3) {
                // window 1 slot written last iteration holds plane q-2; the
                // pointwise inputs of plane q-3 were written two iterations ago
                const int w1 = cur ^ 1, sl = (s + 1) % 3;
                const bool fin = s >= 5;
#pragma unroll
                for (int rr = 0; rr < 2; ++rr) {
                    const int r = r0 + rr;
                    float A[NF], B[NF], ctr[NF], dj[NF], dk[NF], nw[NPF], cn = 0.0f;
                    plane_parts<NF>(S.p1[w1], S.c1[w1], r + 1, tx + 1, A, B, ctr, dj, dk);
                    PointData<L> pd{};
                    if (fin) {
                        const VP z = S.pp[sl][r][tx];
#pragma unroll
                        for (int f = 0; f < NPF; ++f) pd.prev[f] = fld(z, f);
                        pd.chip = S.cp[sl][r][tx];
                        pd.m = S.mp[sl][r][tx];
                    }
                    qb[rr].feed(kpB, A, B, ctr, dj, dk, fin, pd, nw, cn);
                    if (fin && r < nj && tx < nk) {
                        const Off32T<NT> o{uint32_t(q - 3) * planeB + uint32_t(j0 + r) * rowB +
                                           uint32_t(k0 + tx) * 4u};
#pragma unroll
                        for (int f = 0; f < NPF; ++f)
                            o.st(field_out<L>(P2.pr_out2, P2.pi_out2, f, n3), nw[f]);
                        o.st(P2.chi_out2, cn);
                    }
                }
            }
            __syncthreads();
        };
        for (int s = 0; s < cnt; s += 2) {
            iter(s, 0);
            if (s + 1 < cnt) iter(s + 1, 1);
        }
    }
}

template <int L, bool NT>
__global__ __launch_bounds__(T2T) void step2_tiled(Ptrs2 P2, KParams kp, int nkt, int njt, int CH) {
    const int nb = gridDim.x, b = blockIdx.x;
    const int lb = (nb & 7) == 0 ? (b & 7) * (nb >> 3) + (b >> 3) : b;
    const int kt = lb % nkt;
    const int rest = lb / nkt;
    const int jt = rest % njt;
    const int i0 = (rest / njt) * CH;
    tile2_body<L, NT>(P2, kp, kt * T2K, jt * T2J, i0, min(i0 + CH, kp.N));
}

template <int L, bool NT>
int launch_step2_t(const Ptrs2& P2, const KParams& kp, hipStream_t st) {
    const int N = kp.N;
    const int nkt = (N + T2K - 1) / T2K, njt = (N + T2J - 1) / T2J;
    // one 960-lane workgroup per CU: ~2048 workgroups, and chunks of at least
    // 8 planes (each chunk also feeds 4 planes around it)
    const int CH = kp.chunk > 0 ? min(kp.chunk, N) : choose_chunk(N, nkt * njt, 8, 2048);
    const int nch = (N + CH - 1) / CH;
    const long long nblocks = (long long)nkt * njt * nch;
    hipLaunchKernelGGL((step2_tiled<L, NT>), dim3((unsigned)nblocks), dim3(T2T), 0, st, P2, kp, nkt,
                       njt, CH);
    LFM_LAUNCH_CHECK("step2_tiled");
    return LFM_OK;
}

template <int L>
int launch_step2(const Ptrs2& P2, const KParams& kp, hipStream_t st) {
    return streaming<L>(P2.P.n3) ? launch_step2_t<L, true>(P2, kp, st)
                                 : launch_step2_t<L, false>(P2, kp, st);
}

// Batched two steps per launch: the descriptors of a batch plan whose
// column tiles are 64 x 8 (tile variants 2 and 4), one 960-lane workgroup per
// (lattice, tile, plane chunk), tile2_body on the lattice's arena spans.  The
// arenas are the single step's (outputs of step n+1 in pr_out / pi_out /
// chi_out) plus the outputs of step n+2.
struct BatchArenas2 {
    BatchArenas a;
    float* pr_out2; float* pi_out2; float* chi_out2;
};

template <int L, bool NT>
__global__ __launch_bounds__(T2T) void step2_tiled_batch(BatchArenas2 A2,
                                                         const BatchLattice* __restrict__ lat,
                                                         const int4* __restrict__ desc,
                                                         int flags_or) {
    const int4 d = desc[blockIdx.x];
    if (d.x < 0) return;   // padding descriptor (whole workgroup exits together)
    const BatchLattice& Lb = lat[d.x];
    KParams kp = Lb.kp;
    kp.flags |= flags_or;
    const BatchArenas& A = A2.a;
    const long long o = Lb.off;
    Ptrs2 P2{};
    Ptrs& P = P2.P;
    P.pr = A.pr + o; P.prp = A.prp + o; P.chi = A.chi + o; P.chip = A.chip + o;
    if (A.pi) { P.pi = A.pi + o; P.pip = A.pip + o; P.pi_out = A.pi_out + o; P2.pi_out2 = A2.pi_out2 + o; }
    P.mask = (kp.flags & LFM_F_MASK) ? A.mask + o : nullptr;
    P.pr_out = A.pr_out + o; P.chi_out = A.chi_out + o;
    P2.pr_out2 = A2.pr_out2 + o; P2.chi_out2 = A2.chi_out2 + o;
    P.n3 = size_t(Lb.N) * Lb.N * Lb.N;
    tile2_body<L, NT>(P2, kp, d.y, d.z, d.w, min(d.w + Lb.CH, Lb.N));
}

constexpr int tile_tj(int tile) {
    return tile == 1 ? 4 : tile == 2 ? 8 : tile == 3 ? 16 : tile == 4 ? 8 : tile == 5 ? 16 : 16;
}

template <int L>
int launch_batch(const BatchArenas& A, const BatchLattice* lat, const int4* desc, int ndesc,
                 int tile, int flags_or, bool nt, hipStream_t st) {
    const dim3 g(ndesc);
#define LFM_BTILES(NT)                                                                                  \
    switch (tile) {                                                                                     \
        case 1: hipLaunchKernelGGL((step_tiled_batch<L, 4, 1, NT>), g, dim3(256), 0, st, A, lat, desc, flags_or); break; \
        case 2: hipLaunchKernelGGL((step_tiled_batch<L, 4, 2, NT>), g, dim3(256), 0, st, A, lat, desc, flags_or); break; \
        case 3: hipLaunchKernelGGL((step_tiled_batch<L, 4, 4, NT>), g, dim3(256), 0, st, A, lat, desc, flags_or); break; \
        case 4: hipLaunchKernelGGL((step_tiled_batch<L, 8, 1, NT>), g, dim3(512), 0, st, A, lat, desc, flags_or); break; \
        case 5: hipLaunchKernelGGL((step_tiled_batch<L, 8, 2, NT>), g, dim3(512), 0, st, A, lat, desc, flags_or); break; \
        case 6: hipLaunchKernelGGL((step_tiled_batch<L, 16, 1, NT>), g, dim3(1024), 0, st, A, lat, desc, flags_or); break; \
        default: set_error("tile variant %d unknown", tile); return LFM_E_ARG;                         \
    }
    if (nt) { LFM_BTILES(true) } else { LFM_BTILES(false) }
#undef LFM_BTILES
    LFM_LAUNCH_CHECK("step_tiled_batch");
    return LFM_OK;
}

struct BatchPlan {
    int level = 0, nlat = 0, tile = 0, ndesc = 0;
    bool scv = false;            // some lattice has kappa_tube > 0
    long long total = 0;
    BatchLattice* d_lat = nullptr;
    int4* d_desc = nullptr;
};

KParams to_k(const lfm_params_t* p) {
    KParams k;
    k.N = p->N; k.flags = p->flags; k.dt2 = p->dt2; k.kappa = p->kappa; k.lam = p->lambda_self;
    k.chi0 = p->chi0; k.e0sq = p->e0_sq; k.epsw = p->epsilon_w; k.kc = p->kappa_c;
    k.ecc = p->epsilon_cc; k.ks = p->kappa_string; k.kt = p->kappa_tube;
    k.tile = p->tile; k.chunk = p->chunk;
    // IEEE single division, round to nearest: bit-equal to a device division
    k.kc0 = p->kappa / p->chi0;
    k.kcc0 = p->kappa_c / p->chi0;
    k.ib = 0;
    k.ie = p->N;
    const bool src = (p->flags & LFM_F_SOURCE) != 0;
    const bool det = (p->flags & LFM_F_DETECTOR) != 0;
    k.step_off = (src || det) ? p->step_off : 0;
    k.src_axis = src ? p->src_axis : 0;
    k.src_plane = src ? p->src_plane : -1;
    k.src_len = src ? p->src_len : 0;
    k.det_axis = det ? p->det_axis : 0;
    k.det_plane = det ? p->det_plane : -1;
    k.det_every = det ? p->det_every : 1;
    k.det_field = det ? p->det_field : 0;
    k.det_first = det ? p->det_first : 0;
    k.det_len = det ? p->det_len : 0;
    return k;
}

// LFM_F_SOURCE / LFM_F_DETECTOR arguments: pointers given, axes, planes,
// lengths and interval in range.  Fills the source inputs (envelope, amounts,
// counter: NULL without the flags) and the detector frames, with their float
// counts for the overlap checks.
struct Extras {
    const float* in[3] = {nullptr, nullptr, nullptr};
    size_t nin[3] = {0, 0, 0};
    float* frames = nullptr;
    size_t nframes = 0;
};
int extras_check(const lfm_params_t* p, Extras& x, const char* who) {
    const bool src = (p->flags & LFM_F_SOURCE) != 0;
    const bool det = (p->flags & LFM_F_DETECTOR) != 0;
    if (!src && !det) return LFM_OK;
    if (!p->step_counter) {
        set_error("%s: LFM_F_SOURCE / LFM_F_DETECTOR set but step_counter is NULL", who);
        return LFM_E_ARG;
    }
    x.in[2] = reinterpret_cast<const float*>(p->step_counter);
    x.nin[2] = 2;
    if (src) {
        if (!p->src_env || !p->src_amp) {
            set_error("%s: LFM_F_SOURCE set but src_env or src_amp is NULL", who);
            return LFM_E_ARG;
        }
        if (p->src_axis < 0 || p->src_axis > 2 || p->src_plane < 0 || p->src_plane >= p->N) {
            set_error("%s: source axis %d / plane %d out of range (axis 0..2, plane 0..N-1)",
                      who, p->src_axis, p->src_plane);
            return LFM_E_ARG;
        }
        if (p->src_len < 0 || p->src_len > (int64_t(1) << 40)) {
            set_error("%s: src_len %lld out of range", who, (long long)p->src_len);
            return LFM_E_ARG;
        }
        x.in[0] = p->src_env; x.nin[0] = size_t(p->N) * p->N;
        x.in[1] = p->src_amp; x.nin[1] = size_t(2 * p->src_len);
    }
    if (det) {
        if (!p->det_frames) {
            set_error("%s: LFM_F_DETECTOR set but det_frames is NULL", who);
            return LFM_E_ARG;
        }
        if (p->det_axis < 0 || p->det_axis > 2 || p->det_plane < 0 || p->det_plane >= p->N) {
            set_error("%s: detector axis %d / plane %d out of range (axis 0..2, plane 0..N-1)",
                      who, p->det_axis, p->det_plane);
            return LFM_E_ARG;
        }
        if (p->det_every < 1 || p->det_field < 0 || p->det_field > 1 || p->det_len < 0 ||
            p->det_len > (int64_t(1) << 40) / (int64_t(p->N) * p->N)) {
            set_error("%s: detector interval %d / field %d / frames %lld out of range", who,
                      p->det_every, p->det_field, (long long)p->det_len);
            return LFM_E_ARG;
        }
        x.frames = p->det_frames;
        x.nframes = size_t(p->det_len) * p->N * p->N;
    }
    return LFM_OK;
}

// Argument validation common to all levels.  `ins`/`outs` list (pointer,
// float count) pairs; NULL outputs are skipped.
int validate(const lfm_params_t* p, const float* const* ins, const size_t* nin, int ni,
             float* const* outs, const size_t* nout, int no, const char* who) {
    if (!p) { set_error("%s: params is NULL", who); return LFM_E_ARG; }
    if (p->N < 8 || p->N > 1290) {
        set_error("%s: N=%d out of range [8, 1290]", who, p->N);
        return LFM_E_ARG;
    }
    for (int a = 0; a < no; ++a) {
        if (!outs[a]) continue;
        for (int b = 0; b < ni; ++b)
            if (overlaps(outs[a], nout[a] * 4, ins[b], nin[b] * 4)) {
                set_error("%s: output %d overlaps input %d (in-place update is not allowed)",
                          who, a, b);
                return LFM_E_ARG;
            }
        for (int b = a + 1; b < no; ++b)
            if (overlaps(outs[a], nout[a] * 4, outs[b], nout[b] * 4)) {
                set_error("%s: outputs %d and %d overlap", who, a, b);
                return LFM_E_ARG;
            }
    }
    return LFM_OK;
}

}  // namespace
}  // namespace lfm

using namespace lfm;

extern "C" int lfm_step_real_f32(const float* E, const float* E_prev, const float* chi,
                                 const float* chi_prev, const float* mask, float* E_out,
                                 float* E_prev_out, float* chi_out, float* chi_prev_out,
                                 const lfm_params_t* p, void* stream) {
    if (!E || !E_prev || !chi || !chi_prev || !E_out || !chi_out || !p) {
        set_error("lfm_step_real_f32: required pointer is NULL");
        return LFM_E_ARG;
    }
    if ((p->flags & LFM_F_MASK) && !mask) {
        set_error("lfm_step_real_f32: LFM_F_MASK set but mask is NULL");
        return LFM_E_ARG;
    }
    if ((p->flags & LFM_F_KEEP_PSI) && !p->keep) {
        set_error("lfm_step_real_f32: LFM_F_KEEP_PSI set but params.keep is NULL");
        return LFM_E_ARG;
    }
    const size_t n3 = size_t(p->N) * p->N * p->N;
    const float* keep = (p->flags & LFM_F_KEEP_PSI) ? p->keep : nullptr;
    Extras x;
    if (int rc = extras_check(p, x, "lfm_step_real_f32")) return rc;
    const float* ins[] = {E, E_prev, chi, chi_prev, mask, keep, x.in[0], x.in[1], x.in[2]};
    const size_t nin[] = {n3, n3, n3, n3, n3, n3, x.nin[0], x.nin[1], x.nin[2]};
    float* outs[] = {E_out, E_prev_out, chi_out, chi_prev_out, x.frames};
    const size_t nout[] = {n3, n3, n3, n3, x.nframes};
    int rc = validate(p, ins, nin, 9, outs, nout, 5, "lfm_step_real_f32");
    if (rc) return rc;
    Ptrs P{};
    P.pr = E; P.prp = E_prev; P.chi = chi; P.chip = chi_prev;
    P.mask = (p->flags & LFM_F_MASK) ? mask : nullptr;
    P.keep = keep;
    P.src_env = x.in[0]; P.src_amp = x.in[1]; P.det_frames = x.frames;
    P.counter = reinterpret_cast<const long long*>(x.in[2]);
    P.pr_out = E_out; P.prp_out = E_prev_out; P.chi_out = chi_out; P.chip_out = chi_prev_out;
    P.n3 = n3;
    return launch<REAL>(P, to_k(p), (hipStream_t)stream);
}

extern "C" int lfm_step_complex_f32(const float* pr, const float* pr_prev, const float* pi,
                                    const float* pi_prev, const float* chi,
                                    const float* chi_prev, const float* mask, float* pr_out,
                                    float* pr_prev_out, float* pi_out, float* pi_prev_out,
                                    float* chi_out, float* chi_prev_out, const lfm_params_t* p,
                                    void* stream) {
    if (!pr || !pr_prev || !pi || !pi_prev || !chi || !chi_prev || !pr_out || !pi_out ||
        !chi_out || !p) {
        set_error("lfm_step_complex_f32: required pointer is NULL");
        return LFM_E_ARG;
    }
    if ((p->flags & LFM_F_MASK) && !mask) {
        set_error("lfm_step_complex_f32: LFM_F_MASK set but mask is NULL");
        return LFM_E_ARG;
    }
    if ((pr_prev_out == nullptr) != (pi_prev_out == nullptr)) {
        set_error("lfm_step_complex_f32: pr_prev_out and pi_prev_out must both be given or both NULL");
        return LFM_E_ARG;
    }
    if ((p->flags & LFM_F_KEEP_PSI) && !p->keep) {
        set_error("lfm_step_complex_f32: LFM_F_KEEP_PSI set but params.keep is NULL");
        return LFM_E_ARG;
    }
    const size_t n3 = size_t(p->N) * p->N * p->N;
    const float* keep = (p->flags & LFM_F_KEEP_PSI) ? p->keep : nullptr;
    Extras x;
    if (int rc = extras_check(p, x, "lfm_step_complex_f32")) return rc;
    const float* ins[] = {pr, pr_prev, pi, pi_prev, chi, chi_prev, mask, keep,
                          x.in[0], x.in[1], x.in[2]};
    const size_t nin[] = {n3, n3, n3, n3, n3, n3, n3, n3, x.nin[0], x.nin[1], x.nin[2]};
    float* outs[] = {pr_out, pr_prev_out, pi_out, pi_prev_out, chi_out, chi_prev_out, x.frames};
    const size_t nout[] = {n3, n3, n3, n3, n3, n3, x.nframes};
    int rc = validate(p, ins, nin, 11, outs, nout, 7, "lfm_step_complex_f32");
    if (rc) return rc;
    Ptrs P{};
    P.pr = pr; P.pi = pi; P.prp = pr_prev; P.pip = pi_prev; P.chi = chi; P.chip = chi_prev;
    P.mask = (p->flags & LFM_F_MASK) ? mask : nullptr;
    P.keep = keep;
    P.src_env = x.in[0]; P.src_amp = x.in[1]; P.det_frames = x.frames;
    P.counter = reinterpret_cast<const long long*>(x.in[2]);
    P.pr_out = pr_out; P.pi_out = pi_out; P.prp_out = pr_prev_out; P.pip_out = pi_prev_out;
    P.chi_out = chi_out; P.chip_out = chi_prev_out;
    P.n3 = n3;
    return launch<COMPLEX>(P, to_k(p), (hipStream_t)stream);
}

static int step2_check(const lfm_params_t* p, const char* fn) {
    if (p->flags &
        (LFM_F_DIRECT | LFM_F_FREEZE_CHI | LFM_F_KEEP_PSI | LFM_F_SOURCE | LFM_F_DETECTOR)) {
        set_error("%s: LFM_F_DIRECT / LFM_F_FREEZE_CHI / LFM_F_KEEP_PSI / LFM_F_SOURCE / "
                  "LFM_F_DETECTOR not supported", fn);
        return LFM_E_ARG;
    }
    if (p->N > 1024) {
        set_error("%s: N=%d above 1024", fn, p->N);
        return LFM_E_ARG;
    }
    return LFM_OK;
}

extern "C" int lfm_step2_real_f32(const float* E, const float* E_prev, const float* chi,
                                  const float* chi_prev, const float* mask, float* E_out1,
                                  float* chi_out1, float* E_out2, float* chi_out2,
                                  const lfm_params_t* p, void* stream) {
    if (!E || !E_prev || !chi || !chi_prev || !E_out1 || !chi_out1 || !E_out2 || !chi_out2 || !p) {
        set_error("lfm_step2_real_f32: required pointer is NULL");
        return LFM_E_ARG;
    }
    if ((p->flags & LFM_F_MASK) && !mask) {
        set_error("lfm_step2_real_f32: LFM_F_MASK set but mask is NULL");
        return LFM_E_ARG;
    }
    if (int rc = step2_check(p, "lfm_step2_real_f32")) return rc;
    const size_t n3 = size_t(p->N) * p->N * p->N;
    const float* ins[] = {E, E_prev, chi, chi_prev, mask};
    const size_t nin[] = {n3, n3, n3, n3, n3};
    float* outs[] = {E_out1, chi_out1, E_out2, chi_out2};
    const size_t nout[] = {n3, n3, n3, n3};
    int rc = validate(p, ins, nin, 5, outs, nout, 4, "lfm_step2_real_f32");
    if (rc) return rc;
    Ptrs2 P2{};
    Ptrs& P = P2.P;
    P.pr = E; P.prp = E_prev; P.chi = chi; P.chip = chi_prev;
    P.mask = (p->flags & LFM_F_MASK) ? mask : nullptr;
    P.pr_out = E_out1; P.chi_out = chi_out1;
    P.n3 = n3;
    P2.pr_out2 = E_out2; P2.chi_out2 = chi_out2;
    return launch_step2<REAL>(P2, to_k(p), (hipStream_t)stream);
}

extern "C" int lfm_step2_complex_f32(const float* pr, const float* pr_prev, const float* pi,
                                     const float* pi_prev, const float* chi,
                                     const float* chi_prev, const float* mask, float* pr_out1,
                                     float* pi_out1, float* chi_out1, float* pr_out2,
                                     float* pi_out2, float* chi_out2, const lfm_params_t* p,
                                     void* stream) {
    if (!pr || !pr_prev || !pi || !pi_prev || !chi || !chi_prev || !pr_out1 || !pi_out1 ||
        !chi_out1 || !pr_out2 || !pi_out2 || !chi_out2 || !p) {
        set_error("lfm_step2_complex_f32: required pointer is NULL");
        return LFM_E_ARG;
    }
    if ((p->flags & LFM_F_MASK) && !mask) {
        set_error("lfm_step2_complex_f32: LFM_F_MASK set but mask is NULL");
        return LFM_E_ARG;
    }
    if (int rc = step2_check(p, "lfm_step2_complex_f32")) return rc;
    const size_t n3 = size_t(p->N) * p->N * p->N;
    const float* ins[] = {pr, pr_prev, pi, pi_prev, chi, chi_prev, mask};
    const size_t nin[] = {n3, n3, n3, n3, n3, n3, n3};
    float* outs[] = {pr_out1, pi_out1, chi_out1, pr_out2, pi_out2, chi_out2};
    const size_t nout[] = {n3, n3, n3, n3, n3, n3};
    int rc = validate(p, ins, nin, 7, outs, nout, 6, "lfm_step2_complex_f32");
    if (rc) return rc;
    Ptrs2 P2{};
    Ptrs& P = P2.P;
    P.pr = pr; P.pi = pi; P.prp = pr_prev; P.pip = pi_prev; P.chi = chi; P.chip = chi_prev;
    P.mask = (p->flags & LFM_F_MASK) ? mask : nullptr;
    P.pr_out = pr_out1; P.pi_out = pi_out1; P.chi_out = chi_out1;
    P.n3 = n3;
    P2.pr_out2 = pr_out2; P2.pi_out2 = pi_out2; P2.chi_out2 = chi_out2;
    return launch_step2<COMPLEX>(P2, to_k(p), (hipStream_t)stream);
}

extern "C" int lfm_step_color_f32(const float* pr, const float* pr_prev, const float* pi,
                                  const float* pi_prev, const float* chi, const float* chi_prev,
                                  const float* mask, float* pr_out, float* pr_prev_out,
                                  float* pi_out, float* pi_prev_out, float* chi_out,
                                  float* chi_prev_out, const float* sa, float* sa_clip_out,
                                  float* psisq_out, const lfm_params_t* p, void* stream) {
    if (!pr || !pr_prev || !pi || !pi_prev || !chi || !chi_prev || !pr_out || !pi_out ||
        !chi_out || !p) {
        set_error("lfm_step_color_f32: required pointer is NULL");
        return LFM_E_ARG;
    }
    if ((p->flags & LFM_F_MASK) && !mask) {
        set_error("lfm_step_color_f32: LFM_F_MASK set but mask is NULL");
        return LFM_E_ARG;
    }
    if ((pr_prev_out == nullptr) != (pi_prev_out == nullptr)) {
        set_error("lfm_step_color_f32: pr_prev_out and pi_prev_out must both be given or both NULL");
        return LFM_E_ARG;
    }
    if (p->kappa_tube > 0.0f && !sa) {
        set_error("lfm_step_color_f32: kappa_tube > 0 needs the S_a fields (sa is NULL)");
        return LFM_E_ARG;
    }
    if (sa_clip_out && !sa) {
        set_error("lfm_step_color_f32: sa_clip_out given without sa");
        return LFM_E_ARG;
    }
    const bool sa_scv = (p->flags & LFM_F_SA_SCV) != 0;
    if (sa_scv && (sa_clip_out || !(p->kappa_tube > 0.0f))) {
        set_error("lfm_step_color_f32: LFM_F_SA_SCV needs kappa_tube > 0 and no sa_clip_out "
                  "(the S_a fields themselves are not passed)");
        return LFM_E_ARG;
    }
    if (p->flags & (LFM_F_SOURCE | LFM_F_DETECTOR)) {
        set_error("lfm_step_color_f32: LFM_F_SOURCE / LFM_F_DETECTOR are for the REAL and "
                  "COMPLEX levels");
        return LFM_E_ARG;
    }
    if ((p->flags & LFM_F_KEEP_PSI) && !p->keep) {
        set_error("lfm_step_color_f32: LFM_F_KEEP_PSI set but params.keep is NULL");
        return LFM_E_ARG;
    }
    const size_t n3 = size_t(p->N) * p->N * p->N;
    const float* keep = (p->flags & LFM_F_KEEP_PSI) ? p->keep : nullptr;
    const float* ins[] = {pr, pr_prev, pi, pi_prev, chi, chi_prev, mask, sa, keep};
    const size_t nin[] = {3 * n3, 3 * n3, 3 * n3, 3 * n3, n3, n3, n3, sa_scv ? n3 : 3 * n3, n3};
    float* outs[] = {pr_out, pr_prev_out, pi_out, pi_prev_out, chi_out, chi_prev_out,
                     sa_clip_out, psisq_out};
    const size_t nout[] = {3 * n3, 3 * n3, 3 * n3, 3 * n3, n3, n3, 3 * n3, 3 * n3};
    int rc = validate(p, ins, nin, 9, outs, nout, 8, "lfm_step_color_f32");
    if (rc) return rc;
    Ptrs P{};
    P.pr = pr; P.pi = pi; P.prp = pr_prev; P.pip = pi_prev; P.chi = chi; P.chip = chi_prev;
    P.mask = (p->flags & LFM_F_MASK) ? mask : nullptr;
    P.keep = keep;
    P.sa = p->kappa_tube > 0.0f ? sa : nullptr;
    P.pr_out = pr_out; P.pi_out = pi_out; P.prp_out = pr_prev_out; P.pip_out = pi_prev_out;
    P.chi_out = chi_out; P.chip_out = chi_prev_out;
    P.sa_clip = P.sa ? sa_clip_out : nullptr;
    P.psisq = psisq_out;
    P.n3 = n3;
    return launch<COLOR>(P, to_k(p), (hipStream_t)stream);
}

extern "C" int lfm_step_slab_f32(int32_t level, int32_t planes, const float* pr,
                                 const float* pr_prev, const float* pi, const float* pi_prev,
                                 const float* chi, const float* chi_prev, const float* mask,
                                 const float* sa, float* pr_out, float* pi_out, float* chi_out,
                                 float* psisq_out, const lfm_params_t* p, void* stream) {
    if (!p || !pr || !pr_prev || !chi || !chi_prev || !pr_out || !chi_out) {
        set_error("lfm_step_slab_f32: required pointer is NULL");
        return LFM_E_ARG;
    }
    if (level < 0 || level > 2) {
        set_error("lfm_step_slab_f32: level %d is not 0 (REAL), 1 (COMPLEX) or 2 (COLOR)", level);
        return LFM_E_ARG;
    }
    if (p->N < 8 || p->N > 1024) {
        set_error("lfm_step_slab_f32: N=%d out of range [8, 1024]", p->N);
        return LFM_E_ARG;
    }
    if (planes < 3 || planes > p->N) {
        set_error("lfm_step_slab_f32: planes=%d out of range [3, N=%d] (a slab stores its "
                  "interior planes and one halo plane on each side)", planes, p->N);
        return LFM_E_ARG;
    }
    if (level > 0 && (!pi || !pi_prev || !pi_out)) {
        set_error("lfm_step_slab_f32: complex levels need pi, pi_prev and pi_out");
        return LFM_E_ARG;
    }
    if ((p->flags & LFM_F_MASK) && !mask) {
        set_error("lfm_step_slab_f32: LFM_F_MASK set but mask is NULL");
        return LFM_E_ARG;
    }
    if (p->flags & (LFM_F_DIRECT | LFM_F_KEEP_PSI | LFM_F_SOURCE | LFM_F_DETECTOR)) {
        set_error("lfm_step_slab_f32: LFM_F_DIRECT / LFM_F_KEEP_PSI / LFM_F_SOURCE / "
                  "LFM_F_DETECTOR not supported");
        return LFM_E_ARG;
    }
    const bool scv = level == 2 && p->kappa_tube > 0.0f;
    if (scv && !sa) {
        set_error("lfm_step_slab_f32: kappa_tube > 0 needs the S_a input (sa is NULL)");
        return LFM_E_ARG;
    }
    if ((p->flags & LFM_F_SA_SCV) && !scv) {
        set_error("lfm_step_slab_f32: LFM_F_SA_SCV is for COLOR with kappa_tube > 0");
        return LFM_E_ARG;
    }
    const size_t ns = size_t(planes) * p->N * p->N;     // one field of the slab
    const size_t np = level == 2 ? 3 * ns : ns;
    const size_t nsa = (p->flags & LFM_F_SA_SCV) ? ns : np;
    const float* ins[] = {pr, pr_prev, level ? pi : nullptr, level ? pi_prev : nullptr, chi,
                          chi_prev, mask, scv ? sa : nullptr};
    const size_t nin[] = {np, np, np, np, ns, ns, ns, nsa};
    float* outs[] = {pr_out, level ? pi_out : nullptr, chi_out, scv ? psisq_out : nullptr};
    const size_t nout[] = {np, np, ns, np};
    for (int a = 0; a < 4; ++a) {
        if (!outs[a]) continue;
        for (int b = 0; b < 8; ++b)
            if (ins[b] && overlaps(outs[a], nout[a] * 4, ins[b], nin[b] * 4)) {
                set_error("lfm_step_slab_f32: output %d overlaps input %d", a, b);
                return LFM_E_ARG;
            }
        for (int b = a + 1; b < 4; ++b)
            if (outs[b] && overlaps(outs[a], nout[a] * 4, outs[b], nout[b] * 4)) {
                set_error("lfm_step_slab_f32: outputs %d and %d overlap", a, b);
                return LFM_E_ARG;
            }
    }
    Ptrs P{};
    P.pr = pr; P.prp = pr_prev; P.chi = chi; P.chip = chi_prev;
    if (level) { P.pi = pi; P.pip = pi_prev; P.pi_out = pi_out; }
    P.mask = (p->flags & LFM_F_MASK) ? mask : nullptr;
    P.sa = scv ? sa : nullptr;
    P.pr_out = pr_out; P.chi_out = chi_out;
    P.psisq = scv ? psisq_out : nullptr;
    P.n3 = ns;
    KParams k = to_k(p);
    k.ib = 1;
    k.ie = planes - 1;
    const hipStream_t st = (hipStream_t)stream;
    const int tile = k.tile > 0 ? k.tile : 1;
    if (tile > LFM_TILE_VARIANTS) {
        set_error("tile variant %d unknown (1..%d)", tile, LFM_TILE_VARIANTS);
        return LFM_E_ARG;
    }
    switch (level) {
        case 0: return launch<REAL>(P, k, st);
        case 1: return launch<COMPLEX>(P, k, st);
        default: return launch<COLOR>(P, k, st);
    }
}

extern "C" int lfm_step_color_fused_f32(const float* pr, const float* pr_prev, const float* pi,
                                        const float* pi_prev, const float* chi,
                                        const float* chi_prev, const float* mask, float* pr_out,
                                        float* pi_out, float* chi_out, void* scv_plan,
                                        const lfm_params_t* p, void* stream) {
    if (!pr || !pr_prev || !pi || !pi_prev || !chi || !chi_prev || !pr_out || !pi_out ||
        !chi_out || !p || !scv_plan) {
        set_error("lfm_step_color_fused_f32: required pointer is NULL");
        return LFM_E_ARG;
    }
    if ((p->flags & LFM_F_MASK) && !mask) {
        set_error("lfm_step_color_fused_f32: LFM_F_MASK set but mask is NULL");
        return LFM_E_ARG;
    }
    if (!(p->kappa_tube > 0.0f)) {
        set_error("lfm_step_color_fused_f32: needs kappa_tube > 0 (use lfm_step_color_f32)");
        return LFM_E_ARG;
    }
    if (p->flags & (LFM_F_DIRECT | LFM_F_FREEZE_CHI | LFM_F_SA_SCV | LFM_F_KEEP_PSI |
                    LFM_F_SOURCE | LFM_F_DETECTOR)) {
        set_error("lfm_step_color_fused_f32: LFM_F_DIRECT / LFM_F_FREEZE_CHI / LFM_F_SA_SCV / "
                  "LFM_F_KEEP_PSI / LFM_F_SOURCE / LFM_F_DETECTOR not supported");
        return LFM_E_ARG;
    }
    ScvView v{};
    scv_view(scv_plan, &v);
    if (v.method != 2 || v.nc != 3 || v.N != p->N || !kfft_supported(p->N)) {
        set_error("lfm_step_color_fused_f32: needs a direct-method S_a plan with 3 colours of the "
                  "same N, N in {64, 128, 256, 512} (plan N=%d colours=%d method=%d, step N=%d)",
                  v.N, v.nc, v.method, p->N);
        return LFM_E_ARG;
    }
    const size_t n3 = size_t(p->N) * p->N * p->N;
    const float* ins[] = {pr, pr_prev, pi, pi_prev, chi, chi_prev, mask};
    const size_t nin[] = {3 * n3, 3 * n3, 3 * n3, 3 * n3, n3, n3, n3};
    float* outs[] = {pr_out, pi_out, chi_out};
    const size_t nout[] = {3 * n3, 3 * n3, n3};
    int rc = validate(p, ins, nin, 7, outs, nout, 3, "lfm_step_color_fused_f32");
    if (rc) return rc;
    Ptrs P{};
    P.pr = pr; P.pi = pi; P.prp = pr_prev; P.pip = pi_prev; P.chi = chi; P.chip = chi_prev;
    P.mask = (p->flags & LFM_F_MASK) ? mask : nullptr;
    // S_a comes from the plan's spectrum inside the kernel; a non-NULL marker
    // switches the SCV term on (it is never dereferenced)
    P.sa = reinterpret_cast<const float*>(v.spec);
    P.pr_out = pr_out; P.pi_out = pi_out; P.chi_out = chi_out;
    P.n3 = n3;
    return launch_kfft_any(P, to_k(p), v, (hipStream_t)stream);
}

extern "C" int lfm_batch_plan_create(int32_t level, int32_t nlat, const lfm_params_t* params,
                                     int32_t tile, int32_t chunk, void** plan) {
    if (!plan || !params || nlat < 1 || level < 0 || level > 2) {
        set_error("lfm_batch_plan_create: bad argument");
        return LFM_E_ARG;
    }
    *plan = nullptr;
    if (tile <= 0) tile = 1;
    if (tile > LFM_TILE_VARIANTS) { set_error("tile variant %d unknown", tile); return LFM_E_ARG; }
    const int TJ = tile_tj(tile);
    const int ncomp = level == 2 ? 3 : 1;
    std::vector<BatchLattice> lat(nlat);
    std::vector<int4> desc;
    long long off = 0;
    for (int b = 0; b < nlat; ++b) {
        const lfm_params_t& p = params[b];
        if (p.N < 8 || p.N > 1023) {
            set_error("lfm_batch_plan_create: lattice %d has N=%d outside [8, 1023]", b, p.N);
            return LFM_E_ARG;
        }
        BatchLattice& L = lat[b];
        L.off = off;
        L.N = p.N;
        L.ncomp = ncomp;
        L.CH = chunk > 0 ? std::min<int>(chunk, p.N) : std::min<int>(16, p.N);
        L.kp = to_k(&p);
        L.kp.flags &= LFM_F_MASK;   // per-lattice flags: the mask only (the rest is per call)
        off += (long long)p.N * p.N * p.N;
    }
    // Largest lattices first (their tiles are the fullest), tiles of one
    // lattice chunk contiguous.
    std::vector<int> order(nlat);
    for (int b = 0; b < nlat; ++b) order[b] = b;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return lat[a].N > lat[b].N; });
    for (int b : order) {
        const int N = lat[b].N, CH = lat[b].CH;
        for (int i0 = 0; i0 < N; i0 += CH)
            for (int j0 = 0; j0 < N; j0 += TJ)
                for (int k0 = 0; k0 < N; k0 += TK) desc.push_back(make_int4(b, k0, j0, i0));
    }
    // XCD grouping: physical block p runs on XCD p % 8; place logical tile l at
    // p so that each XCD takes a contiguous run of tiles.
    const int n = int((desc.size() + 7) / 8 * 8);
    std::vector<int4> phys(n, make_int4(-1, 0, 0, 0));
    for (int l = 0; l < int(desc.size()); ++l) {
        const int x = l / (n / 8), m = l % (n / 8);
        phys[m * 8 + x] = desc[l];
    }
    auto* P = new BatchPlan();
    P->level = level; P->nlat = nlat; P->tile = tile; P->ndesc = n; P->total = off;
    for (int b = 0; b < nlat; ++b) P->scv = P->scv || (level == 2 && params[b].kappa_tube > 0.0f);
    hipError_t e = hipMalloc((void**)&P->d_lat, sizeof(BatchLattice) * nlat);
    if (e == hipSuccess) e = hipMalloc((void**)&P->d_desc, sizeof(int4) * n);
    if (e == hipSuccess) e = hipMemcpy(P->d_lat, lat.data(), sizeof(BatchLattice) * nlat, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(P->d_desc, phys.data(), sizeof(int4) * n, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        set_error("lfm_batch_plan_create: %s", hipGetErrorString(e));
        if (P->d_lat) (void)hipFree(P->d_lat);
        if (P->d_desc) (void)hipFree(P->d_desc);
        delete P;
        return LFM_E_HIP;
    }
    *plan = P;
    return LFM_OK;
}

extern "C" int64_t lfm_batch_total_sites(void* plan) {
    return plan ? static_cast<BatchPlan*>(plan)->total : -1;
}

extern "C" int32_t lfm_batch_workgroups(void* plan) {
    return plan ? static_cast<BatchPlan*>(plan)->ndesc : -1;
}

extern "C" int lfm_batch_plan_destroy(void* plan) {
    auto* P = static_cast<BatchPlan*>(plan);
    if (!P) return LFM_OK;
    (void)hipFree(P->d_lat);
    (void)hipFree(P->d_desc);
    delete P;
    return LFM_OK;
}

extern "C" int lfm_batch_step_f32(void* plan, const float* pr, const float* pr_prev,
                                  const float* pi, const float* pi_prev, const float* chi,
                                  const float* chi_prev, const float* mask, float* pr_out,
                                  float* pi_out, float* chi_out, const float* sa, float* psisq_out,
                                  int32_t flags, void* stream) {
    auto* P = static_cast<BatchPlan*>(plan);
    if (!P || !pr || !pr_prev || !chi || !chi_prev || !mask || !pr_out || !chi_out) {
        set_error("lfm_batch_step_f32: required pointer is NULL");
        return LFM_E_ARG;
    }
    if (P->level > 0 && (!pi || !pi_prev || !pi_out)) {
        set_error("lfm_batch_step_f32: complex levels need pi, pi_prev and pi_out");
        return LFM_E_ARG;
    }
    if (P->scv && !sa) {
        set_error("lfm_batch_step_f32: lattices with kappa_tube > 0 need the S_a arena");
        return LFM_E_ARG;
    }
    if ((flags & LFM_F_SA_SCV) && P->level != 2) {
        set_error("lfm_batch_step_f32: LFM_F_SA_SCV is for COLOR batches");
        return LFM_E_ARG;
    }
    if (flags & (LFM_F_KEEP_PSI | LFM_F_SOURCE | LFM_F_DETECTOR)) {
        set_error("lfm_batch_step_f32: LFM_F_KEEP_PSI / LFM_F_SOURCE / LFM_F_DETECTOR are not "
                  "supported");
        return LFM_E_ARG;
    }
    const size_t ns = size_t(P->total) * (P->level == 2 ? 3 : 1);
    const float* sa_in = P->scv ? sa : nullptr;
    float* sq_out = P->scv ? psisq_out : nullptr;
    const size_t nsa = (flags & LFM_F_SA_SCV) ? size_t(P->total) : ns;
    const float* ins[] = {pr, pr_prev, pi, pi_prev, chi, chi_prev, mask, sa_in};
    const size_t nin[] = {ns, ns, ns, ns, size_t(P->total), size_t(P->total), size_t(P->total), nsa};
    float* outs[] = {pr_out, pi_out, chi_out, sq_out};
    const size_t nout[] = {ns, ns, size_t(P->total), ns};
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 8; ++b)
            if (outs[a] && overlaps(outs[a], nout[a] * 4, ins[b], nin[b] * 4)) {
                set_error("lfm_batch_step_f32: output %d overlaps input %d", a, b);
                return LFM_E_ARG;
            }
    BatchArenas A{pr, pr_prev, P->level ? pi : nullptr, P->level ? pi_prev : nullptr, chi, chi_prev,
                  mask, pr_out, P->level ? pi_out : nullptr, chi_out, sa_in, sq_out};
    const int fo = flags & (LFM_F_DAMP_PREV | LFM_F_FREEZE_CHI | LFM_F_CHI_DELTA | LFM_F_SA_SCV);
    hipStream_t st = (hipStream_t)stream;
    switch (P->level) {
        case 0: return launch_batch<REAL>(A, P->d_lat, P->d_desc, P->ndesc, P->tile, fo,
                                          streaming<REAL>(size_t(P->total)), st);
        case 1: return launch_batch<COMPLEX>(A, P->d_lat, P->d_desc, P->ndesc, P->tile, fo,
                                             streaming<COMPLEX>(size_t(P->total)), st);
        default: return launch_batch<COLOR>(A, P->d_lat, P->d_desc, P->ndesc, P->tile, fo,
                                            streaming<COLOR>(size_t(P->total)), st);
    }
}

extern "C" int lfm_batch_step2_f32(void* plan, const float* pr, const float* pr_prev,
                                   const float* pi, const float* pi_prev, const float* chi,
                                   const float* chi_prev, const float* mask, float* pr_out1,
                                   float* pi_out1, float* chi_out1, float* pr_out2,
                                   float* pi_out2, float* chi_out2, int32_t flags,
                                   void* stream) {
    auto* P = static_cast<BatchPlan*>(plan);
    if (!P || !pr || !pr_prev || !chi || !chi_prev || !mask || !pr_out1 || !chi_out1 ||
        !pr_out2 || !chi_out2) {
        set_error("lfm_batch_step2_f32: required pointer is NULL");
        return LFM_E_ARG;
    }
    if (P->level == 2) {
        set_error("lfm_batch_step2_f32: REAL and COMPLEX batches only");
        return LFM_E_ARG;
    }
    if (P->level == 1 && (!pi || !pi_prev || !pi_out1 || !pi_out2)) {
        set_error("lfm_batch_step2_f32: COMPLEX batches need pi, pi_prev and both pi outputs");
        return LFM_E_ARG;
    }
    if (tile_tj(P->tile) != T2J) {
        set_error("lfm_batch_step2_f32: the plan's column tiles must be 64 x %d (tile variant 2 "
                  "or 4), plan has tile variant %d", T2J, P->tile);
        return LFM_E_ARG;
    }
    if (flags & ~(LFM_F_DAMP_PREV | LFM_F_CHI_DELTA)) {
        set_error("lfm_batch_step2_f32: only LFM_F_DAMP_PREV and LFM_F_CHI_DELTA are supported");
        return LFM_E_ARG;
    }
    const size_t n = size_t(P->total);
    const bool cx = P->level == 1;
    const float* ins[] = {pr, pr_prev, cx ? pi : nullptr, cx ? pi_prev : nullptr, chi, chi_prev,
                          mask};
    float* outs[] = {pr_out1, cx ? pi_out1 : nullptr, chi_out1, pr_out2, cx ? pi_out2 : nullptr,
                     chi_out2};
    for (int a = 0; a < 6; ++a) {
        if (!outs[a]) continue;
        for (int b = 0; b < 7; ++b)
            if (ins[b] && overlaps(outs[a], n * 4, ins[b], n * 4)) {
                set_error("lfm_batch_step2_f32: output %d overlaps input %d", a, b);
                return LFM_E_ARG;
            }
        for (int b = a + 1; b < 6; ++b)
            if (outs[b] && overlaps(outs[a], n * 4, outs[b], n * 4)) {
                set_error("lfm_batch_step2_f32: outputs %d and %d overlap", a, b);
                return LFM_E_ARG;
            }
    }
    BatchArenas2 A2{};
    A2.a = BatchArenas{pr, pr_prev, cx ? pi : nullptr, cx ? pi_prev : nullptr, chi, chi_prev, mask,
                       pr_out1, cx ? pi_out1 : nullptr, chi_out1, nullptr, nullptr};
    A2.pr_out2 = pr_out2;
    A2.pi_out2 = cx ? pi_out2 : nullptr;
    A2.chi_out2 = chi_out2;
    const dim3 g(P->ndesc), b(T2T);
    hipStream_t st = (hipStream_t)stream;
    if (P->level == 0) {
        if (streaming<REAL>(n)) hipLaunchKernelGGL((step2_tiled_batch<REAL, true>), g, b, 0, st, A2, P->d_lat, P->d_desc, flags);
        else hipLaunchKernelGGL((step2_tiled_batch<REAL, false>), g, b, 0, st, A2, P->d_lat, P->d_desc, flags);
    } else {
        if (streaming<COMPLEX>(n)) hipLaunchKernelGGL((step2_tiled_batch<COMPLEX, true>), g, b, 0, st, A2, P->d_lat, P->d_desc, flags);
        else hipLaunchKernelGGL((step2_tiled_batch<COMPLEX, false>), g, b, 0, st, A2, P->d_lat, P->d_desc, flags);
    }
    LFM_LAUNCH_CHECK("step2_tiled_batch");
    return LFM_OK;
}

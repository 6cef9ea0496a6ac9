// Device-side building blocks of the length-N line FFTs (Stockham autosort,
// radix-8 stages plus one radix-2/4 stage; thread t of a line's TC = N/V
// threads holds the V positions t + e*TC, V = 8 or 16), shared by the
// transform passes and the fused COLOR step.
#pragma once

#include <hip/hip_runtime.h>

namespace lfm {

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

// Every pass touches each value once.  For N >= 512 (spectra of 1 GB and
// more) the HBM accesses are streaming (non-temporal); for smaller N the
// plain accesses let the 256 MB last-level cache hold the spectrum between
// passes.
typedef float f2v __attribute__((ext_vector_type(2)));
template <int N>
__device__ __forceinline__ float2 ld2nt(const float2* p) {
    if constexpr (N >= 512) {
        const f2v v = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(p));
        return make_float2(v.x, v.y);
    } else {
        return *p;
    }
}
template <int N>
__device__ __forceinline__ void st2nt(float2* p, float2 v) {
    if constexpr (N >= 512) {
        const f2v w = {v.x, v.y};
        __builtin_nontemporal_store(w, reinterpret_cast<f2v*>(p));
    } else {
        *p = v;
    }
}
template <int N>
__device__ __forceinline__ float ld1nt(const float* p) {
    if constexpr (N >= 512) return __builtin_nontemporal_load(p);
    else return *p;
}
template <int N>
__device__ __forceinline__ void st1nt(float* p, float v) {
    if constexpr (N >= 512) __builtin_nontemporal_store(v, p);
    else *p = v;
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

// One radix-R stage (span Ns) on the V register values of thread t.
// Butterfly u (of NB = V/R) takes positions j + r*(N/R), j = t + u*TC,
// which are the register slots u + r*NB.
template <int N, int R, int V = 8>
__device__ __forceinline__ void stage_regs(float2 (&x)[V], int t, int Ns, int sign,
                                           const float2* __restrict__ twl) {
    constexpr int TC = N / V;
    constexpr int NB = V / R;
#pragma unroll
    for (int u = 0; u < NB; ++u) {
        const int k = (t + u * TC) & (Ns - 1);
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            float2 a = x[u + r * NB];
            if (r > 0 && Ns > 1) {
                float2 w = twl[(r * k * (N / (Ns * R))) & (N - 1)];
                if (sign > 0) w.y = -w.y;
                a = cmul(a, w);
            }
            v[r] = a;
        }
        dft<R>(v, sign);
#pragma unroll
        for (int r = 0; r < R; ++r) x[u + r * NB] = v[r];
    }
}

// How a stage boundary orders its LDS writes and reads.
//   Block: the line spans several wavefronts.  A workgroup barrier before the
//          writes (every thread has read the previous boundary's values)
//          and one between the writes and the reads.
//   Wave:  the line lies in one wavefront (N/8 <= 64 threads per line).  A
//          wavefront's LDS accesses complete in issue order, so only the
//          compiler has to be kept from moving LDS accesses across the
//          boundary (wave barrier, no instruction).
//   Alt:   two LDS buffers taken in turn at successive boundaries.  The
//          barrier between one boundary's writes and reads also orders every
//          thread's reads of the other buffer (at the previous boundary)
//          before the next boundary's writes to it: one barrier a boundary.
enum class Sync { Block, Wave, Alt };

template <Sync S>
__device__ __forceinline__ void line_sync() {
    if constexpr (S == Sync::Wave) __builtin_amdgcn_wave_barrier();
    else __syncthreads();
}

// Stage boundary through LDS: results of the radix-R stage of span Ns go to
// their Stockham positions (j/Ns)*Ns*R + j%Ns + r*Ns, then every thread
// reads back positions t + e*TC.  `at(pos)` maps a line position to its LDS
// slot.
template <int N, int R, Sync S, int V = 8, class At>
__device__ __forceinline__ void exchange(float2 (&x)[V], float2* lds, int t, int Ns, At at) {
    constexpr int TC = N / V;
    constexpr int NB = V / R;
    if constexpr (S != Sync::Alt) line_sync<S>();
#pragma unroll
    for (int u = 0; u < NB; ++u) {
        const int j = t + u * TC;
        const int k = j & (Ns - 1);
        const int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) lds[at(base + r * Ns)] = x[u + r * NB];
    }
    line_sync<S>();
#pragma unroll
    for (int e = 0; e < V; ++e) x[e] = lds[at(t + e * TC)];
}

// Full length-N transform of the register line; result in the same slots.
// With Sync::Alt, boundaries alternate between `lds` and `lds_alt`, and
// `*par` counts the boundaries the workgroup has passed (kept across calls).
template <int N, Sync S = Sync::Block, int V = 8, class At>
__device__ __forceinline__ void fft_line(float2 (&x)[V], float2* lds, int t, int sign,
                                         const float2* __restrict__ twl, At at,
                                         float2* lds_alt = nullptr, int* par = nullptr) {
    constexpr int L = ilog2(N);
    constexpr int n8 = L / 3;
    constexpr int rem = L % 3;
    int Ns = 1;
#pragma unroll
    for (int s = 0; s < n8; ++s) {
        stage_regs<N, 8, V>(x, t, Ns, sign, twl);
        if (s + 1 < n8 || rem != 0) {
            if constexpr (S == Sync::Alt) {
                exchange<N, 8, S, V>(x, (*par & 1) ? lds_alt : lds, t, Ns, at);
                ++*par;
            } else {
                exchange<N, 8, S, V>(x, lds, t, Ns, at);
            }
        }
        Ns *= 8;
    }
    if constexpr (rem == 1) stage_regs<N, 2, V>(x, t, Ns, sign, twl);
    if constexpr (rem == 2) stage_regs<N, 4, V>(x, t, Ns, sign, twl);
}

template <int N>
__device__ __forceinline__ void load_twiddles(float2* twl, const float2* __restrict__ tw) {
    for (int m = threadIdx.x; m < N; m += blockDim.x) twl[m] = tw[m];
}


// As fft_line, for a workgroup in which only some threads hold a line
// (`active`): every thread takes part in the barriers, only active threads
// touch LDS.  `sign` may differ between lines (forward and inverse side by
// side).
template <int N, int R, class At>
__device__ __forceinline__ void exchange_m(float2 (&x)[8], float2* lds, int t, int Ns, At at,
                                          bool active) {
    constexpr int TC = N / 8;
    constexpr int NB = 8 / R;
    __syncthreads();
    if (active) {
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const int j = t + u * TC;
            const int k = j & (Ns - 1);
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) lds[at(base + r * Ns)] = x[u + r * NB];
        }
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int e = 0; e < 8; ++e) x[e] = lds[at(t + e * TC)];
    }
}

template <int N, class At>
__device__ __forceinline__ void fft_line_m(float2 (&x)[8], float2* lds, int t, int sign,
                                           const float2* __restrict__ twl, At at, bool active) {
    constexpr int L = ilog2(N);
    constexpr int n8 = L / 3;
    constexpr int rem = L % 3;
    int Ns = 1;
#pragma unroll
    for (int s = 0; s < n8; ++s) {
        stage_regs<N, 8>(x, t, Ns, sign, twl);
        if (s + 1 < n8 || rem != 0) exchange_m<N, 8>(x, lds, t, Ns, at, active);
        Ns *= 8;
    }
    if constexpr (rem == 1) stage_regs<N, 2>(x, t, Ns, sign, twl);
    if constexpr (rem == 2) stage_regs<N, 4>(x, t, Ns, sign, twl);
}

}  // namespace lfm

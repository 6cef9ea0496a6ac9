// The 19-point Laplacian (lfm/core/stencils.py:26-71, weights
// lfm/constants.py:135-142) in the plane-partial form every kernel of the
// library uses, with each rounding spelled out (fmaf and the mul_rn / add_rn /
// sub_rn below) so kernels evaluating the same site in different code round
// identically:
//   A(q) = 1/3 (4 in-plane faces) + 1/6 (4 in-plane edges) - 4 f
//   B(q) = 1/3 f + 1/6 (4 in-plane faces)
//   lap(i) = (A(i) + B(i-1)) + B(i+1)
#pragma once

#include <hip/hip_runtime.h>

namespace lfm {

// One float32 operation, rounded to nearest, never fused with a neighbouring
// product or sum into a multiply-add: contraction is off inside these
// functions, and the compiler fuses a product and a sum only when both allow
// it.  (HIP's __fmul_rn / __fadd_rn / __fsub_rn are plain operators in
// contractible code under the default -ffp-contract=fast-honor-pragmas, so
// they do not give this guarantee.)
__device__ __forceinline__ float mul_rn(float a, float b) {
#pragma clang fp contract(off)
    return a * b;
}
__device__ __forceinline__ float add_rn(float a, float b) {
#pragma clang fp contract(off)
    return a + b;
}
__device__ __forceinline__ float sub_rn(float a, float b) {
#pragma clang fp contract(off)
    return a - b;
}

constexpr float WF = 1.0f / 3.0f;  // face weight   (lfm/constants.py:135)
constexpr float WE = 1.0f / 6.0f;  // edge weight   (lfm/constants.py:138)

__device__ __forceinline__ float lap_part_a(float faces, float diag, float ctr) {
    return fmaf(WF, faces, fmaf(WE, diag, mul_rn(-4.0f, ctr)));
}
__device__ __forceinline__ float lap_part_b(float ctr, float faces) {
    return fmaf(WF, ctr, mul_rn(WE, faces));
}

// |Psi_a|^2 = Pr^2 + Pi^2 (numpy_backend.py:340-341, the S_a transform's
// input) with its roundings fixed, so the transform pass that forms it from
// Psi and any kernel writing it for later use give the same bits.
__device__ __forceinline__ float psi_sq(float r, float m) {
    return fmaf(r, r, mul_rn(m, m));
}

// The S_a colour-variance term of GOV-02 (kernel_source.py:187-197):
// SCV = sum_a S_a^2 - 1/3 (sum_a S_a)^2 of the S_a clipped at 0
// (numpy_backend.py:353), every rounding explicit.  The step kernels apply
// it to S_a they read; the last transform pass evaluates the same function
// on the S_a it produces, so either route gives the same bits.
__device__ __forceinline__ float scv_term(float s0, float s1, float s2, float (&clip)[3]) {
    const float s[3] = {s0, s1, s2};
    float s_sum = 0.0f, s_sq = 0.0f;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const float sv = s[q] > 0.0f ? s[q] : 0.0f;
        clip[q] = sv;
        s_sum = add_rn(s_sum, sv);
        s_sq = fmaf(sv, sv, s_sq);
    }
    return fmaf(-1.0f / 3.0f, mul_rn(s_sum, s_sum), s_sq);
}

}  // namespace lfm

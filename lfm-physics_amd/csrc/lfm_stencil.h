// The 19-point Laplacian (lfm/core/stencils.py:26-71, weights
// lfm/constants.py:135-142) in the plane-partial form every kernel of the
// library uses, with each rounding spelled out (fmaf and non-contractible
// products) so kernels evaluating the same site in different code round
// identically:
//   A(q) = 1/3 (4 in-plane faces) + 1/6 (4 in-plane edges) - 4 f
//   B(q) = 1/3 f + 1/6 (4 in-plane faces)
//   lap(i) = (A(i) + B(i-1)) + B(i+1)
#pragma once

#include <hip/hip_runtime.h>

namespace lfm {

constexpr float WF = 1.0f / 3.0f;  // face weight   (lfm/constants.py:135)
constexpr float WE = 1.0f / 6.0f;  // edge weight   (lfm/constants.py:138)

__device__ __forceinline__ float lap_part_a(float faces, float diag, float ctr) {
    return fmaf(WF, faces, fmaf(WE, diag, __fmul_rn(-4.0f, ctr)));
}
__device__ __forceinline__ float lap_part_b(float ctr, float faces) {
    return fmaf(WF, ctr, __fmul_rn(WE, faces));
}

}  // namespace lfm

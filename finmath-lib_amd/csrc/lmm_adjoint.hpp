// Wavefront reductions and the reciprocal of the adjoint sweep (lmm_spot_adjoint.hpp); the
// single-value DPP sum is in lmm_step.hpp.
#pragma once

#include "lmm_step.hpp"

namespace fm {

// Wavefront sums of four per-lane values at once (a transposed butterfly): two exchange
// steps within each quad (DPP quad permutes) halve the values a lane carries, two DPP row
// rotations sum the four lanes of each row that carry the same value, and two lane permutes
// sum across the rows. Lane l < 4 returns the total of value 2*(l & 1) + ((l >> 1) & 1).
__device__ __forceinline__ double wave_sum4(const double (&v)[4], int lane) {
    const bool b0 = (lane & 1) != 0, b1 = (lane & 2) != 0;
    double a2[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const double send = b0 ? v[i] : v[i + 2];
        const double keep = b0 ? v[i + 2] : v[i];
        a2[i] = keep + dpp_move<0xb1>(send);      // quad_perm [1,0,3,2]: lane ^ 1
    }
    const double send = b1 ? a2[0] : a2[1];
    const double keep = b1 ? a2[1] : a2[0];
    double a = keep + dpp_move<0x4e>(send);       // quad_perm [2,3,0,1]: lane ^ 2
    a = a + dpp_move<0x124>(a);                   // row_ror:4
    a = a + dpp_move<0x128>(a);                   // row_ror:8
    a = a + __shfl_xor(a, 16, 64);
    a = a + __shfl_xor(a, 32, 64);
    return a;
}

// Reciprocal for the adjoint-only terms: the hardware estimate in float (about 1 ulp),
// the correctly rounded division in double.
template <typename real> __device__ __forceinline__ real adj_rcp(real x);
template <> __device__ __forceinline__ float adj_rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <> __device__ __forceinline__ double adj_rcp<double>(double x) { return 1.0 / x; }

} // namespace fm

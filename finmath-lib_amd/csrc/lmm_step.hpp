// Definitions shared by the path kernels (lmm_spot.hpp: the spot-measure Euler kernel,
// lmm_general.hpp: the other measures and schemes, lmm_spot_adjoint.hpp: the reverse sweep):
// the model records, the increments, the state-space transform and the wavefront reductions
// of the per-path payoffs.
//
// Model data layout: one record of rec_size_of(FCAP) reals per (time step,
// component): the factor loadings padded with zeros to FCAP, then variance*(-0.5),
// d/d and 1/d. A zero loading adds exact zeros, so the padding changes no value;
// a record is one aligned scalar load.
#pragma once

#include <hip/hip_runtime.h>
#include "fm_internal.hpp"
#include "fm_math.hpp"

namespace fm {

constexpr int kStepThreads = 256;

// Wavefront sum by DPP row operations (quad_perm [1,0,3,2], [2,3,0,1], row_ror 4 and 8,
// row_bcast 15 and 31): the total is valid in lane 63 only. Each step moves the two
// 32-bit halves of the double; lanes without a source read 0. Call with every lane of the
// wavefront active.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

__device__ __forceinline__ double wave_sum_dpp(double v) {
    v = v + dpp_move<0xb1>(v);
    v = v + dpp_move<0x4e>(v);
    v = v + dpp_move<0x124>(v);
    v = v + dpp_move<0x128>(v);
    v = v + dpp_move<0x142>(v);
    v = v + dpp_move<0x143>(v);
    return v;
}

// Wavefront sum, returned in every lane: the DPP sum read from lane 63. Unlike a butterfly of
// lane permutes it keeps no per-lane permute addresses live across the kernel.
__device__ __forceinline__ double wave_reduce_sum(double v) {
    const long long b = __double_as_longlong(wave_sum_dpp(v));
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), 63);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

template <typename real> __device__ __forceinline__ real dev_exp(real x);
template <typename real> __device__ __forceinline__ real dev_log(real x);
template <> __device__ __forceinline__ double dev_exp<double>(double x) { return exp(x); }
template <> __device__ __forceinline__ double dev_log<double>(double x) { return fm_log(x); }
// FP32 form: the hardware base-2 exponential and logarithm (about 1 ulp each).
template <> __device__ __forceinline__ float dev_exp<float>(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }
template <> __device__ __forceinline__ float dev_log<float>(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }

// Java Math.max(x, 0.0) / Math.min(x, c): NaN propagates.
template <typename real> __device__ __forceinline__ real jmax0(real x) { return (x != x) ? x : (x >= (real)0 ? x : (real)0); }
template <typename real> __device__ __forceinline__ real jmin(real x, real c) { return (x > c) ? c : x; }

// Reduce one per-path value into the wave partial of `slot`.
__device__ __forceinline__ void reduce_slot(double x, bool active, int cnt, int lane, int64_t wave,
                                            double* __restrict__ wave_sum, double* __restrict__ wave_m2,
                                            int stride, int slot) {
    const double s = wave_reduce_sum(active ? x : 0.0);
    const double mean = (cnt > 0) ? s / (double)cnt : 0.0;
    const double d = active ? (x - mean) : 0.0;
    const double m2 = wave_reduce_sum(d * d);
    if (lane == 0 && cnt > 0) {
        wave_sum[(size_t)wave * stride + slot] = s;
        wave_m2[(size_t)wave * stride + slot] = m2;
    }
}

// Record of one (time step, component).
template <int FCAP, typename real>
struct Rec {
    static constexpr int size = rec_size_of(FCAP);
    const real* __restrict__ r;
    __device__ __forceinline__ real lam(int k) const { return r[k]; }
    __device__ __forceinline__ real hv() const { return r[FCAP]; }
    __device__ __forceinline__ real c1() const { return r[FCAP + 1]; }
    __device__ __forceinline__ real c2() const { return r[FCAP + 2]; }
};

template <int FCAP, typename real>
__device__ __forceinline__ Rec<FCAP, real> rec_at(const real* __restrict__ rec, int n, int i, int j) {
    const real* r = rec + ((size_t)i * n + j) * Rec<FCAP, real>::size;
    return Rec<FCAP, real>{(const real*)__builtin_assume_aligned(r, Rec<FCAP, real>::size * sizeof(real))};
}

// Value of a state: applyStateSpaceTransform (LMM:1200-1212), exp in the log-normal
// state space, then the cap.
template <typename real>
__device__ __forceinline__ real to_value(real y, bool lognormal, bool capped, real cap) {
    const real v = lognormal ? dev_exp<real>(y) : y;
    return capped ? jmin(v, cap) : v;
}

// The increments dW_k(t_i) of one path, zero above the F factors (path-major layout
// [path][time index][factor]).
template <int FCAP, typename real>
__device__ __forceinline__ void load_increments(real (&dW)[FCAP], const real* __restrict__ dw, int64_t pc, int nT,
                                                int i, int F) {
#pragma unroll
    for (int k = 0; k < FCAP; k++) dW[k] = (k < F) ? dw[((size_t)pc * nT + i) * F + k] : (real)0;
}

} // namespace fm

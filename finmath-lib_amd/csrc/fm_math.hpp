// Double-precision natural logarithm for the path kernels.
//
// x = m 2^e with m in [sqrt(1/2), sqrt(2)), f = m - 1 (exact), s = f / (2 + f):
//   log(1 + f) = 2 atanh(s) = f - s (f - R),   R = sum_{k>=1} 2/(2k+1) s^(2k)
// evaluated as f - (h - s (h + R)) with h = f^2/2 (the identity f (1 - s) = 2 s turns
// f - s f into f - h + s h), and log(x) = e ln2_hi + (log(1 + f) + e ln2_lo) with ln2_hi
// holding 21 significant bits so that e ln2_hi is exact. |s| <= 0.1716, so ten terms of
// the series leave a truncation below 1e-18 relative; the error of the result is about
// one unit in the last place (checked against the C library on the CPU).
//
// Fused multiply-adds are written explicitly, so the result does not depend on the
// translation unit's contraction setting.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>

namespace fm {

__host__ __device__ inline double fm_log(double x) {
    constexpr double kLn2Hi = 0x1.62e42p-1;                  // 21 significant bits of ln 2
    constexpr double kLn2Lo = 0x1.fdf473de6af28p-22;         // ln 2 - kLn2Hi
    constexpr double kSqrtHalf = 0.70710678118654752440;
    int e = 0;
    double m = frexp(x, &e);                                 // x = m 2^e, m in [0.5, 1)
    if (m < kSqrtHalf) { m = m + m; e = e - 1; }
    const double f = m - 1.0;
    const double d = 2.0 + f;
    double s = f / d;
    const double z = s * s;
    // R(z) = z (2/3 + z (2/5 + z (2/7 + ... + z 2/21)))
    double p = 0.09523809523809523;                          // 2/21
    p = fma(p, z, 0.10526315789473684);                      // 2/19
    p = fma(p, z, 0.11764705882352941);                      // 2/17
    p = fma(p, z, 0.13333333333333333);                      // 2/15
    p = fma(p, z, 0.15384615384615385);                      // 2/13
    p = fma(p, z, 0.18181818181818182);                      // 2/11
    p = fma(p, z, 0.22222222222222222);                      // 2/9
    p = fma(p, z, 0.2857142857142857);                       // 2/7
    p = fma(p, z, 0.4);                                      // 2/5
    p = fma(p, z, 0.6666666666666666);                       // 2/3
    const double R = p * z;
    const double h = 0.5 * f * f;
    const double de = (double)e;
    const double r = de * kLn2Hi + (f - (h - fma(s, h + R, de * kLn2Lo)));
    // x <= 0, infinities and NaN: log(0) = -inf, log(x < 0) = NaN, log(+inf) = +inf
    if (!(x > 0.0) || x == HUGE_VAL) return (x == 0.0) ? -HUGE_VAL : ((x > 0.0) ? x : (x - x) / (x - x));
    return r;
}

// a / b by the hardware reciprocal estimate refined with two Newton-Raphson steps
// (r <- r + r (1 - b r)) and one correction of the quotient (q <- q + r (a - b q)). For finite
// b of moderate exponent (the path kernels divide by 1 + L d or L + 1/d) the result is
// within one unit in the last place of the correctly rounded quotient and almost always
// equal to it; b = +-inf gives NaN instead of zero. Seven instructions against the eleven
// of the scaled IEEE division sequence.
// fm_rcp_nr(b) is the refined reciprocal, fm_div_nr_r(a, b, r) the quotient from it: divisions of
// several numerators by one b share the reciprocal and give the same results as fm_div_nr.
__device__ __forceinline__ double fm_rcp_nr(double b) {
    double r = __builtin_amdgcn_rcp(b);
    double e = fma(-b, r, 1.0);
    r = fma(r, e, r);
    e = fma(-b, r, 1.0);
    return fma(r, e, r);
}

__device__ __forceinline__ double fm_div_nr_r(double a, double b, double r) {
    const double q = a * r;
    return fma(r, fma(-b, q, a), q);
}

__device__ __forceinline__ double fm_div_nr(double a, double b) { return fm_div_nr_r(a, b, fm_rcp_nr(b)); }

} // namespace fm

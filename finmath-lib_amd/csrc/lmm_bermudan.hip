// Bermudan swaption by regression: the backward induction of
// BermudanSwaption.getValues (montecarlo/interestrate/products/BermudanSwaption.java:91-157)
// over per-path columns written by the fused path kernel.
//
// The path kernel leaves, per path, the period payoffs divided by the unadjusted
// numeraire of their payment date and the raw regression basis of every exercise
// date. The induction runs backwards over the exercise dates; each date takes two
// launches:
//   berm_step_kernel   apply the decision of the later date (trigger = sum_k x_k b_k,
//                      values = trigger >= 0 ? values : underlying), add the payoffs of
//                      the periods down to this date, and accumulate per workgroup the
//                      Gram matrix X^T X and X^T y of y = values - underlying;
//   berm_solve_kernel  reduce the workgroup partials in a fixed order, average them,
//                      and solve X^T X x = X^T y with the pseudo-inverse of the
//                      symmetric eigen decomposition (the SVD solver of
//                      MonteCarloConditionalExpectationRegression.java:127-147 on a
//                      symmetric positive semi-definite matrix).
// The final step applies the first date's decision, adds the remaining periods and
// reduces values * N(0) / w into the product's reduction slot.

#include <hip/hip_runtime.h>
#include "fm_internal.hpp"
#include "hip_check.hpp"
#include "lmm_step.hpp"

namespace fm {

namespace {

constexpr int kBermThreads = 256;

// Numeraire adjusted by the discount curve, in the operation order of getNumeraire
// (LIBORMarketModelFromCovarianceModel.java:869-874).
__device__ __forceinline__ double adjusted_numeraire(const BermBasisRef& r, double N) {
    if (r.nmode == 1) return (1.0 / r.adj) * (N * r.zb);
    if (r.nmode == 2) return (N * r.zb) / r.adj;
    return N;
}

// Basis functions of one exercise date for path p; returns their number.
__device__ __forceinline__ int berm_basis(const BermBasisRef& r, const double* __restrict__ contrib, int P,
                                          int64_t p, double (&b)[kBermMaxBasis]) {
#pragma unroll
    for (int k = 0; k < kBermMaxBasis; k++) b[k] = 0.0;
    b[0] = 1.0;
    if (r.kind == FM_BASIS_REFERENCE) {
        const double ds = contrib[(size_t)r.col0 * P + p];
        const double dl = contrib[(size_t)(r.col0 + 1) * P + p];
        const double N = adjusted_numeraire(r, contrib[(size_t)(r.col0 + 2) * P + p]);
        b[1] = ds; b[2] = ds * ds; b[3] = dl; b[4] = dl * dl; b[5] = 1.0 / N;
        return 6;
    }
    const double S = contrib[(size_t)r.col0 * P + p];
    b[1] = S; b[2] = S * S; b[3] = adjusted_numeraire(r, contrib[(size_t)(r.col0 + 1) * P + p]);
    return 4;
}

__global__ void __launch_bounds__(kBermThreads)
berm_step_kernel(const BermStep s)
{
    const int P = s.P;
    const int64_t p = (int64_t)blockIdx.x * kBermThreads + threadIdx.x;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int64_t wave = p >> 6;
    const bool active = p < P;
    const int64_t pc = active ? p : (int64_t)P - 1;
    const int cnt = (int)((int64_t)P - wave * 64 > 64 ? 64 : ((int64_t)P - wave * 64));

    double v = s.first ? 0.0 : s.values[pc];
    double u = s.first ? 0.0 : s.under[pc];
    if (s.apply) {
        double b[kBermMaxBasis];
        const int K = berm_basis(s.apply_basis, s.contrib, P, pc, b);
        // getConditionalExpectation (MonteCarloConditionalExpectationRegression.java:79-91)
        double t = b[0] * s.apply_coef[0];
        for (int k = 1; k < K; k++) t = t + b[k] * s.apply_coef[k];
        // choose (RandomVariableFromDoubleArray.java:1342-1363): trigger >= 0 keeps the values
        if (!(t >= 0.0)) {
            v = u;
            if (s.ex_index && active) s.ex_index[p] = s.apply_period;
        }
    }
    for (int q = 0; q < s.num_per; q++) {
        double x = s.contrib[(size_t)s.per_col[q] * P + pc] * s.per_scale[q];
        x = x * s.w;
        if (s.callable) u = u + x; else v = v + x;
    }
    if (s.regress) {
        __shared__ double part[kBermThreads / 64][kBermParts];
        double b[kBermMaxBasis];
        const int K = berm_basis(s.reg_basis, s.contrib, P, pc, b);
        const double y = v - u;
        int e = 0;
        for (int i = 0; i < K; i++, e++) {
            const double r = wave_reduce_sum(active ? y * b[i] : 0.0);
            if (lane == 0) part[wid][e] = r;
        }
        for (int i = 0; i < K; i++)
            for (int j = i; j < K; j++, e++) {
                const double r = wave_reduce_sum(active ? b[i] * b[j] : 0.0);
                if (lane == 0) part[wid][e] = r;
            }
        __syncthreads();
        if ((int)threadIdx.x < e) {
            double acc = 0.0;
            for (int w = 0; w < kBermThreads / 64; w++) acc += part[w][threadIdx.x];
            s.partials[(size_t)blockIdx.x * kBermParts + threadIdx.x] = acc;
        }
    }
    if (s.final_) {
        const double val = (v * s.n0) / s.w;
        reduce_slot(val, active, cnt, lane, wave, s.wave_sum, s.wave_m2, s.stride, s.slot);
    } else if (active) {
        s.values[p] = v;
        s.under[p] = u;
    }
}

// One workgroup: fixed-order reduction of the workgroup partials, then (thread 0)
// the regression solve in FP64.
__global__ void __launch_bounds__(kBermThreads)
berm_solve_kernel(const double* __restrict__ partials, int num_blocks, int K, int P, double* __restrict__ coef)
{
    __shared__ double red[kBermThreads];
    __shared__ double tot[kBermParts];
    __shared__ double A[kBermMaxBasis * kBermMaxBasis], V[kBermMaxBasis * kBermMaxBasis];
    const int nparts = K + K * (K + 1) / 2;
    const int t = threadIdx.x;
    for (int e = 0; e < nparts; e++) {
        // Neumaier sum of a strided subset, then a fixed tree
        double sum = 0.0, c = 0.0;
        for (int b = t; b < num_blocks; b += kBermThreads) {
            const double x = partials[(size_t)b * kBermParts + e];
            const double tt = sum + x;
            if (fabs(sum) >= fabs(x)) c += (sum - tt) + x; else c += (x - tt) + sum;
            sum = tt;
        }
        red[t] = sum + c;
        __syncthreads();
        for (int m = kBermThreads / 2; m >= 1; m >>= 1) {
            if (t < m) red[t] = red[t] + red[t + m];
            __syncthreads();
        }
        if (t == 0) tot[e] = red[0];
        __syncthreads();
    }
    if (t != 0) return;
    double rhs[kBermMaxBasis], x[kBermMaxBasis], ev[kBermMaxBasis];
    int e = 0;
    for (int i = 0; i < K; i++, e++) rhs[i] = tot[e] / (double)P;
    for (int i = 0; i < K; i++)
        for (int j = i; j < K; j++, e++) {
            const double g = tot[e] / (double)P;
            A[i * K + j] = g;
            A[j * K + i] = g;
        }
    // cyclic Jacobi rotations (Golub & Van Loan 8.5); V accumulates the eigenvectors
    for (int i = 0; i < K; i++)
        for (int j = 0; j < K; j++) V[i * K + j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0;
        for (int i = 0; i < K; i++)
            for (int j = i + 1; j < K; j++) off += A[i * K + j] * A[i * K + j];
        if (off < 1e-300) break;
        for (int pp = 0; pp < K; pp++)
            for (int q = pp + 1; q < K; q++) {
                const double apq = A[pp * K + q];
                if (fabs(apq) < 1e-300) continue;
                const double app = A[pp * K + pp], aqq = A[q * K + q];
                const double theta = (aqq - app) / (2.0 * apq);
                const double tn = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double cs = 1.0 / sqrt(tn * tn + 1.0), sn = tn * cs;
                for (int k = 0; k < K; k++) {
                    const double akp = A[k * K + pp], akq = A[k * K + q];
                    A[k * K + pp] = cs * akp - sn * akq;
                    A[k * K + q] = sn * akp + cs * akq;
                }
                for (int k = 0; k < K; k++) {
                    const double apk = A[pp * K + k], aqk = A[q * K + k];
                    A[pp * K + k] = cs * apk - sn * aqk;
                    A[q * K + k] = sn * apk + cs * aqk;
                }
                for (int k = 0; k < K; k++) {
                    const double vkp = V[k * K + pp], vkq = V[k * K + q];
                    V[k * K + pp] = cs * vkp - sn * vkq;
                    V[k * K + q] = sn * vkp + cs * vkq;
                }
            }
    }
    double smax = 0.0;
    for (int i = 0; i < K; i++) { ev[i] = A[i * K + i]; smax = fmax(smax, fabs(ev[i])); x[i] = 0.0; }
    // singular values at or below the SVD solver's threshold are dropped
    const double tol = fmax(K * smax * 0x1.0p-52, sqrt(0x1.0p-1022));
    for (int i = 0; i < K; i++) {
        if (!(fabs(ev[i]) > tol)) continue;
        double d = 0.0;
        for (int r = 0; r < K; r++) d += V[r * K + i] * rhs[r];
        d /= ev[i];
        for (int r = 0; r < K; r++) x[r] += V[r * K + i] * d;
    }
    for (int i = 0; i < kBermMaxBasis; i++) coef[i] = (i < K) ? x[i] : 0.0;
}

} // namespace

int berm_num_blocks(int P) { return (int)(((int64_t)P + kBermThreads - 1) / kBermThreads); }

void launch_berm_step(const BermStep& s, void* stream) {
    hipLaunchKernelGGL(berm_step_kernel, dim3((unsigned)berm_num_blocks(s.P)), dim3(kBermThreads), 0,
                       (hipStream_t)stream, s);
    HIP_CHECK(hipGetLastError());
}

void launch_berm_solve(const double* partials, int num_blocks, int K, int P, double* coef, void* stream) {
    hipLaunchKernelGGL(berm_solve_kernel, dim3(1), dim3(kBermThreads), 0, (hipStream_t)stream,
                       partials, num_blocks, K, P, coef);
    HIP_CHECK(hipGetLastError());
}

} // namespace fm

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
//
// The adjoint of the induction (berm_adjoint_step_kernel) runs the exercise dates forwards with
// the per-path adjoints of the values and of the underlying, and leaves the weight of every period
// payoff per path for the reverse sweep of the paths. At an exercise date, choose(trigger, values,
// underlying) passes the adjoint to the side it chose; the trigger's own derivative is the
// DiracDeltaApproximationMethod of RandomVariableDifferentiableAADFactory.java:26-67 (AAD:396-424:
// ONE (Y - Z), ZERO 0, DISCRETE_DELTA (Y - Z) 1{-eps/2 <= X < eps/2} / eps, eps = width x sd(X));
// the trigger is a conditional expectation, self-adjoint (AAD:188-192), so its adjoint is regressed
// on the same basis, whose functions are not differentiated, and added to values and subtracted
// from the underlying.

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

// Basis functions of one exercise date for path p.
template <int KIND>
__device__ __forceinline__ void berm_basis(const BermBasisRef& r, const double* __restrict__ contrib, int P,
                                           int64_t p, double (&b)[KIND == FM_BASIS_REFERENCE ? 6 : 4]) {
    b[0] = 1.0;
    if (KIND == FM_BASIS_REFERENCE) {
        const double ds = contrib[(size_t)r.col0 * P + p];
        const double dl = contrib[(size_t)(r.col0 + 1) * P + p];
        const double N = adjusted_numeraire(r, contrib[(size_t)(r.col0 + 2) * P + p]);
        b[1] = ds; b[2] = ds * ds; b[3] = dl; b[4] = dl * dl; b[5] = 1.0 / N;
    } else {
        const double S = contrib[(size_t)r.col0 * P + p];
        b[1] = S; b[2] = S * S; b[3] = adjusted_numeraire(r, contrib[(size_t)(r.col0 + 1) * P + p]);
    }
}

// Regression partials of one path: X^T y and the upper triangle of X^T X.
template <int K>
__device__ __forceinline__ void regression_accumulate(double (&acc)[K + K * (K + 1) / 2], double y,
                                                      const double (&b)[K]) {
    int e = 0;
#pragma unroll
    for (int i = 0; i < K; i++, e++) acc[e] += y * b[i];
#pragma unroll
    for (int i = 0; i < K; i++)
#pragma unroll
        for (int j = i; j < K; j++, e++) acc[e] += b[i] * b[j];
}

// One row of NP regression partials per workgroup (fixed-order wave and workgroup sums).
template <int NP>
__device__ __forceinline__ void regression_partials(const double (&acc)[NP], double* __restrict__ partials, int lane,
                                                    int wid) {
    __shared__ double part[kBermThreads / 64][NP];
#pragma unroll
    for (int e = 0; e < NP; e++) {
        const double r = wave_reduce_sum(acc[e]);
        if (lane == 0) part[wid][e] = r;
    }
    __syncthreads();
    if ((int)threadIdx.x < NP) {
        double t = 0.0;
        for (int w = 0; w < kBermThreads / 64; w++) t += part[w][threadIdx.x];
        partials[(size_t)blockIdx.x * kBermParts + threadIdx.x] = t;
    }
}

// Grid-stride over paths: every thread handles the paths p = g, g + G, ... (G threads in
// the grid) and keeps its regression partials in registers; each workgroup writes one
// row of partials (fixed-order wave and workgroup reductions).
template <int KIND>
__global__ void __launch_bounds__(kBermThreads)
berm_step_kernel(const BermStep s)
{
    constexpr int K = KIND == FM_BASIS_REFERENCE ? 6 : 4;
    constexpr int NP = K + K * (K + 1) / 2;
    const int P = s.P;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    double acc[NP];
#pragma unroll
    for (int e = 0; e < NP; e++) acc[e] = 0.0;
    const int64_t G = (int64_t)gridDim.x * kBermThreads;
    // whole wavefronts stay together so the slot reduction sees aligned groups of 64 paths
    for (int64_t p0 = (int64_t)blockIdx.x * kBermThreads; p0 < P; p0 += G) {
        const int64_t p = p0 + threadIdx.x;
        const int64_t wave = p >> 6;
        const bool active = p < P;
        const int64_t pc = active ? p : (int64_t)P - 1;
        const int cnt = (int)((int64_t)P - wave * 64 > 64 ? 64 : ((int64_t)P - wave * 64));
        double v = s.first ? 0.0 : s.values[pc];
        double u = s.first ? 0.0 : s.under[pc];
        if (s.apply) {
            double b[K];
            berm_basis<KIND>(s.apply_basis, s.contrib, P, pc, b);
            // getConditionalExpectation (MonteCarloConditionalExpectationRegression.java:79-91)
            double t = b[0] * s.apply_coef[0];
#pragma unroll
            for (int k = 1; k < K; k++) t = t + b[k] * s.apply_coef[k];
            // choose (RandomVariableFromDoubleArray.java:1342-1363): trigger >= 0 keeps the values
            if (s.rec_trig && active) {
                s.rec_trig[p] = t;
                s.rec_diff[p] = v - u;
            }
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
        if (s.regress && active) {
            double b[K];
            berm_basis<KIND>(s.reg_basis, s.contrib, P, pc, b);
            regression_accumulate<K>(acc, v - u, b);
        }
        if (s.final_) {
            const double val = (v * s.n0) / s.w;
            reduce_slot(val, active, cnt, lane, wave, s.wave_sum, s.wave_m2, s.stride, s.slot);
        } else if (active) {
            s.values[p] = v;
            s.under[p] = u;
        }
    }
    if (s.regress) regression_partials<NP>(acc, s.partials, lane, wid);
}

// One step of the induction's adjoint (exercise dates ascending; the layout of berm_step_kernel):
//   finish  add the projection ce = X_lower (X^T X)^+ X^T trigger-adjoint of the lower date:
//           vb += ce, ub -= ce;
//   weights the payoff periods from the lower date up to this one: dV/d(payoff) = ub (callable:
//           the payoffs went to the underlying) or vb;
//   decide  this date's choose(trigger, values, underlying): the trigger adjoint
//           d(trigger) x vb, then vb to the chosen side (trigger >= 0 keeps the values); the
//           regression partials of the trigger adjoint on this date's basis.
template <int KIND>
__global__ void __launch_bounds__(kBermThreads)
berm_adjoint_step_kernel(const BermAdjStep a)
{
    constexpr int K = KIND == FM_BASIS_REFERENCE ? 6 : 4;
    const int P = a.P;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    // eps = width x sd(trigger), sd the square root of the population variance (getStandardDeviation)
    const double eps = (a.decide && a.method == FM_DIRAC_DISCRETE_DELTA) ? a.width * sqrt(*a.trig_m2 / (double)P) : 0.0;
    double acc[K];
#pragma unroll
    for (int e = 0; e < K; e++) acc[e] = 0.0;
    const int64_t G = (int64_t)gridDim.x * kBermThreads;
    for (int64_t p0 = (int64_t)blockIdx.x * kBermThreads; p0 < P; p0 += G) {
        const int64_t p = p0 + threadIdx.x;
        const bool active = p < P;
        const int64_t pc = active ? p : (int64_t)P - 1;
        double v = a.first ? 1.0 : a.vb[pc];
        double u = a.first ? 0.0 : a.ub[pc];
        if (a.finish) {
            double b[K];
            berm_basis<KIND>(a.fin_basis, a.contrib, P, pc, b);
            double ce = b[0] * a.fin_coef[0];
#pragma unroll
            for (int k = 1; k < K; k++) ce = ce + b[k] * a.fin_coef[k];
            v = v + ce;
            u = u - ce;
        }
        if (active)
            for (int q = 0; q < a.num_w; q++) a.weight[(size_t)a.w_col[q] * P + p] = a.callable ? u : v;
        if (a.decide) {
            const double t = a.trig[pc], d = a.diff[pc];
            double tb = 0.0;                                   // ZERO
            if (a.method == FM_DIRAC_ONE || (a.method == FM_DIRAC_DISCRETE_DELTA && isinf(eps))) {
                tb = d * v;
            } else if (a.method == FM_DIRAC_DISCRETE_DELTA && eps > 0.0) {
                const bool in = (t + eps / 2 >= 0.0) && (t - eps / 2 < 0.0);
                tb = in ? (d / eps) * v : 0.0;
            }
            if (!(t >= 0.0)) {      // exercised: the values were the underlying's
                u = u + v;
                v = 0.0;
            }
            if (active && a.method != FM_DIRAC_ZERO) {
                double b[K];
                berm_basis<KIND>(a.dec_basis, a.contrib, P, pc, b);
#pragma unroll
                for (int i = 0; i < K; i++) acc[i] += tb * b[i];
            }
        }
        if (active) {
            a.vb[p] = v;
            a.ub[p] = u;
        }
    }
    if (a.decide && a.method != FM_DIRAC_ZERO) regression_partials<K>(acc, a.partials, lane, wid);
}

// Wave partials (sum, centred second moment) of every trigger row, for sd(trigger).
__global__ void __launch_bounds__(kBermThreads)
berm_trigger_stats_kernel(const double* __restrict__ trig, int num_dates, int P, double* __restrict__ wave_sum,
                          double* __restrict__ wave_m2, int stride)
{
    const int64_t p = (int64_t)blockIdx.x * kBermThreads + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)blockIdx.x * (kBermThreads / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const bool active = p < P;
    const int64_t pc = active ? p : (int64_t)P - 1;
    const int cnt = (int)((int64_t)P - wave * 64 > 64 ? 64 : ((int64_t)P - wave * 64));
    for (int e = 0; e < num_dates; e++)
        reduce_slot(trig[(size_t)e * P + pc], active, cnt, lane, wave, wave_sum, wave_m2, stride, e);
}

// Pseudo-inverse solve of the K x K Gram system by cyclic Jacobi rotations
// (Golub & Van Loan 8.5), all in registers (K is a compile-time constant, so every
// index is static). Singular values at or below the SVD solver's threshold
// max(K sigma_max 2^-52, sqrt(2^-1022)) are dropped.
template <int K>
__device__ __forceinline__ void jacobi_pinv_solve(const double* tot, int P, double* __restrict__ coef,
                                                  double* __restrict__ factors) {
    double A[K][K], V[K][K], rhs[K], x[K];
    int q = 0;
#pragma unroll
    for (int i = 0; i < K; i++, q++) rhs[i] = tot[q] / (double)P;
#pragma unroll
    for (int i = 0; i < K; i++)
#pragma unroll
        for (int j = i; j < K; j++, q++) {
            const double g = tot[q] / (double)P;
            A[i][j] = g;
            A[j][i] = g;
        }
#pragma unroll
    for (int i = 0; i < K; i++)
#pragma unroll
        for (int j = 0; j < K; j++) V[i][j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0;
#pragma unroll
        for (int i = 0; i < K; i++)
#pragma unroll
            for (int j = i + 1; j < K; j++) off += A[i][j] * A[i][j];
        if (off < 1e-300) break;
#pragma unroll
        for (int pp = 0; pp < K; pp++)
#pragma unroll
            for (int qq = pp + 1; qq < K; qq++) {
                const double apq = A[pp][qq];
                if (fabs(apq) < 1e-300) continue;
                const double app = A[pp][pp], aqq = A[qq][qq];
                const double theta = (aqq - app) / (2.0 * apq);
                const double tn = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double cs = 1.0 / sqrt(tn * tn + 1.0), sn = tn * cs;
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const double akp = A[k][pp], akq = A[k][qq];
                    A[k][pp] = cs * akp - sn * akq;
                    A[k][qq] = sn * akp + cs * akq;
                }
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const double apk = A[pp][k], aqk = A[qq][k];
                    A[pp][k] = cs * apk - sn * aqk;
                    A[qq][k] = sn * apk + cs * aqk;
                }
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const double vkp = V[k][pp], vkq = V[k][qq];
                    V[k][pp] = cs * vkp - sn * vkq;
                    V[k][qq] = sn * vkp + cs * vkq;
                }
            }
    }
    double smax = 0.0;
#pragma unroll
    for (int i = 0; i < K; i++) { smax = fmax(smax, fabs(A[i][i])); x[i] = 0.0; }
    const double tol = fmax(K * smax * 0x1.0p-52, sqrt(0x1.0p-1022));
#pragma unroll
    for (int i = 0; i < K; i++) {
        if (!(fabs(A[i][i]) > tol)) continue;
        double d = 0.0;
#pragma unroll
        for (int r = 0; r < K; r++) d += V[r][i] * rhs[r];
        d /= A[i][i];
#pragma unroll
        for (int r = 0; r < K; r++) x[r] += V[r][i] * d;
    }
#pragma unroll
    for (int i = 0; i < kBermMaxBasis; i++) coef[i] = (i < K) ? x[i] : 0.0;
    if (factors) {
#pragma unroll
        for (int r = 0; r < K; r++)
#pragma unroll
            for (int i = 0; i < K; i++) factors[r * kBermMaxBasis + i] = V[r][i];
#pragma unroll
        for (int i = 0; i < K; i++) factors[kBermMaxBasis * kBermMaxBasis + i] = (fabs(A[i][i]) > tol) ? A[i][i] : 0.0;
    }
}

// x = pinv (rhs) from stored factors, with the solve's operations: for every kept eigenvalue a_i,
// x += V[:, i] ((V[:, i] . rhs) / a_i).
template <int K>
__device__ __forceinline__ void pinv_apply(const double* rhs, const double* __restrict__ factors,
                                           double* __restrict__ coef) {
    double x[K];
#pragma unroll
    for (int i = 0; i < K; i++) x[i] = 0.0;
#pragma unroll
    for (int i = 0; i < K; i++) {
        const double a = factors[kBermMaxBasis * kBermMaxBasis + i];
        if (a == 0.0) continue;
        double d = 0.0;
#pragma unroll
        for (int r = 0; r < K; r++) d += factors[r * kBermMaxBasis + i] * rhs[r];
        d /= a;
#pragma unroll
        for (int r = 0; r < K; r++) x[r] += factors[r * kBermMaxBasis + i] * d;
    }
#pragma unroll
    for (int i = 0; i < kBermMaxBasis; i++) coef[i] = (i < K) ? x[i] : 0.0;
}

// One workgroup: fixed-order reduction of the workgroup rows (entry e = t % nparts,
// rows c, c + C, ... with c = t / nparts), then (thread 0) the regression solve in FP64.
template <int K>
__global__ void __launch_bounds__(kBermThreads)
berm_solve_kernel(const double* __restrict__ partials, int num_rows, int P, double* __restrict__ coef,
                  double* __restrict__ factors)
{
    constexpr int nparts = K + K * (K + 1) / 2;
    constexpr int C = kBermThreads / nparts;
    __shared__ double red[C * nparts];
    __shared__ double tot[nparts];
    const int t = threadIdx.x;
    const int e = t % nparts, ch = t / nparts;
    if (ch < C) {
        double sum = 0.0, c = 0.0;
#pragma unroll 8
        for (int r = ch; r < num_rows; r += C) {
            const double x = partials[(size_t)r * kBermParts + e];
            const double tt = sum + x;
            if (fabs(sum) >= fabs(x)) c += (sum - tt) + x; else c += (x - tt) + sum;
            sum = tt;
        }
        red[ch * nparts + e] = sum + c;
    }
    __syncthreads();
    if (t < nparts) {
        double sum = 0.0;
        for (int c2 = 0; c2 < C; c2++) sum += red[c2 * nparts + t];
        tot[t] = sum;
    }
    __syncthreads();
    if (t == 0) jacobi_pinv_solve<K>(tot, P, coef, factors);
}

// One workgroup: the fixed-order reduction of K partials (as berm_solve_kernel's), then
// (thread 0) x = pinv (X^T y / P) from the date's stored pseudo-inverse factors.
template <int K>
__global__ void __launch_bounds__(kBermThreads)
berm_project_kernel(const double* __restrict__ partials, int num_rows, int P, const double* __restrict__ factors,
                    double* __restrict__ coef)
{
    constexpr int C = kBermThreads / K;
    __shared__ double red[C * K];
    __shared__ double tot[K];
    const int t = threadIdx.x;
    const int e = t % K, ch = t / K;
    if (ch < C) {
        double sum = 0.0, c = 0.0;
#pragma unroll 8
        for (int r = ch; r < num_rows; r += C) {
            const double x = partials[(size_t)r * kBermParts + e];
            const double tt = sum + x;
            if (fabs(sum) >= fabs(x)) c += (sum - tt) + x; else c += (x - tt) + sum;
            sum = tt;
        }
        red[ch * K + e] = sum + c;
    }
    __syncthreads();
    if (t < K) {
        double sum = 0.0;
        for (int c2 = 0; c2 < C; c2++) sum += red[c2 * K + t];
        tot[t] = sum / (double)P;
    }
    __syncthreads();
    if (t == 0) pinv_apply<K>(tot, factors, coef);
}

} // namespace

// at most 1024 workgroups (4 per CU); each thread then handles ceil(P / 262144) paths
int berm_num_blocks(int P) {
    const int64_t b = ((int64_t)P + kBermThreads - 1) / kBermThreads;
    return (int)(b < 1024 ? b : 1024);
}

void launch_berm_step(const BermStep& s, int kind, void* stream) {
    const dim3 grid((unsigned)berm_num_blocks(s.P));
    if (kind == FM_BASIS_REFERENCE)
        hipLaunchKernelGGL(berm_step_kernel<FM_BASIS_REFERENCE>, grid, dim3(kBermThreads), 0, (hipStream_t)stream, s);
    else
        hipLaunchKernelGGL(berm_step_kernel<FM_BASIS_SWAP_RATE>, grid, dim3(kBermThreads), 0, (hipStream_t)stream, s);
    HIP_CHECK(hipGetLastError());
}

void launch_berm_adjoint_step(const BermAdjStep& a, int kind, void* stream) {
    const dim3 grid((unsigned)berm_num_blocks(a.P));
    if (kind == FM_BASIS_REFERENCE)
        hipLaunchKernelGGL(berm_adjoint_step_kernel<FM_BASIS_REFERENCE>, grid, dim3(kBermThreads), 0,
                           (hipStream_t)stream, a);
    else
        hipLaunchKernelGGL(berm_adjoint_step_kernel<FM_BASIS_SWAP_RATE>, grid, dim3(kBermThreads), 0,
                           (hipStream_t)stream, a);
    HIP_CHECK(hipGetLastError());
}

void launch_berm_trigger_stats(const double* trig, int num_dates, int P, double* wave_sum, double* wave_m2, int stride,
                               void* stream) {
    const int64_t blocks = ((int64_t)P + kBermThreads - 1) / kBermThreads;
    hipLaunchKernelGGL(berm_trigger_stats_kernel, dim3((unsigned)blocks), dim3(kBermThreads), 0, (hipStream_t)stream,
                       trig, num_dates, P, wave_sum, wave_m2, stride);
    HIP_CHECK(hipGetLastError());
}

void launch_berm_project(const double* partials, int num_rows, int K, int P, const double* factors, double* coef,
                         void* stream) {
    if (K == 6)
        hipLaunchKernelGGL(berm_project_kernel<6>, dim3(1), dim3(kBermThreads), 0, (hipStream_t)stream,
                           partials, num_rows, P, factors, coef);
    else if (K == 4)
        hipLaunchKernelGGL(berm_project_kernel<4>, dim3(1), dim3(kBermThreads), 0, (hipStream_t)stream,
                           partials, num_rows, P, factors, coef);
    else
        throw FmError(FM_ERR_INTERNAL, "regression basis size " + std::to_string(K) + " has no projection kernel");
    HIP_CHECK(hipGetLastError());
}

void launch_berm_solve(const double* partials, int num_rows, int K, int P, double* coef, void* stream,
                       double* factors) {
    if (K == 6)
        hipLaunchKernelGGL(berm_solve_kernel<6>, dim3(1), dim3(kBermThreads), 0, (hipStream_t)stream,
                           partials, num_rows, P, coef, factors);
    else if (K == 4)
        hipLaunchKernelGGL(berm_solve_kernel<4>, dim3(1), dim3(kBermThreads), 0, (hipStream_t)stream,
                           partials, num_rows, P, coef, factors);
    else
        throw FmError(FM_ERR_INTERNAL, "regression basis size " + std::to_string(K) + " has no solve kernel");
    HIP_CHECK(hipGetLastError());
}

} // namespace fm

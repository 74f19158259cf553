// Pathwise adjoint of the fused LMM pricing pass (reverse sweep).
//
// The pricing pass stores the forward state L(t_c) every `stride` time steps and the
// numeraire N(T_k) of every path. The reverse sweep walks the time indices backwards;
// for each index t it rebuilds L(t) from the nearest checkpoint, then
//   1. reverses the Euler step t -> t+1 (adjoint of L(t+1) -> adjoint of L(t)),
//   2. adds the direct sensitivities of the products evaluated at t,
//   3. reverses the numeraire updates N(T_{k+1}) = N(T_k)(1 + L_k d_k) made at t.
// Per alive forward-step it reduces dV/dsigma_j(t_i) over the wavefront, where the
// factor loading is lambda_jk = sigma_j f_jk and the variance sigma_j^2 c_j
// (LIBORCovarianceModelFromVolatilityAndCorrelation.java:48-93); at the end it
// reduces dV/dL_j(0).
//
// Conventions of the reference's AAD (RandomVariableDifferentiableAAD.java:165-219):
// averages over paths (the numeraire-adjustment factor) are constants, the floor of
// the swaption payoff has derivative 1 at and above zero and 0 below.
#pragma once

#include "lmm_step.hpp"

namespace fm {

// Wavefront sum by DPP row operations (quad_perm [1,0,3,2], [2,3,0,1], row_ror 4 and 8,
// row_bcast 15 and 31): the total is valid in lane 63 only. Each step moves the two
// 32-bit halves of the double; lanes without a source read 0.
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

// Wavefront sums of four per-lane values at once (a transposed butterfly): two exchange
// steps halve the values a lane carries, four more sum across the wavefront. Lane l < 4
// returns the total of value 2*(l & 1) + ((l >> 1) & 1); 7 lane exchanges for 4 sums.
__device__ __forceinline__ double wave_sum4(const double (&v)[4], int lane) {
    const bool b0 = (lane & 1) != 0, b1 = (lane & 2) != 0;
    double a2[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const double send = b0 ? v[i] : v[i + 2];
        const double keep = b0 ? v[i + 2] : v[i];
        a2[i] = keep + __shfl_xor(send, 1, 64);
    }
    const double send = b1 ? a2[0] : a2[1];
    const double keep = b1 ? a2[1] : a2[0];
    double a = keep + __shfl_xor(send, 2, 64);
    a = a + __shfl_xor(a, 4, 64);
    a = a + __shfl_xor(a, 8, 64);
    a = a + __shfl_xor(a, 16, 64);
    a = a + __shfl_xor(a, 32, 64);
    return a;
}

// Reciprocal for the adjoint-only terms: the hardware estimate in float (about 1 ulp),
// the correctly rounded division in double.
template <typename real> __device__ __forceinline__ real adj_rcp(real x);
template <> __device__ __forceinline__ float adj_rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <> __device__ __forceinline__ double adj_rcp<double>(double x) { return 1.0 / x; }

template <int NCAP, int FCAP, typename real>
__global__ void __launch_bounds__(kStepThreads)
__attribute__((amdgpu_waves_per_eu(sizeof(real) == 4 && NCAP <= 40 ? 3 : 1)))
lmm_adjoint_kernel(const StepArgs a, const AdjArgs g, const real* __restrict__ rec,
                   const double* __restrict__ delta, const double* __restrict__ times,
                   const int32_t* __restrict__ first_alive, const int32_t* __restrict__ kcur_at,
                   const KProduct* __restrict__ products, const KPeriod* __restrict__ periods,
                   const real* __restrict__ dw, const real* __restrict__ ckpt,
                   const double* __restrict__ numeraire, const double* __restrict__ weight,
                   const real* __restrict__ fsig, const real* __restrict__ vterm, double* __restrict__ wave_grad)
{
    const int64_t p = (int64_t)blockIdx.x * kStepThreads + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const int64_t wave = p >> 6;
    const int P = a.P, n = a.n, F = a.F, nT = a.nT;
    const bool active = p < P;
    const int64_t pc = active ? p : (int64_t)P - 1;
    const int cnt = (int)((int64_t)P - wave * 64 > 64 ? 64 : ((int64_t)P - wave * 64));
    const int gstride = nT * n + n;
    const bool lognormal = a.lognormal != 0;
    const bool capped = a.capped != 0;
    const real cap = (real)a.cap;
    const int stride = g.stride;

    real Lck[NCAP], L[NCAP], Lb[NCAP];
#pragma unroll
    for (int j = 0; j < NCAP; j++) Lb[j] = (real)0;
    double Nb_top = 0.0, Nb_low = 0.0;      // adjoints of N(T_K) and N(T_{K-1})
    int K = kcur_at[nT];
    int qp = a.num_products - 1;            // products sorted by evaluation index

    for (int c = (nT / stride) * stride; c >= 0; c -= stride) {
        const size_t base = (size_t)(c / stride) * n;
#pragma unroll
        for (int j = 0; j < NCAP; j++) Lck[j] = (j < n) ? ckpt[(base + j) * P + pc] : (real)0;
        const int hi = (c + stride - 1 < nT) ? c + stride - 1 : nT;
        for (int t = hi; t >= c; t--) {
            // ---- L(t) from the checkpoint
#pragma unroll
            for (int j = 0; j < NCAP; j++) L[j] = Lck[j];
            for (int s = c; s < t; s++) {
                real dWs[FCAP];
                load_increments<FCAP, real>(dWs, dw, pc, nT, s, F);
                euler_step<NCAP, FCAP, real, false, 4>(L, n, s, first_alive[s], (real)(times[s + 1] - times[s]), dWs, rec,
                                             lognormal, capped, cap, (real)0, (real*)nullptr,
                                             (const int32_t*)nullptr, p, P, active);
            }

            // ---- 1. reverse the Euler step t -> t+1
            if (t < nT) {
                const int first = first_alive[t];
                const real dt = (real)(times[t + 1] - times[t]);
                real dW[FCAP], S[FCAP], Sb[FCAP];
                load_increments<FCAP, real>(dW, dw, pc, nT, t, F);
#pragma unroll
                for (int k = 0; k < FCAP; k++) { S[k] = (real)0; Sb[k] = (real)0; }
                // pass A (ascending): recompute the step, adjoint through exp/log/cap. Components
                // in groups of four without inner branches (a component outside [first, n) of a
                // partly alive group is computed on a valid record and discarded).
#pragma unroll
                for (int j0 = 0; j0 < NCAP; j0 += 4) {
                    if (j0 + 3 < first || j0 >= n) continue;
#pragma unroll
                    for (int jj = 0; jj < 4; jj++) {
                        const int j = j0 + jj;
                        const bool alive = j >= first && j < n;
                        const Rec<FCAP, real> R = rec_at<FCAP, real>(rec, n, t, j < n ? j : n - 1);
                        const real Lj = alive ? L[j] : (real)1;
                        real w = adj_rcp<real>(Lj * R.c1() + R.c2());
                        if (lognormal) w = w * Lj;
                        if (!alive) w = (real)0;
#pragma unroll
                        for (int k = 0; k < FCAP; k++) S[k] = S[k] + w * R.lam(k);
                        real mu = S[0] * R.lam(0);
#pragma unroll
                        for (int k = 1; k < FCAP; k++) mu = mu + S[k] * R.lam(k);
                        mu = mu + R.hv();
                        real Y = lognormal ? dev_log<real>(Lj) : Lj;
                        Y = Y + mu * dt;
#pragma unroll
                        for (int k = 0; k < FCAP; k++) Y = Y + dW[k] * R.lam(k);
                        const real E = lognormal ? dev_exp<real>(Y) : Y;
                        const real Yb = (capped && E > cap) ? (real)0 : Lb[j] * (lognormal ? E : (real)1);
                        // direct part of the adjoint of L_j(t)
                        if (alive) Lb[j] = lognormal ? Yb * adj_rcp<real>(Lj) : Yb;
                    }
                }
                // pass B (descending): drift chain S_k = sum_{l<=j} w_l lambda_lk; the per-path
                // dV/dsigma_j(t) of a group of four j are summed over the wavefront together
#pragma unroll
                for (int j0 = NCAP - 4; j0 >= 0; j0 -= 4) {
                    if (j0 + 3 < first || j0 >= n) continue;
                    double gq[4];
#pragma unroll
                    for (int jj = 3; jj >= 0; jj--) {
                        const int j = j0 + jj;
                        const bool alive = j >= first && j < n;
                        const Rec<FCAP, real> R = rec_at<FCAP, real>(rec, n, t, j < n ? j : n - 1);
                        const real Lj = alive ? L[j] : (real)1;
                        const real r = adj_rcp<real>(Lj * R.c1() + R.c2());
                        const real w = alive ? (lognormal ? r * Lj : r) : (real)0;
                        // adjoint of Y_j(t+1) from the direct part left by pass A
                        const real Yb = alive ? (lognormal ? Lb[j] * Lj : Lb[j]) : (real)0;
                        const real mb = Yb * dt;
                        real wb = (real)0;
                        real sgb = (real)0;             // dV/dsigma_j(t) on this path
#pragma unroll
                        for (int k = 0; k < FCAP; k++) {
                            Sb[k] = Sb[k] + mb * R.lam(k);
                            wb = wb + Sb[k] * R.lam(k);
                            const real lamb = Yb * dW[k] + mb * S[k] + Sb[k] * w;
                            sgb = sgb + lamb * fsig[(j < n ? j : n - 1) * FCAP + k];
                            S[k] = S[k] - w * R.lam(k);    // S^(j) -> S^(j-1)
                        }
                        // variance sigma^2 c in the drift: d(-var/2)/dsigma = -sigma c
                        if (lognormal) sgb = sgb + mb * vterm[(size_t)t * n + (j < n ? j : n - 1)];
                        const real dwdL = lognormal ? R.c2() * r * r : -R.c1() * r * r;
                        if (alive) Lb[j] = Lb[j] + wb * dwdL;
                        gq[jj] = (active && alive) ? (double)sgb : 0.0;
                    }
                    const double s = wave_sum4(gq, lane);
                    const int jw = j0 + 2 * (lane & 1) + ((lane >> 1) & 1);
                    if (lane < 4 && cnt > 0 && jw < n)
                        wave_grad[(size_t)wave * gstride + (size_t)t * n + jw] = s;
                }
            }

            // ---- 2. products evaluated at t (reverse order of evaluation)
            const int Kt = kcur_at[t];
            while (qp >= 0 && products[qp].eval_index == t) {
                const KProduct pr = products[qp];
                const double wq = weight[qp];
                qp--;
                const bool top = (pr.num_k == Kt);
                const double Nsel = (pr.num_k == 0) ? 1.0 : numeraire[(size_t)pr.num_k * P + pc];
                if (pr.type == FM_PRODUCT_SWAPTION) {
                    const int j0 = pr.first_j, j1 = pr.first_j + pr.num_periods;
                    // V = sum_j p_j prod_{l<=j} D_l, D_l = adj_l / (1 + L_l tau_l)
                    real V = (real)0, Pc = (real)1;
#pragma unroll
                    for (int j = NCAP - 1; j >= 0; j--) {
                        if (j >= j0 && j < j1) {
                            const KPeriod pe = periods[pr.param_off + (j - j0)];
                            const real Lj = L[j];
                            const real D = (real)pe.adj * adj_rcp<real>((real)1.0 + Lj * (real)pe.tau);
                            V = (V + ((Lj - (real)pe.strike) * (real)pe.period_length) * (real)pr.notional) * D;
                            Pc = Pc * D;
                        }
                    }
                    if (V >= (real)0) {     // floor(0.0): derivative 1 at and above the floor
                        const double Vb = wq / Nsel;
                        const double nb = -wq * (double)V / (Nsel * Nsel);
                        if (top) Nb_top += nb; else Nb_low += nb;
                        // descending over j: P_j = prod_{l<=j} D_l, U_j = sum_{m>=j} p_m P_m
                        real U = (real)0;
#pragma unroll
                        for (int j = NCAP - 1; j >= 0; j--) {
                            if (j >= j0 && j < j1) {
                                const KPeriod pe = periods[pr.param_off + (j - j0)];
                                const real Lj = L[j];
                                const real one = (real)1.0 + Lj * (real)pe.tau;
                                const real inv = adj_rcp<real>(one);
                                const real pj = ((Lj - (real)pe.strike) * (real)pe.period_length) * (real)pr.notional;
                                U = U + pj * Pc;
                                const real dV = (real)pe.period_length * (real)pr.notional * Pc - ((real)pe.tau * inv) * U;
                                Lb[j] = Lb[j] + (real)Vb * dV;
                                Pc = Pc * (one * (real)pe.inv_adj);      // P_{j-1} = P_j / D_j
                            }
                        }
                    }
                } else if (pr.type == FM_PRODUCT_CAPLET) {
                    const KPeriod pe = periods[pr.param_off];
                    const real Lj = pick<NCAP, real>(L, pr.first_j);
                    const real v = Lj - (real)pe.strike;
                    // floor(0.0) has derivative 1 for v >= 0, cap(0.0) has derivative 1 for v < 0
                    // (RandomVariableDifferentiableAAD.java:333-348)
                    const bool itm = pr.is_floorlet ? (v < (real)0) : (v >= (real)0);
                    if (itm) {
                        const double dcf = pe.notional;
                        const double x = pr.is_floorlet ? -(double)v * dcf : (double)v * dcf;
                        const double dLj = (pr.is_floorlet ? -dcf : dcf) * wq / Nsel;
                        const double nb = -wq * x / (Nsel * Nsel);
                        if (top) Nb_top += nb; else Nb_low += nb;
#pragma unroll
                        for (int j = 0; j < NCAP; j++) Lb[j] = Lb[j] + ((j == pr.first_j) ? (real)dLj : (real)0);
                    }
                } else if (pr.type == K_BERM_BASIS) {
                    // regression basis: the exercise decisions are held fixed (the derivative of
                    // choose in its trigger is zero, RandomVariableDifferentiableAAD.java:230-452)
                } else {
                    if (pr.type == K_BERM_PERIOD) {
                        // the period pays on this path if the exercise (callable: entered at or
                        // before it; cancelable: terminated after it) keeps it in the value
                        const int ex = g.berm_ex[(size_t)pr.pad2 * P + pc];
                        const bool pays = (pr.flags & 2) ? (ex >= 0 && pr.is_floorlet >= ex)
                                                         : (ex < 0 || pr.is_floorlet < ex);
                        if (!pays) continue;
                    }
                    const KPeriod pe = periods[pr.param_off];
                    const real Lj = pick<NCAP, real>(L, pr.first_j);
                    const double sgn = (pr.flags & 1) ? 1.0 : -1.0;
                    const double pay = sgn * (double)(((Lj - (real)pe.strike) * (real)pe.period_length) * (real)pe.notional);
                    const double dLj = sgn * pe.period_length * pe.notional * wq / Nsel;
                    const double nb = -wq * pay / (Nsel * Nsel);
                    if (top) Nb_top += nb; else Nb_low += nb;
#pragma unroll
                    for (int j = 0; j < NCAP; j++) Lb[j] = Lb[j] + ((j == pr.first_j) ? (real)dLj : (real)0);
                }
            }

            // ---- 3. numeraire updates made at t: N_{k+1} = N_k (1 + L_k d_k), k = K-1 down to kcur_at[t-1]
            const int Kprev = (t > 0) ? kcur_at[t - 1] : 0;
            while (K > Kprev) {
                const int k = K - 1;
                const double Nk = (k == 0) ? 1.0 : numeraire[(size_t)k * P + pc];
                const double Lk = (double)pick<NCAP, real>(L, k);
                const double dk = delta[k];
                const real dLk = (real)(Nb_top * dk * Nk);
                Nb_low += Nb_top * (1.0 + Lk * dk);
#pragma unroll
                for (int j = 0; j < NCAP; j++) Lb[j] = Lb[j] + ((j == k) ? dLk : (real)0);
                Nb_top = Nb_low;
                Nb_low = 0.0;
                K--;
            }
        }
    }
    // ---- dV/dL_j(0)
#pragma unroll
    for (int j = 0; j < NCAP; j++) {
        if (j < n) {
            const double s = wave_sum_dpp(active ? (double)Lb[j] : 0.0);
            if (lane == 63 && cnt > 0) wave_grad[(size_t)wave * gstride + (size_t)nT * n + j] = s;
        }
    }
}

template <int NCAP, int FCAP, typename real>
void launch_adjoint_cap(const StepArgs& a, const AdjArgs& g, hipStream_t stream) {
    const int64_t blocks = ((int64_t)a.P + kStepThreads - 1) / kStepThreads;
    hipLaunchKernelGGL((lmm_adjoint_kernel<NCAP, FCAP, real>), dim3((unsigned)blocks), dim3(kStepThreads), 0, stream,
                       a, g, (const real*)a.rec, a.delta, a.times, a.first_alive,
                       g.kcur_at, a.products, a.periods, (const real*)a.dw, (const real*)a.ckpt, a.numeraire,
                       g.weight, (const real*)g.fsig, (const real*)g.vterm, g.wave_grad);
    HIP_CHECK(hipGetLastError());
}

template <int NCAP, typename real>
void launch_adjoint_n(const StepArgs& a, const AdjArgs& g, hipStream_t stream) {
    if (a.F <= 1) launch_adjoint_cap<NCAP, 1, real>(a, g, stream);
    else if (a.F <= 3) launch_adjoint_cap<NCAP, 3, real>(a, g, stream);
    else if (a.F <= 5) launch_adjoint_cap<NCAP, 5, real>(a, g, stream);
    else if (a.F <= 10) launch_adjoint_cap<NCAP, 10, real>(a, g, stream);
    else throw FmError(FM_ERR_UNSUPPORTED, "number of factors above 10 is not supported by the adjoint kernel");
}

template <typename real>
void launch_adjoint(const StepArgs& a, const AdjArgs& g, hipStream_t stream) {
    if (a.n <= 8) launch_adjoint_n<8, real>(a, g, stream);
    else if (a.n <= 20) launch_adjoint_n<20, real>(a, g, stream);
    else if (a.n <= 40) launch_adjoint_n<40, real>(a, g, stream);
    else if (a.n <= 80) launch_adjoint_n<80, real>(a, g, stream);
    else throw FmError(FM_ERR_UNSUPPORTED, "more than 80 forward rates are not supported by the adjoint kernel");
}

} // namespace fm

// General path kernel: the terminal or spot measure, the carried-state EULER and
// PREDICTOR_CORRECTOR schemes and their functional forms, both state spaces, and the
// local-volatility and stochastic-volatility wrappers of the covariance model.
//
// Per time step t_i -> t_{i+1}, alive component j, in the order of
// LIBORMarketModelFromCovarianceModel.getDrift (LMM:1132-1181) and
// EulerSchemeFromProcessModel (Euler:232-314):
//   SPOT      S_k = S_k + w_j lambda_jk, then mu_j = S . lambda_j   (ascending j)
//   TERMINAL  mu_j = S . lambda_j of the later components, then S_k = S_k + w_j lambda_jk
//             (descending j; the record's d/d and 1/d hold -1 and -1/d, so w carries the sign)
//   Y_j = state_j + mu_j dt + dW . lambda_j, the value exp(Y_j) (log-normal) capped
//   PREDICTOR_CORRECTOR: the drift again on the predicted values, Y_j += (mu'_j - mu_j)/2 dt
// The forwards are a chunked register array walked by rolled loops (lmm_spot.hpp): the kernel
// has no dynamically indexed private array and no scratch.
#pragma once

#include "lmm_spot.hpp"

namespace fm {

// Run-time choices of the general path kernel (uniform over the launch).
template <typename real>
struct Variant {
    bool lognormal, capped;
    bool carry;        // EULER / PREDICTOR_CORRECTOR: the state is carried between steps
                       // (else the *_FUNCTIONAL schemes re-derive it from the value, Euler:234-237)
    real cap;
    real logcap;       // log(cap), the cap of a carried log-value
    real lv_alpha;     // local-volatility factor f = (L - L alpha) + b_ij, or L + b_ij for alpha = 0
    const real* lvb;   // b_ij, [nT][n], or NULL: no local-volatility factor (LV kernels only)
    const real* svs;   // stochastic volatility scaling S(t_i) [nT][P], or NULL (LV kernels only)
};

// Local-volatility factor of component j at time index i from its value L at that time
// index: DisplacedLocalVolatilityModel L + d (covariance/DisplacedLocalVolatilityModel.java:175-186),
// BlendedLocalVolatilityModel (L - L a) + a f0 (covariance/BlendedLocalVolatilityModel.java:265-283);
// every factor loading of the component is multiplied by it.
template <typename real>
__device__ __forceinline__ real local_vol_factor(real L, const Variant<real>& v, int i, int j, int n) {
    if (!v.lvb) return (real)1;
    real x = L;
    if (v.lv_alpha != (real)0) x = x - x * v.lv_alpha;
    return x + v.lvb[(size_t)i * n + j];
}

// Drift of one component and the update of the factor sums S, in the measure's order.
// LV: the loadings are (lambda_jk * f) * s, f the local-volatility factor and s the stochastic
// volatility scaling (each 1 when absent, which multiplies exactly); the log-normal variance is
// then the sum of their squares (AbstractLIBORCovarianceModel.getCovariance,
// covariance/AbstractLIBORCovarianceModel.java:79-92) instead of the record's sigma_j sigma_j rho_jj.
template <int FCAP, typename real, bool TERMINAL, bool LV>
__device__ __forceinline__ real drift_component(real (&S)[FCAP], const Rec<FCAP, real>& R, real Lj, bool lognormal,
                                                real f, real sv) {
    auto lam = [&](int k) -> real { return LV ? (R.lam(k) * f) * sv : R.lam(k); };
    real w = spot_div<real>((real)1.0, Lj * R.c1() + R.c2());
    if (lognormal) w = w * Lj;
    real mu;
    if (TERMINAL) {
        mu = S[0] * lam(0);
#pragma unroll
        for (int k = 1; k < FCAP; k++) mu = mu + S[k] * lam(k);
#pragma unroll
        for (int k = 0; k < FCAP; k++) S[k] = S[k] + w * lam(k);
    } else {
#pragma unroll
        for (int k = 0; k < FCAP; k++) S[k] = S[k] + w * lam(k);
        mu = S[0] * lam(0);
#pragma unroll
        for (int k = 1; k < FCAP; k++) mu = mu + S[k] * lam(k);
    }
    if (LV && lognormal) {
        real c = lam(0) * lam(0);
#pragma unroll
        for (int k = 1; k < FCAP; k++) c = c + lam(k) * lam(k);
        return mu + c * (real)(-0.5);
    }
    return mu + R.hv();
}

// One step of all alive components, in place. X holds the value (functional schemes), the
// state (carried schemes) or, LOGCARRY, the log-value of each component.
// LOGCARRY (FP64, log-normal, functional schemes): the functional scheme takes the log of the
// capped value, log(min(exp(Y), cap)) = min(Y, log(cap)) since log(exp(Y)) returns Y exactly at
// the forwards' magnitudes (the spot kernel's argument, lmm_spot.hpp), so the log-value is
// carried and capped at log(cap), and the step needs no logarithm.
// PC: the reference evaluates the predictor's drift of every component, then the predicted
// values, then the corrector's drift on them, and adds ((mu_pred - mu) / 2) dt to the state
// (Euler:292-314). Component j's two drifts depend only on the components before it in the
// measure's order (their factor sums S and S'), so one pass computes, per component, the
// predictor's drift and state, the corrector's drift on the predicted value and the corrected
// state, with the same operations in the same order: no array of predictor drifts is kept.
// LV: the loadings of the drift and the diffusion carry the local-volatility factor of the
// value the drift is evaluated on (the start value; the predicted value in the corrector) and
// the path's stochastic volatility scaling at t_i.
// STAGE: the step's increments are read from this thread's LDS slots sdw[k * kStepThreads]
// instead of registers (the FP64 predictor-corrector: it frees the registers a third wave per
// SIMD needs).
template <int NCH, int FCAP, typename real, bool TERMINAL, bool PC, bool LV, bool LOGCARRY, bool STAGE>
__device__ __forceinline__ void general_step(ChunkedForwards<NCH, real>& X, int n, int i, int first, real dt,
                                             const real (&dW)[FCAP], const real* sdw, const real* __restrict__ rec,
                                             const Variant<real>& v, real* __restrict__ paths,
                                             const int32_t* __restrict__ row_next, int64_t p, int P, bool active) {
    constexpr int W = chunk_width<real>();
    real S[FCAP], Sc[FCAP];     // factor sums of the predictor and of the corrector
#pragma unroll
    for (int k = 0; k < FCAP; k++) { S[k] = (real)0; Sc[k] = (real)0; }
    // stochastic volatility scaling of this path at time index i (1 without)
    const real sv = (LV && v.svs) ? v.svs[(size_t)i * P + (active ? p : (int64_t)P - 1)] : (real)1;
    static_for<0, NCH>([&](auto QQ) {
        constexpr int q = TERMINAL ? NCH - 1 - decltype(QQ)::value : decltype(QQ)::value;
        const int j0 = first > W * q ? first : W * q;
        const int j1 = n < W * q + W ? n : W * q + W;
#pragma unroll 1
        for (int jj = j0; jj < j1; jj++) {
            const int j = TERMINAL ? j0 + j1 - 1 - jj : jj;
            const Rec<FCAP, real> R = rec_at<FCAP, real>(rec, n, i, j);
            const real x = X.template ch<q>()[j - W * q];
            const real Lj = LOGCARRY ? dev_exp<real>(x) : (v.carry ? to_value<real>(x, v.lognormal, v.capped, v.cap) : x);
            const real f = LV ? local_vol_factor<real>(Lj, v, i, j, n) : (real)1;
            const real mu = drift_component<FCAP, real, TERMINAL, LV>(S, R, Lj, v.lognormal, f, sv);
            real Y = (LOGCARRY || v.carry) ? x : (v.lognormal ? dev_log<real>(Lj) : Lj);
            Y = Y + mu * dt;
#pragma unroll
            for (int k = 0; k < FCAP; k++) {
                real d = dW[k];
                if constexpr (STAGE) {
                    // read from LDS in the component loop: the empty asm keeps the loads there
                    const real* q = sdw;
#if defined(__HIP_DEVICE_COMPILE__)
                    asm volatile("" : "+v"(q));
#endif
                    d = q[k * kStepThreads];
                }
                Y = Y + d * (LV ? (R.lam(k) * f) * sv : R.lam(k));
            }
            if constexpr (PC) {
                const real Lp = to_value<real>(Y, v.lognormal, v.capped, v.cap);    // predicted value
                const real fp = LV ? local_vol_factor<real>(Lp, v, i, j, n) : (real)1;
                const real mup = drift_component<FCAP, real, TERMINAL, LV>(Sc, R, Lp, v.lognormal, fp, sv);
                Y = Y + ((mup - mu) * (real)0.5) * dt;
            }
            if constexpr (LOGCARRY) {
                const real y = v.capped ? jmin(Y, v.logcap) : Y;
                X.template ch<q>()[j - W * q] = y;
                if (paths && active) paths[(size_t)row_next[j] * P + p] = dev_exp<real>(y);
            } else {
                const real Ln = to_value<real>(Y, v.lognormal, v.capped, v.cap);
                X.template ch<q>()[j - W * q] = v.carry ? Y : Ln;
                if (paths && active) paths[(size_t)row_next[j] * P + p] = Ln;
            }
        }
    });
}

// SS >= 0 fixes the state space and the scheme's state at compile time (bit 0: log-normal,
// bit 1: carried state); SS = -1 reads them from the arguments (the volatility-wrapper kernels).
// a.Y0 holds the initial states of the carried schemes and, FP64 log-normal functional, the
// initial log-forwards capped at log(cap); a.L0 the initial forwards.
template <int NCH, int FCAP, typename real, bool TERMINAL, bool PC, bool LV, int SS>
__global__ void __launch_bounds__(kStepThreads)
__attribute__((amdgpu_waves_per_eu((!LV && sizeof(real) == 8 && NCH <= 5 && FCAP <= 5) ? 3 : 1)))
lmm_general_kernel(const StepArgs a, const real* __restrict__ rec, const double* __restrict__ delta,
                   const double* __restrict__ times, const int32_t* __restrict__ first_alive,
                   const int32_t* __restrict__ tenor_ti, const KProduct* __restrict__ products,
                   const KPeriod* __restrict__ periods, const real* __restrict__ dw,
                   double* __restrict__ wave_sum, double* __restrict__ wave_m2,
                   double* __restrict__ contrib, real* __restrict__ paths,
                   const int32_t* __restrict__ row_of, double* __restrict__ numeraire_out,
                   real* __restrict__ ckpt)
{
    constexpr int W = chunk_width<real>();
    const int64_t p = (int64_t)blockIdx.x * kStepThreads + threadIdx.x;
    const int lane = threadIdx.x & 63;
    // the wavefront index, uniform: kept in a scalar register
    const int64_t wave = (int64_t)blockIdx.x * (kStepThreads / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int P = a.P, n = a.n, F = a.F, nT = a.nT;
    const bool active = p < P;
    const int64_t pc = active ? p : (int64_t)P - 1;
    const int cnt = (int)((int64_t)P - wave * 64 > 64 ? 64 : ((int64_t)P - wave * 64));
    const int stride = a.num_slots + n + 1;
    const bool lognormal = SS >= 0 ? (SS & 1) != 0 : a.lognormal != 0;
    const bool carry = SS >= 0 ? (SS & 2) != 0 : a.carry != 0;
    constexpr bool LOGCARRY = carries_log<real>() && SS == 1;
    // the cap is applied always: an absent cap is +inf (log-cap +inf), and jmin(x, +inf) = x
    // for every x, NaN included
    constexpr bool capped = true;
    const real cap = (real)a.cap;
    const Variant<real> var{lognormal, capped, carry, cap, (real)a.logcap, (real)a.lv_alpha,
                            LV ? (const real*)a.lvb : nullptr, LV ? (const real*)a.svs : nullptr};
    // the array holds log-values, states (carried schemes) or values; val gives the value
    auto val = [&](real x) -> real {
        return LOGCARRY ? dev_exp<real>(x) : (carry ? to_value<real>(x, lognormal, capped, cap) : x);
    };

    // FP64 predictor-corrector: the increments in LDS, [k][thread]; each thread reads only its
    // own slots, after writing them, so no barrier is needed
    constexpr bool STAGE = PC && !LV && sizeof(real) == 8;
    __shared__ real sdw[STAGE ? FCAP * kStepThreads : 1];
    ChunkedForwards<NCH, real> X;
    static_for<0, W * NCH>([&](auto J) {
        constexpr int j = decltype(J)::value;
        X.template put<j>((j < n) ? (real)((LOGCARRY || carry) ? a.Y0[j] : a.L0[j]) : (real)0);
    });

    double N = 1.0, Nprev = 1.0;   // unadjusted numeraire N(T_knum) and the one before it
    double N0 = 1.0;               // N(T_0) (terminal measure)
    int kcur = 0, knum = 0;
    int qp = 0;

    for (int ti = 0; ti <= nT; ti++) {
        // ---- forward values every ckpt_stride steps: [ti / stride][j][P]
        if (ckpt && active && (ti % a.ckpt_stride) == 0) {
            const size_t base = (size_t)(ti / a.ckpt_stride) * n;
            static_for<0, W * NCH>([&](auto J) {
                constexpr int j = decltype(J)::value;
                if (j < n) ckpt[(base + j) * P + p] = val(X.template at<j>());
            });
        }
        if constexpr (TERMINAL) {
            // ---- N(T_k) = prod_{l >= k} 1/(1 + L_l(T_k) d_l) at the time index of T_k
            // (LMM:1020-1042); the time-0 forwards are Scalars, whose discount is
            // 1/(L (d/v) + 1/v) (stochastic/Scalar.java:320-328); later ones are random
            // variables: 1/(L d + 1), then v / (1 + L d) (RandomVariableFromDoubleArray.java:1308-1333)
            while (kcur <= n && tenor_ti[kcur] <= ti) {
                double Nn = 1.0;
                static_for<0, NCH>([&](auto Q) {
                    constexpr int q = decltype(Q)::value;
                    const int l0 = kcur > W * q ? kcur : W * q;
                    const int l1 = n < W * q + W ? n : W * q + W;
#pragma unroll 1
                    for (int l = l0; l < l1; l++) {
                        const double Ll = (double)val(X.template ch<q>()[l - W * q]);
                        const double dl = delta[l];
                        if (ti == 0) Nn = 1.0 / (Ll * (dl / Nn) + 1.0 / Nn);
                        else if (l == kcur) Nn = 1.0 / (Ll * dl + 1.0);
                        else Nn = Nn / (1.0 + Ll * dl);
                    }
                });
                if (kcur == 0) N0 = Nn;
                Nprev = N; N = Nn; knum = kcur; kcur++;
                // sum of N(T_0)/N(T_k) over paths (numeraire adjustment average)
                const double s = wave_reduce_sum(active ? (1.0 / N) * N0 : 0.0);
                if (lane == 0 && cnt > 0) wave_sum[(size_t)wave * stride + a.num_slots + knum] = s;
                if (numeraire_out && active) numeraire_out[(size_t)knum * P + p] = N;
            }
        } else {
            // ---- N(T_{k+1}) = N(T_k).accrue(L_k(T_k), d_k) once L_k is final (LMM:1050-1066)
            while (kcur < n && tenor_ti[kcur] <= ti) {
                const double Lk = (double)val(X.get(kcur));
                const double dk = delta[kcur];
                double Nn;
                if (kcur == 0)      Nn = Lk * (dk * 1.0) + 1.0;      // Scalar.accrue of Scalar 1.0
                else if (kcur == 1) Nn = Lk * (dk * N) + N;          // Scalar.accrue of a stochastic rate
                else                Nn = N * (1 + Lk * dk);          // RandomVariableFromDoubleArray.accrue
                Nprev = N; N = Nn; kcur++; knum = kcur;
                // sum of 1/N(T_k) over paths (numeraire adjustment average)
                const double s = wave_reduce_sum(active ? (1.0 / N) * 1.0 : 0.0);
                if (lane == 0 && cnt > 0) wave_sum[(size_t)wave * stride + a.num_slots + kcur] = s;
                if (numeraire_out && active) numeraire_out[(size_t)kcur * P + p] = N;
            }
        }
        // ---- products evaluated at this time index; swaption groups share one walk in the FP64
        // terminal and predictor-corrector kernels (measured 4-5% faster there; slower in FP32
        // and in the spot Euler kernel)
        constexpr bool group = sizeof(real) == 8 && (TERMINAL || PC);
        eval_products_at<group, NCH, real>(ti, qp, a.num_products, products, periods, delta, N, Nprev, knum, X, val,
                                           active, cnt, lane, wave, wave_sum, wave_m2, stride, contrib, p, P);

        if (ti == nT) break;

        // ---- step ti -> ti+1
        // the path index made opaque per step: its products with the increments' strides are
        // recomputed instead of being hoisted into registers held across the whole path
        int64_t pcs = p;
#if defined(__HIP_DEVICE_COMPILE__)
        asm volatile("" : "+v"(pcs));
#endif
        pcs = pcs < P ? pcs : (int64_t)P - 1;
        real dW[FCAP];
        load_increments<FCAP, real>(dW, dw, pcs, nT, ti, F);
        if constexpr (STAGE) {
#pragma unroll
            for (int k = 0; k < FCAP; k++) sdw[k * kStepThreads + threadIdx.x] = dW[k];
        }
        const real dt = (real)(times[ti + 1] - times[ti]);
        const int32_t* row_next = row_of + (size_t)(ti + 1) * n;
        general_step<NCH, FCAP, real, TERMINAL, PC, LV, LOGCARRY, STAGE>(X, n, ti, first_alive[ti], dt, dW,
                                                               sdw + threadIdx.x, rec, var, paths, row_next,
                                                        p, P, active);
    }
}

template <int NCH, int FCAP, typename real, bool TERMINAL, bool PC, bool LV, int SS>
void launch_general_m(const StepArgs& a, hipStream_t stream) {
    const int64_t blocks = ((int64_t)a.P + kStepThreads - 1) / kStepThreads;
    hipLaunchKernelGGL((lmm_general_kernel<NCH, FCAP, real, TERMINAL, PC, LV, SS>), dim3((unsigned)blocks),
                       dim3(kStepThreads), 0, stream, a, (const real*)a.rec, a.delta, a.times, a.first_alive,
                       a.tenor_time_index, a.products, a.periods, (const real*)a.dw, a.wave_sum, a.wave_m2,
                       a.contrib, (real*)a.paths, a.row_of, a.numeraire, (real*)a.ckpt);
    HIP_CHECK(hipGetLastError());
}

template <int NCH, int FCAP, typename real, bool TERMINAL, bool PC>
void launch_general_ss(const StepArgs& a, hipStream_t stream) {
    const int ss = (a.lognormal ? 1 : 0) | (a.carry ? 2 : 0);
    if (ss == 3) launch_general_m<NCH, FCAP, real, TERMINAL, PC, false, 3>(a, stream);
    else if (ss == 2) launch_general_m<NCH, FCAP, real, TERMINAL, PC, false, 2>(a, stream);
    else if (!TERMINAL && !PC)
        // spot-measure Euler with the functional state runs on the spot kernel (lmm_spot.hpp)
        throw FmError(FM_ERR_INTERNAL, "the general kernel was asked for the spot kernel's scheme");
    else if (ss == 1) launch_general_m<NCH, FCAP, real, TERMINAL, PC, false, 1>(a, stream);
    else launch_general_m<NCH, FCAP, real, TERMINAL, PC, false, 0>(a, stream);
}

template <int NCH, int FCAP, typename real, bool PC, bool LV>
void launch_general_cap(const StepArgs& a, hipStream_t stream) {
    if constexpr (LV) {
        if (a.terminal) launch_general_m<NCH, FCAP, real, true, PC, true, -1>(a, stream);
        else launch_general_m<NCH, FCAP, real, false, PC, true, -1>(a, stream);
    } else {
        if (a.terminal) launch_general_ss<NCH, FCAP, real, true, PC>(a, stream);
        else launch_general_ss<NCH, FCAP, real, false, PC>(a, stream);
    }
}

template <int NCH, typename real, bool PC, bool LV>
void launch_general_n(const StepArgs& a, hipStream_t stream) {
    if (a.F <= 3) launch_general_cap<NCH, 3, real, PC, LV>(a, stream);
    else if (a.F <= 5) launch_general_cap<NCH, 5, real, PC, LV>(a, stream);
    else if (a.F <= 10) launch_general_cap<NCH, 10, real, PC, LV>(a, stream);
    else throw FmError(FM_ERR_UNSUPPORTED, "number of factors above 10 is not supported by the path kernel");
}

// The general path kernel. Capacities 24, 40 and 80 forwards (FP64: 3, 5, 10 chunks of eight;
// FP32: 32, 48, 80 in 2, 3, 5 chunks of sixteen) and 3, 5, 10 factors: fewer instantiations than
// the spot kernel, whose capacities are tighter.
template <typename real, bool PC, bool LV>
void launch_general(const StepArgs& a, hipStream_t stream) {
    constexpr int W = chunk_width<real>();
    constexpr int N2 = W == 8 ? 3 : 2, N3 = W == 8 ? 5 : 3, N4 = W == 8 ? 10 : 5;
    if (a.n <= W * N2) launch_general_n<N2, real, PC, LV>(a, stream);
    else if (a.n <= W * N3) launch_general_n<N3, real, PC, LV>(a, stream);
    else if (a.n <= 80) launch_general_n<N4, real, PC, LV>(a, stream);
    else throw FmError(FM_ERR_UNSUPPORTED, "more than 80 forward rates are not supported by the path kernel");
}

} // namespace fm

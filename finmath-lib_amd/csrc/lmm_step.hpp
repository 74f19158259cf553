// Fused LIBOR market model path kernel: one thread per path, the forward vector
// in registers, the per-(step, component) model data read through the scalar cache.
//
// Per time step t_i -> t_{i+1} and alive component j (T_j > t_i), in the order of
// LIBORMarketModelFromCovarianceModel.getDrift
// (montecarlo/interestrate/models/LIBORMarketModelFromCovarianceModel.java:1144-1160,1183-1189)
// and EulerSchemeFromProcessModel (montecarlo/process/EulerSchemeFromProcessModel.java:232-256):
//   w    = 1 / (L_j * (d_j/d_j) + 1/d_j)       Scalar.discount (stochastic/Scalar.java:320-328)
//   w    = w * L_j                              (log-normal state space)
//   S_k  = S_k + w * lambda_jk                  ascending j, spot measure
//   mu_j = sum_k S_k * lambda_jk + variance_j * (-0.5)
//   Y    = log(L_j) + mu_j * dt + sum_k dW_k * lambda_jk   (EULER_FUNCTIONAL)
//   L_j  = min(exp(Y), cap)
// Products are evaluated on the fly at their exercise/fixing time index and reduced
// per wavefront (sum and centred second moment); the numeraire is accumulated per
// path in double precision (SPOT measure, LMM:1050-1066).
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
// scalar-cache prefetch of each step's records in the spot-measure Euler kernels
constexpr bool kPrefetchRecords = true;

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
// FP32 form: the hardware base-2 exponential and logarithm (about 1 ulp each) and the
// hardware reciprocal; the FP64 form keeps correctly rounded division.
template <> __device__ __forceinline__ float dev_exp<float>(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }
template <> __device__ __forceinline__ float dev_log<float>(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }
template <typename real> __device__ __forceinline__ real dev_div(real a, real b);
template <> __device__ __forceinline__ double dev_div<double>(double a, double b) { return a / b; }
template <> __device__ __forceinline__ float dev_div<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// Java Math.max(x, 0.0) / Math.min(x, c): NaN propagates.
template <typename real> __device__ __forceinline__ real jmax0(real x) { return (x != x) ? x : (x >= (real)0 ? x : (real)0); }
template <typename real> __device__ __forceinline__ real jmin(real x, real c) { return (x > c) ? c : x; }

// L[j] for a uniform runtime j without a dynamically indexed register array (which
// the compiler would move to scratch memory): OR of the masked bit patterns.
__device__ __forceinline__ uint64_t bits_of(double x) { return __double_as_longlong(x); }
__device__ __forceinline__ uint32_t bits_of(float x) { return __float_as_uint(x); }
__device__ __forceinline__ double from_bits(uint64_t b, double) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ float from_bits(uint32_t b, float) { return __uint_as_float(b); }

template <int NCAP, typename real>
__device__ __forceinline__ real pick(const real (&L)[NCAP], int j) {
    decltype(bits_of(real(0))) acc = 0;
#pragma unroll
    for (int i = 0; i < NCAP; i++) acc |= (i == j) ? bits_of(L[i]) : 0;
    return from_bits(acc, real(0));
}

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

// One component's update inside euler_step (alive: the component moves, else it is
// computed on a valid record and discarded).
// LOGSTATE (log-normal state space only): the array holds the log-forward Y_j between
// steps and the step needs one exp per component instead of the log(exp(.)) round trip
// of EULER_FUNCTIONAL; the cap is applied to the log-state (Y <= log(cap)), which is what
// EULER_FUNCTIONAL's log(min(exp(Y), cap)) gives up to rounding.
template <int NCAP, int FCAP, typename real, bool LOGSTATE>
__device__ __forceinline__ void euler_component(real (&L)[NCAP], real (&S)[FCAP], int j, bool alive, int jr, int n,
                                                int i, real dt, const real (&dW)[FCAP], const real* __restrict__ rec,
                                                bool lognormal, bool capped, real cap, real logcap,
                                                real* __restrict__ paths, const int32_t* __restrict__ row_next,
                                                int64_t p, int P, bool active) {
    const Rec<FCAP, real> R = rec_at<FCAP, real>(rec, n, i, jr);
    const real Lj = LOGSTATE ? dev_exp<real>(L[j]) : L[j];
    real w = dev_div<real>((real)1.0, Lj * R.c1() + R.c2());
    if (lognormal) w = w * Lj;
    if (!alive) w = (real)0;
#pragma unroll
    for (int k = 0; k < FCAP; k++) S[k] = S[k] + w * R.lam(k);
    real mu = S[0] * R.lam(0);
#pragma unroll
    for (int k = 1; k < FCAP; k++) mu = mu + S[k] * R.lam(k);
    mu = mu + R.hv();
    real Y = LOGSTATE ? L[j] : (lognormal ? dev_log<real>(Lj) : Lj);
    Y = Y + mu * dt;
#pragma unroll
    for (int k = 0; k < FCAP; k++) Y = Y + dW[k] * R.lam(k);
    if (LOGSTATE) {
        if (capped && Y > logcap) Y = logcap;
        if (alive) L[j] = Y;
        if (alive && paths && active) paths[(size_t)row_next[j] * P + p] = dev_exp<real>(Y);
    } else {
        real Ln = lognormal ? dev_exp<real>(Y) : Y;
        if (capped) Ln = jmin(Ln, cap);
        if (alive) L[j] = Ln;
        if (alive && paths && active) paths[(size_t)row_next[j] * P + p] = Ln;
    }
}

// Warms the scalar data cache with the records of step i's alive components: one scalar load
// per 64-byte line into a scratch SGPR, then one wait, so that the component loop's record
// loads hit the cache instead of each waiting on L2 in turn.
template <int FCAP, typename real>
__device__ __forceinline__ void prefetch_step_records(const real* __restrict__ rec, int n, int i, int first) {
    constexpr int rs = Rec<FCAP, real>::size;
    const uintptr_t b = (uintptr_t)(rec + ((size_t)i * n + first) * rs) & ~(uintptr_t)63;
    const uintptr_t e = (uintptr_t)(rec + ((size_t)i * n + n) * rs);
    uint32_t sink = 0;
    for (uintptr_t a = b; a < e; a += 64) asm volatile("s_load_dword %0, %1, 0x0" : "+s"(sink) : "s"(a));
    asm volatile("s_waitcnt lgkmcnt(0)" : "+s"(sink) : : "memory");
}

// One Euler step t_i -> t_{i+1} of all alive components, in place. Shared by the
// pricing pass and the adjoint recompute; both group sizes give bit-identical states.
// GROUP 1: a uniform branch per component. GROUP 4: components in groups of four
// without branches inside a group, so the loads and transcendentals of four
// components overlap; a component outside [first, n) of a partly alive group is
// computed and discarded (w = 0 adds exact zeros to S).
// LOGSTATE: the array holds the log-forwards (see euler_component).
template <int NCAP, int FCAP, typename real, bool LOGSTATE = false, int GROUP = 1, bool PREFETCH = false>
__device__ __forceinline__ void euler_step(real (&L)[NCAP], int n, int i, int first, real dt, const real (&dW)[FCAP],
                                           const real* __restrict__ rec, bool lognormal, bool capped, real cap,
                                           real logcap, real* __restrict__ paths,
                                           const int32_t* __restrict__ row_next, int64_t p, int P, bool active) {
    static_assert(NCAP % GROUP == 0, "component capacity must be a multiple of the group size");
    if (PREFETCH && first < n) prefetch_step_records<FCAP, real>(rec, n, i, first);
    real S[FCAP];
#pragma unroll
    for (int k = 0; k < FCAP; k++) S[k] = (real)0;
#pragma unroll
    for (int j0 = 0; j0 < NCAP; j0 += GROUP) {
        if (j0 + GROUP - 1 < first || j0 >= n) continue;  // uniform: the whole group is frozen
#pragma unroll
        for (int jj = 0; jj < GROUP; jj++) {
            const int j = j0 + jj;
            const bool alive = j >= first && j < n;        // uniform
            euler_component<NCAP, FCAP, real, LOGSTATE>(L, S, j, alive, j < n ? j : n - 1, n, i, dt, dW, rec,
                                                       lognormal, capped, cap, logcap, paths, row_next, p, P,
                                                       active);
        }
    }
}

template <int FCAP, typename real>
__device__ __forceinline__ void load_increments(real (&dW)[FCAP], const real* __restrict__ dw, int64_t pc, int nT,
                                                int i, int F) {
#pragma unroll
    for (int k = 0; k < FCAP; k++) dW[k] = (k < F) ? dw[((size_t)pc * nT + i) * F + k] : (real)0;
}

// Variants of this path kernel (template argument VAR), both spot measure, Euler:
//   kVarFunctional  EULER_FUNCTIONAL (the array holds the values)
//   kVarLogState    log-normal, the log-forwards carried between steps and capped as
//                   log-states, contraction allowed
// The spot-measure pricing kernels are in lmm_spot.hpp, the terminal measure, the carried
// schemes, the predictor-corrector and the volatility wrappers in lmm_general.hpp.
constexpr int kVarFunctional = 0, kVarLogState = 1;

template <int NCAP, int FCAP, typename real, int VAR = kVarFunctional>
__global__ void __launch_bounds__(kStepThreads)
__attribute__((amdgpu_waves_per_eu((VAR == kVarLogState && NCAP == 40) ? 3 : 1)))
lmm_fused_kernel(const StepArgs a, const real* __restrict__ rec, const double* __restrict__ delta,
                 const double* __restrict__ times, const int32_t* __restrict__ first_alive,
                 const int32_t* __restrict__ tenor_ti, const KProduct* __restrict__ products,
                 const KPeriod* __restrict__ periods, const real* __restrict__ dw,
                 double* __restrict__ wave_sum, double* __restrict__ wave_m2,
                 double* __restrict__ contrib, real* __restrict__ paths,
                 const int32_t* __restrict__ row_of, double* __restrict__ numeraire_out,
                 real* __restrict__ ckpt)
{
    constexpr bool LOGSTATE = VAR == kVarLogState;
    const int64_t p = (int64_t)blockIdx.x * kStepThreads + threadIdx.x;
    const int lane = threadIdx.x & 63;
    // the wavefront index, uniform: kept in a scalar register
    const int64_t wave = (int64_t)blockIdx.x * (kStepThreads / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int P = a.P, n = a.n, F = a.F, nT = a.nT;
    const bool active = p < P;
    const int64_t pc = active ? p : (int64_t)P - 1;
    const int cnt = (int)((int64_t)P - wave * 64 > 64 ? 64 : ((int64_t)P - wave * 64));
    const int stride = a.num_slots + n + 1;
    const bool lognormal = a.lognormal != 0;
    const bool capped = a.capped != 0;
    const real cap = (real)a.cap;
    // the array holds log-forwards (LOGSTATE) or values; val gives the value
    const real logcap = (real)a.logcap;
    auto val = [&](real x) -> real { return LOGSTATE ? dev_exp<real>(x) : x; };

    real L[NCAP];
#pragma unroll
    for (int j = 0; j < NCAP; j++) L[j] = (j < n) ? (real)(LOGSTATE ? a.Y0[j] : a.L0[j]) : (real)0;

    double N = 1.0, Nprev = 1.0;   // unadjusted numeraire N(T_knum) and the one before it
    int kcur = 0, knum = 0;
    int qp = 0;

    for (int ti = 0; ti <= nT; ti++) {
        // ---- forward-state checkpoint for the adjoint sweep
        if (ckpt && active && (ti % a.ckpt_stride) == 0) {
            const size_t base = (size_t)(ti / a.ckpt_stride) * n;
#pragma unroll
            for (int j = 0; j < NCAP; j++) if (j < n) ckpt[(base + j) * P + p] = val(L[j]);
        }
        // ---- spot measure: N(T_{k+1}) = N(T_k).accrue(L_k(T_k), d_k) once L_k is final
        while (kcur < n && tenor_ti[kcur] <= ti) {
            const double Lk = (double)val(pick<NCAP, real>(L, kcur));
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

        // ---- products evaluated at this time index
        while (qp < a.num_products && products[qp].eval_index == ti) {
            const KProduct pr = products[qp];
            qp++;
            const double Nsel = (pr.num_k == knum) ? N : Nprev;
            if (pr.type == FM_PRODUCT_SWAPTION) {
                real V = (real)0;
#pragma unroll
                for (int j = NCAP - 1; j >= 0; j--) {
                    if (j >= pr.first_j && j < pr.first_j + pr.num_periods) {
                        const KPeriod pe = periods[pr.param_off + (j - pr.first_j)];
                        const real Lj = val(L[j]);
                        const real payoff = ((Lj - (real)pe.strike) * (real)pe.period_length) * (real)pr.notional;
                        V = V + payoff;
                        V = dev_div<real>(V, (real)1.0 + Lj * (real)pe.tau);
                        V = V * (real)pe.adj;
                    }
                }
                const double x = (double)jmax0(V) / Nsel;
                reduce_slot(x, active, cnt, lane, wave, wave_sum, wave_m2, stride, pr.out_slot);
            } else if (pr.type == FM_PRODUCT_CAPLET) {
                const KPeriod pe = periods[pr.param_off];
                const real Lj = val(pick<NCAP, real>(L, pr.first_j));
                real v = Lj - (real)pe.strike;
                if (!pr.is_floorlet) v = jmax0(v) * (real)pe.notional;
                else v = ((v != v) ? v : (v > (real)0 ? (real)0 : v)) * ((real)(-1.0) * (real)pe.notional);
                const double x = (double)v / Nsel;
                reduce_slot(x, active, cnt, lane, wave, wave_sum, wave_m2, stride, pr.out_slot);
            } else if (pr.type == K_BERM_BASIS) {
                // raw regression basis of a Bermudan exercise date T_e (BermudanSwaption.java:215-252)
                if (pr.flags == FM_BASIS_REFERENCE) {
                    // L_short over tenor periods [first_j, pe.j), L_long over [first_j, pe.pay_k):
                    // a single period is the forward itself, several accrue (LMM:1290-1305)
                    const KPeriod pe = periods[pr.param_off];
                    real accS = (real)0, accL = (real)0, rs = (real)0, rl = (real)0;
#pragma unroll
                    for (int j = 0; j < NCAP; j++) {
                        if (j >= pr.first_j && j < pe.pay_k) {
                            const real Lj = val(L[j]);
                            const real d = (real)delta[j];
                            if (j == pr.first_j) { rs = Lj; rl = Lj; }
                            accL = (j == pr.first_j) ? Lj * d + (real)1.0 : accL * ((real)1 + Lj * d);
                            if (j < pe.j) accS = (j == pr.first_j) ? Lj * d + (real)1.0 : accS * ((real)1 + Lj * d);
                        }
                    }
                    if (pe.j > pr.first_j + 1) rs = (accS - (real)1.0) / (real)pe.tau;
                    if (pe.pay_k > pr.first_j + 1) rl = (accL - (real)1.0) / (real)pe.period_length;
                    const real ds = (real)1.0 / (rs * (real)pe.tau + (real)1.0);
                    const real dl = (real)1.0 / (rl * (real)pe.period_length + (real)1.0);
                    if (active) {
                        contrib[(size_t)pr.out_slot * P + p] = (double)ds;
                        contrib[(size_t)(pr.out_slot + 1) * P + p] = (double)dl;
                        contrib[(size_t)(pr.out_slot + 2) * P + p] = Nsel;
                    }
                } else {
                    // par rate of the remaining periods: P = 1/(L pl + 1), A = P pl, then
                    // P = P/(L pl + 1), A = A + P pl; S = (1 - P)/A
                    real Pd = (real)1, A = (real)0;
#pragma unroll
                    for (int j = 0; j < NCAP; j++) {
                        if (j >= pr.first_j && j < pr.first_j + pr.num_periods) {
                            const real pl = (real)periods[pr.param_off + (j - pr.first_j)].period_length;
                            const real Lj = val(L[j]);
                            if (j == pr.first_j) { Pd = (real)1.0 / (Lj * pl + (real)1.0); A = Pd * pl; }
                            else { Pd = Pd / (Lj * pl + (real)1.0); A = A + Pd * pl; }
                        }
                    }
                    const real S = (Pd * (real)(-1.0) + (real)1.0) / A;
                    if (active) {
                        contrib[(size_t)pr.out_slot * P + p] = (double)S;
                        contrib[(size_t)(pr.out_slot + 1) * P + p] = Nsel;
                    }
                }
            } else {   // one period of a swap: contribution column, combined after the adjustment is known
                const KPeriod pe = periods[pr.param_off];
                const real Lj = val(pick<NCAP, real>(L, pr.first_j));
                double payoff = (double)(((Lj - (real)pe.strike) * (real)pe.period_length) * (real)pe.notional);
                if (!(pr.flags & 1)) payoff = payoff * -1.0;
                if (active) contrib[(size_t)pr.out_slot * P + p] = payoff / Nsel;
            }
        }

        if (ti == nT) break;

        // ---- Euler step ti -> ti+1
        real dW[FCAP];
        load_increments<FCAP, real>(dW, dw, pc, nT, ti, F);
        const real dt = (real)(times[ti + 1] - times[ti]);
        const int32_t* row_next = row_of + (size_t)(ti + 1) * n;
        euler_step<NCAP, FCAP, real, LOGSTATE, 1, kPrefetchRecords>(L, n, ti, first_alive[ti], dt, dW, rec, lognormal,
                                                                    capped, cap, logcap, paths, row_next, p, P, active);
    }
}

} // namespace fm

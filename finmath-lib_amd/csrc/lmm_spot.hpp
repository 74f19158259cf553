// Spot-measure Euler path kernel (EULER_FUNCTIONAL, and EULER where it is the same scheme):
// the pricing and simulation pass of the plain model, FP64 or FP32-mixed.
//
// Per time step t_i -> t_{i+1} and alive component j (T_j > t_i), in the order of
// LIBORMarketModelFromCovarianceModel.getDrift
// (montecarlo/interestrate/models/LIBORMarketModelFromCovarianceModel.java:1144-1160,1183-1189)
// and EulerSchemeFromProcessModel (montecarlo/process/EulerSchemeFromProcessModel.java:232-256):
//   w    = 1 / (L_j * (d_j/d_j) + 1/d_j)        Scalar.discount (stochastic/Scalar.java:320-328)
//   w    = w * L_j                               (log-normal state space)
//   S_k  = S_k + w * lambda_jk                   ascending j
//   mu_j = sum_k S_k * lambda_jk + variance_j * (-0.5)
//   Y_j  = Y_j + mu_j * dt + sum_k dW_k * lambda_jk
//   L_j  = min(exp(Y_j), cap)
// FP64: the state Y_j (the log-forward; the forward itself in the normal state space) is
// carried from step to step. EULER_FUNCTIONAL re-derives it from the value as
// log(min(exp(Y), cap)), which gives back Y up to one rounding (exactly whenever the
// log-forward is at least 2 in magnitude, that is for forwards below 13.5 %); the cap is
// applied to the state as Y <= log(cap). Products read L_j = exp(Y_j) where they are evaluated.
// FP32-mixed: the value is carried and the state re-derived each step with the hardware
// logarithm. The adjoint's forward pass is this kernel, so it sees the priced float paths.
//
// Layout: the forwards of a path live in registers as chunks of sixteen VGPRs (eight FP64 or
// sixteen FP32 forwards). Components are processed by a rolled loop per chunk that reads and
// writes its element with an indexed register move (s_set_gpr_idx_on), so the kernel body holds
// one copy of the component update per chunk instead of one per forward: a small instruction
// footprint and about 160 VGPRs at 40 forwards x 5 factors in FP64 (three waves per SIMD). The
// step's model records are read through the scalar cache, one 64-byte record per component.
#pragma once

#include <type_traits>

#include "lmm_step.hpp"
#include "hip_check.hpp"

namespace fm {

template <int I, int N, typename Fn>
__device__ __forceinline__ void static_for(Fn&& fn) {
    if constexpr (I < N) {
        fn(std::integral_constant<int, I>{});
        static_for<I + 1, N>(fn);
    }
}

// Forwards per chunk: a chunk is sixteen VGPRs, eight FP64 or sixteen FP32 forwards. At this
// width the compiler reads and writes an element at a uniform run-time index with one indexed
// move (s_set_gpr_idx_on); narrower FP32 chunks are lowered to chains of compares and selects.
template <typename real> constexpr int chunk_width() { return 64 / (int)sizeof(real); }
template <typename real> struct VecCh;
template <> struct VecCh<double> { typedef double t __attribute__((ext_vector_type(8))); };
template <> struct VecCh<float> { typedef float t __attribute__((ext_vector_type(16))); };

// NCH chunks of forwards in registers (NCH <= 10).
template <int NCH, typename real>
struct ChunkedForwards {
    using V = typename VecCh<real>::t;
    static constexpr int W = chunk_width<real>();
    static_assert(NCH >= 1 && NCH <= 10, "one to ten chunks of forwards");
    V c0, c1, c2, c3, c4, c5, c6, c7, c8, c9;
    template <int Q> __device__ __forceinline__ V& ch() {
        if constexpr (Q == 0) return c0; else if constexpr (Q == 1) return c1; else if constexpr (Q == 2) return c2;
        else if constexpr (Q == 3) return c3; else if constexpr (Q == 4) return c4; else if constexpr (Q == 5) return c5;
        else if constexpr (Q == 6) return c6; else if constexpr (Q == 7) return c7; else if constexpr (Q == 8) return c8;
        else return c9;
    }
    template <int J> __device__ __forceinline__ real at() { return ch<J / W>()[J % W]; }
    template <int J> __device__ __forceinline__ void put(real v) { ch<J / W>()[J % W] = v; }
    // Element j for a uniform run-time j: one indexed register read in the chunk of j. The
    // empty asm keeps each chunk a register value, so the per-chunk reads are not merged into
    // a load from a memory copy of the forwards.
    __device__ __forceinline__ real get(int j) {
        const int e = j % W, qj = j / W;
        real r = (real)0;
        static_for<0, NCH>([&](auto Q) {
            constexpr int q = decltype(Q)::value;
            if (qj == q) {
                const V v = ch<q>();
#if defined(__HIP_DEVICE_COMPILE__)
                asm volatile("" :: "v"(v));
#endif
                r = v[e];
            }
        });
        return r;
    }
    // Element j += x for a uniform run-time j: one indexed register write in the chunk of j. The
    // asm takes the chunk as an input and output, which keeps it a register value.
    __device__ __forceinline__ void add(int j, real x) {
        const int e = j % W, qj = j / W;
        static_for<0, NCH>([&](auto Q) {
            constexpr int q = decltype(Q)::value;
            if (qj == q) {
                V v = ch<q>();
#if defined(__HIP_DEVICE_COMPILE__)
                asm volatile("" : "+v"(v));
#endif
                v[e] = v[e] + x;
                ch<q>() = v;
            }
        });
    }
};

template <typename real> __device__ __forceinline__ real spot_div(real a, real b);
template <> __device__ __forceinline__ double spot_div<double>(double a, double b) { return fm_div_nr(a, b); }
template <> __device__ __forceinline__ float spot_div<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
// spot_div split at the reciprocal of b, for numerators sharing one divisor (same results)
template <typename real> __device__ __forceinline__ real spot_rcp(real b);
template <> __device__ __forceinline__ double spot_rcp<double>(double b) { return fm_rcp_nr(b); }
template <> __device__ __forceinline__ float spot_rcp<float>(float b) { return __builtin_amdgcn_rcpf(b); }
template <typename real> __device__ __forceinline__ real spot_div_r(real a, real b, real r);
template <> __device__ __forceinline__ double spot_div_r<double>(double a, double b, double r) { return fm_div_nr_r(a, b, r); }
template <> __device__ __forceinline__ float spot_div_r<float>(float a, float, float r) { return a * r; }

// The spot kernel carries log-forwards in FP64 and values in FP32 (see above).
template <typename real> constexpr bool carries_log() { return std::is_same<real, double>::value; }

// One Euler step of all alive components, in place. Y holds log-forwards (FP64, log-normal) or
// forwards; ycap is log(cap) for log-forwards, cap otherwise. The state space is a template
// argument: a run-time branch around the exponential would split the component's block and
// expose the latency of its record load.
template <int NCH, int FCAP, typename real, bool lognormal>
__device__ __forceinline__ void spot_step(ChunkedForwards<NCH, real>& Y, int n, int i, int first, real dt,
                                          const real (&dW)[FCAP], const real* __restrict__ rec,
                                          bool capped, real ycap, real* __restrict__ paths,
                                          const int32_t* __restrict__ row_next, int64_t p, int P, bool active) {
    constexpr bool LOGCARRY = carries_log<real>();
    constexpr int W = chunk_width<real>();
    real S[FCAP];
#pragma unroll
    for (int k = 0; k < FCAP; k++) S[k] = (real)0;
    static_for<0, NCH>([&](auto Q) {
        constexpr int q = decltype(Q)::value;
        const int j0 = first > W * q ? first : W * q;
        const int j1 = n < W * q + W ? n : W * q + W;
#pragma unroll 1
        for (int j = j0; j < j1; j++) {
            const Rec<FCAP, real> R = rec_at<FCAP, real>(rec, n, i, j);
            const real x = Y.template ch<q>()[j - W * q];
            const real Lj = (LOGCARRY && lognormal) ? dev_exp<real>(x) : x;
            real w = spot_div<real>((real)1.0, Lj * R.c1() + R.c2());
            if (lognormal) w = w * Lj;
#pragma unroll
            for (int k = 0; k < FCAP; k++) S[k] = S[k] + w * R.lam(k);
            real mu = S[0] * R.lam(0);
#pragma unroll
            for (int k = 1; k < FCAP; k++) mu = mu + S[k] * R.lam(k);
            mu = mu + R.hv();
            real y = (!LOGCARRY && lognormal) ? dev_log<real>(Lj) : x;
            y = y + mu * dt;
#pragma unroll
            for (int k = 0; k < FCAP; k++) y = y + dW[k] * R.lam(k);
            if (LOGCARRY) {
                if (capped) y = jmin(y, ycap);
                Y.template ch<q>()[j - W * q] = y;
                if (paths && active) paths[(size_t)row_next[j] * P + p] = lognormal ? dev_exp<real>(y) : y;
            } else {
                real v = lognormal ? dev_exp<real>(y) : y;
                if (capped) v = jmin(v, ycap);
                Y.template ch<q>()[j - W * q] = v;
                if (paths && active) paths[(size_t)row_next[j] * P + p] = v;
            }
        }
    });
}

// Products evaluated at time index ti, in the order they were lowered (qp advances over them):
// swaptions and caplets are reduced into their wave partials, swap periods and Bermudan
// regression bases go to per-path contribution columns. Y holds the path's states and val(y)
// gives the forward L_j(t_ti) of state y; N and Nprev are the numeraires N(T_knum) and the one
// before it. GROUP values a swaption group in one walk; without it every member walks its own
// copy of the periods (the FP32 kernels: the shared walk measured slower there).
template <bool GROUP, int NCH, typename real, typename Val>
__device__ __forceinline__ void eval_products_at(int ti, int& qp, int num_products, const KProduct* __restrict__ products,
                                                 const KPeriod* __restrict__ periods, const double* __restrict__ delta,
                                                 double N, double Nprev, int knum, ChunkedForwards<NCH, real>& Y,
                                                 Val&& val, bool active, int cnt, int lane, int64_t wave,
                                                 double* __restrict__ wave_sum, double* __restrict__ wave_m2,
                                                 int stride, double* __restrict__ contrib, int64_t p, int P) {
    constexpr int W = chunk_width<real>();
    auto value_of = [&](int j) -> real { return val(Y.get(j)); };
    while (qp < num_products && products[qp].eval_index == ti) {
        const KProduct pr = products[qp];
        qp++;
        const double Nsel = (pr.num_k == knum) ? N : Nprev;
        if (pr.type == FM_PRODUCT_SWAPTION) {
            // Swaption.java:150-190: the periods backwards, value = (value + payoff) / (1 + L tau) * adj.
            // The periods are walked chunk by chunk, each element read with one indexed move. With
            // GROUP, the pr.pad2 swaptions that follow differ from this one only in strike and
            // notional: the walk values them too, sharing L_j and the reciprocal of 1 + L_j tau;
            // member s's periods follow this one's at s * num_periods and carry its notional.
            constexpr int G = GROUP ? kSwaptionGroup : 1;
            const int m = GROUP ? 1 + pr.pad2 : 1;
            real V[G];
#pragma unroll
            for (int g = 0; g < G; g++) V[g] = (real)0;
            const int j0 = pr.first_j, j1 = pr.first_j + pr.num_periods;
            static_for<0, NCH>([&](auto QR) {
                constexpr int q = NCH - 1 - decltype(QR)::value;
                const int ja = j0 > W * q ? j0 : W * q;
                const int jb = j1 < W * q + W ? j1 : W * q + W;
#pragma unroll 1
                for (int j = jb - 1; j >= ja; j--) {
                    const KPeriod pe = periods[pr.param_off + (j - j0)];
                    const real Lj = val(Y.template ch<q>()[j - W * q]);
                    const real b = (real)1.0 + Lj * (real)pe.tau;
                    const real r = spot_rcp<real>(b);
                    // every member's strike and notional are read up front (indices clamped to the
                    // group) so their scalar loads issue together
                    real Kg[G], Ng[G];
#pragma unroll
                    for (int g = 0; g < G; g++) {
                        const KPeriod& pg = periods[pr.param_off + (g < m ? g : 0) * pr.num_periods + (j - j0)];
                        Kg[g] = (real)pg.strike;
                        Ng[g] = GROUP ? (real)pg.notional : (real)pr.notional;
                    }
#pragma unroll
                    for (int g = 0; g < G; g++) {
                        if (g < m) {
                            const real payoff = ((Lj - Kg[g]) * (real)pe.period_length) * Ng[g];
                            V[g] = V[g] + payoff;
                            V[g] = spot_div_r<real>(V[g], b, r);
                            V[g] = V[g] * (real)pe.adj;
                        }
                    }
                }
            });
#pragma unroll
            for (int g = 0; g < G; g++)
                if (g < m)
                    reduce_slot((double)jmax0(V[g]) / Nsel, active, cnt, lane, wave, wave_sum, wave_m2, stride,
                                products[qp - 1 + g].out_slot);
            qp += m - 1;
        } else if (pr.type == FM_PRODUCT_CAPLET) {
            const KPeriod pe = periods[pr.param_off];
            const real Lj = value_of(pr.first_j);
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
#pragma unroll 1
                for (int j = pr.first_j; j < pe.pay_k; j++) {
                    const real Lj = value_of(j);
                    const real d = (real)delta[j];
                    if (j == pr.first_j) { rs = Lj; rl = Lj; }
                    accL = (j == pr.first_j) ? Lj * d + (real)1.0 : accL * ((real)1 + Lj * d);
                    if (j < pe.j) accS = (j == pr.first_j) ? Lj * d + (real)1.0 : accS * ((real)1 + Lj * d);
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
#pragma unroll 1
                for (int j = pr.first_j; j < pr.first_j + pr.num_periods; j++) {
                    const real pl = (real)periods[pr.param_off + (j - pr.first_j)].period_length;
                    const real Lj = value_of(j);
                    if (j == pr.first_j) { Pd = (real)1.0 / (Lj * pl + (real)1.0); A = Pd * pl; }
                    else { Pd = Pd / (Lj * pl + (real)1.0); A = A + Pd * pl; }
                }
                const real S = (Pd * (real)(-1.0) + (real)1.0) / A;
                if (active) {
                    contrib[(size_t)pr.out_slot * P + p] = (double)S;
                    contrib[(size_t)(pr.out_slot + 1) * P + p] = Nsel;
                }
            }
        } else {   // one period of a swap: contribution column, combined after the adjustment is known
            const KPeriod pe = periods[pr.param_off];
            const real Lj = value_of(pr.first_j);
            double payoff = (double)(((Lj - (real)pe.strike) * (real)pe.period_length) * (real)pe.notional);
            if (!(pr.flags & 1)) payoff = payoff * -1.0;
            if (active) contrib[(size_t)pr.out_slot * P + p] = payoff / Nsel;
        }
    }
}

// BUILD tells apart instantiations compiled with different floating-point flags (0: without
// contraction, the FP64 and FP32-mixed builds; 1: with fused multiply-add, FP64_LOGSTATE). Two
// translation units that instantiated one kernel under one name with different flags would
// give one definition to both: the linker keeps a single copy of a template instantiation.
template <int NCH, int FCAP, typename real, bool lognormal, int BUILD = 0>
__global__ void __launch_bounds__(kStepThreads)
lmm_spot_kernel(const StepArgs a, const real* __restrict__ rec, const double* __restrict__ delta,
                const double* __restrict__ times, const int32_t* __restrict__ first_alive,
                const int32_t* __restrict__ tenor_ti, const KProduct* __restrict__ products,
                const KPeriod* __restrict__ periods, const real* __restrict__ dw,
                double* __restrict__ wave_sum, double* __restrict__ wave_m2,
                double* __restrict__ contrib, real* __restrict__ paths,
                const int32_t* __restrict__ row_of, double* __restrict__ numeraire_out,
                real* __restrict__ ckpt)
{
    const int64_t p = (int64_t)blockIdx.x * kStepThreads + threadIdx.x;
    const int lane = threadIdx.x & 63;
    // the wavefront index, uniform: kept in a scalar register
    const int64_t wave = (int64_t)blockIdx.x * (kStepThreads / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int P = a.P, n = a.n, F = a.F, nT = a.nT;
    const bool active = p < P;
    const int64_t pc = active ? p : (int64_t)P - 1;
    const int cnt = (int)((int64_t)P - wave * 64 > 64 ? 64 : ((int64_t)P - wave * 64));
    const int stride = a.num_slots + n + 1;
    const bool capped = a.capped != 0;
    constexpr bool LOGCARRY = carries_log<real>();
    constexpr int W = chunk_width<real>();
    const real ycap = (LOGCARRY && lognormal) ? (real)a.logcap : (real)a.cap;
    auto val = [&](real y) -> real { return (LOGCARRY && lognormal) ? dev_exp<real>(y) : y; };

    ChunkedForwards<NCH, real> Y;
    static_for<0, W * NCH>([&](auto J) {
        constexpr int j = decltype(J)::value;
        Y.template put<j>((j < n) ? (real)a.Y0[j] : (real)0);
    });

    double N = 1.0, Nprev = 1.0;   // unadjusted numeraire N(T_knum) and the one before it
    int kcur = 0, knum = 0;
    int qp = 0;

    for (int ti = 0; ti <= nT; ti++) {
        // ---- the carried states every ckpt_stride steps, for the adjoint's reverse sweep:
        // [ti / stride][j][P], stored non-temporal (read back only by the sweep, after this pass)
        if (ckpt && active && (ti % a.ckpt_stride) == 0) {
            const size_t base = (size_t)(ti / a.ckpt_stride) * n;
            const int cfirst = a.ckpt_first[ti / a.ckpt_stride];
            static_for<0, W * NCH>([&](auto J) {
                constexpr int j = decltype(J)::value;
                if (j < n && j >= cfirst) __builtin_nontemporal_store(Y.template at<j>(), &ckpt[(base + j) * P + p]);
            });
        }
        // ---- N(T_{k+1}) = N(T_k).accrue(L_k(T_k), d_k) once L_k is final (LMM:1050-1066)
        while (kcur < n && tenor_ti[kcur] <= ti) {
            const double Lk = (double)val(Y.get(kcur));
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
        constexpr bool group = sizeof(real) == 8;
        eval_products_at<group, NCH, real>(ti, qp, a.num_products, products, periods, delta, N, Nprev, knum, Y, val,
                                           active, cnt, lane, wave, wave_sum, wave_m2, stride, contrib, p, P);

        if (ti == nT) break;

        // ---- Euler step ti -> ti+1
        real dW[FCAP];
        load_increments<FCAP, real>(dW, dw, pc, nT, ti, F);
        const real dt = (real)(times[ti + 1] - times[ti]);
        const int32_t* row_next = row_of + (size_t)(ti + 1) * n;
        spot_step<NCH, FCAP, real, lognormal>(Y, n, ti, first_alive[ti], dt, dW, rec, capped, ycap, paths, row_next,
                                              p, P, active);
    }
}

template <int NCH, int FCAP, typename real, bool lognormal, int BUILD>
void launch_spot_ss(const StepArgs& a, hipStream_t stream) {
    const int64_t blocks = ((int64_t)a.P + kStepThreads - 1) / kStepThreads;
    hipLaunchKernelGGL((lmm_spot_kernel<NCH, FCAP, real, lognormal, BUILD>), dim3((unsigned)blocks), dim3(kStepThreads), 0,
                       stream, a, (const real*)a.rec, a.delta, a.times, a.first_alive, a.tenor_time_index, a.products,
                       a.periods, (const real*)a.dw, a.wave_sum, a.wave_m2, a.contrib, (real*)a.paths, a.row_of,
                       a.numeraire, (real*)a.ckpt);
    HIP_CHECK(hipGetLastError());
}

template <int NCH, int FCAP, typename real, int BUILD>
void launch_spot_cap(const StepArgs& a, hipStream_t stream) {
    if (a.lognormal) launch_spot_ss<NCH, FCAP, real, true, BUILD>(a, stream);
    else launch_spot_ss<NCH, FCAP, real, false, BUILD>(a, stream);
}

template <int NCH, typename real, int BUILD>
void launch_spot_n(const StepArgs& a, hipStream_t stream) {
    if (a.F <= 1) launch_spot_cap<NCH, 1, real, BUILD>(a, stream);
    else if (a.F <= 3) launch_spot_cap<NCH, 3, real, BUILD>(a, stream);
    else if (a.F <= 5) launch_spot_cap<NCH, 5, real, BUILD>(a, stream);
    else if (a.F <= 10) launch_spot_cap<NCH, 10, real, BUILD>(a, stream);
    else throw FmError(FM_ERR_UNSUPPORTED, "number of factors above 10 is not supported by the path kernel");
}

// The spot Euler kernel of the pricing and simulation passes. a.Y0 holds the initial states:
// FP64 log-normal the log-forwards capped at log(cap) (a.logcap = log(cap)), otherwise the
// capped initial forwards.
template <typename real, int BUILD = 0>
void launch_spot(const StepArgs& a, hipStream_t stream) {
    // 1, 3, 5, 10 chunks of eight FP64 forwards; 1, 2, 3, 5 chunks of sixteen FP32 forwards
    constexpr int W = chunk_width<real>();
    constexpr int N2 = W == 8 ? 3 : 2, N3 = W == 8 ? 5 : 3, N4 = W == 8 ? 10 : 5;
    if (a.n <= W) launch_spot_n<1, real, BUILD>(a, stream);
    else if (a.n <= W * N2) launch_spot_n<N2, real, BUILD>(a, stream);
    else if (a.n <= W * N3) launch_spot_n<N3, real, BUILD>(a, stream);
    else if (a.n <= 80) launch_spot_n<N4, real, BUILD>(a, stream);
    else throw FmError(FM_ERR_UNSUPPORTED, "more than 80 forward rates are not supported by the path kernel");
}

} // namespace fm

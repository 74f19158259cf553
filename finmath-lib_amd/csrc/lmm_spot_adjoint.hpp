// Pathwise adjoint (reverse sweep) of the spot Euler pricing pass.
//
// The pricing pass (lmm_spot_kernel) stores the carried state of every path every `stride`
// time steps and the numeraire N(T_k). The reverse sweep walks the segments between
// checkpoints backwards. At the top index of a segment it reloads the segment's checkpoint and
// rebuilds the state with the pricing pass's own step (spot_step), so the sweep differentiates
// the paths that were priced; the states in between go to a per-path segment buffer
// ([stride-2][n][P]), so each step is recomputed once. For each index t it then
//   1. reverses the Euler step t -> t+1 (adjoint of y(t+1) -> adjoint of L(t)),
//   2. adds the direct sensitivities of the products evaluated at t,
//   3. reverses the numeraire updates N(T_{k+1}) = N(T_k)(1 + L_k d_k) made at t,
//   4. turns the adjoint of L(t) into that of the state y(t) of the step t-1 -> t while the
//      state at t is at hand, so step 1 at t-1 needs only the drift sums of its step.
// Per alive forward-step it reduces dV/dsigma_j(t_i) over the wavefront, where the factor
// loading is lambda_jk = sigma_j f_jk and the variance sigma_j^2 c_j
// (LIBORCovarianceModelFromVolatilityAndCorrelation.java:48-93); at the end it reduces dV/dL_j(0).
//
// Per path the sweep holds two chunked register arrays, the state at t and the adjoint of the
// forwards, and processes components in rolled loops (lmm_spot.hpp).
//
// Conventions of the reference's AAD (RandomVariableDifferentiableAAD.java:165-219): averages
// over paths (the numeraire-adjustment factor) are constants, the floor of the swaption payoff
// has derivative 1 at and above zero and 0 below. A Bermudan period's payoff enters with a
// per-path weight: 1 or 0 by the recorded exercise (the exercise decisions held fixed), or the
// weights of the induction's own adjoint (lmm_bermudan.hip) when given.
#pragma once

#include "lmm_spot.hpp"
#include "lmm_adjoint.hpp"

namespace fm {

template <int NCH, int FCAP, typename real, bool lognormal>
__global__ void __launch_bounds__(kStepThreads)
lmm_spot_adjoint_kernel(const StepArgs a, const AdjArgs g, const real* __restrict__ rec,
                        const double* __restrict__ delta, const double* __restrict__ times,
                        const int32_t* __restrict__ first_alive, const int32_t* __restrict__ kcur_at,
                        const KProduct* __restrict__ products, const KPeriod* __restrict__ periods,
                        const real* __restrict__ dw, const real* __restrict__ ckpt,
                        const double* __restrict__ numeraire, const double* __restrict__ weight,
                        const real* __restrict__ fsig, const real* __restrict__ vterm,
                        real* __restrict__ seg, double* __restrict__ wave_grad)
{
    constexpr bool LOGCARRY = carries_log<real>() && lognormal;   // the state is the log-forward
    constexpr int W = chunk_width<real>();
    const int64_t p = (int64_t)blockIdx.x * kStepThreads + threadIdx.x;
    const int lane = threadIdx.x & 63;
    // the wavefront index, uniform: kept in a scalar register
    const int64_t wave = (int64_t)blockIdx.x * (kStepThreads / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int P = a.P, n = a.n, F = a.F, nT = a.nT;
    const bool active = p < P;
    const int64_t pc = active ? p : (int64_t)P - 1;
    const int cnt = (int)((int64_t)P - wave * 64 > 64 ? 64 : ((int64_t)P - wave * 64));
    const int gstride = nT * n + n;
    const bool capped = a.capped != 0;
    const real ycap = LOGCARRY ? (real)a.logcap : (real)a.cap;
    const int stride = g.stride;
    auto val = [&](real x) -> real { return LOGCARRY ? dev_exp<real>(x) : x; };

    ChunkedForwards<NCH, real> X;     // state at t
    ChunkedForwards<NCH, real> Lb;    // adjoint of the forwards L_j(t)
    static_for<0, W * NCH>([&](auto J) {
        constexpr int j = decltype(J)::value;
        Lb.template put<j>((real)0);
    });
    double Nb_top = 0.0, Nb_low = 0.0;      // adjoints of N(T_K) and N(T_{K-1})
    int K = kcur_at[nT];
    int qp = a.num_products - 1;            // products sorted by evaluation index

    for (int c = (nT / stride) * stride; c >= 0; c -= stride) {
        const int hi = (c + stride - 1 < nT) ? c + stride - 1 : nT;
        for (int t = hi; t >= c; t--) {
            // ---- the state at t: the checkpoint at c for t = c and t = hi, the segment's stored
            // states in between. At t = hi the pricing pass's steps c .. hi-1 rebuild the state
            // and store the states c+1 .. hi-1 for the later t of the segment. The empty asm
            // makes the source's offset opaque, so the compiler does not hoist the loads out of
            // the t loop and keep another register array of forwards live across the segment.
            size_t row = (t > c && t < hi) ? (size_t)(t - c - 1) * n : (size_t)(c / stride) * n;
            const real* src = (t > c && t < hi) ? seg : ckpt;
#if defined(__HIP_DEVICE_COMPILE__)
            asm volatile("" : "+s"(row));
#endif
            // components below cfirst are dead and read by nothing in this segment. Checkpoints and
            // segment states are read and written non-temporal: a wavefront's segment (two states of
            // n components at stride 4) outgrows its share of the L2 before it is read back.
            const int cfirst = a.ckpt_first[c / stride];
            static_for<0, W * NCH>([&](auto J) {
                constexpr int j = decltype(J)::value;
                X.template put<j>((j < n && j >= cfirst) ? __builtin_nontemporal_load(&src[(row + j) * P + pc]) : (real)0);
            });
            if (t == hi) {
                for (int s = c; s < t; s++) {
                    real dWs[FCAP];
                    load_increments<FCAP, real>(dWs, dw, pc, nT, s, F);
                    spot_step<NCH, FCAP, real, lognormal>(X, n, s, first_alive[s], (real)(times[s + 1] - times[s]),
                                                          dWs, rec, capped, ycap, (real*)nullptr,
                                                          (const int32_t*)nullptr, p, P, active);
                    if (s + 1 < hi && active) {
                        const size_t srow = (size_t)(s - c) * n;
                        static_for<0, W * NCH>([&](auto J) {
                            constexpr int j = decltype(J)::value;
                            if (j < n && j >= cfirst) __builtin_nontemporal_store(X.template at<j>(), &seg[(srow + j) * P + p]);
                        });
                    }
                }
            }

            // ---- 1. reverse the Euler step t -> t+1. Lb holds the adjoint of the new state
            // y_j(t+1) for the components the step moves (converted at the end of t+1).
            if (t < nT) {
                const int first = first_alive[t];
                const real dt = (real)(times[t + 1] - times[t]);
                real dW[FCAP], S[FCAP], Sb[FCAP];
                load_increments<FCAP, real>(dW, dw, pc, nT, t, F);
#pragma unroll
                for (int k = 0; k < FCAP; k++) { S[k] = (real)0; Sb[k] = (real)0; }
                // pass A: Sb_k = sum_j mb_j lambda_jk with mb_j = Yb_j dt (Yb_j the adjoint of
                // y_j(t+1)), the adjoint of the step's drift sums before any component's share is
                // taken off; it reads only Lb and the factor loadings
                static_for<0, NCH>([&](auto Q) {
                    constexpr int q = decltype(Q)::value;
                    const int j0 = first > W * q ? first : W * q;
                    const int j1 = n < W * q + W ? n : W * q + W;
#pragma unroll 1
                    for (int j = j0; j < j1; j++) {
                        const Rec<FCAP, real> R = rec_at<FCAP, real>(rec, n, t, j);
                        const real mb = Lb.template ch<q>()[j - W * q] * dt;
#pragma unroll
                        for (int k = 0; k < FCAP; k++) Sb[k] = Sb[k] + mb * R.lam(k);
                    }
                });
                // pass B (ascending): the drift sums S^(j) = S^(j-1) + w_j lambda_j as the step built
                // them, and Sb^(j) = sum_{i >= j} mb_i lambda_i, pass A's total with the shares of the
                // components below j taken off; the per-path dV/dsigma_j(t) of groups of four j are
                // summed over the wavefront together
                real gq0 = (real)0, gq1 = (real)0, gq2 = (real)0, gq3 = (real)0;
                static_for<0, NCH>([&](auto Q) {
                    constexpr int q = decltype(Q)::value;
                    const int j0 = first > W * q ? first : W * q;
                    const int j1 = n < W * q + W ? n : W * q + W;
#pragma unroll 1
                    for (int j = j0; j < j1; j++) {
                        const Rec<FCAP, real> R = rec_at<FCAP, real>(rec, n, t, j);
                        const real Lj = val(X.template ch<q>()[j - W * q]);
                        const real r = spot_div<real>((real)1.0, Lj * R.c1() + R.c2());
                        const real w = lognormal ? r * Lj : r;
                        const real Yb = Lb.template ch<q>()[j - W * q];     // adjoint of y_j(t+1)
                        const real mb = Yb * dt;
                        real wb = (real)0;
                        // dV/dsigma_j(t) on this path = sum_k lambda-bar_jk f_jk with lambda-bar_jk =
                        // Yb dW_k + mb S_k + w Sb_k, gathered as Yb (dW.f) + mb (S.f) + w (Sb.f)
                        real dwf = (real)0, sf = (real)0, sbf = (real)0;
#pragma unroll
                        for (int k = 0; k < FCAP; k++) {
                            const real f = fsig[j * FCAP + k];
                            S[k] = S[k] + w * R.lam(k);    // S^(j-1) -> S^(j)
                            wb = wb + Sb[k] * R.lam(k);
                            dwf = dwf + dW[k] * f;
                            sf = sf + S[k] * f;
                            sbf = sbf + Sb[k] * f;
                            Sb[k] = Sb[k] - mb * R.lam(k); // Sb^(j) -> Sb^(j+1)
                        }
                        real sgb = Yb * dwf + mb * sf + w * sbf;
                        // variance sigma^2 c in the drift: d(-var/2)/dsigma = -sigma c
                        if (lognormal) sgb = sgb + mb * vterm[(size_t)t * n + j];
                        // y_j(t+1) = log L_j(t) + ... (or L_j(t) + ...), and w_j depends on L_j(t)
                        const real dwdL = lognormal ? R.c2() * r * r : -R.c1() * r * r;
                        const real direct = lognormal ? Yb * adj_rcp<real>(Lj) : Yb;
                        Lb.template ch<q>()[j - W * q] = direct + wb * dwdL;
                        const real gs = active ? sgb : (real)0;
                        const int m = j & 3;
                        gq0 = (m == 0) ? gs : gq0;
                        gq1 = (m == 1) ? gs : gq1;
                        gq2 = (m == 2) ? gs : gq2;
                        gq3 = (m == 3) ? gs : gq3;
                        if (m == 3 || j == n - 1) {     // uniform: the group of four is complete
                            const double v4[4] = {(double)gq0, (double)gq1, (double)gq2, (double)gq3};
                            const double s = wave_sum4(v4, lane);
                            const int jw = (j & ~3) + 2 * (lane & 1) + ((lane >> 1) & 1);
                            if (lane < 4 && cnt > 0 && jw < n && jw >= first)
                                wave_grad[(size_t)wave * gstride + (size_t)t * n + jw] = s;
                            gq0 = gq1 = gq2 = gq3 = (real)0;
                        }
                    }
                });
                // the components the step does not move have dV/dsigma_j(t) = 0: this row of the
                // wavefront's gradient is then written in full and needs no clearing beforehand
                if (cnt > 0)
                    for (int j = lane; j < first; j += 64) wave_grad[(size_t)wave * gstride + (size_t)t * n + j] = 0.0;
            }

            // ---- 2. products evaluated at t (reverse order of evaluation). The per-path conditions
            // (in the money, exercised) scale the contributions instead of branching around them, and
            // the swaption walks the chunks of its periods in place: the adjoint forwards then stay
            // one register array through the product loop.
            const int Kt = kcur_at[t];
            while (qp >= 0 && products[qp].eval_index == t) {
                const KProduct pr = products[qp];
                const double wq = weight[qp];
                qp--;
                const bool top = (pr.num_k == Kt);
                const double Nsel = (pr.num_k == 0) ? 1.0 : numeraire[(size_t)pr.num_k * P + pc];
                double nb = 0.0;            // adjoint of the selected numeraire
                real xadd = (real)0;        // adjoint of L_{first_j}(t) (caplets and swap periods)
                if (pr.type == FM_PRODUCT_SWAPTION) {
                    const int j0 = pr.first_j, j1 = pr.first_j + pr.num_periods;
                    // V = sum_j p_j prod_{l<=j} D_l, D_l = adj_l / (1 + L_l tau_l)
                    real V = (real)0, Pc = (real)1;
                    static_for<0, NCH>([&](auto QR) {
                        constexpr int q = NCH - 1 - decltype(QR)::value;
                        const int ja = j0 > W * q ? j0 : W * q;
                        const int jb = j1 < W * q + W ? j1 : W * q + W;
#pragma unroll 1
                        for (int j = jb - 1; j >= ja; j--) {
                            const KPeriod pe = periods[pr.param_off + (j - j0)];
                            const real Lj = val(X.template ch<q>()[j - W * q]);
                            const real D = (real)pe.adj * adj_rcp<real>((real)1.0 + Lj * (real)pe.tau);
                            V = (V + ((Lj - (real)pe.strike) * (real)pe.period_length) * (real)pr.notional) * D;
                            Pc = Pc * D;
                        }
                    });
                    // floor(0.0): derivative 1 at and above the floor, 0 below
                    const double Vb = (V >= (real)0) ? wq / Nsel : 0.0;
                    nb = (V >= (real)0) ? -wq * (double)V / (Nsel * Nsel) : 0.0;
                    // descending over j: P_j = prod_{l<=j} D_l, U_j = sum_{m>=j} p_m P_m
                    real U = (real)0;
                    static_for<0, NCH>([&](auto QR) {
                        constexpr int q = NCH - 1 - decltype(QR)::value;
                        const int ja = j0 > W * q ? j0 : W * q;
                        const int jb = j1 < W * q + W ? j1 : W * q + W;
#pragma unroll 1
                        for (int j = jb - 1; j >= ja; j--) {
                            const KPeriod pe = periods[pr.param_off + (j - j0)];
                            const real Lj = val(X.template ch<q>()[j - W * q]);
                            const real one = (real)1.0 + Lj * (real)pe.tau;
                            const real inv = adj_rcp<real>(one);
                            const real pj = ((Lj - (real)pe.strike) * (real)pe.period_length) * (real)pr.notional;
                            U = U + pj * Pc;
                            const real dV = (real)pe.period_length * (real)pr.notional * Pc - ((real)pe.tau * inv) * U;
                            Lb.template ch<q>()[j - W * q] = Lb.template ch<q>()[j - W * q] + (real)Vb * dV;
                            Pc = Pc * (one * (real)pe.inv_adj);      // P_{j-1} = P_j / D_j
                        }
                    });
                } else if (pr.type == FM_PRODUCT_CAPLET) {
                    const KPeriod pe = periods[pr.param_off];
                    const real Lj = val(X.get(pr.first_j));
                    const real v = Lj - (real)pe.strike;
                    // floor(0.0) has derivative 1 for v >= 0, cap(0.0) has derivative 1 for v < 0
                    // (RandomVariableDifferentiableAAD.java:333-348)
                    const bool itm = pr.is_floorlet ? (v < (real)0) : (v >= (real)0);
                    const double dcf = pe.notional;
                    const double x = pr.is_floorlet ? -(double)v * dcf : (double)v * dcf;
                    xadd = itm ? (real)((pr.is_floorlet ? -dcf : dcf) * wq / Nsel) : (real)0;
                    nb = itm ? -wq * x / (Nsel * Nsel) : 0.0;
                } else if (pr.type == K_BERM_BASIS) {
                    // regression basis: not differentiated (the reference's conditional
                    // expectation treats its basis functions as constants)
                } else {
                    double pw = 1.0;       // the payoff's weight on this path
                    if (pr.type == K_BERM_PERIOD) {
                        if (g.berm_weight) {
                            pw = g.berm_weight[(size_t)pr.out_slot * P + pc];
                        } else {
                            // the period pays on this path if the exercise (callable: entered at or
                            // before it; cancelable: terminated after it) keeps it in the value
                            const int ex = g.berm_ex[(size_t)pr.pad2 * P + pc];
                            pw = ((pr.flags & 2) ? (ex >= 0 && pr.is_floorlet >= ex)
                                                 : (ex < 0 || pr.is_floorlet < ex)) ? 1.0 : 0.0;
                        }
                    }
                    const KPeriod pe = periods[pr.param_off];
                    const real Lj = val(X.get(pr.first_j));
                    const double sgn = (pr.flags & 1) ? 1.0 : -1.0;
                    const double pay = sgn * (double)(((Lj - (real)pe.strike) * (real)pe.period_length) * (real)pe.notional);
                    xadd = (real)(pw * sgn * pe.period_length * pe.notional * wq / Nsel);
                    nb = -pw * wq * pay / (Nsel * Nsel);
                }
                if (top) Nb_top += nb; else Nb_low += nb;
                Lb.add(pr.first_j, xadd);
            }

            // ---- 3. numeraire updates made at t: N_{k+1} = N_k (1 + L_k d_k), k = K-1 down to kcur_at[t-1]
            const int Kprev = (t > 0) ? kcur_at[t - 1] : 0;
            while (K > Kprev) {
                const int k = K - 1;
                const double Nk = (k == 0) ? 1.0 : numeraire[(size_t)k * P + pc];
                const double Lk = (double)val(X.get(k));
                const double dk = delta[k];
                Lb.add(k, (real)(Nb_top * dk * Nk));
                Nb_low += Nb_top * (1.0 + Lk * dk);
                Nb_top = Nb_low;
                Nb_low = 0.0;
                K--;
            }

            // ---- 4. Lb now holds the full adjoint of L(t). For the components the step t-1 -> t
            // moved, turn it into the adjoint of the new state y_j(t): times L_j(t) = exp(y_j(t))
            // (log-normal), and zero where the cap binds (cap(x, c) has derivative 0 for x >= c,
            // RandomVariableDifferentiableAAD.java:316-323; x >= c exactly where the capped state
            // equals the cap).
            if (t > 0) {
                const int first = first_alive[t - 1];
                static_for<0, NCH>([&](auto Q) {
                    constexpr int q = decltype(Q)::value;
                    const int j0 = first > W * q ? first : W * q;
                    const int j1 = n < W * q + W ? n : W * q + W;
#pragma unroll 1
                    for (int j = j0; j < j1; j++) {
                        const real x = X.template ch<q>()[j - W * q];
                        const real lb = Lb.template ch<q>()[j - W * q];
                        const real yb = lognormal ? lb * val(x) : lb;
                        Lb.template ch<q>()[j - W * q] = (capped && x >= ycap) ? (real)0 : yb;
                    }
                });
            }
        }
    }
    // ---- dV/dL_j(0)
    static_for<0, W * NCH>([&](auto J) {
        constexpr int j = decltype(J)::value;
        if (j < n) {
            const double s = wave_sum_dpp(active ? (double)Lb.template at<j>() : 0.0);
            if (lane == 63 && cnt > 0) wave_grad[(size_t)wave * gstride + (size_t)nT * n + j] = s;
        }
    });
}

template <int NCH, int FCAP, typename real, bool lognormal>
void launch_spot_adjoint_ss(const StepArgs& a, const AdjArgs& g, hipStream_t stream) {
    const int64_t blocks = ((int64_t)a.P + kStepThreads - 1) / kStepThreads;
    hipLaunchKernelGGL((lmm_spot_adjoint_kernel<NCH, FCAP, real, lognormal>), dim3((unsigned)blocks),
                       dim3(kStepThreads), 0, stream, a, g, (const real*)a.rec, a.delta, a.times, a.first_alive,
                       g.kcur_at, a.products, a.periods, (const real*)a.dw, (const real*)a.ckpt, a.numeraire,
                       g.weight, (const real*)g.fsig, (const real*)g.vterm, (real*)g.seg, g.wave_grad);
    HIP_CHECK(hipGetLastError());
}

template <int NCH, int FCAP, typename real>
void launch_spot_adjoint_cap(const StepArgs& a, const AdjArgs& g, hipStream_t stream) {
    if (a.lognormal) launch_spot_adjoint_ss<NCH, FCAP, real, true>(a, g, stream);
    else launch_spot_adjoint_ss<NCH, FCAP, real, false>(a, g, stream);
}

template <int NCH, typename real>
void launch_spot_adjoint_n(const StepArgs& a, const AdjArgs& g, hipStream_t stream) {
    if (a.F <= 1) launch_spot_adjoint_cap<NCH, 1, real>(a, g, stream);
    else if (a.F <= 3) launch_spot_adjoint_cap<NCH, 3, real>(a, g, stream);
    else if (a.F <= 5) launch_spot_adjoint_cap<NCH, 5, real>(a, g, stream);
    else if (a.F <= 10) launch_spot_adjoint_cap<NCH, 10, real>(a, g, stream);
    else throw FmError(FM_ERR_UNSUPPORTED, "number of factors above 10 is not supported by the adjoint kernel");
}

// The reverse sweep of the spot pricing pass; a.ckpt holds its checkpointed states.
template <typename real>
void launch_spot_adjoint(const StepArgs& a, const AdjArgs& g, hipStream_t stream) {
    // the chunk counts of the pricing pass (lmm_spot.hpp)
    constexpr int W = chunk_width<real>();
    constexpr int N2 = W == 8 ? 3 : 2, N3 = W == 8 ? 5 : 3, N4 = W == 8 ? 10 : 5;
    if (a.n <= W) launch_spot_adjoint_n<1, real>(a, g, stream);
    else if (a.n <= W * N2) launch_spot_adjoint_n<N2, real>(a, g, stream);
    else if (a.n <= W * N3) launch_spot_adjoint_n<N3, real>(a, g, stream);
    else if (a.n <= 80) launch_spot_adjoint_n<N4, real>(a, g, stream);
    else throw FmError(FM_ERR_UNSUPPORTED, "more than 80 forward rates are not supported by the adjoint kernel");
}

} // namespace fm

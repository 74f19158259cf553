// CPU oracle for the LIBOR Market Model Monte Carlo hot path.
//
// TEST INFRASTRUCTURE ONLY. This file is a plain, sequential restatement of the
// reference algorithm (finmath-lib, Java) used to check the HIP implementation.
// Only tests/, the smoke check and the CPU-baseline leg of bench.py may load it.
// The product path never links, loads or calls anything in this directory.
//
// Every function names the reference file:line it restates. Paths are relative to
// src/main/java/net/finmath/ of the reference.
//
// Floating point: compiled with -O2 -ffp-contract=off (no fused multiply-add, as
// the JVM never fuses), sequential loops in the reference's order, Kahan averages.
//
// Parity status of each part (what pins it) is listed in the header comment of
// each section. Anything without a pin says "parity unpinned".

#include <cstdint>
#include <cstring>
#include <cmath>
#include <vector>
#include <memory>
#include <algorithm>
#include <string>
#include <stdexcept>
#include <map>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "orc_rv.hpp"

extern "C" {

// ---------------------------------------------------------------------------
// MT19937 with the seeding of Apache commons-math3 3.6.1 MersenneTwister(long)
// (third-party dependency of the reference: pom.xml:22,157-160; wrapped by
// randomnumbers/MersenneTwister.java:26-29,42-51).
//
// Published algorithm restated: Matsumoto & Nishimura (1998), mt19937ar.c of
// 2002-01-26 (init_genrand, init_by_array, genrand_int32). commons-math3's
// setSeed(long s) passes the key {(int)(s>>>32), (int)(s & 0xffffffff)} to
// init_by_array; nextDouble() composes ((long)next(26) << 26 | next(26)) * 2^-52,
// i.e. two 32-bit outputs per double, the first word giving the high bits.
// Pins: numpy's MT19937 (init_by_array seeding, raw 32-bit output) for the word
// stream; the commons seeding key and 26+26 composition are "parity unpinned"
// (no reference test pins them; restated from the published commons source).
// ---------------------------------------------------------------------------

struct orc_mt_state { uint32_t mt[624]; int mti; };

static void mt_init_genrand(orc_mt_state* s, uint32_t seed) {
    s->mt[0] = seed;
    for (int i = 1; i < 624; i++)
        s->mt[i] = 1812433253u * (s->mt[i - 1] ^ (s->mt[i - 1] >> 30)) + (uint32_t)i;
    s->mti = 624;
}

static void mt_init_by_array(orc_mt_state* s, const uint32_t* key, int len) {
    mt_init_genrand(s, 19650218u);
    int i = 1, j = 0;
    for (int k = (624 > len ? 624 : len); k; k--) {
        s->mt[i] = (s->mt[i] ^ ((s->mt[i - 1] ^ (s->mt[i - 1] >> 30)) * 1664525u)) + key[j] + (uint32_t)j;
        i++; j++;
        if (i >= 624) { s->mt[0] = s->mt[623]; i = 1; }
        if (j >= len) j = 0;
    }
    for (int k = 623; k; k--) {
        s->mt[i] = (s->mt[i] ^ ((s->mt[i - 1] ^ (s->mt[i - 1] >> 30)) * 1566083941u)) - (uint32_t)i;
        i++;
        if (i >= 624) { s->mt[0] = s->mt[623]; i = 1; }
    }
    s->mt[0] = 0x80000000u;
    s->mti = 624;
}

static uint32_t mt_next(orc_mt_state* s) {
    if (s->mti >= 624) {
        for (int k = 0; k < 624; k++) {
            uint32_t y = (s->mt[k] & 0x80000000u) | (s->mt[(k + 1) % 624] & 0x7fffffffu);
            s->mt[k] = s->mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        s->mti = 0;
    }
    uint32_t y = s->mt[s->mti++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

// commons-math3 MersenneTwister.setSeed(long): key {(int)(seed>>>32), (int)seed}.
static void mt_seed_commons(orc_mt_state* s, int64_t seed) {
    uint64_t u = (uint64_t)seed;
    uint32_t key[2] = { (uint32_t)(u >> 32), (uint32_t)(u & 0xffffffffu) };
    mt_init_by_array(s, key, 2);
}

// commons-math3 BitsStreamGenerator.nextDouble(): 26 high bits then 26 low bits.
static double mt_next_double(orc_mt_state* s) {
    uint64_t hi = (uint64_t)(mt_next(s) >> 6);
    uint64_t lo = (uint64_t)(mt_next(s) >> 6);
    return (double)((hi << 26) | lo) * 0x1.0p-52;
}

// Seeded state (624 words, before the first twist) for a reference seed.
int orc_mt_seeded_state(int64_t seed, uint32_t* out624) {
    orc_mt_state s; mt_seed_commons(&s, seed);
    std::memcpy(out624, s.mt, sizeof(s.mt));
    return 0;
}

// Tempered output words [offset, offset+count) of the stream for a seed.
int orc_mt_words(int64_t seed, uint64_t offset, uint64_t count, uint32_t* out) {
    orc_mt_state s; mt_seed_commons(&s, seed);
    for (uint64_t i = 0; i < offset; i++) (void)mt_next(&s);
    for (uint64_t i = 0; i < count; i++) out[i] = mt_next(&s);
    return 0;
}

// Same generator from an explicit init_by_array key (for published known answers).
int orc_mt_words_by_array(const uint32_t* key, int len, uint64_t count, uint32_t* out) {
    orc_mt_state s; mt_init_by_array(&s, key, len);
    for (uint64_t i = 0; i < count; i++) out[i] = mt_next(&s);
    return 0;
}

// Continue the generator from an explicit 624-word state and read position (the
// state layout of numpy's MT19937 bit generator, used to replay its test set).
int orc_mt_words_from_state(const uint32_t* state624, int pos, uint64_t count, uint32_t* out) {
    orc_mt_state s; std::memcpy(s.mt, state624, sizeof(s.mt)); s.mti = pos;
    for (uint64_t i = 0; i < count; i++) out[i] = mt_next(&s);
    return 0;
}

int orc_mt_doubles(int64_t seed, uint64_t offset_doubles, uint64_t count, double* out) {
    orc_mt_state s; mt_seed_commons(&s, seed);
    for (uint64_t i = 0; i < 2 * offset_doubles; i++) (void)mt_next(&s);
    for (uint64_t i = 0; i < count; i++) out[i] = mt_next_double(&s);
    return 0;
}

// ---------------------------------------------------------------------------
// Wichura AS241 (PPND16) inverse normal CDF.
// Restates functions/NormalDistribution.java:67-162 (published algorithm:
// M. J. Wichura, Appl. Statist. 37(3), 1988). The coefficient table is the
// published one. Special case of the reference: p <= 0 or p >= 1 in the tail
// branch returns 0 (NormalDistribution.java:141-143).
// Pins: scipy.special.ndtri (relative 1e-15 over the central and tail ranges),
// symmetry icdf(1-p) = -icdf(p), icdf(0.5) = 0, icdf(0) = 0 (reference quirk).
// ---------------------------------------------------------------------------
double orc_icdf(double p) {
    const double split1 = 0.425, split2 = 5.0, const1 = 0.180625, const2 = 1.6;
    static const double A[8] = {3.3871328727963666080e+00, 1.3314166789178437745e+02, 1.9715909503065514427e+03,
        1.3731693765509461125e+04, 4.5921953931549871457e+04, 6.7265770927008700853e+04,
        3.3430575583588128105e+04, 2.5090809287301226727e+03};
    static const double B[8] = {1.0, 4.2313330701600911252e+01, 6.8718700749205790830e+02, 5.3941960214247511077e+03,
        2.1213794301586595867e+04, 3.9307895800092710610e+04, 2.8729085735721942674e+04,
        5.2264952788528545610e+03};
    static const double C[8] = {1.42343711074968357734e+00, 4.63033784615654529590e+00, 5.76949722146069140550e+00,
        3.64784832476320460504e+00, 1.27045825245236838258e+00, 2.41780725177450611770e-01,
        2.27238449892691845833e-02, 7.74545014278341407640e-04};
    static const double D[8] = {1.0, 2.05319162663775882187e+00, 1.67638483018380384940e+00, 6.89767334985100004550e-01,
        1.48103976427480074590e-01, 1.51986665636164571966e-02, 5.47593808499534494600e-04,
        1.05075007164441684324e-09};
    static const double E[8] = {6.65790464350110377720e+00, 5.46378491116411436990e+00, 1.78482653991729133580e+00,
        2.96560571828504891230e-01, 2.65321895265761230930e-02, 1.24266094738807843860e-03,
        2.71155556874348757815e-05, 2.01033439929228813265e-07};
    static const double Fc[8] = {1.0, 5.99832206555887937690e-01, 1.36929880922735805310e-01, 1.48753612908506148525e-02,
        7.86869131145613259100e-04, 1.84631831751005468180e-05, 1.42151175831644588870e-07,
        2.04426310338993978564e-15};
    // Horner forms exactly as the published algorithm: ((c7 r + c6) r + ...) r + c0.
    auto horner = [](const double* c, double r) {
        double v = c[7];
        for (int i = 6; i >= 0; i--) v = v * r + c[i];
        return v;
    };
    double q = p - 0.5;
    if (std::fabs(q) <= split1) {
        double r = const1 - q * q;
        return q * horner(A, r) / horner(B, r);
    }
    double r = (q < 0.0) ? p : 1.0 - p;
    if (r <= 0.0) return 0.0;
    r = std::sqrt(-std::log(r));
    double v;
    if (r <= split2) { r -= const2; v = horner(C, r) / horner(D, r); }
    else { r -= split2; v = horner(E, r) / horner(Fc, r); }
    if (q < 0.0) v = -v;
    return v;
}

void orc_icdf_array(const double* p, int64_t n, double* out) {
    for (int64_t i = 0; i < n; i++) out[i] = orc_icdf(p[i]);
}

// ---------------------------------------------------------------------------
// TimeDiscretizationFromArray tick rounding (time/TimeDiscretizationFromArray.java:39,387-389).
// Pins: closed form (k*0.5 survives rounding to the 1/8760 tick exactly).
// ---------------------------------------------------------------------------
double orc_round_time(double t) {
    const double tick = 1.0 / (365.0 * 24.0);
    return std::rint(t / tick) * tick;
}

// java_binary_search: orc_rv.hpp
using orc::java_binary_search;

int orc_time_index(const double* times, int n, double t) {
    return java_binary_search(times, n, orc_round_time(t));
}

// ---------------------------------------------------------------------------
// BrownianMotionFromMersenneRandomNumbers.doGenerateBrownianMotion
// (montecarlo/BrownianMotionFromMersenneRandomNumbers.java:141-191): loop path,
// then time, then factor; dW = ICDF(nextDouble) * sqrt(dt_t).
// Output layout here: [nT][F][P] as the reference's brownianIncrementsArray.
// Pins: composition of the pinned MT words and AS241; statistical moments.
// ---------------------------------------------------------------------------
int orc_brownian(const double* times, int nT, int F, int P, int32_t seed, double* out) {
    orc_mt_state s; mt_seed_commons(&s, (int64_t)seed);
    std::vector<double> sq(nT);
    for (int t = 0; t < nT; t++) sq[t] = std::sqrt(times[t + 1] - times[t]);
    for (int p = 0; p < P; p++)
        for (int t = 0; t < nT; t++)
            for (int k = 0; k < F; k++) {
                double u = mt_next_double(&s);
                out[((size_t)t * F + k) * P + p] = orc_icdf(u) * sq[t];
            }
    return 0;
}

// ---------------------------------------------------------------------------
// Symmetric eigen decomposition by cyclic Jacobi rotations (textbook method,
// Golub & Van Loan 8.5). The reference calls commons-math3 EigenDecomposition
// (functions/LinearAlgebra.java:406); eigenvectors of distinct eigenvalues agree
// up to sign, which the factor reduction fixes. Pins: numpy.linalg.eigh.
// evecs: column c is the eigenvector of evals[c]; row-major [n][n].
// ---------------------------------------------------------------------------
int orc_eigen_sym(const double* a_in, int n, double* evals, double* evecs) {
    std::vector<double> a(a_in, a_in + (size_t)n * n);
    for (int i = 0; i < n; i++) for (int j = 0; j < n; j++) evecs[i * n + j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0;
        for (int i = 0; i < n; i++) for (int j = i + 1; j < n; j++) off += a[i * n + j] * a[i * n + j];
        if (off < 1e-300) break;
        for (int p = 0; p < n; p++) {
            for (int q = p + 1; q < n; q++) {
                double apq = a[p * n + q];
                if (std::fabs(apq) < 1e-300) continue;
                double app = a[p * n + p], aqq = a[q * n + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; k++) {
                    double akp = a[k * n + p], akq = a[k * n + q];
                    a[k * n + p] = c * akp - s * akq;
                    a[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    double apk = a[p * n + k], aqk = a[q * n + k];
                    a[p * n + k] = c * apk - s * aqk;
                    a[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    double vkp = evecs[k * n + p], vkq = evecs[k * n + q];
                    evecs[k * n + p] = c * vkp - s * vkq;
                    evecs[k * n + q] = s * vkp + c * vkq;
                }
            }
        }
    }
    for (int i = 0; i < n; i++) evals[i] = a[i * n + i];
    return 0;
}

// LinearAlgebra.getFactorMatrixUsingCommonsMath (functions/LinearAlgebra.java:392-445):
// sort eigenvalues descending, sign so that v[0] > 0, scale by sqrt(max(lambda,0)/|v|^2).
static void factor_matrix(const double* corr, int n, int m, double* out) {
    std::vector<double> ev(n), V((size_t)n * n);
    orc_eigen_sym(corr, n, ev.data(), V.data());
    std::vector<int> idx(n);
    for (int i = 0; i < n; i++) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return ev[a] > ev[b]; });
    for (int f = 0; f < m; f++) {
        int c = idx[f];
        double lambda = ev[c];
        double sign = V[0 * n + c] > 0.0 ? 1.0 : -1.0;
        double norm2 = 0.0;
        for (int r = 0; r < n; r++) norm2 += V[r * n + c] * V[r * n + c];
        lambda = std::max(lambda, 0.0);
        for (int r = 0; r < n; r++) out[r * m + f] = sign * std::sqrt(lambda / norm2) * V[r * n + c];
    }
}

// LinearAlgebra.factorReductionUsingCommonsMath (functions/LinearAlgebra.java:454-482):
// PCA, renormalise rows, PCA again of F F^T.
int orc_factor_reduction(const double* corr, int n, int m, double* out) {
    std::vector<double> f((size_t)n * m);
    factor_matrix(corr, n, m, f.data());
    for (int r = 0; r < n; r++) {
        double ss = 0.0;
        for (int k = 0; k < m; k++) ss += f[r * m + k] * f[r * m + k];
        if (ss != 0.0) { for (int k = 0; k < m; k++) f[r * m + k] = f[r * m + k] / std::sqrt(ss); }
        else { for (int k = 0; k < m; k++) f[r * m + k] = 1.0; }
    }
    std::vector<double> red((size_t)n * n);
    for (int r = 0; r < n; r++)
        for (int c = 0; c < n; c++) {
            double s = 0.0;
            for (int k = 0; k < m; k++) s += f[r * m + k] * f[c * m + k];
            red[r * n + c] = s;
        }
    factor_matrix(red.data(), n, m, out);
    return 0;
}

// LIBORCorrelationModelExponentialDecay.initialize
// (montecarlo/interestrate/models/covariance/LIBORCorrelationModelExponentialDecay.java:102-134):
// rho_ij = exp(-max(a,0)|T_i-T_j|), factor reduction, correlation reset to F F^T with unit diagonal.
int orc_corr_expdecay(const double* tenor, int n, double a, int m, double* factors, double* corr_out) {
    a = std::max(a, 0.0);
    std::vector<double> rho((size_t)n * n);
    for (int r = 0; r < n; r++) for (int c = 0; c < n; c++) rho[r * n + c] = std::exp(-a * std::fabs(tenor[r] - tenor[c]));
    orc_factor_reduction(rho.data(), n, m, factors);
    if (corr_out) {
        for (int c1 = 0; c1 < n; c1++) {
            for (int c2 = 0; c2 < c1; c2++) {
                double s = 0.0;
                for (int k = 0; k < m; k++) s += factors[c1 * m + k] * factors[c2 * m + k];
                corr_out[c1 * n + c2] = s; corr_out[c2 * n + c1] = s;
            }
            corr_out[c1 * n + c1] = 1.0;
        }
    }
    return 0;
}

// LIBORVolatilityModelFourParameterExponentialForm.getVolatility
// (.../covariance/LIBORVolatilityModelFourParameterExponentialForm.java:168-195), Scalar
// arithmetic: a.addProduct(b,tau) = b*tau + a (stochastic/Scalar.java:350-352),
// .mult(c.mult(-tau).exp()).add(d). tau <= 0 gives 0.
int orc_vol_fourparam(const double* times, int nT, const double* tenor, int n,
                      double a, double b, double c, double d, double* vol) {
    for (int i = 0; i < nT; i++)
        for (int j = 0; j < n; j++) {
            double tau = tenor[j] - times[i];
            vol[i * n + j] = (tau <= 0) ? 0.0 : (b * tau + a) * std::exp(c * (-tau)) + d;
        }
    return 0;
}

// ForwardCurveInterpolation (LINEAR on VALUE, CONSTANT extrapolation) value at a time
// (marketdata/model/curves/ForwardCurveInterpolation.java:298-347,
//  interpolation/RationalFunctionInterpolation.java:184-226). At a node the node value.
double orc_forward_curve(const double* nodeT, const double* nodeF, int nn, double t) {
    int idx = java_binary_search(nodeT, nn, t);
    if (idx >= 0) return nodeF[idx];
    int iv = -idx - 2;
    if (iv < 0) return nodeF[0];
    if (iv > nn - 2) return nodeF[nn - 1];
    // linear rational function: c0 + c1*(x-x_i)
    double c0 = nodeF[iv], c1 = (nodeF[iv + 1] - nodeF[iv]) / (nodeT[iv + 1] - nodeT[iv]);
    double x = t - nodeT[iv];
    double num = 0.0, pw = 1.0;
    num += c0 * pw; pw *= x;
    num += c1 * pw;
    return num;
}

// DiscountCurveFromForwardCurve.getDiscountFactor
// (marketdata/model/curves/DiscountCurveFromForwardCurve.java:115-143), timeScaling 1.
double orc_df_from_forward_curve(const double* nodeT, const double* nodeF, int nn, double paymentOffset, double maturity) {
    double time = 0.0, df = 1.0;
    while (time < maturity) {
        df /= 1.0 + orc_forward_curve(nodeT, nodeF, nn, time) * std::min(paymentOffset, maturity - time) * 1.0;
        time += paymentOffset;
    }
    return df;
}

} // extern "C"

// ---------------------------------------------------------------------------
// RandomVariable semantics: a value that is either deterministic or an array of
// path realisations, with the type priority of Scalar (0) and
// RandomVariableFromDoubleArray (1). Each op restates the one of
// stochastic/Scalar.java:200-371 or montecarlo/RandomVariableFromDoubleArray.java
// (line numbers at each op), including which operand's implementation runs.
// ---------------------------------------------------------------------------
namespace orc {

// (RV and its operations: orc_rv.hpp)


// ---------------------------------------------------------------------------
// LIBOR market model + Euler scheme.
// LIBORMarketModelFromCovarianceModel (montecarlo/interestrate/models/
// LIBORMarketModelFromCovarianceModel.java, "LMM" below) with
// LIBORCovarianceModelFromVolatilityAndCorrelation, driven by
// EulerSchemeFromProcessModel.doPrecalculateProcess (montecarlo/process/
// EulerSchemeFromProcessModel.java:170-326, "Euler" below).
// Pins: closed forms (zero volatility gives deterministic forwards and
// numeraires; bonds reprice the discount curve by construction of the numeraire
// adjustment), brute force on tiny inputs, invariants in tests/.
// ---------------------------------------------------------------------------
enum { SCHEME_EULER = 0, SCHEME_PC = 1, SCHEME_EULER_FUNCTIONAL = 2, SCHEME_PC_FUNCTIONAL = 3 };
enum { MEASURE_SPOT = 0, MEASURE_TERMINAL = 1 };
enum { SS_NORMAL = 0, SS_LOGNORMAL = 1 };

struct Model {
    int nT, n, F, P;
    std::vector<double> times, tenor, L0, vol, fac, pd;
    double capv;
    int scheme, measure, ss;
    bool has_disc;
    std::vector<double> fc_t, fc_f;               // forward curve nodes (analytic interpolation adjustment)
    double fc_offset = 0.0;                       // the curve's payment offset
    int interp = 1;                               // InterpolationMethod ordinal: LINEAR, LOG_LINEAR_UNCORRECTED, LOG_LINEAR_CORRECTED
    bool has_lv = false;                          // local-volatility wrapper of the covariance model
    double lv_alpha = 0.0;
    std::vector<double> lvb;                      // [nT][n] offsets
    bool has_sv = false;                          // stochastic volatility scaling of the loadings
    double sv_nu = 0.0, sv_rho = 0.0;
    int32_t sv_seed = 0;
    int sv_F = 2, sv_a = 0, sv_b = 1;
    int sv_kind = 0;                              // 0: log-normal scaling, 1: Heston variance
    double sv_kappa = 0.0, sv_theta = 0.0, sv_xi = 0.0;
    std::vector<RV> svs;                          // S(t_i), i < nT
    std::vector<std::vector<RV>> proc;            // [nT+1][n]
    std::vector<std::vector<RV>> dW;              // [nT][F]
    std::map<int, RV> numeraires;                 // unadjusted at tenor index
    std::vector<double> adjchain;                 // deterministic numeraire adjustment at tenor dates
    std::vector<std::string> err;

    int libor_period_index(double t) const { return java_binary_search(tenor.data(), n + 1, orc_round_time(t)); }
    int process_time_index(double t) const { return java_binary_search(times.data(), nT + 1, orc_round_time(t)); }

    RV lambda(int ti, int j, int k) const { return mult(scalar(vol[ti * n + j]), fac[j * F + k]); }   // FromVolAndCorr:48-57
    RV variance_(int ti, int j) const { return mult(mult(scalar(vol[ti * n + j]), scalar(vol[ti * n + j])), 1.0); } // :83-93

    // Factor loadings of component j at time index ti on the realisation `real`: the
    // covariance model's, times the local-volatility factor of L_j when a wrapper is set:
    // DisplacedLocalVolatilityModel L + d (DisplacedLocalVolatilityModel.java:175-186),
    // BlendedLocalVolatilityModel L - L a + a f0 (BlendedLocalVolatilityModel.java:265-283).
    std::vector<RV> loadings(int ti, int j, const std::vector<RV>& real) const {
        std::vector<RV> l(F);
        for (int k = 0; k < F; k++) l[k] = lambda(ti, j, k);
        if (has_lv) {
            RV f = real[j];
            if (lv_alpha != 0.0) f = sub(f, mult(f, lv_alpha));
            f = add(f, lvb[(size_t)ti * n + j]);
            for (int k = 0; k < F; k++) l[k] = mult(l[k], f);
        }
        if (has_sv)                                                              // StochasticVolatility.java:247-252
            for (int k = 0; k < F; k++) l[k] = mult(l[k], svs[ti]);
        return l;
    }
    // Variance of the log-coordinate drift adjustment: sigma_j sigma_j rho_jj of the
    // covariance model, or, under a local-volatility wrapper, the sum of the squared
    // loadings (AbstractLIBORCovarianceModel.getCovariance, AbstractLIBORCovarianceModel.java:79-92).
    RV variance_of(int ti, int j, const std::vector<RV>& l) const {
        if (!has_lv && !has_sv) return variance_(ti, j);
        RV c = mult(l[0], l[0]);
        for (int k = 1; k < F; k++) c = addProduct(c, l[k], l[k]);
        return c;
    }

    RV transform(const RV& y) const {                                            // LMM:1200-1212
        RV v = y;
        if (ss == SS_LOGNORMAL) v = expv(v);
        if (!std::isinf(capv)) v = cap(v, capv);
        return v;
    }
    RV transform_inv(const RV& v) const { return ss == SS_LOGNORMAL ? logv(v) : v; }   // LMM:1215-1223

    int first_alive(int ti) const {                                              // LMM:1126-1130
        int f = libor_period_index(times[ti]) + 1;
        if (f < 0) f = -f - 1 + 1;
        return f;
    }

    std::vector<RV> drift(int ti, const std::vector<RV>& real) const {            // LMM:1125-1191
        std::vector<RV> d(n);
        std::vector<bool> has(n, false);
        int first = first_alive(ti);
        RV zero = scalar(0.0);
        for (int j = first; j < n; j++) { d[j] = zero; has[j] = true; }
        std::vector<RV> S(F, zero);
        std::vector<std::vector<RV>> lam(n);
        for (int j = first; j < n; j++) lam[j] = loadings(ti, j, real);
        if (measure == MEASURE_SPOT) {
            for (int j = first; j < n; j++) {
                double dl = tenor[j + 1] - tenor[j];
                RV w = discount(scalar(dl), real[j], dl);
                if (ss == SS_LOGNORMAL) w = mult(w, real[j]);
                for (int k = 0; k < F; k++) S[k] = addProduct(S[k], w, lam[j][k]);
                RV r = d[j];
                for (int k = 0; k < F; k++) r = addProduct(r, S[k], lam[j][k]);   // RandomVariable.java:659-666
                d[j] = r;
            }
        } else {
            for (int j = n - 1; j >= first; j--) {
                double dl = tenor[j + 1] - tenor[j];
                RV w = discount(scalar(-dl), real[j], dl);
                if (ss == SS_LOGNORMAL) w = mult(w, real[j]);
                RV r = d[j];
                for (int k = 0; k < F; k++) r = addProduct(r, S[k], lam[j][k]);
                d[j] = r;
                for (int k = 0; k < F; k++) S[k] = addProduct(S[k], w, lam[j][k]);
            }
        }
        if (ss == SS_LOGNORMAL)
            for (int j = first; j < n; j++) d[j] = addProduct(d[j], variance_of(ti, j, lam[j]), -0.5);
        std::vector<RV> out(n);
        for (int j = 0; j < n; j++) if (has[j]) out[j] = d[j];
        (void)has;
        return out;
    }

    // The scaling process of LIBORCovarianceModelStochasticVolatility (:174-236): S = exp(X),
    // functional Euler (Euler:234-256) with drift nu^2 (-0.5) and loadings (rho nu,
    // sqrt((rho^2 - 1)(-1)) nu) on factors a, b of its own Brownian motion; S(0) = exp(0.0).
    void simulate_scaling() {
        std::vector<double> bm((size_t)nT * sv_F * P);
        orc_brownian(times.data(), nT, sv_F, P, sv_seed, bm.data());
        auto inc = [&](int t, int k) {
            std::vector<double> v(bm.begin() + ((size_t)t * sv_F + k) * P, bm.begin() + ((size_t)t * sv_F + k + 1) * P);
            return rvarr(std::move(v));
        };
        const RV drift = mult(mult(scalar(sv_nu), scalar(sv_nu)), -0.5);
        const RV l0 = mult(scalar(sv_rho), scalar(sv_nu));
        const RV l1 = mult(scalar(std::sqrt((sv_rho * sv_rho - 1.0) * -1.0)), scalar(sv_nu));
        svs.assign(nT, RV());
        RV S = expv(scalar(0.0));
        for (int t = 0; t < nT; t++) {
            svs[t] = S;
            RV X = logv(S);
            X = addProduct(X, drift, times[t + 1] - times[t]);
            X = addProduct(X, l0, inc(t, sv_a));
            X = addProduct(X, l1, inc(t, sv_b));
            S = expv(X);
        }
    }

    // The variance process of LIBORCovarianceModelStochasticHestonVolatility (:172-256): V(0) = 1,
    // identity transform, drift (V - theta)(-kappa), loading sqrt(max(V, 0)) xi on factor a;
    // the loadings are scaled by sqrt(max(V, 0)).
    void simulate_heston_scaling() {
        std::vector<double> bm((size_t)nT * sv_F * P);
        orc_brownian(times.data(), nT, sv_F, P, sv_seed, bm.data());
        const RV kappa = scalar(sv_kappa), theta = scalar(sv_theta), xi = scalar(sv_xi);
        svs.assign(nT, RV());
        RV V = scalar(1.0);
        for (int t = 0; t < nT; t++) {
            svs[t] = sqrtv(floor_(V, 0.0));
            const RV drift = mult(sub(V, theta), mult(kappa, -1.0));
            const RV load = mult(sqrtv(floor_(V, 0.0)), xi);
            std::vector<double> v(bm.begin() + ((size_t)t * sv_F + sv_a) * P, bm.begin() + ((size_t)t * sv_F + sv_a + 1) * P);
            RV X = V;
            X = addProduct(X, drift, times[t + 1] - times[t]);
            X = addProduct(X, load, rvarr(std::move(v)));
            V = X;
        }
    }

    void simulate(int32_t seed, int threads) {
        if (has_sv) {
            if (sv_kind == 1) simulate_heston_scaling();
            else simulate_scaling();
        }
        // Brownian increments (BrownianMotionFromMersenneRandomNumbers.java:141-191)
        std::vector<double> bm((size_t)nT * F * P);
        orc_brownian(times.data(), nT, F, P, seed, bm.data());
        dW.assign(nT, std::vector<RV>(F));
        for (int t = 0; t < nT; t++)
            for (int k = 0; k < F; k++) {
                std::vector<double> v(bm.begin() + ((size_t)t * F + k) * P, bm.begin() + ((size_t)t * F + k + 1) * P);
                dW[t][k] = rvarr(std::move(v));
            }
        proc.assign(nT + 1, std::vector<RV>(n));
        std::vector<RV> state(n);
        for (int j = 0; j < n; j++) {                                            // Euler:187-192, LMM:1081-1093
            double y0 = (ss == SS_LOGNORMAL) ? std::log(std::max(L0[j], 0.0)) : L0[j];
            state[j] = scalar(y0);
            proc[0][j] = transform(state[j]);
        }
        for (int ti = 1; ti <= nT; ti++) {                                       // Euler:202-318
            double dt = times[ti] - times[ti - 1];
            int first = first_alive(ti - 1);
            std::vector<RV> d = drift(ti - 1, proc[ti - 1]);
            std::vector<RV> next(n);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic) num_threads(threads > 0 ? threads : 1)
#endif
            for (int j = 0; j < n; j++) {
                if (j < first) continue;
                RV cur = (scheme == SCHEME_EULER_FUNCTIONAL || scheme == SCHEME_PC_FUNCTIONAL) ? transform_inv(proc[ti - 1][j]) : state[j];
                cur = addProduct(cur, d[j], dt);                                 // Euler:248
                const std::vector<RV> lam = loadings(ti - 1, j, proc[ti - 1]);       // Euler:240
                for (int k = 0; k < F; k++) cur = addProduct(cur, lam[k], dW[ti - 1][k]);   // Euler:252
                state[j] = cur;
                next[j] = transform(cur);                                        // Euler:255
            }
            for (int j = 0; j < n; j++) proc[ti][j] = (j < first) ? proc[ti - 1][j] : next[j];   // Euler:279-290
            if (scheme == SCHEME_PC || scheme == SCHEME_PC_FUNCTIONAL) {         // Euler:292-314
                std::vector<RV> dp = drift(ti - 1, proc[ti]);
                for (int j = first; j < n; j++) {
                    RV adj = mult(divd(sub(dp[j], d[j]), 2.0), dt);
                    state[j] = add(state[j], adj);
                    proc[ti][j] = transform(state[j]);
                }
            }
        }
        // deterministic numeraire adjustment chain (LMM:915-944)
        if (has_disc) {
            adjchain.assign(n + 1, 0.0);
            RV a = scalar(pd[0]);
            adjchain[0] = a.val;
            for (int i = 0; i < n; i++) {
                double dfp = pd[i], dfn = pd[i + 1], ts = tenor[i + 1] - tenor[i];
                RV fr = scalar((dfp / dfn - 1.0) / ts);
                a = discount(a, fr, ts);
                adjchain[i + 1] = a.val;
            }
        }
    }

    RV libor_at(int ti, int j) const { return proc[ti][j]; }                   // LMM:1307-1312

    // ForwardCurveInterpolation.getForward(model, fixingTime, paymentOffset)
    // (marketdata/model/curves/ForwardCurveInterpolation.java:337-347)
    double curve_forward(double fixing, double payment_offset) const {
        if (fc_t.empty()) throw std::runtime_error("interpolated forwards need the model's forward curve");
        double f = orc_forward_curve(fc_t.data(), fc_f.data(), (int)fc_t.size(), fixing);
        if (payment_offset != fc_offset)
            f = (std::exp(std::log(1 + f * fc_offset) * payment_offset / fc_offset) - 1.0) / payment_offset;
        return f;
    }

    // getInterpolationDriftAdjustmentEvaluated (LMM:1432-1499): trapezoid integral over the
    // simulation times up to index e of |lambda_j(t_s)|^2 / (1 + L_j(t_s) dT)^2 (times L_j^2
    // log-normal), with L_j read at most at its fixing index.
    RV interpolation_drift_adjustment(int e, int j) const {
        const double tdt = tenor[j + 1] - tenor[j];
        int fix = process_time_index(tenor[j]);
        if (fix < 0) fix = -fix - 2;
        RV adj = scalar(0.0);
        RV prev;
        std::vector<RV> real(n);
        for (int s_ = 0; s_ <= e; s_++) {
            if (s_ >= nT) throw std::runtime_error("loading beyond the simulation grid");
            real[j] = libor_at(s_ == 0 ? 0 : std::min(s_, fix), j);
            const std::vector<RV> l = loadings(s_, j, real);
            RV g = scalar(0.0);
            for (int k = 0; k < F; k++) g = add(g, mult(l[k], l[k]));
            RV den = add(mult(real[j], tdt), 1.0);
            g = div(g, mult(den, den));
            if (ss == SS_LOGNORMAL) g = mult(g, mult(real[j], real[j]));
            if (s_ > 0) adj = add(adj, mult(add(g, prev), 0.5 * (times[s_] - times[s_ - 1])));
            prev = g;
        }
        return adj;
    }

    // getOnePlusInterpolatedLIBORDt (LMM:1324-1381): the long forward of tenor period j at time
    // index ti (its fixing time index once fixed); LOG_LINEAR_*: its one-plus raised to the
    // power alpha = (T_{j+1} - S) / (T_{j+1} - T_j) (CORRECTED: minus the drift adjustment times
    // S dT (T_j - S) / 2 inside the exponential), LINEAR: the linear blend; times the curve's
    // short/interpolated ratio.
    RV one_plus_interpolated(int ti, double S, int j) const {
        const double ts = tenor[j], te = tenor[j + 1], tdt = te - ts;
        if (ts < times[ti]) {
            const int fi = process_time_index(ts);
            if (fi < 0) throw std::runtime_error("tenor discretization not part of time discretization");
            ti = std::min(ti, fi);
        }
        RV long_ = add(mult(libor_at(ti, j), tdt), 1.0);
        const double small = te - S, alpha = small / tdt;
        RV r;
        if (interp == 0) r = add(mult(long_, alpha), 1 - alpha);
        else if (interp == 1) r = expv(mult(logv(long_), alpha));
        else {
            const double coef = 0.5 * small * (ts - S);
            RV a = mult(interpolation_drift_adjustment(ti, j), coef);
            r = expv(sub(mult(logv(long_), alpha), a));
        }
        const double a_long = 1 + curve_forward(ts, tdt) * tdt;
        const double a_short = 1 + curve_forward(S, small) * small;
        const double a_interp = interp == 0 ? a_long * alpha + (1 - alpha) : std::exp(std::log(a_long) * alpha);
        return mult(r, a_short / a_interp);
    }

    // LMM.getForwardRate (LMM:1237-1305): on the grid, several periods, off-grid ends.
    RV forward_rate(double time, double start, double end) const {
        int si = libor_period_index(start), ei = libor_period_index(end);
        time = std::min(time, start);
        int ti = process_time_index(time);
        if (ti < 0) {
            ti = -ti - 2;
            if (time - times[ti] > times[ti + 1] - time) ti++;                   // ROUND_NEAREST
        }
        if (ei < 0) {                                                            // LMM:1258-1265
            const int pe = (-ei - 1) - 1;
            const double next_end = tenor[pe + 1];
            RV long_ = add(mult(forward_rate(time, start, next_end), next_end - start), 1.0);
            RV interp = one_plus_interpolated(ti, end, pe);
            return divd(sub(div(long_, interp), 1.0), end - start);
        }
        if (si < 0) {                                                            // LMM:1268-1280
            const int ps = (-si - 1) - 1;
            const double next_start = tenor[ps + 1];
            if (next_start > end) throw std::runtime_error("interpolation not possible");
            if (next_start == end) return divd(sub(one_plus_interpolated(ti, start, ps), 1.0), end - start);
            RV long_ = add(mult(forward_rate(time, next_start, end), end - next_start), 1.0);
            RV interp = one_plus_interpolated(ti, start, ps);
            return divd(sub(mult(long_, interp), 1.0), end - start);
        }
        if (si + 1 == ei) return libor_at(ti, si);
        RV acc; bool have = false;
        for (int pi = si; pi < ei; pi++) {
            double sub_ = tenor[pi + 1] - tenor[pi];
            RV l = libor_at(ti, pi);
            acc = have ? accrue(acc, l, sub_) : add(mult(l, sub_), 1.0);
            have = true;
        }
        return divd(sub(acc, 1.0), end - start);
    }

    // getNumerairetUnAdjustedAtLIBORIndex (LMM:1017-1074)
    RV numeraire_unadjusted_at(int k) {
        auto it = numeraires.find(k);
        if (it != numeraires.end()) return it->second;
        RV N;
        if (measure == MEASURE_TERMINAL) {
            int ti = process_time_index(tenor[k]);
            if (ti < 0) ti = -ti - 1;
            N = scalar(1.0);
            for (int li = k; li <= n - 1; li++) N = discount(N, libor_at(ti, li), tenor[li + 1] - tenor[li]);
        } else {
            if (k != 0) {
                int ti = process_time_index(tenor[k - 1]);
                if (ti < 0) ti = -ti - 1;
                double pl = tenor[k] - tenor[k - 1];
                N = accrue(numeraire_unadjusted_at(k - 1), libor_at(ti, k - 1), pl);
            } else {
                N = scalar(1.0);
            }
        }
        numeraires[k] = N;
        return N;
    }

    // getNumerairetUnAdjusted (LMM:968-1015): off the tenor grid, the numeraire of the next
    // tenor date (spot) or the bond chain from the next tenor date (terminal, which needs the
    // time on the simulation grid), discounted by the interpolated short forward.
    RV numeraire_unadjusted(double T) {
        const int k = libor_period_index(T);
        if (k >= 0) return numeraire_unadjusted_at(k);
        const int upper = -k - 1, lower = upper - 1;
        if (lower < 0) throw std::runtime_error("numeraire requested before the first tenor date");
        RV N;
        if (measure == MEASURE_TERMINAL) {
            N = scalar(1.0);
            for (int l = upper; l <= n - 1; l++) {
                const int ti = process_time_index(std::min(T, tenor[l]));
                if (ti < 0) throw std::runtime_error("time not on the simulation grid");
                N = discount(N, libor_at(ti, l), tenor[l + 1] - tenor[l]);
            }
        } else {
            N = numeraire_unadjusted_at(upper);
        }
        return discount(N, forward_rate(T, T, tenor[upper]), tenor[upper] - T);
    }

    // getNumeraireDefaultableZeroBondAsOfTimeZero (LMM:883-906): the adjustment chain on the
    // tenor grid, log-linear in between.
    double adjustment_at(double T) const {
        const int k = libor_period_index(T);
        if (k >= 0) return adjchain[k];
        const int prev = std::min(-k - 2, n + 1 - 2), next = prev + 1;
        const double tp = tenor[prev], tn = tenor[next];
        return adjchain[prev] * std::pow(adjchain[next] / adjchain[prev], (T - tp) / (tn - tp));
    }

    // getNumeraire (LMM:860-876).
    RV numeraire(double T) {
        RV N = numeraire_unadjusted(T);
        if (has_disc) {
            RV adj = scalar(adjustment_at(T));
            double zb = average(mult(invert(N), numeraire_unadjusted_at(0)));
            N = div(mult(N, zb), adj);
        }
        return N;
    }
    RV weights() const { return scalar(1.0 / P); }                               // Euler:184,317
};

} // namespace orc

extern "C" {

typedef struct orc_lmm_spec {
    int nT; const double* times;       // nT+1 simulation times
    int n; const double* tenor;        // n+1 tenor dates
    int F; int P; int32_t seed;
    const double* L0;                  // n initial forwards
    const double* vol;                 // [nT][n] sigma_j(t_i)
    const double* factors;             // [n][F]
    double cap;
    int scheme, measure, state_space;
    const double* pd;                  // n+1 discount-curve factors at tenor dates, or NULL
    int threads;
    double lv_alpha;                   // local-volatility factor (L - L alpha) + b, or L + b for alpha = 0
    const double* lv_offset;           // [nT][n] b, or NULL (no local-volatility wrapper)
    int fc_n;                          // forward curve nodes (interpolated forwards), 0 = none
    const double* fc_t;
    const double* fc_f;
    double fc_offset;                  // the curve's payment offset
    int interpolation;                 // InterpolationMethod ordinal (LMM:168)
    int sv_on;                         // stochastic volatility scaling
    double sv_nu, sv_rho;
    int32_t sv_seed;                   // its Brownian motion: seed, factors, the two factors used
    int sv_F, sv_a, sv_b;
    int sv_kind;                       // 0: log-normal scaling (nu, rho), 1: Heston (kappa, theta, xi)
    double sv_kappa, sv_theta, sv_xi;
} orc_lmm_spec;

static std::string g_err;
const char* orc_last_error(void) { return g_err.c_str(); }

void* orc_lmm_create(const orc_lmm_spec* s) {
    try {
        auto* m = new orc::Model();
        m->nT = s->nT; m->n = s->n; m->F = s->F; m->P = s->P;
        m->times.assign(s->times, s->times + s->nT + 1);
        m->tenor.assign(s->tenor, s->tenor + s->n + 1);
        m->L0.assign(s->L0, s->L0 + s->n);
        m->vol.assign(s->vol, s->vol + (size_t)s->nT * s->n);
        m->fac.assign(s->factors, s->factors + (size_t)s->n * s->F);
        m->capv = s->cap; m->scheme = s->scheme; m->measure = s->measure; m->ss = s->state_space;
        m->has_disc = s->pd != nullptr;
        if (m->has_disc) m->pd.assign(s->pd, s->pd + s->n + 1);
        if (s->fc_n > 0) {
            m->fc_t.assign(s->fc_t, s->fc_t + s->fc_n);
            m->fc_f.assign(s->fc_f, s->fc_f + s->fc_n);
            m->fc_offset = s->fc_offset;
        }
        m->interp = s->interpolation;
        m->has_sv = s->sv_on != 0;
        m->sv_nu = s->sv_nu; m->sv_rho = s->sv_rho; m->sv_seed = s->sv_seed;
        m->sv_F = s->sv_F; m->sv_a = s->sv_a; m->sv_b = s->sv_b;
        m->sv_kind = s->sv_kind; m->sv_kappa = s->sv_kappa; m->sv_theta = s->sv_theta; m->sv_xi = s->sv_xi;
        m->has_lv = s->lv_offset != nullptr;
        if (m->has_lv) {
            m->lv_alpha = s->lv_alpha;
            m->lvb.assign(s->lv_offset, s->lv_offset + (size_t)s->nT * s->n);
        }
        m->simulate(s->seed, s->threads);
        return m;
    } catch (const std::exception& e) { g_err = e.what(); return nullptr; }
}

void orc_lmm_destroy(void* h) { delete (orc::Model*)h; }

static void fill(const orc::RV& v, int P, double* out) {
    for (int p = 0; p < P; p++) out[p] = v.get((size_t)p);
}

int orc_lmm_process_value(void* h, int t, int j, double* out) {
    auto* m = (orc::Model*)h; fill(m->proc[t][j], m->P, out); return 0;
}
int orc_lmm_brownian_increment(void* h, int t, int k, double* out) {
    auto* m = (orc::Model*)h; fill(m->dW[t][k], m->P, out); return 0;
}
int orc_lmm_numeraire(void* h, double T, double* out) {
    try { auto* m = (orc::Model*)h; fill(m->numeraire(T), m->P, out); return 0; }
    catch (const std::exception& e) { g_err = e.what(); return -1; }
}
int orc_lmm_forward_rate(void* h, double time, double start, double end, double* out) {
    try { auto* m = (orc::Model*)h; fill(m->forward_rate(time, start, end), m->P, out); return 0; }
    catch (const std::exception& e) { g_err = e.what(); return -1; }
}
int orc_lmm_numeraire_unadjusted(void* h, int k, double* out) {
    auto* m = (orc::Model*)h; fill(m->numeraire_unadjusted_at(k), m->P, out); return 0;
}
double orc_lmm_numeraire_adjustment(void* h, int k) {
    auto* m = (orc::Model*)h; return m->has_disc ? m->adjchain[k] : 1.0;
}

// Swaption.getValue (montecarlo/interestrate/products/Swaption.java:137-199), evaluationTime 0.
// disc_adj: per-period discounting adjustment (Swaption.java:165-175), NULL = 1.0
// (the case where the discount curve is the forward curve's own).
int orc_lmm_swaption(void* h, double exercise, int nper, const double* fix, const double* pay,
                     const double* periodLengths, const double* strikes, double notional,
                     const double* disc_adj, double* value, double* error, double* per_path) {
    try {
        auto* m = (orc::Model*)h;
        using namespace orc;
        RV v = scalar(0.0);
        for (int per = nper - 1; per >= 0; per--) {
            double fd = fix[per], pd_ = pay[per], K = strikes[per];
            if (pd_ <= 0.0) break;
            double pl = periodLengths ? periodLengths[per] : pd_ - fd;
            RV libor = m->forward_rate(exercise, fd, pd_);
            RV payoff = mult(mult(sub(libor, K), pl), notional);
            double dd = std::max(fd, exercise);
            double da = disc_adj ? disc_adj[per] : 1.0;
            v = add(v, payoff);
            v = mult(discount(v, libor, pd_ - dd), da);
        }
        RV vals = floor_(v, 0.0);
        vals = mult(div(vals, m->numeraire(exercise)), m->weights());
        vals = div(mult(vals, m->numeraire(0.0)), m->weights());
        *value = average(vals);
        *error = stderror(vals);
        if (per_path) fill(vals, m->P, per_path);
        return 0;
    } catch (const std::exception& e) { g_err = e.what(); return -1; }
}

// SimpleSwap.getValue (montecarlo/interestrate/products/SimpleSwap.java:137-170).
int orc_lmm_swap(void* h, int nper, const double* fix, const double* pay, const double* strikes,
                 const double* notionals, int isPayFix, double* value, double* error, double* per_path) {
    try {
        auto* m = (orc::Model*)h;
        using namespace orc;
        RV vals = scalar(0.0);
        for (int per = 0; per < nper; per++) {
            double fd = fix[per], pd_ = pay[per], K = strikes[per], pl = pd_ - fd;
            RV fr = m->forward_rate(fd, fd, pd_);
            RV payoff = mult(mult(sub(fr, K), pl), notionals ? notionals[per] : 1.0);
            if (!isPayFix) payoff = mult(payoff, -1.0);
            payoff = mult(div(payoff, m->numeraire(pd_)), m->weights());
            vals = add(vals, payoff);
        }
        vals = div(mult(vals, m->numeraire(0.0)), m->weights());
        *value = average(vals);
        *error = stderror(vals);
        if (per_path) fill(vals, m->P, per_path);
        return 0;
    } catch (const std::exception& e) { g_err = e.what(); return -1; }
}

// Caplet.getValue, value unit VALUE (montecarlo/interestrate/products/Caplet.java:115-140).
int orc_lmm_caplet(void* h, double maturity, double periodLength, double strike, double daycount,
                   int isFloorlet, double* value, double* error) {
    try {
        auto* m = (orc::Model*)h;
        using namespace orc;
        double payDate = maturity + periodLength;
        RV fr = m->forward_rate(maturity, maturity, payDate);
        RV N = m->numeraire(payDate);
        RV vals = fr;
        if (!isFloorlet) vals = mult(floor_(sub(vals, strike), 0.0), daycount);
        else vals = mult(cap(sub(vals, strike), 0.0), -1.0 * daycount);
        vals = mult(div(vals, N), m->weights());
        vals = div(mult(vals, m->numeraire(0.0)), m->weights());
        *value = average(vals);
        *error = stderror(vals);
        return 0;
    } catch (const std::exception& e) { g_err = e.what(); return -1; }
}


// ---------------------------------------------------------------------------
// Bermudan swaption by regression (American Monte Carlo).
// BermudanSwaption.getValues (montecarlo/interestrate/products/BermudanSwaption.java:91-157),
// default basis getBasisFunctions (:215-252), the conditional expectation of
// MonteCarloConditionalExpectationRegression (montecarlo/conditionalexpectation/
// MonteCarloConditionalExpectationRegression.java:79-150).
//
// The reference solves X^T X x = X^T y with commons-math3's SingularValueDecomposition
// solver: singular values at or below tol = max(K * sigma_max * 2^-52, sqrt(2^-1022))
// are dropped (pseudo-inverse). X^T X is symmetric positive semi-definite, so its
// singular values are |eigenvalues| and the pseudo-inverse solve is
// x = sum_{|lambda_i| > tol} v_i (v_i . X^T y) / lambda_i, computed here from the
// Jacobi eigen decomposition above. Pins: numpy's pinv/lstsq on the same Gram
// matrices; the zero-volatility closed form (rank-1 Gram matrix). The last ulp of
// commons' SVD is "parity unpinned".
//
// basis 0: the reference's default basis (BermudanSwaption.java:215-252):
//   1, 1/(1 + L_short dt), its square, 1/(1 + L_long dt), its square, 1/N(T_e).
// basis 1: a RegressionBasisFunctionsProvider of swap rate and numeraire:
//   1, S, S*S, N(T_e), with S the par rate of the remaining periods at T_e:
//   P = 1/(L_e pl_e + 1), A = P pl_e; P = P/(L_p pl_p + 1), A = A + P pl_p;
//   S = (1 - P)/A (forward rates L_p = getForwardRate(T_e, fix_p, pay_p)).
// ---------------------------------------------------------------------------
namespace orc {

// pinv_solve: orc_rv.hpp

static std::vector<RV> bermudan_basis(Model* m, int basis, double fd, int e, int nper, const double* fix,
                                      const double* pay, const double* pl) {
    std::vector<RV> b;
    b.push_back(rvdet(1.0));
    if (basis == 0) {
        int idx = java_binary_search(fix, nper, fd);
        if (idx < 0) idx = -idx;
        if (idx >= nper) idx = nper - 1;
        RV rs = m->forward_rate(fd, fd, pay[idx]);
        RV ds = invert(add(mult(rs, pay[idx] - fd), 1.0));
        b.push_back(ds);
        b.push_back(powv(ds, 2.0));
        RV rl = m->forward_rate(fd, fix[idx], pay[nper - 1]);
        RV dl = invert(add(mult(rl, pay[nper - 1] - fix[idx]), 1.0));
        b.push_back(dl);
        b.push_back(powv(dl, 2.0));
        b.push_back(invert(m->numeraire(fd)));
    } else {
        RV P, A;
        for (int p = e; p < nper; p++) {
            RV L = m->forward_rate(fd, fix[p], pay[p]);
            if (p == e) { P = invert(add(mult(L, pl[p]), 1.0)); A = mult(P, pl[p]); }
            else { P = div(P, add(mult(L, pl[p]), 1.0)); A = add(A, mult(P, pl[p])); }
        }
        RV S = div(add(mult(P, -1.0), 1.0), A);
        b.push_back(S);
        b.push_back(mult(S, S));
        b.push_back(m->numeraire(fd));
    }
    return b;
}

} // namespace orc

// coefficients (optional): [number of exercise dates][K] regression coefficients in
// the order the backward induction solves them (last exercise date first).
int orc_lmm_bermudan(void* h, int nper, const int* isExercise, const double* fix, const double* periodLength,
                     const double* pay, const double* notionals, const double* swaprates, int isCallable,
                     int basis, double* value, double* error, double* per_path, double* exercise_time,
                     double* coefficients) {
    try {
        auto* m = (orc::Model*)h;
        using namespace orc;
        RV values = scalar(0.0), under = scalar(0.0), ex = scalar(INFINITY);
        int ncoef = 0;
        for (int per = nper - 1; per >= 0; per--) {
            const double fd = fix[per], pl = periodLength[per], pd_ = pay[per];
            const double notional = notionals ? notionals[per] : 1.0, K = swaprates[per];
            RV libor = m->forward_rate(fd, fd, fd + pl);
            RV payoff = mult(mult(sub(libor, K), pl), notional);
            payoff = mult(div(payoff, m->numeraire(pd_)), m->weights());
            if (isCallable) under = add(under, payoff);
            else values = add(values, payoff);
            if (isExercise[per]) {
                RV y = sub(values, under);
                std::vector<RV> b = bermudan_basis(m, basis, fd, per, nper, fix, pay, periodLength);
                const int Kb = (int)b.size();
                std::vector<double> XTX((size_t)Kb * Kb), XTy((size_t)Kb);
                for (int i = 0; i < Kb; i++)
                    for (int j = i; j < Kb; j++) {
                        XTX[(size_t)i * Kb + j] = average(mult(b[(size_t)i], b[(size_t)j]));
                        XTX[(size_t)j * Kb + i] = XTX[(size_t)i * Kb + j];
                    }
                for (int i = 0; i < Kb; i++) XTy[(size_t)i] = average(mult(y, b[(size_t)i]));
                std::vector<double> x = pinv_solve(XTX, XTy, Kb);
                if (coefficients) for (int i = 0; i < Kb; i++) coefficients[(size_t)ncoef * Kb + i] = x[(size_t)i];
                ncoef++;
                RV trig = mult(b[0], x[0]);
                for (int i = 1; i < Kb; i++) trig = addProduct(trig, b[(size_t)i], x[(size_t)i]);
                values = choose(trig, values, under);
                ex = choose(trig, ex, scalar(fd));
            }
        }
        values = div(mult(values, m->numeraire(0.0)), m->weights());
        *value = average(values);
        *error = stderror(values);
        if (per_path) fill(values, m->P, per_path);
        if (exercise_time) fill(ex, m->P, exercise_time);
        return 0;
    } catch (const std::exception& e) { g_err = e.what(); return -1; }
}

} // extern "C"

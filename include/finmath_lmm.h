/*
 * finmath_lmm.h -- C ABI of the MI355X LIBOR Market Model Monte Carlo engine.
 *
 * The engine replaces, for one GPU, the hot path of finmath-lib's
 *   EulerSchemeFromProcessModel over LIBORMarketModelFromCovarianceModel
 * driven by BrownianMotionFromMersenneRandomNumbers, together with the payoff
 * valuation of the interest-rate products that read that process.
 *
 * Reference paths below are relative to src/main/java/net/finmath/ of finmath-lib.
 *
 * Conventions
 *  - Plain C: opaque handles, plain pointers and sizes; no C++ or torch types.
 *  - All array arguments are host memory owned by the caller unless a function
 *    says "device pointer". Device buffers are owned by the context.
 *  - Every function returns an int status (FM_OK = 0). On error a message is
 *    available from fm_last_error() (thread-local). Status codes map onto the
 *    reference's exceptions: FM_ERR_INVALID_ARGUMENT -> IllegalArgumentException
 *    (the reference validates with Validate.isTrue, e.g.
 *    montecarlo/BrownianMotionFromMersenneRandomNumbers.java:77-78); the device,
 *    memory and numerical codes -> CalculationException
 *    (exception/CalculationException.java) or the RuntimeException the Euler
 *    scheme throws (montecarlo/process/EulerSchemeFromProcessModel.java:212-214).
 *  - Contexts are independent: distinct contexts may be used concurrently from
 *    different threads (one HIP stream per context). Calls on one context are
 *    serialised by an internal mutex.
 */
#ifndef FINMATH_LMM_H
#define FINMATH_LMM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------- */
#define FM_OK                   0
#define FM_ERR_INVALID_ARGUMENT 1
#define FM_ERR_OUT_OF_MEMORY    2
#define FM_ERR_DEVICE           3
#define FM_ERR_NUMERICAL        4
#define FM_ERR_STATE            5
#define FM_ERR_INTERNAL         6
#define FM_ERR_UNSUPPORTED      7

/* Enumerations use the ordinals of the reference enums. */
/* LIBORMarketModelFromCovarianceModel.Measure (models/LIBORMarketModelFromCovarianceModel.java:165) */
#define FM_MEASURE_SPOT      0
#define FM_MEASURE_TERMINAL  1
/* LIBORMarketModelFromCovarianceModel.StateSpace (models/LIBORMarketModelFromCovarianceModel.java:166) */
#define FM_STATE_SPACE_NORMAL    0
#define FM_STATE_SPACE_LOGNORMAL 1
/* EulerSchemeFromProcessModel.Scheme (montecarlo/process/EulerSchemeFromProcessModel.java:68-73) */
#define FM_SCHEME_EULER                           0
#define FM_SCHEME_PREDICTOR_CORRECTOR             1
#define FM_SCHEME_EULER_FUNCTIONAL                2
#define FM_SCHEME_PREDICTOR_CORRECTOR_FUNCTIONAL  3
/* Arithmetic of the path simulation. FP64: everything in double (the plain form,
 * RandomVariableFromDoubleArray semantics). FP32_MIXED: drift, exp/log, loadings
 * and payoffs in float; numeraire, payoff/numeraire and all sums in double. */
#define FM_PRECISION_FP64        0
#define FM_PRECISION_FP32_MIXED  1
/* FP64 arithmetic that carries the log-forwards between steps (one exp per alive
 * forward-step instead of the reference's log(exp(.)) round trip) and lets the
 * compiler fuse multiply-adds: the same values as FP64 up to a few units in the last
 * place per step. Applies to the spot measure, log-normal state space and the Euler
 * schemes (EULER only with an infinite cap: the log-forward kernel caps log-states);
 * every other configuration runs as FM_PRECISION_FP64. No adjoint. */
#define FM_PRECISION_FP64_LOGSTATE 2

typedef struct fm_lmm fm_lmm;   /* one simulation: model + Brownian driver + Euler process */

/* Shape and options of a simulation. Every measure, state space and scheme of the
 * reference is supported: SPOT and TERMINAL (LIBORMarketModelFromCovarianceModel.java:
 * 1132-1181, numeraire 1020-1066), NORMAL and LOGNORMAL, EULER / PREDICTOR_CORRECTOR
 * (state carried between steps) and their _FUNCTIONAL forms (state re-derived from the
 * value), EulerSchemeFromProcessModel.java:232-314. Under TERMINAL, payoffs of forwards
 * fixed before their payment date are evaluated at the payment date's time index.
 * Replaces the constructor arguments of
 *   TimeDiscretizationFromArray (time/TimeDiscretizationFromArray.java:210-217),
 *   BrownianMotionFromMersenneRandomNumbers(td, factors, paths, seed)
 *     (montecarlo/BrownianMotionFromMersenneRandomNumbers.java:70-104),
 *   LIBORMarketModelFromCovarianceModel properties measure/stateSpace/liborCap
 *     (models/LIBORMarketModelFromCovarianceModel.java:400-414) and
 *   EulerSchemeFromProcessModel(model, driver[, scheme])
 *     (montecarlo/process/EulerSchemeFromProcessModel.java:96-123). */
typedef struct fm_lmm_config {
    int32_t       num_time_steps;   /* nT >= 1 */
    const double* times;            /* nT+1 simulation times, increasing, times[0] = 0 */
    int32_t       num_libors;       /* n >= 1 (number of forward rates = components) */
    const double* tenor;            /* n+1 tenor dates T_0 = 0 < T_1 < ... < T_n */
    int32_t       num_factors;      /* F >= 1 */
    int32_t       num_paths;        /* P >= 1 */
    int32_t       seed;             /* Java int seed of the Mersenne Twister */
    int32_t       measure;          /* FM_MEASURE_* */
    int32_t       state_space;      /* FM_STATE_SPACE_* */
    int32_t       scheme;           /* FM_SCHEME_* */
    int32_t       precision;        /* FM_PRECISION_* */
    double        libor_cap;        /* liborCap; +infinity disables the cap (default 1e5) */
    int32_t       store_paths;      /* 1: fm_lmm_simulate materialises the process */
    int32_t       checkpoint_stride;/* adjoint forward-state checkpoint stride (0 -> 4) */
    int32_t       regenerate_increments; /* 0: the Brownian increments are generated once and
                                            kept in device memory, as the reference caches them
                                            (BrownianMotionFromMersenneRandomNumbers.java:123-136);
                                            1: regenerated by every pricing/simulation call */
    int64_t       path_offset;      /* first path of this context in the seed's stream: the context
                                       simulates paths [path_offset, path_offset + num_paths) of a
                                       run with this seed (path p uses the words from 2 p nT F on,
                                       BrownianMotionFromMersenneRandomNumbers.java:167-178), so
                                       path ranges can be resumed or split across devices */
} fm_lmm_config;

/* Create a context (allocates device memory, seeds the generator).
 * Replaces the constructors listed at fm_lmm_config. */
int fm_lmm_create(const fm_lmm_config* config, fm_lmm** out);
void fm_lmm_destroy(fm_lmm* ctx);

/* Initial forwards L_j(0), j < n, in value space.
 * Replaces LIBORMarketModelFromCovarianceModel.getInitialState
 * (models/LIBORMarketModelFromCovarianceModel.java:1081-1093). */
int fm_lmm_set_initial_forwards(fm_lmm* ctx, const double* forwards);

/* Covariance tables for every simulation time index i < nT and component j < n:
 *   factor_loading[(i*n + j)*F + k] = lambda_{j,k}(t_i)
 *   variance[i*n + j]               = instantaneous variance used by the
 *                                     log-coordinate drift adjustment.
 * Replaces LIBORMarketModelFromCovarianceModel.getFactorLoading
 * (models/LIBORMarketModelFromCovarianceModel.java:1194-1197) -> covariance
 * model getFactorLoading/getCovariance
 * (models/covariance/AbstractLIBORCovarianceModel.java:45-76,
 *  models/covariance/LIBORCovarianceModelFromVolatilityAndCorrelation.java:48-93). */
int fm_lmm_set_covariance(fm_lmm* ctx, const double* factor_loading, const double* variance);

/* Local-volatility wrapper of the covariance model: every factor loading of component j
 * at simulation time index i is multiplied by a per-path factor of its value L_j(t_i),
 *   f = (L_j - L_j * alpha) + offsets[i*n + j]   (alpha != 0)
 *   f = L_j + offsets[i*n + j]                   (alpha == 0),
 * in the drift (with the predicted value in the corrector step) and in the diffusion;
 * in the log-normal state space the drift adjustment then uses the variance
 * sum_k (lambda_jk f)^2 instead of the covariance table. offsets == NULL removes the
 * wrapper. The adjoint refuses a context with the wrapper.
 * Replaces DisplacedLocalVolatilityModel.getFactorLoading (displacement d: alpha = 0,
 * offsets = d; models/covariance/DisplacedLocalVolatilityModel.java:175-186) and
 * BlendedLocalVolatilityModel.getFactorLoading (blend a, forward curve f0: alpha = a,
 * offsets = a * f0(max(T_j - t_i, 0)), or a without curve;
 * models/covariance/BlendedLocalVolatilityModel.java:265-283), with the variance of
 * AbstractLIBORCovarianceModel.getCovariance (models/covariance/AbstractLIBORCovarianceModel.java:79-92). */
int fm_lmm_set_local_volatility(fm_lmm* ctx, double alpha, const double* offsets);

/* Stochastic volatility scaling of the covariance model: every factor loading at simulation
 * time index i is multiplied (after a local-volatility factor) by the path's S(t_i) of the
 * scalar process
 *   S(t_0) = 1,  log S(t_{i+1}) = log S(t_i) - (nu^2 / 2) dt_i + rho nu dB_a
 *                                + sqrt(1 - rho^2) nu dB_b,
 * with dB_a, dB_b the increments of factors factor_a and factor_b of the Brownian motion of the
 * context `driver` (same simulation times, paths and path offset, FP64; it may be ctx itself,
 * which shares those factors with the forwards as a BrownianMotionView would). S is computed
 * on the device when this is called. driver == NULL removes the scaling. A context with the
 * scaling is not serialised and the adjoint refuses it.
 * Replaces LIBORCovarianceModelStochasticVolatility.getFactorLoading and its scaling process
 * (models/covariance/LIBORCovarianceModelStochasticVolatility.java:174-256), run by
 * EulerSchemeFromProcessModel (montecarlo/process/EulerSchemeFromProcessModel.java:232-256). */
int fm_lmm_set_stochastic_volatility(fm_lmm* ctx, double nu, double rho, fm_lmm* driver, int32_t factor_a,
                                     int32_t factor_b);

/* Heston scaling of the covariance model: every factor loading at simulation time index i is
 * multiplied (after a local-volatility factor) by sqrt(max(V(t_i), 0)) of the variance process
 *   V(t_0) = 1,  V(t_{i+1}) = V(t_i) + (V(t_i) - theta)(-kappa) dt_i + sqrt(max(V(t_i), 0)) xi dB,
 * with dB the increments of factor `factor` of the driver context's Brownian motion (the same
 * conditions as fm_lmm_set_stochastic_volatility, whose scaling this replaces; driver == NULL
 * removes either).
 * Replaces LIBORCovarianceModelStochasticHestonVolatility.getFactorLoading and its variance
 * process (models/covariance/LIBORCovarianceModelStochasticHestonVolatility.java:172-256). */
int fm_lmm_set_heston_volatility(fm_lmm* ctx, double kappa, double theta, double xi, fm_lmm* driver, int32_t factor);

/* Discount factors P^d(T_k), k <= n, of the model's discount curve, or NULL for a
 * model without discount curve (no numeraire adjustment).
 * Replaces the deterministic numeraire adjustment of
 * LIBORMarketModelFromCovarianceModel.getNumeraireDefaultableZeroBondAsOfTimeZero
 * (models/LIBORMarketModelFromCovarianceModel.java:915-944). */
int fm_lmm_set_discount_curve(fm_lmm* ctx, const double* discount_factors);

/* ---- process (materialised) --------------------------------------------------
 * Simulate and keep the process in device memory (requires store_paths = 1).
 * Replaces EulerSchemeFromProcessModel.doPrecalculateProcess
 * (montecarlo/process/EulerSchemeFromProcessModel.java:170-326). */
int fm_lmm_simulate(fm_lmm* ctx);

/* P realisations of L_j(t_i). Replaces getProcessValue(timeIndex, componentIndex)
 * (montecarlo/process/EulerSchemeFromProcessModel.java:131-146) and
 * LIBORMarketModelFromCovarianceModel.getLIBOR (models/...CovarianceModel.java:1307-1312). */
int fm_lmm_get_process_value(fm_lmm* ctx, int32_t time_index, int32_t component, double* out);

/* Device pointer (double[P], or float[P] for FP32_MIXED) of the same values, valid
 * until the next simulate or destroy. For a GPU-resident RandomVariable. */
int fm_lmm_process_value_device_ptr(fm_lmm* ctx, int32_t time_index, int32_t component, const void** out);

/* P realisations of the numeraire N(T) at a tenor date (adjusted by the discount
 * curve when one is set). Replaces LIBORMarketModelFromCovarianceModel.getNumeraire
 * (models/LIBORMarketModelFromCovarianceModel.java:860-876,1017-1074). */
int fm_lmm_get_numeraire(fm_lmm* ctx, double time, double* out);

/* P realisations of the Brownian increment dW_k(t_i).
 * Replaces BrownianMotionFromMersenneRandomNumbers.getBrownianIncrement
 * (montecarlo/BrownianMotionFromMersenneRandomNumbers.java:123-136). */
int fm_lmm_get_brownian_increment(fm_lmm* ctx, int32_t time_index, int32_t factor, double* out);

/* ---- products ---------------------------------------------------------------- */
#define FM_PRODUCT_SWAPTION  1  /* interestrate/products/Swaption.java:137-199 */
#define FM_PRODUCT_SWAP      2  /* interestrate/products/SimpleSwap.java:137-170 */
#define FM_PRODUCT_CAPLET    3  /* interestrate/products/Caplet.java:115-140 (value unit VALUE) */
#define FM_PRODUCT_BERMUDAN_SWAPTION 4  /* interestrate/products/BermudanSwaption.java:91-157 */

/* Regression basis of a Bermudan swaption's exercise decision. */
#define FM_BASIS_REFERENCE 0  /* BermudanSwaption.getBasisFunctions (BermudanSwaption.java:215-252):
                                 1, 1/(1+L_short dt), its square, 1/(1+L_long dt), its square, 1/N(T_e) */
#define FM_BASIS_SWAP_RATE 1  /* a RegressionBasisFunctionsProvider
                                 (montecarlo/conditionalexpectation/RegressionBasisFunctionsProvider.java):
                                 1, S, S*S, N(T_e), S = par rate of the remaining periods at T_e */

/* One product. Dates are model times; arrays have num_periods entries.
 *  SWAPTION: exercise_date, fixing/payment dates, period_lengths (NULL ->
 *            payment - fixing), swaprates, notional, discounting_adjustment
 *            (per period, NULL -> 1.0 as when the discount curve is the forward
 *            curve's own, Swaption.java:165-175).
 *  SWAP:     fixing/payment dates, swaprates, notionals (NULL -> 1.0), is_pay_fix.
 *  CAPLET:   exercise_date = maturity, period_lengths[0] = period length,
 *            swaprates[0] = strike, notional = daycount fraction, is_floorlet.
 *  BERMUDAN_SWAPTION: fixing/payment dates, period_lengths, swaprates, notionals
 *            (NULL -> 1.0), exercise_flags (isPeriodStartDateExerciseDate, nonzero =
 *            exercisable at the period's fixing date), is_cancelable (0: the right
 *            to enter the swap, isCallable = true; 1: the right to terminate it),
 *            basis (FM_BASIS_*). The exercise decision regresses the continuation
 *            value on the basis (MonteCarloConditionalExpectationRegression.java:79-150):
 *            Gram matrix and right-hand side averaged over all paths, solved by the
 *            pseudo-inverse of the reference's SVD solver. */
typedef struct fm_product {
    int32_t       type;
    int32_t       num_periods;
    double        exercise_date;
    const double* fixing_dates;
    const double* payment_dates;
    const double* period_lengths;
    const double* swaprates;
    const double* notionals;
    const double* discounting_adjustment;
    double        notional;
    int32_t       is_pay_fix;
    int32_t       is_floorlet;
    const int32_t* exercise_flags;
    int32_t       is_cancelable;
    int32_t       basis;
} fm_product;

/* Value products at evaluation time 0 in one fused simulation pass:
 * values[q] = getValue(model) (AbstractMonteCarloProduct.java:82-84),
 * errors[q] = MC standard error of getValues (AbstractMonteCarloProduct.java:87-104). */
int fm_lmm_value_products(fm_lmm* ctx, const fm_product* products, int32_t count,
                          double* values, double* errors);

/* fm_lmm_value_products without device allocation. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (tight pricing loops, harnesses that forbid
 * hipMalloc). Same inputs, outputs and results as fm_lmm_value_products. The per-call device
 * buffers (wave partials, product tables, reduction output) come from a 64 MiB array that
 * every GPU holds from the library's first use until the process ends; a pass needing more
 * (about 16 B x (products + forwards + 1) per 64 paths) fails with FM_ERR_INVALID_ARGUMENT.
 * Only single-date products (swaptions, caplets): swaps and Bermudans need per-path columns
 * (FM_ERR_UNSUPPORTED). The context's own buffers (model tables, generator states, Brownian
 * increments) must already have their size: call fm_lmm_value_products once on the context
 * first, and again after a change that adds a table (a first local-volatility offset table);
 * a call that would have to grow them fails with FM_ERR_STATE. Calls on one GPU share the
 * array and are serialised by the library; each returns after its kernels have finished. */
int fm_lmm_value_products_noalloc(fm_lmm* ctx, const fm_product* products, int32_t count,
                                  double* values, double* errors);

/* Exercise time of every path of Bermudan swaption number `bermudan` (counting only the
 * Bermudan products, in order) of the last fm_lmm_value_products / fm_lmm_adjoint call:
 * the fixing date of the period at which the backward induction exercised, +infinity if
 * it never did. Replaces the "exerciseTime" entry of BermudanSwaption.getValues
 * (montecarlo/interestrate/products/BermudanSwaption.java:140-156). */
int fm_lmm_bermudan_exercise_times(fm_lmm* ctx, int32_t bermudan, double* out);

/* ---- pathwise adjoint (Greeks) -----------------------------------------------
 * Volatility factorization of the covariance tables, needed by the adjoint:
 *   factor_loading[(i*n+j)*F+k] = volatility[i*n+j] * factors[j*F+k],
 *   variance[i*n+j]              = volatility[i*n+j]^2 * variance_factor[j]
 * (variance_factor NULL -> 1.0), i.e. the structure of
 * LIBORCovarianceModelFromVolatilityAndCorrelation.getFactorLoading/getCovariance
 * (models/covariance/LIBORCovarianceModelFromVolatilityAndCorrelation.java:48-93). */
int fm_lmm_set_volatility_factorization(fm_lmm* ctx, const double* volatility, const double* factors,
                                        const double* variance_factor);

/* Gradient of V = sum_q weights[q] * value_q (weights NULL -> all 1) by a
 * hand-written reverse sweep of the Euler scheme with forward-state checkpoints.
 * Spot measure, EULER_FUNCTIONAL (or EULER where it coincides: normal state space
 * without cap) only; other measures and schemes give FM_ERR_UNSUPPORTED. Checkpoints
 * are taken every checkpoint_stride steps (the sweep recomputes the states in between):
 *   value                      = V
 *   d_initial_forwards[j]      = dV/dL_j(0)              (n entries, may be NULL)
 *   d_volatility[i*n + j]      = dV/dsigma_j(t_i)        (nT*n entries, may be NULL)
 * Greeks of volatility parameters follow by the chain rule through sigma (the
 * per-forward vega of forward j is sum_i d_volatility[i*n+j]). Averages over paths
 * (the numeraire adjustment) are held constant, as the reference's AAD treats
 * getAverage results (montecarlo/automaticdifferentiation/backward/
 * RandomVariableDifferentiableAAD.java:165-219).
 * Replaces RandomVariableDifferentiable.getGradient
 * (montecarlo/automaticdifferentiation/RandomVariableDifferentiable.java:64-75)
 * for the leaves FORWARD(T_j) and the volatility parameters of
 * LIBORMarketModelFromCovarianceModel.getModelParameters
 * (models/LIBORMarketModelFromCovarianceModel.java:1699-1732). */
int fm_lmm_adjoint(fm_lmm* ctx, const fm_product* products, int32_t count, const double* weights,
                   double* value, double* d_initial_forwards, double* d_volatility);

/* Derivative of a Bermudan's exercise indicator choose(trigger, values, underlying) in its
 * trigger within fm_lmm_adjoint: the DiracDeltaApproximationMethod and
 * diracDeltaApproximationWidthPerStdDev properties of the reference's AAD factory
 * (montecarlo/automaticdifferentiation/backward/RandomVariableDifferentiableAADFactory.java:26-67,
 * RandomVariableDifferentiableAAD.java:396-424), with the enum's ordinals:
 *   FM_DIRAC_DISCRETE_DELTA (the default, width 0.05): (Y - Z) 1{-eps/2 <= X < eps/2} / eps,
 *                           eps = width * sd(X) over the paths (eps infinite: Y - Z; eps <= 0: 0)
 *   FM_DIRAC_ONE:           Y - Z
 *   FM_DIRAC_ZERO:          0 (the exercise decisions held fixed)
 * X is the trigger, the regression of values - underlying on the date's basis; Y - Z the values
 * and underlying it chooses between. The trigger's adjoint is regressed on the same basis (the
 * conditional expectation is self-adjoint, AAD:188-192; its basis functions are constants).
 * The REGRESSION_ON_DENSITY and REGRESSION_ON_DISTRIBUITON methods give FM_ERR_UNSUPPORTED;
 * width must be positive. Products other than Bermudans have no trigger. */
#define FM_DIRAC_DISCRETE_DELTA 0
#define FM_DIRAC_REGRESSION_ON_DENSITY 1
#define FM_DIRAC_REGRESSION_ON_DISTRIBUTION 2
#define FM_DIRAC_ONE 3
#define FM_DIRAC_ZERO 4
int fm_lmm_set_dirac_delta(fm_lmm* ctx, int32_t method, double width);

/* ---- generator ----------------------------------------------------------------
 * The GPU MT19937 stream of commons-math3 MersenneTwister(long seed) as used by
 * the reference (randomnumbers/MersenneTwister.java:26-51): 32-bit output words
 * [offset, offset+count) and nextDouble() values [offset, offset+count). */
int fm_mt19937_words(int64_t seed, uint64_t offset, uint64_t count, uint32_t* out);
int fm_mt19937_doubles(int64_t seed, uint64_t offset, uint64_t count, double* out);

/* ---- serialisation --------------------------------------------------------------
 * The reference serialises a model and its Brownian motion by their parameters and
 * seed and regenerates the transient state (increments, process) on demand
 * (models/LIBORMarketModelFromCovarianceModel.java:1734-1745,
 *  montecarlo/BrownianMotionFromMersenneRandomNumbers.java:253-257). Likewise:
 * fm_lmm_serialize writes the configuration (seed, path range, grids, options) and
 * every table set on the context into a versioned little-endian byte string; NULL
 * buffer -> *size receives the needed size. fm_lmm_deserialize creates a context that
 * regenerates the same increments and paths bit for bit. */
int fm_lmm_serialize(fm_lmm* ctx, void* buffer, int64_t* size);
int fm_lmm_deserialize(const void* buffer, int64_t size, fm_lmm** out);

/* ---- diagnostics --------------------------------------------------------------- */
const char* fm_last_error(void);

/* Per-stage device time of the last fm_lmm_value_products / fm_lmm_simulate /
 * fm_lmm_adjoint call, measured with HIP events on the context's stream:
 * ms[0] generator, ms[1] path step, ms[2] reductions, ms[3] adjoint sweep,
 * ms[4] whole call. */
int fm_lmm_last_timings(fm_lmm* ctx, double* ms, int32_t count);

/* Library version string and the device it runs on. */
int fm_device_info(char* buffer, int32_t length);

#ifdef __cplusplus
}
#endif

#endif /* FINMATH_LMM_H */

// Internal declarations shared by the host code and kernels of the LMM engine.
#pragma once

#include <cstdint>
#include <cstddef>
#include <string>
#include <vector>
#include <stdexcept>

#include "finmath_lmm.h"
#include "finmath_rv.h"

namespace fm {

// Error carried to the C ABI as a status code plus message.
struct FmError : std::runtime_error {
    int code;
    FmError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

// ---- MT19937 jump-ahead (mt19937_gf2.cpp) ----------------------------------
int mt_jump_degree();
int mt_jump_words();
const std::vector<uint64_t>& mt_jump_poly_pow2(int e);     // x^(2^e) mod phi
const std::vector<uint64_t>& mt_char_poly();
void mt_jump_window_host(const uint32_t* src624, const std::vector<uint64_t>& c, uint32_t* dst624);
// Advance a 624-word window by `words` outputs (words < 2^41).
void mt_jump_window_by(uint32_t* window624, uint64_t words);

// commons-math3 MersenneTwister(long) seeding (init_by_array with key {hi, lo}).
void mt_seed_window(int64_t seed, uint32_t* window624);

// ---- generator launch (mt19937.hip) -------------------------------------------
struct StreamLayout {
    uint64_t total_words = 0;    // 2 * number of doubles of the whole stream
    int      log2_len = 0;       // sub-stream length L = 2^log2_len words
    int64_t  num_sub = 0;        // number of sub-streams
};
StreamLayout choose_layout(uint64_t total_words);

// Compute the start windows of all sub-streams on the device (jump tree).
// d_states: num_sub * 624 words. d_poly: device scratch for jump polynomials.
void launch_jump_tree(const StreamLayout& lay, const uint32_t* seed_window, uint32_t* d_states,
                      uint64_t* d_poly, void* stream);

// Generate the Brownian increments dW[D] = icdf(u_D) * sqrt_dt[(D / F) % nT] for all
// D < total_words/2 in stream (path-major) order. out is double* or float*.
// uniform_dt: every sqrt_dt entry is the same (one scale for all steps).
void launch_brownian(const StreamLayout& lay, const uint32_t* d_states, int nT, int F,
                     const double* d_sqrt_dt, int uniform_dt, void* out, int out_is_float, void* stream);

// Raw tempered words [0, count) of the stream (count <= lay.total_words).
void launch_words(const StreamLayout& lay, const uint32_t* d_states, uint64_t count,
                  uint32_t* out, void* stream);

// ---- path simulation (lmm_step.hip) ---------------------------------------------
// Factor capacity of the compiled path kernels and the record size per (step, component).
constexpr int fcap_of(int F) { return F <= 1 ? 1 : F <= 3 ? 3 : F <= 5 ? 5 : 10; }
constexpr int rec_size_of(int fcap) { return ((fcap + 3 + 3) / 4) * 4; }

// Product description lowered for the kernel.
struct KProduct {
    int32_t type;         // FM_PRODUCT_*
    int32_t eval_index;   // simulation time index at which it is evaluated (single-date)
    int32_t first_j;      // first libor index of the periods
    int32_t num_periods;
    int32_t out_slot;     // reduction slot (single-date) or contribution column base (swap)
    int32_t is_floorlet;
    int32_t flags;        // bit0: pay fix (swap)
    int32_t num_k;        // tenor index of the numeraire the payoff is divided by
    double  notional;
    double  exercise;
    double  weight;       // adjoint: objective weight of this (lowered) product per path
    int32_t param_off;    // offset into the per-period parameter table
    int32_t pad2;         // swaption: number of following swaptions valued with it (pricing);
                          // adjoint Bermudan period: the Bermudan's index
};

// Swaptions valued together by one walk over their shared periods (KProduct.pad2 of the first
// is the number of those that follow it)
constexpr int kSwaptionGroup = 4;

// Per-period parameters (strike, period length, discounting tau, adjustment, notional)
struct KPeriod {
    double strike, period_length, tau, adj, notional;
    double inv_adj;       // 1 / adj
    int32_t j, pay_k;     // libor index and tenor index of the payment date
};

struct StepArgs {
    int nT, n, F, P;
    int precision;
    int state_space;      // FM_STATE_SPACE_*
    int store_paths;
    const double* times;         // device nT+1
    const int32_t* first_alive;  // device nT
    const int32_t* tenor_time_index; // device n+1: process time index of T_k (rounded up)
    const void* rec;             // device [nT][n][rec_size_of(fcap_of(F))] records (double or float)
    int lognormal, capped;       // state space is LOGNORMAL; the liborCap applies
    const double* delta;         // device n: T_{j+1}-T_j
    const double* L0;            // device n: initial forwards in value space
    const double* Y0;            // device n: initial states (log(max(L0, 0)) or L0) of the carried schemes;
                                 // for the log-forward and spot kernels kept below log(cap)
                                 // (the spot kernel in the normal state space: the capped forwards)
    double logcap;               // log(cap) for the log-forward kernel
    int terminal;                // TERMINAL measure (general kernel)
    int carry;                   // EULER / PREDICTOR_CORRECTOR: the state is carried (general kernel)
    int pc;                      // predictor-corrector (general kernel)
    double lv_alpha;             // local-volatility factor (L - L alpha) + b_ij, or L + b_ij for alpha = 0
    const void* lvb;             // device [nT][n] b_ij (double or float), or NULL; local-volatility kernels only
    const void* svs;             // device [nT][P] stochastic volatility scaling (double or float), or NULL
    const void* dw;              // device dW [P][nT][F]
    double cap;
    int32_t scheme;
    // products
    int num_products;
    const KProduct* products;    // device
    const KPeriod* periods;      // device
    int num_slots;               // single-date product reduction slots
    int num_contrib;             // swap contribution columns
    double* contrib;             // device [num_contrib][P]
    // outputs
    double* wave_sum;            // [num_waves][num_slots + n + 1]
    double* wave_m2;             // [num_waves][num_slots]
    void* paths;                 // device triangle rows [rows][P] (store_paths)
    const int32_t* row_of;       // device [(nT+1)*n] row index of (t,j), -1 = initial value
    double* numeraire;           // device [(n+1)][P] unadjusted numeraire (store_paths or adjoint)
    void* ckpt;                  // device [ceil((nT+1)/stride)][n][P] forward states (adjoint), or NULL
    int ckpt_stride;
    const int32_t* ckpt_first;   // device [checkpoints]: the first component the reverse sweep reads from
                                 // each checkpoint's segment on (lower rows are neither written nor read)
};
void launch_lmm_step_fp64(const StepArgs& a, void* stream);
void launch_lmm_step_fp32(const StepArgs& a, void* stream);
void launch_lmm_step_fp64_logstate(const StepArgs& a, void* stream);
// General variant (terminal measure, carried state, predictor-corrector), see lmm_step.hpp.
void launch_lmm_step_general_fp64(const StepArgs& a, void* stream);
void launch_lmm_step_general_fp32(const StepArgs& a, void* stream);
void launch_lmm_step_pc_fp64(const StepArgs& a, void* stream);
void launch_lmm_step_pc_fp32(const StepArgs& a, void* stream);
// General variant with a local-volatility covariance wrapper (a.pc selects the corrector).
void launch_lmm_step_lv_fp64(const StepArgs& a, void* stream);
void launch_lmm_step_lv_fp32(const StepArgs& a, void* stream);
// Stochastic volatility scaling S(t_i), i < nT, of every path: S(0) = 1 and
// log S(t_{i+1}) = log S(t_i) + drift dt_i + l0 dW_a + l1 dW_b from the driver's FP64 increments
// [P][nT][Fd] (factors a, b), written as [nT][P] (double, or float when out_fp32).
void launch_stochastic_volatility(const double* dw, int nT, int Fd, int fa, int fb, int P, const double* times,
                                  double drift, double l0, double l1, void* out, int out_fp32, void* stream);
// Heston scaling sqrt(max(V(t_i), 0)), V(0) = 1, V <- V + (V - theta)(-kappa) dt_i + sqrt(max(V, 0)) xi dW_a.
void launch_heston_volatility(const double* dw, int nT, int Fd, int fa, int P, const double* times, double kappa,
                              double theta, double xi, void* out, int out_fp32, void* stream);

// Reverse sweep (lmm_adjoint.hpp): adjoint arguments and launchers.
struct AdjArgs {
    int stride;                      // checkpoint stride
    const int32_t* kcur_at;          // [nT+1] numeraire index after the advance at each time index
    const double* weight;            // [num kernel products] objective weight of each lowered product
    const void* fsig;                // [n][fcap_of(F)] factor matrix f_jk, zero padded (path precision)
    const void* vterm;               // [nT][n] -sigma_j(t_i) c_j, d(variance * -1/2)/dsigma (path precision)
    void* seg;                       // [max(stride-2,1)][n][P] states inside a checkpoint segment (path precision)
    double* wave_grad;               // [num_waves][nT*n + n] partials of dV/dsigma and dV/dL(0)
    const int32_t* berm_ex;          // [num Bermudans][P] exercise period per path (-1: never), or NULL
    const double* berm_weight;       // [contribution column][P] weight of each Bermudan period's payoff on
                                     // each path (the induction's adjoint), or NULL: 1 or 0 by berm_ex
};
void launch_lmm_adjoint_fp64(const StepArgs& a, const AdjArgs& g, void* stream);
void launch_lmm_adjoint_fp32(const StepArgs& a, const AdjArgs& g, void* stream);

// Deterministic reduction of per-wave partials: out_sum[col], out_m2[col] (Chan merge).
// Zero count doubles with a kernel (no memset node: safe inside a captured stream).
void launch_fill_zero(double* p, size_t count, void* stream);
void launch_reduce_partials(const double* wave_sum, const double* wave_m2, int num_waves, int stride,
                            int num_slots, int num_sums, int P, double* out_sum, double* out_m2, void* stream);

// Column sums of a row-major table [num_rows][stride], first `cols` columns, in a fixed
// order; scratch holds reduce_columns_scratch(cols) doubles.
void launch_reduce_columns(const double* rows, int num_rows, int stride, int cols, double* scratch, double* out,
                           void* stream);
int reduce_columns_scratch(int cols);

// Per-path swap values from period contributions and their scales, reduced into wave partials.
void launch_swap_finalize(const double* contrib, int P, int num_swaps, const int32_t* col0, const int32_t* ncol,
                          const int32_t* slot, const double* scale, double* wave_sum, double* wave_m2,
                          int stride, void* stream);

// ---- Bermudan swaption (lmm_bermudan.hip) ----------------------------------------
// Kernel product type of a Bermudan exercise date: the path kernel writes the raw
// regression basis into contribution columns starting at out_slot (flags = FM_BASIS_*):
//   FM_BASIS_REFERENCE: 1/(1+L_short dt), 1/(1+L_long dt), N(T_e) (unadjusted)
//   FM_BASIS_SWAP_RATE: S, N(T_e) (unadjusted)
constexpr int32_t K_BERM_BASIS = 11;
// Adjoint only: one period of a Bermudan, a swap period gated per path by the exercise
// period of the pricing pass (first_j = libor index, is_floorlet = period index,
// pad2 = Bermudan index, flags bit1 = callable).
constexpr int32_t K_BERM_PERIOD = 12;
constexpr int kBermMaxBasis = 8;
constexpr int kBermParts = kBermMaxBasis + kBermMaxBasis * (kBermMaxBasis + 1) / 2;   // X^T y, upper X^T X

struct BermBasisRef {
    int32_t kind;          // FM_BASIS_*
    int32_t col0;          // first raw basis column
    int32_t nmode;         // numeraire adjustment: 0 none, 1 (1/adj)(N zb), 2 (N zb)/adj
    int32_t pad;
    double zb, adj;
};

struct BermStep {
    int P;
    int callable;                // isCallable: payoffs go to the underlying (else to the values)
    double w;                    // Monte Carlo weight 1/P
    const double* contrib;       // device [columns][P]
    const int32_t* per_col;      // device: payoff columns of the periods added in this step (descending)
    const double* per_scale;     // device: their numeraire-adjustment scales adj/zb
    int num_per;
    int first;                   // values and underlying start at zero
    int apply;                   // apply the exercise decision of the later date first
    BermBasisRef apply_basis;
    const double* apply_coef;    // device [kBermMaxBasis]
    int32_t apply_period;        // period index recorded in ex_index when exercised
    int regress;                 // accumulate the regression partials of this date
    BermBasisRef reg_basis;
    double* partials;            // device [blocks][kBermParts]
    int final_;                  // reduce values * N(0) / w into the product's slot
    double n0;
    double* values;              // device [P]
    double* under;               // device [P]
    int32_t* ex_index;           // device [P] exercise period (-1 never), or NULL
    double* wave_sum;
    double* wave_m2;
    int stride, slot;
    double* rec_trig;            // device [P]: the applied decision's trigger per path, or NULL
    double* rec_diff;            // device [P]: values - underlying it chose between, or NULL
};
int berm_num_blocks(int P);
void launch_berm_step(const BermStep& s, int kind, void* stream);
// factors (may be NULL) receives the solve's pseudo-inverse: the eigenvectors V [K][K] (row r,
// column i at r * kBermMaxBasis + i) and the inverted kept eigenvalues [K] (0 where dropped) at
// kBermMaxBasis * kBermMaxBasis, so that another right-hand side projects the same way.
constexpr int kBermPinv = kBermMaxBasis * kBermMaxBasis + kBermMaxBasis;
void launch_berm_solve(const double* partials, int num_rows, int K, int P, double* coef, void* stream,
                       double* factors = nullptr);
// x = pinv (X^T y / P) of one date from its stored factors and the rows of X^T y partials.
void launch_berm_project(const double* partials, int num_rows, int K, int P, const double* factors, double* coef,
                         void* stream);

// One step of the Bermudan induction's adjoint, exercise dates ascending (the reverse of the
// induction). Per path it carries the adjoints vb of the values and ub of the underlying.
struct BermAdjStep {
    int P;
    int callable;
    int method;                  // FM_DIRAC_*: derivative of the exercise indicator in its trigger
    const double* contrib;       // device [columns][P] (the raw regression bases)
    int first;                   // the state starts at vb = 1, ub = 0
    int finish;                  // add the projection of the lower date's trigger adjoint:
    BermBasisRef fin_basis;      //   its basis
    const double* fin_coef;      //   device [kBermMaxBasis]: (X^T X)^+ X^T trigger-adjoint
    const int32_t* w_col;        // device: payoff columns whose weight is the state after finish
    int num_w;
    double* weight;              // device [columns][P]: dV/d(payoff) of each period per path
    int decide;                  // the exercise decision of this date:
    BermBasisRef dec_basis;      //   its basis (regression partials of the trigger adjoint)
    const double* trig;          //   device [P]: the trigger (the regressed continuation)
    const double* diff;          //   device [P]: values - underlying the decision chose between
    const double* trig_m2;       //   device: the trigger's centred second moment over the paths
    double width;                //   DISCRETE_DELTA: eps = width x sqrt(m2 / P)
    double* partials;            // device [blocks][kBermParts]: X^T trigger-adjoint
    double* vb;                  // device [P]
    double* ub;                  // device [P]
};
void launch_berm_adjoint_step(const BermAdjStep& a, int kind, void* stream);
// Per-date sums and centred second moments of trigger rows [num_dates][P] into the wave
// partials [num_waves][stride] (columns 0 .. num_dates-1).
void launch_berm_trigger_stats(const double* trig, int num_dates, int P, double* wave_sum, double* wave_m2, int stride,
                               void* stream);

// ---- device random variables (rv.hip) ---------------------------------------------------
// Copy of P device realisations (double, or float widened to double) as a random variable.
fm_rv* rv_from_device(int device, const void* src, bool is_float, int64_t size, double time);
fm_rv* rv_constant(double time, double value);
// Numeraire adjustment (x zb) / adj, or (1/adj) (x zb) in the Scalar form (new random variable).
fm_rv* rv_scale_numeraire(fm_rv* x, double zb, double adj, bool scalar_form);

// ---- entry points over caller scratch (capi.cpp) -------------------------------------------
int context_device(const fm_lmm* c);
// fm_lmm_value_products; with scratch != NULL the per-call device buffers (wave partials, product
// tables, reduction output) are taken from [scratch, scratch + scratch_bytes) and no context
// buffer may grow (FM_ERR_STATE), so the call allocates no device memory.
int value_products_in(fm_lmm* c, const fm_product* products, int32_t count, double* values, double* errors,
                      void* scratch, size_t scratch_bytes);

} // namespace fm

/*
 * finmath_rv.h -- C ABI of the device-resident random variables of the MI355X engine.
 *
 * A random variable is an immutable vector of P double realisations in device memory,
 * or a constant (deterministic). It replaces, for products that are not fused into the
 * path kernel, finmath-lib's RandomVariableFromDoubleArray (montecarlo/
 * RandomVariableFromDoubleArray.java) created through a RandomVariableFactory
 * (montecarlo/RandomVariableFactory.java:22-95), and its lazily evaluated form
 * RandomVariableLazyEvaluation (montecarlo/RandomVariableLazyEvaluation.java):
 * element operations are recorded as a short register program and evaluated in one
 * pass over the realisations, optionally reduced to an average in the same pass.
 *
 * Conventions as in finmath_lmm.h: plain C, caller-owned host arrays, int status codes
 * (FM_OK = 0), messages from fm_last_error(). Realisations are evaluated on one stream
 * per device; every function that returns host data synchronises it.
 */
#ifndef FINMATH_RV_H
#define FINMATH_RV_H

#include <stdint.h>
#include "finmath_lmm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fm_rv fm_rv;

/* RandomVariableFromDoubleArray(time, realisations) (RandomVariableFromDoubleArray.java:96-110):
 * copies `size` host values to the device. */
int fm_rv_from_host(double time, int64_t size, const double* values, fm_rv** out);
/* RandomVariableFromDoubleArray(time, value): a deterministic random variable (no device memory). */
int fm_rv_constant(double time, double value, fm_rv** out);
void fm_rv_release(fm_rv* rv);

/* getFiltrationTime, size (1 for a constant), isDeterministic, doubleValue (constants). */
int fm_rv_info(const fm_rv* rv, double* time, int64_t* size, int32_t* is_deterministic, double* value);
/* getRealizations(size): `size` values (a constant is repeated). */
int fm_rv_to_host(const fm_rv* rv, int64_t size, double* out);
/* Device pointer of the realisations (NULL for a constant), valid while rv is alive. */
int fm_rv_device_ptr(const fm_rv* rv, const double** out);

/* ---- register programs ---------------------------------------------------------
 * An instruction writes register dst (0..15) from registers a, b, c and the immediate:
 *   LOAD      dst = input[a]                    CONST     dst = imm
 *   ADD       dst = a + b                       SUB       dst = a - b
 *   MULT      dst = a * b                       DIV       dst = a / b
 *   CAP       dst = Math.min(a, b)              FLOOR     dst = Math.max(a, b)
 *   POW       dst = Math.pow(a, b)              EXP, LOG, SQRT, ABS, EXPM1, SIN, COS (of a)
 *   INVERT    dst = 1.0 / a                     SQUARED   dst = a * a
 *   ADD_PRODUCT dst = a + b * c                 ADD_RATIO dst = a + b / c
 *   SUB_RATIO dst = a - b / c                   ACCRUE    dst = a * (1 + b * imm)
 *   DISCOUNT  dst = a / (1.0 + b * imm)         CHOOSE    dst = a >= 0 ? b : c
 *   ISNAN     dst = a is NaN ? 1 : 0
 * with the element formulas of RandomVariableFromDoubleArray (cap :743-756,1217-1244,
 * accrue :1279-1305, discount :1308-1333, choose :1342-1363, addProduct :1366-1428,
 * addRatio/subRatio :1440-1484, isNaN :1486-1498) and Java's Math.min/max NaN and
 * signed-zero rules. */
#define FM_RV_LOAD         0
#define FM_RV_CONST        1
#define FM_RV_ADD          2
#define FM_RV_SUB          3
#define FM_RV_MULT         4
#define FM_RV_DIV          5
#define FM_RV_CAP          6
#define FM_RV_FLOOR        7
#define FM_RV_POW          8
#define FM_RV_EXP          9
#define FM_RV_LOG          10
#define FM_RV_SQRT         11
#define FM_RV_ABS          12
#define FM_RV_INVERT       13
#define FM_RV_SQUARED      14
#define FM_RV_ADD_PRODUCT  15
#define FM_RV_ADD_RATIO    16
#define FM_RV_SUB_RATIO    17
#define FM_RV_ACCRUE       18
#define FM_RV_DISCOUNT     19
#define FM_RV_CHOOSE       20
#define FM_RV_ISNAN        21
#define FM_RV_EXPM1        22
#define FM_RV_SIN          23
#define FM_RV_COS          24

#define FM_RV_MAX_INSTRUCTIONS 64
#define FM_RV_MAX_INPUTS       16
#define FM_RV_REGISTERS        16

typedef struct fm_rv_instr {
    int32_t op;                 /* FM_RV_* */
    int16_t dst, a, b, c;       /* registers (LOAD: a = input index) */
    double  imm;                /* CONST value, ACCRUE/DISCOUNT period length */
} fm_rv_instr;

/* Evaluate `program` element-wise over the inputs (all of one size P, or constants)
 * and return register `result` as a new random variable with filtration time `time`.
 * All inputs constant: the result is a constant evaluated on the host. */
int fm_rv_evaluate(int32_t num_inputs, const fm_rv* const* inputs, int32_t num_instr, const fm_rv_instr* program,
                   int32_t result, double time, fm_rv** out);

/* getAverage of the program's result without materialising it (one pass): the
 * compensated sum over the realisations divided by P (RandomVariableFromDoubleArray
 * .getAverage, Kahan summation, :287-307; here a fixed-order tree of Neumaier sums, so
 * the result is reproducible and within an ulp of the sequential sum). With
 * `variance` non-NULL also getVariance (:334-356): a second pass over (x - average)^2.
 * probabilities non-NULL: getAverage(probabilities) = sum(x p) / P (:310-331) and
 * getVariance(probabilities) = sum((x - average)^2 p) (:359-380). */
int fm_rv_evaluate_average(int32_t num_inputs, const fm_rv* const* inputs, int32_t num_instr,
                           const fm_rv_instr* program, int32_t result, const fm_rv* probabilities,
                           double* average, double* variance);

/* getAverage / getVariance of a random variable. */
int fm_rv_average(const fm_rv* rv, double* average);
int fm_rv_variance(const fm_rv* rv, double* variance);

/* ---- process values as random variables ------------------------------------------
 * getProcessValue(timeIndex, component) (montecarlo/process/EulerSchemeFromProcessModel.java:131-146)
 * of a simulated context (store_paths = 1) as a device random variable (a copy, in
 * double precision; time-0 values are constants). */
int fm_lmm_process_value_rv(fm_lmm* ctx, int32_t time_index, int32_t component, fm_rv** out);
/* getNumeraire(time) (models/LIBORMarketModelFromCovarianceModel.java:860-876) of a
 * simulated context, adjusted like fm_lmm_get_numeraire. */
int fm_lmm_numeraire_rv(fm_lmm* ctx, double time, fm_rv** out);
/* getMonteCarloWeights (EulerSchemeFromProcessModel.java:184,317): the constant 1/P. */
int fm_lmm_weights_rv(fm_lmm* ctx, fm_rv** out);

#ifdef __cplusplus
}
#endif

#endif /* FINMATH_RV_H */

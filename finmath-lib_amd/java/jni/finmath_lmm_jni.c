/*
 * JNI side of net.finmath.montecarlo.interestrate.hip.FinmathHipJNI (JDK 8-21), linked against
 * libfinmath_lmm.so:
 *
 *   cc -O2 -shared -fPIC -I$JAVA_HOME/include -I$JAVA_HOME/include/linux -I include \
 *      finmath-lib_amd/java/jni/finmath_lmm_jni.c -L finmath-lib_amd/lib -lfinmath_lmm \
 *      -Wl,-rpath,'$ORIGIN' -o libfinmath_lmm_jni.so
 *
 * Every native pins its arrays, calls one or two fm_lmm_* entry points (include/finmath_lmm.h) and
 * maps a status other than FM_OK onto the reference's exceptions.
 */
#include <jni.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "finmath_lmm.h"

#define JNI_NAME(name) Java_net_finmath_montecarlo_interestrate_hip_FinmathHipJNI_##name

static int throw_status(JNIEnv* env, int status) {
    if (status == FM_OK) return 0;
    const char* cls = status == FM_ERR_INVALID_ARGUMENT ? "java/lang/IllegalArgumentException"
                    : status == FM_ERR_UNSUPPORTED      ? "java/lang/UnsupportedOperationException"
                                                        : "net/finmath/exception/CalculationException";
    jclass c = (*env)->FindClass(env, cls);
    if (c) (*env)->ThrowNew(env, c, fm_last_error());
    return 1;
}

/* A pinned double[] (NULL stays NULL). */
typedef struct { jdoubleArray array; jdouble* data; } pinned_d;

static pinned_d pin_d(JNIEnv* env, jdoubleArray a) {
    pinned_d p = { a, NULL };
    if (a) p.data = (*env)->GetDoubleArrayElements(env, a, NULL);
    return p;
}

static void unpin_d(JNIEnv* env, pinned_d p, int copy_back) {
    if (p.array) (*env)->ReleaseDoubleArrayElements(env, p.array, p.data, copy_back ? 0 : JNI_ABORT);
}

static fm_lmm* handle(jlong context) { return (fm_lmm*)(intptr_t)context; }

JNIEXPORT jlong JNICALL JNI_NAME(create)(JNIEnv* env, jclass cls, jdoubleArray times, jdoubleArray tenor,
                                         jint factors, jint paths, jint seed, jint measure, jint state_space,
                                         jint scheme, jint precision, jdouble libor_cap, jboolean store_paths,
                                         jint checkpoint_stride, jboolean regenerate, jlong path_offset) {
    (void)cls;
    pinned_d t = pin_d(env, times), k = pin_d(env, tenor);
    fm_lmm_config c;
    memset(&c, 0, sizeof(c));
    c.num_time_steps = (*env)->GetArrayLength(env, times) - 1;
    c.times = t.data;
    c.num_libors = (*env)->GetArrayLength(env, tenor) - 1;
    c.tenor = k.data;
    c.num_factors = factors;
    c.num_paths = paths;
    c.seed = seed;
    c.measure = measure;
    c.state_space = state_space;
    c.scheme = scheme;
    c.precision = precision;
    c.libor_cap = libor_cap;
    c.store_paths = store_paths ? 1 : 0;
    c.checkpoint_stride = checkpoint_stride;
    c.regenerate_increments = regenerate ? 1 : 0;
    c.path_offset = path_offset;
    fm_lmm* ctx = NULL;
    const int s = fm_lmm_create(&c, &ctx);
    unpin_d(env, t, 0);
    unpin_d(env, k, 0);
    if (throw_status(env, s)) return 0;
    return (jlong)(intptr_t)ctx;
}

JNIEXPORT void JNICALL JNI_NAME(setTables)(JNIEnv* env, jclass cls, jlong context, jdoubleArray l0,
                                           jdoubleArray lambda, jdoubleArray variance, jdoubleArray df) {
    (void)cls;
    pinned_d a = pin_d(env, l0), b = pin_d(env, lambda), c = pin_d(env, variance), d = pin_d(env, df);
    int s = fm_lmm_set_initial_forwards(handle(context), a.data);
    if (s == FM_OK) s = fm_lmm_set_covariance(handle(context), b.data, c.data);
    if (s == FM_OK) s = fm_lmm_set_discount_curve(handle(context), d.data);
    unpin_d(env, a, 0); unpin_d(env, b, 0); unpin_d(env, c, 0); unpin_d(env, d, 0);
    throw_status(env, s);
}

JNIEXPORT void JNICALL JNI_NAME(setVolatilityFactorization)(JNIEnv* env, jclass cls, jlong context,
                                                            jdoubleArray volatility, jdoubleArray factors,
                                                            jdoubleArray variance_factor) {
    (void)cls;
    pinned_d a = pin_d(env, volatility), b = pin_d(env, factors), c = pin_d(env, variance_factor);
    const int s = fm_lmm_set_volatility_factorization(handle(context), a.data, b.data, c.data);
    unpin_d(env, a, 0); unpin_d(env, b, 0); unpin_d(env, c, 0);
    throw_status(env, s);
}

JNIEXPORT void JNICALL JNI_NAME(processValue)(JNIEnv* env, jclass cls, jlong context, jint t, jint j,
                                              jdoubleArray out) {
    (void)cls;
    int s = fm_lmm_simulate(handle(context));
    if (s == FM_OK) {
        jdouble* p = (*env)->GetPrimitiveArrayCritical(env, out, NULL);
        s = fm_lmm_get_process_value(handle(context), t, j, p);
        (*env)->ReleasePrimitiveArrayCritical(env, out, p, 0);
    }
    throw_status(env, s);
}

/* Products given column-wise (see FinmathHipJNI.valueProducts), lowered into fm_product with
 * copies of the period arrays that live until free_products. */
typedef struct {
    int count;
    fm_product* p;
    double** owned;
    int32_t** flags;
} lowered;

static double* copy_row(JNIEnv* env, jobjectArray rows, int q) {
    if (!rows) return NULL;
    jdoubleArray row = (jdoubleArray)(*env)->GetObjectArrayElement(env, rows, q);
    if (!row) return NULL;
    const jsize m = (*env)->GetArrayLength(env, row);
    double* out = (double*)malloc((size_t)(m > 0 ? m : 1) * sizeof(double));
    if (out) (*env)->GetDoubleArrayRegion(env, row, 0, m, out);
    (*env)->DeleteLocalRef(env, row);
    return out;
}

static int row_length(JNIEnv* env, jobjectArray rows, int q) {
    jobject row = (*env)->GetObjectArrayElement(env, rows, q);
    const int m = row ? (int)(*env)->GetArrayLength(env, (jarray)row) : 0;
    if (row) (*env)->DeleteLocalRef(env, row);
    return m;
}

static void free_products(lowered* L) {
    if (L->owned) for (int i = 0; i < 5 * L->count; i++) free(L->owned[i]);
    if (L->flags) for (int i = 0; i < L->count; i++) free(L->flags[i]);
    free(L->owned); free(L->flags); free(L->p);
    memset(L, 0, sizeof(*L));
}

static int lower_products(JNIEnv* env, lowered* L, jintArray type, jdoubleArray exercise, jdoubleArray notional,
                          jintArray flags, jobjectArray fixing, jobjectArray payment, jobjectArray length,
                          jobjectArray rate, jobjectArray period_notional, jobjectArray exercise_flags) {
    memset(L, 0, sizeof(*L));
    L->count = (*env)->GetArrayLength(env, type);
    L->p = (fm_product*)calloc((size_t)(L->count > 0 ? L->count : 1), sizeof(fm_product));
    L->owned = (double**)calloc((size_t)(5 * L->count + 1), sizeof(double*));
    L->flags = (int32_t**)calloc((size_t)(L->count + 1), sizeof(int32_t*));
    if (!L->p || !L->owned || !L->flags) { free_products(L); return FM_ERR_OUT_OF_MEMORY; }
    jint* ty = (*env)->GetIntArrayElements(env, type, NULL);
    jint* fl = (*env)->GetIntArrayElements(env, flags, NULL);
    jdouble* ex = (*env)->GetDoubleArrayElements(env, exercise, NULL);
    jdouble* no = (*env)->GetDoubleArrayElements(env, notional, NULL);
    for (int q = 0; q < L->count; q++) {
        fm_product* p = &L->p[q];
        p->type = ty[q];
        p->exercise_date = ex[q];
        p->notional = no[q];
        p->is_pay_fix = fl[q] & 1;
        p->is_floorlet = (fl[q] >> 1) & 1;
        p->is_cancelable = (fl[q] >> 2) & 1;
        p->basis = (fl[q] >> 8) & 0xff;
        p->num_periods = fixing ? row_length(env, fixing, q) : 0;
        if (p->num_periods == 0) p->num_periods = row_length(env, rate, q);
        double** own = &L->owned[5 * q];
        p->fixing_dates = own[0] = copy_row(env, fixing, q);
        p->payment_dates = own[1] = copy_row(env, payment, q);
        p->period_lengths = own[2] = copy_row(env, length, q);
        p->swaprates = own[3] = copy_row(env, rate, q);
        p->notionals = own[4] = copy_row(env, period_notional, q);
        if (exercise_flags) {
            jintArray row = (jintArray)(*env)->GetObjectArrayElement(env, exercise_flags, q);
            if (row) {
                const jsize m = (*env)->GetArrayLength(env, row);
                L->flags[q] = (int32_t*)malloc((size_t)(m > 0 ? m : 1) * sizeof(int32_t));
                if (L->flags[q]) (*env)->GetIntArrayRegion(env, row, 0, m, (jint*)L->flags[q]);
                p->exercise_flags = L->flags[q];
                (*env)->DeleteLocalRef(env, row);
            }
        }
    }
    (*env)->ReleaseIntArrayElements(env, type, ty, JNI_ABORT);
    (*env)->ReleaseIntArrayElements(env, flags, fl, JNI_ABORT);
    (*env)->ReleaseDoubleArrayElements(env, exercise, ex, JNI_ABORT);
    (*env)->ReleaseDoubleArrayElements(env, notional, no, JNI_ABORT);
    return FM_OK;
}

JNIEXPORT void JNICALL JNI_NAME(valueProducts)(JNIEnv* env, jclass cls, jlong context, jintArray type,
                                               jdoubleArray exercise, jdoubleArray notional, jintArray flags,
                                               jobjectArray fixing, jobjectArray payment, jobjectArray length,
                                               jobjectArray rate, jobjectArray period_notional,
                                               jobjectArray exercise_flags, jdoubleArray values,
                                               jdoubleArray errors) {
    (void)cls;
    lowered L;
    int s = lower_products(env, &L, type, exercise, notional, flags, fixing, payment, length, rate,
                           period_notional, exercise_flags);
    if (s == FM_OK) {
        pinned_d v = pin_d(env, values), e = pin_d(env, errors);
        s = fm_lmm_value_products(handle(context), L.p, L.count, v.data, e.data);
        unpin_d(env, v, 1);
        unpin_d(env, e, 1);
    }
    free_products(&L);
    throw_status(env, s);
}

JNIEXPORT jdouble JNICALL JNI_NAME(adjoint)(JNIEnv* env, jclass cls, jlong context, jintArray type,
                                            jdoubleArray exercise, jdoubleArray notional, jintArray flags,
                                            jobjectArray fixing, jobjectArray payment, jobjectArray length,
                                            jobjectArray rate, jobjectArray period_notional,
                                            jobjectArray exercise_flags, jdoubleArray weights,
                                            jdoubleArray d_initial_forwards, jdoubleArray d_volatility) {
    (void)cls;
    lowered L;
    double value = 0.0;
    int s = lower_products(env, &L, type, exercise, notional, flags, fixing, payment, length, rate,
                           period_notional, exercise_flags);
    if (s == FM_OK) {
        pinned_d w = pin_d(env, weights), dl = pin_d(env, d_initial_forwards), ds = pin_d(env, d_volatility);
        s = fm_lmm_adjoint(handle(context), L.p, L.count, w.data, &value, dl.data, ds.data);
        unpin_d(env, w, 0);
        unpin_d(env, dl, 1);
        unpin_d(env, ds, 1);
    }
    free_products(&L);
    throw_status(env, s);
    return value;
}

JNIEXPORT void JNICALL JNI_NAME(setDiracDelta)(JNIEnv* env, jclass cls, jlong context, jint method, jdouble width) {
    (void)cls;
    throw_status(env, fm_lmm_set_dirac_delta(handle(context), (int32_t)method, (double)width));
}

JNIEXPORT void JNICALL JNI_NAME(destroy)(JNIEnv* env, jclass cls, jlong context) {
    (void)env; (void)cls;
    fm_lmm_destroy(handle(context));
}

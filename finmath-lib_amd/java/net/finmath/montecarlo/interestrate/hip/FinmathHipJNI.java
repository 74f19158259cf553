/*
 * JNI binding (JDK 8-21) of libfinmath_lmm.so: the natives of finmath-lib_amd/java/jni/finmath_lmm_jni.c.
 * Arrays are passed as Java arrays; the C side pins them, calls the fm_lmm_* entry point and maps a
 * status other than FM_OK onto IllegalArgumentException (FM_ERR_INVALID_ARGUMENT),
 * UnsupportedOperationException (FM_ERR_UNSUPPORTED: keep the reference's Euler scheme) or
 * net.finmath.exception.CalculationException. The process and simulation classes of this package
 * read the model exactly as EulerSchemeFromProcessModelHIP does and call these instead of the
 * Panama downcalls.
 */
package net.finmath.montecarlo.interestrate.hip;

public final class FinmathHipJNI {

	static {
		System.loadLibrary("finmath_lmm_jni");
	}

	private FinmathHipJNI() { }

	/** fm_lmm_create; returns the fm_lmm* handle. */
	public static native long create(double[] times, double[] tenor, int factors, int paths, int seed, int measure,
			int stateSpace, int scheme, int precision, double liborCap, boolean storePaths, int checkpointStride,
			boolean regenerateIncrements, long pathOffset);

	/** fm_lmm_set_initial_forwards, fm_lmm_set_covariance and fm_lmm_set_discount_curve (df may be null). */
	public static native void setTables(long context, double[] initialForwards, double[] factorLoading,
			double[] variance, double[] discountFactors);

	/** fm_lmm_set_volatility_factorization (varianceFactor may be null). */
	public static native void setVolatilityFactorization(long context, double[] volatility, double[] factors,
			double[] varianceFactor);

	/** fm_lmm_simulate then fm_lmm_get_process_value into out[numberOfPaths]. */
	public static native void processValue(long context, int timeIndex, int component, double[] out);

	/**
	 * fm_lmm_value_products of products given column-wise: type[q], exerciseDate[q], notional[q],
	 * flags[q] (is_pay_fix | is_floorlet << 1 | is_cancelable << 2 | basis << 8), and per product
	 * the period arrays fixing[q], payment[q], periodLength[q], swaprate[q], periodNotional[q]
	 * (entries may be null) and exerciseFlags[q] (Bermudan only).
	 */
	public static native void valueProducts(long context, int[] type, double[] exerciseDate, double[] notional,
			int[] flags, double[][] fixing, double[][] payment, double[][] periodLength, double[][] swaprate,
			double[][] periodNotional, int[][] exerciseFlags, double[] values, double[] errors);

	/** fm_lmm_set_dirac_delta: method is the ordinal of RandomVariableDifferentiableAADFactory.DiracDeltaApproximationMethod. */
	public static native void setDiracDelta(long context, int method, double widthPerStdDev);

	/** fm_lmm_adjoint over products given as for valueProducts; returns the value. */
	public static native double adjoint(long context, int[] type, double[] exerciseDate, double[] notional,
			int[] flags, double[][] fixing, double[][] payment, double[][] periodLength, double[][] swaprate,
			double[][] periodNotional, int[][] exerciseFlags, double[] weights, double[] dInitialForwards,
			double[] dVolatility);

	/** fm_lmm_destroy. */
	public static native void destroy(long context);
}

/*
 * Panama FFM (JDK 22+) binding of libfinmath_lmm.so, the MI355X LIBOR market model engine
 * (C ABI: include/finmath_lmm.h). No native glue is needed: every entry point is a downcall
 * handle. The byte offsets of the two structs the ABI takes are named constants below, each the
 * C compiler's offsetof on x86-64 (natural alignment); tests/test_java_binding.py checks them
 * against the library's own struct definitions.
 */
package net.finmath.montecarlo.interestrate.hip;

import java.lang.foreign.AddressLayout;
import java.lang.foreign.Arena;
import java.lang.foreign.FunctionDescriptor;
import java.lang.foreign.Linker;
import java.lang.foreign.MemorySegment;
import java.lang.foreign.SymbolLookup;
import java.lang.invoke.MethodHandle;

import net.finmath.exception.CalculationException;

import static java.lang.foreign.ValueLayout.ADDRESS;
import static java.lang.foreign.ValueLayout.JAVA_DOUBLE;
import static java.lang.foreign.ValueLayout.JAVA_INT;
import static java.lang.foreign.ValueLayout.JAVA_LONG;

public final class FinmathHip {

	private FinmathHip() { }

	// ---- status codes (FM_*) -----------------------------------------------------------------
	public static final int FM_OK = 0;
	public static final int FM_ERR_INVALID_ARGUMENT = 1;
	public static final int FM_ERR_UNSUPPORTED = 7;

	// ---- product types and regression bases ----------------------------------------------------
	public static final int FM_PRODUCT_SWAPTION = 1;
	public static final int FM_PRODUCT_SWAP = 2;
	public static final int FM_PRODUCT_CAPLET = 3;
	public static final int FM_PRODUCT_BERMUDAN_SWAPTION = 4;
	public static final int FM_BASIS_REFERENCE = 0;
	public static final int FM_BASIS_SWAP_RATE = 1;

	// ---- precision ----------------------------------------------------------------------------
	public static final int FM_PRECISION_FP64 = 0;
	public static final int FM_PRECISION_FP32_MIXED = 1;

	// ---- struct fm_lmm_config: byte offsets and size -------------------------------------------
	public static final long CONFIG_SIZE = 96;
	public static final long OFF_NUM_TIME_STEPS = 0;
	public static final long OFF_TIMES = 8;
	public static final long OFF_NUM_LIBORS = 16;
	public static final long OFF_TENOR = 24;
	public static final long OFF_NUM_FACTORS = 32;
	public static final long OFF_NUM_PATHS = 36;
	public static final long OFF_SEED = 40;
	public static final long OFF_MEASURE = 44;
	public static final long OFF_STATE_SPACE = 48;
	public static final long OFF_SCHEME = 52;
	public static final long OFF_PRECISION = 56;
	public static final long OFF_LIBOR_CAP = 64;
	public static final long OFF_STORE_PATHS = 72;
	public static final long OFF_CHECKPOINT_STRIDE = 76;
	public static final long OFF_REGENERATE_INCREMENTS = 80;
	public static final long OFF_PATH_OFFSET = 88;

	// ---- struct fm_product: byte offsets and size ----------------------------------------------
	public static final long PRODUCT_SIZE = 96;
	public static final long P_TYPE = 0;
	public static final long P_NUM_PERIODS = 4;
	public static final long P_EXERCISE_DATE = 8;
	public static final long P_FIXING_DATES = 16;
	public static final long P_PAYMENT_DATES = 24;
	public static final long P_PERIOD_LENGTHS = 32;
	public static final long P_SWAPRATES = 40;
	public static final long P_NOTIONALS = 48;
	public static final long P_DISCOUNTING_ADJUSTMENT = 56;
	public static final long P_NOTIONAL = 64;
	public static final long P_IS_PAY_FIX = 72;
	public static final long P_IS_FLOORLET = 76;
	public static final long P_EXERCISE_FLAGS = 80;
	public static final long P_IS_CANCELABLE = 88;
	public static final long P_BASIS = 92;

	static final Linker LINKER = Linker.nativeLinker();
	static final SymbolLookup LIB = SymbolLookup.libraryLookup(
			System.getProperty("finmath.hip.library", "libfinmath_lmm.so"), Arena.global());
	static final AddressLayout C_STRING = ADDRESS.withTargetLayout(java.lang.foreign.MemoryLayout.sequenceLayout(4096, java.lang.foreign.ValueLayout.JAVA_BYTE));

	private static MethodHandle fn(String name, FunctionDescriptor d) {
		return LINKER.downcallHandle(LIB.find(name).orElseThrow(() -> new UnsatisfiedLinkError(name)), d);
	}

	static final MethodHandle CREATE = fn("fm_lmm_create", FunctionDescriptor.of(JAVA_INT, ADDRESS, ADDRESS));
	static final MethodHandle DESTROY = fn("fm_lmm_destroy", FunctionDescriptor.ofVoid(ADDRESS));
	static final MethodHandle SET_INITIAL_FORWARDS = fn("fm_lmm_set_initial_forwards",
			FunctionDescriptor.of(JAVA_INT, ADDRESS, ADDRESS));
	static final MethodHandle SET_COVARIANCE = fn("fm_lmm_set_covariance",
			FunctionDescriptor.of(JAVA_INT, ADDRESS, ADDRESS, ADDRESS));
	static final MethodHandle SET_DISCOUNT_CURVE = fn("fm_lmm_set_discount_curve",
			FunctionDescriptor.of(JAVA_INT, ADDRESS, ADDRESS));
	static final MethodHandle SET_VOLATILITY_FACTORIZATION = fn("fm_lmm_set_volatility_factorization",
			FunctionDescriptor.of(JAVA_INT, ADDRESS, ADDRESS, ADDRESS, ADDRESS));
	static final MethodHandle SET_DIRAC_DELTA = fn("fm_lmm_set_dirac_delta",
			FunctionDescriptor.of(JAVA_INT, ADDRESS, JAVA_INT, JAVA_DOUBLE));
	static final MethodHandle SIMULATE = fn("fm_lmm_simulate", FunctionDescriptor.of(JAVA_INT, ADDRESS));
	static final MethodHandle GET_PROCESS_VALUE = fn("fm_lmm_get_process_value",
			FunctionDescriptor.of(JAVA_INT, ADDRESS, JAVA_INT, JAVA_INT, ADDRESS));
	static final MethodHandle GET_NUMERAIRE = fn("fm_lmm_get_numeraire",
			FunctionDescriptor.of(JAVA_INT, ADDRESS, JAVA_DOUBLE, ADDRESS));
	static final MethodHandle GET_BROWNIAN_INCREMENT = fn("fm_lmm_get_brownian_increment",
			FunctionDescriptor.of(JAVA_INT, ADDRESS, JAVA_INT, JAVA_INT, ADDRESS));
	static final MethodHandle VALUE_PRODUCTS = fn("fm_lmm_value_products",
			FunctionDescriptor.of(JAVA_INT, ADDRESS, ADDRESS, JAVA_INT, ADDRESS, ADDRESS));
	static final MethodHandle BERMUDAN_EXERCISE_TIMES = fn("fm_lmm_bermudan_exercise_times",
			FunctionDescriptor.of(JAVA_INT, ADDRESS, JAVA_INT, ADDRESS));
	static final MethodHandle ADJOINT = fn("fm_lmm_adjoint",
			FunctionDescriptor.of(JAVA_INT, ADDRESS, ADDRESS, JAVA_INT, ADDRESS, ADDRESS, ADDRESS, ADDRESS));
	static final MethodHandle LAST_ERROR = fn("fm_last_error", FunctionDescriptor.of(C_STRING));

	/** Thrown for FM_ERR_UNSUPPORTED: the caller keeps the reference's own Euler scheme. */
	public static final class UnsupportedModelException extends Exception {
		private static final long serialVersionUID = 1L;
		UnsupportedModelException(String message) { super(message); }
	}

	static Object call(MethodHandle h, Object... args) {
		try {
			return h.invokeWithArguments(args);
		}
		catch(final Throwable t) {
			throw new IllegalStateException(t);
		}
	}

	static String lastError() {
		return ((MemorySegment) call(LAST_ERROR)).getString(0);
	}

	/** Maps a status to the reference's exceptions: IllegalArgumentException (Validate.isTrue), else CalculationException. */
	static void check(int status) throws CalculationException {
		if(status == FM_OK) {
			return;
		}
		final String message = "[status " + status + "] " + lastError();
		if(status == FM_ERR_INVALID_ARGUMENT) {
			throw new IllegalArgumentException(message);
		}
		throw new CalculationException(message);
	}

	static void checkSupported(int status) throws CalculationException, UnsupportedModelException {
		if(status == FM_ERR_UNSUPPORTED) {
			throw new UnsupportedModelException(lastError());
		}
		check(status);
	}

	static MemorySegment doubles(Arena arena, double[] values) {
		return values == null ? MemorySegment.NULL : arena.allocateFrom(JAVA_DOUBLE, values);
	}

	static MemorySegment ints(Arena arena, int[] values) {
		return values == null ? MemorySegment.NULL : arena.allocateFrom(JAVA_INT, values);
	}

	/** An fm_lmm_config filled from its fields. */
	static MemorySegment config(Arena arena, double[] times, double[] tenor, int factors, int paths, int seed,
			int measure, int stateSpace, int scheme, int precision, double liborCap, boolean storePaths,
			int checkpointStride, boolean regenerateIncrements, long pathOffset) {
		final MemorySegment c = arena.allocate(CONFIG_SIZE, 8);
		c.set(JAVA_INT, OFF_NUM_TIME_STEPS, times.length - 1);
		c.set(ADDRESS, OFF_TIMES, doubles(arena, times));
		c.set(JAVA_INT, OFF_NUM_LIBORS, tenor.length - 1);
		c.set(ADDRESS, OFF_TENOR, doubles(arena, tenor));
		c.set(JAVA_INT, OFF_NUM_FACTORS, factors);
		c.set(JAVA_INT, OFF_NUM_PATHS, paths);
		c.set(JAVA_INT, OFF_SEED, seed);
		c.set(JAVA_INT, OFF_MEASURE, measure);
		c.set(JAVA_INT, OFF_STATE_SPACE, stateSpace);
		c.set(JAVA_INT, OFF_SCHEME, scheme);
		c.set(JAVA_INT, OFF_PRECISION, precision);
		c.set(JAVA_DOUBLE, OFF_LIBOR_CAP, liborCap);
		c.set(JAVA_INT, OFF_STORE_PATHS, storePaths ? 1 : 0);
		c.set(JAVA_INT, OFF_CHECKPOINT_STRIDE, checkpointStride);
		c.set(JAVA_INT, OFF_REGENERATE_INCREMENTS, regenerateIncrements ? 1 : 0);
		c.set(JAVA_LONG, OFF_PATH_OFFSET, pathOffset);
		return c;
	}
}

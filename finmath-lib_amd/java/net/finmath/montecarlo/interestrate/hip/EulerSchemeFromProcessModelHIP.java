/*
 * Drop-in for EulerSchemeFromProcessModel over LIBORMarketModelFromCovarianceModel
 * (montecarlo/process/EulerSchemeFromProcessModel.java:96-123) on the MI355X engine.
 *
 * The model is read only through the reference's public API:
 *   - times, factors, paths, seed:      BrownianMotionFromMersenneRandomNumbers getters (:194-216)
 *   - tenor, measure:                   getLiborPeriodDiscretization(), getMeasure() (LMM:1533, 1547)
 *   - state space and liborCap:         applyStateSpaceTransform on probe states 0 and Double.MAX_VALUE
 *                                       (LMM:1200-1212; the two are private fields without getters,
 *                                       set from the properties map at LMM:400-414)
 *   - initial forwards:                 getForwardRateCurve().getForward(getAnalyticModel(), T_j, T_{j+1}-T_j),
 *                                       as getInitialState computes them (LMM:1081-1093)
 *   - factor loadings, variances:       getFactorLoading(process, i, j, null) (LMM:1194-1197) and
 *                                       getCovarianceModel().getCovariance(t_i, j, j, null) (LMM:1186)
 *   - discount factors:                 getDiscountCurve().getDiscountFactor(getAnalyticModel(), T_k)
 *   - volatility factorization:         for LIBORCovarianceModelFromVolatilityAndCorrelation, the
 *                                       volatility model's sigma_j(t_ci) and the correlation model's
 *                                       factor matrix (the adjoint's dV/dsigma needs both)
 * A model the engine does not take (local or stochastic volatility wrappers are set through their
 * own entry points; more than 80 forwards or 10 factors) makes of(...) return the reference's
 * own EulerSchemeFromProcessModel.
 */
package net.finmath.montecarlo.interestrate.hip;

import java.lang.foreign.Arena;
import java.lang.foreign.MemorySegment;

import net.finmath.exception.CalculationException;
import net.finmath.marketdata.model.curves.DiscountCurve;
import net.finmath.montecarlo.BrownianMotionFromMersenneRandomNumbers;
import net.finmath.montecarlo.RandomVariableFromDoubleArray;
import net.finmath.montecarlo.interestrate.models.LIBORMarketModelFromCovarianceModel;
import net.finmath.montecarlo.interestrate.models.covariance.LIBORCovarianceModel;
import net.finmath.montecarlo.interestrate.models.covariance.LIBORCovarianceModelFromVolatilityAndCorrelation;
import net.finmath.montecarlo.process.EulerSchemeFromProcessModel;
import net.finmath.stochastic.RandomVariable;
import net.finmath.time.TimeDiscretization;

import static java.lang.foreign.ValueLayout.ADDRESS;
import static java.lang.foreign.ValueLayout.JAVA_DOUBLE;

public class EulerSchemeFromProcessModelHIP extends EulerSchemeFromProcessModel implements AutoCloseable {

	private final Arena arena = Arena.ofShared();
	private final LIBORMarketModelFromCovarianceModel model;
	private final int numberOfPaths;
	private final double[] times;
	private final double[] tenor;
	private MemorySegment context;               // fm_lmm*
	private boolean simulated = false;

	/**
	 * The engine's process for the model, or the reference's own Euler scheme when the engine
	 * does not take the model (status FM_ERR_UNSUPPORTED).
	 */
	public static EulerSchemeFromProcessModel of(LIBORMarketModelFromCovarianceModel model,
			BrownianMotionFromMersenneRandomNumbers brownianMotion, Scheme scheme, int precision) throws CalculationException {
		try {
			return new EulerSchemeFromProcessModelHIP(model, brownianMotion, scheme, precision);
		}
		catch(final FinmathHip.UnsupportedModelException e) {
			return new EulerSchemeFromProcessModel(model, brownianMotion, scheme);
		}
	}

	public EulerSchemeFromProcessModelHIP(LIBORMarketModelFromCovarianceModel model,
			BrownianMotionFromMersenneRandomNumbers brownianMotion, Scheme scheme, int precision)
					throws CalculationException, FinmathHip.UnsupportedModelException {
		super(model, brownianMotion, scheme);        // the reference's getters and clone semantics stay
		this.model = model;
		this.numberOfPaths = brownianMotion.getNumberOfPaths();
		final TimeDiscretization td = brownianMotion.getTimeDiscretization();
		this.times = td.getAsDoubleArray();
		this.tenor = model.getLiborPeriodDiscretization().getAsDoubleArray();
		final int nT = times.length - 1, n = tenor.length - 1, factors = brownianMotion.getNumberOfFactors();

		// state space and cap: LOGNORMAL maps state 0 to exp(0) = 1, NORMAL to 0; the largest
		// state maps to the cap (or stays unbounded without one)
		final RandomVariable probe = applyStateSpaceTransform(0, 0,
				new RandomVariableFromDoubleArray(0.0, new double[] { 0.0, Double.MAX_VALUE }));
		final boolean lognormal = probe.get(0) == 1.0;
		final double top = probe.get(1);
		final double liborCap = (Double.isInfinite(top) || top == Double.MAX_VALUE) ? Double.POSITIVE_INFINITY : top;
		final int stateSpace = lognormal ? LIBORMarketModelFromCovarianceModel.StateSpace.LOGNORMAL.ordinal()
				: LIBORMarketModelFromCovarianceModel.StateSpace.NORMAL.ordinal();

		final MemorySegment config = FinmathHip.config(arena, times, tenor, factors, numberOfPaths,
				brownianMotion.getSeed(), model.getMeasure().ordinal(), stateSpace, scheme.ordinal(), precision,
				liborCap, true /* store paths: getProcessValue */, 4, false /* keep increments, as the reference */, 0L);
		final MemorySegment out = arena.allocate(ADDRESS);
		FinmathHip.checkSupported((int) FinmathHip.call(FinmathHip.CREATE, config, out));
		context = out.get(ADDRESS, 0);
		try {
			FinmathHip.check((int) FinmathHip.call(FinmathHip.SET_INITIAL_FORWARDS, context,
					FinmathHip.doubles(arena, initialForwards())));
			final double[] lambda = new double[nT * n * factors];
			final double[] variance = new double[nT * n];
			final LIBORCovarianceModel covariance = model.getCovarianceModel();
			for(int i = 0; i < nT; i++) {
				for(int j = 0; j < n; j++) {
					final RandomVariable[] loading = model.getFactorLoading(this, i, j, null);
					if(loading.length != factors || !loading[0].isDeterministic()) {
						throw new FinmathHip.UnsupportedModelException("factor loadings depend on the path");
					}
					for(int k = 0; k < factors; k++) {
						lambda[(i * n + j) * factors + k] = loading[k].doubleValue();
					}
					variance[i * n + j] = covariance.getCovariance(times[i], j, j, null).doubleValue();
				}
			}
			FinmathHip.check((int) FinmathHip.call(FinmathHip.SET_COVARIANCE, context,
					FinmathHip.doubles(arena, lambda), FinmathHip.doubles(arena, variance)));
			FinmathHip.check((int) FinmathHip.call(FinmathHip.SET_DISCOUNT_CURVE, context,
					FinmathHip.doubles(arena, discountFactors())));
			if(covariance instanceof LIBORCovarianceModelFromVolatilityAndCorrelation) {
				setVolatilityFactorization((LIBORCovarianceModelFromVolatilityAndCorrelation) covariance, factors);
			}
		}
		catch(final CalculationException | FinmathHip.UnsupportedModelException | RuntimeException e) {
			close();
			throw e;
		}
	}

	private double[] initialForwards() {
		final double[] forwards = new double[tenor.length - 1];
		for(int j = 0; j < forwards.length; j++) {
			forwards[j] = model.getForwardRateCurve().getForward(model.getAnalyticModel(), tenor[j], tenor[j + 1] - tenor[j]);
		}
		return forwards;
	}

	private double[] discountFactors() {
		final DiscountCurve curve = model.getDiscountCurve();
		if(curve == null) {
			return null;
		}
		final double[] df = new double[tenor.length];
		for(int k = 0; k < tenor.length; k++) {
			df[k] = curve.getDiscountFactor(model.getAnalyticModel(), tenor[k]);
		}
		return df;
	}

	// sigma_j(t_ci) on the simulation grid (t_ci: the covariance model's time index of t_i,
	// rounded down, AbstractLIBORCovarianceModel.java:54-60) and f_jk with c_j = rho_jj
	private void setVolatilityFactorization(LIBORCovarianceModelFromVolatilityAndCorrelation covariance, int factors)
			throws CalculationException {
		final int nT = times.length - 1, n = tenor.length - 1;
		final TimeDiscretization covarianceTimes = covariance.getTimeDiscretization();
		final double[] sigma = new double[nT * n];
		final double[] f = new double[n * factors];
		final double[] c = new double[n];
		for(int i = 0; i < nT; i++) {
			int ci = covarianceTimes.getTimeIndex(times[i]);
			if(ci < 0) {
				ci = -ci - 2;
			}
			for(int j = 0; j < n; j++) {
				sigma[i * n + j] = covariance.getVolatilityModel().getVolatility(ci, j).doubleValue();
			}
		}
		for(int j = 0; j < n; j++) {
			for(int k = 0; k < factors; k++) {
				f[j * factors + k] = covariance.getCorrelationModel().getFactorLoading(0, k, j);
			}
			c[j] = covariance.getCorrelationModel().getCorrelation(0, j, j);
		}
		FinmathHip.check((int) FinmathHip.call(FinmathHip.SET_VOLATILITY_FACTORIZATION, context,
				FinmathHip.doubles(arena, sigma), FinmathHip.doubles(arena, f), FinmathHip.doubles(arena, c)));
	}

	/** The engine context (fm_lmm*) for the valuation and gradient calls of this package. */
	MemorySegment context() {
		return context;
	}

	Arena arena() {
		return arena;
	}

	LIBORMarketModelFromCovarianceModel liborModel() {
		return model;
	}

	double[] tenorDates() {
		return tenor;
	}

	@Override
	public RandomVariable getProcessValue(final int timeIndex, final int componentIndex) {
		try {
			if(!simulated) {
				FinmathHip.check((int) FinmathHip.call(FinmathHip.SIMULATE, context));
				simulated = true;
			}
			final MemorySegment buffer = arena.allocate(JAVA_DOUBLE, numberOfPaths);
			FinmathHip.check((int) FinmathHip.call(FinmathHip.GET_PROCESS_VALUE, context, timeIndex, componentIndex, buffer));
			if(timeIndex == 0) {
				return new RandomVariableFromDoubleArray(getTime(0), buffer.getAtIndex(JAVA_DOUBLE, 0));
			}
			return new RandomVariableFromDoubleArray(getTime(timeIndex), buffer.toArray(JAVA_DOUBLE));
		}
		catch(final CalculationException e) {
			throw new RuntimeException(e);
		}
	}

	@Override
	public void close() {
		if(context != null && !context.equals(MemorySegment.NULL)) {
			FinmathHip.call(FinmathHip.DESTROY, context);
			context = MemorySegment.NULL;
		}
		arena.close();
	}
}

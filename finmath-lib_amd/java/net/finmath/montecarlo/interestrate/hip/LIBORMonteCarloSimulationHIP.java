/*
 * LIBORMonteCarloSimulationFromLIBORModel (montecarlo/interestrate/LIBORMonteCarloSimulationFromLIBORModel.java:27-206)
 * over the engine's process, with
 *   - getValues(products): every product the fused kernel knows is valued in one simulation
 *     pass (fm_lmm_value_products); any other product runs its own getValue(this), which reads
 *     the materialised process through getProcessValue;
 *   - getGradient(products, weights): the pathwise adjoint (fm_lmm_adjoint) keyed like
 *     LIBORMarketModelFromCovarianceModel.getModelParameters() (LMM:1699-1732): "FORWARD(T_j,T_{j+1})"
 *     holds dV/dL_j(0), "COVARIANCEMODELPARAMETER(i)" the chain rule of dV/dsigma_j(t_i) through
 *     LIBORVolatilityModelFourParameterExponentialForm (sigma = (a + b tau) exp(-c tau) + d, :168-195)
 *     when that is the volatility model and it is calibrateable (its getParameter() lists a, b, c, d
 *     first, LIBORCovarianceModelFromVolatilityAndCorrelation.java:96-108).
 *
 * Products the fused kernel knows:
 *   - Swaption: lowered through its public getters (products/Swaption.java:245-267);
 *   - CapletHIP, SimpleSwapHIP, BermudanSwaptionHIP below: the reference classes (whose fields have
 *     no getters) extended to remember their constructor arguments.
 */
package net.finmath.montecarlo.interestrate.hip;

import java.lang.foreign.Arena;
import java.lang.foreign.MemorySegment;
import java.util.ArrayList;
import java.util.HashMap;
import java.util.List;
import java.util.Map;

import net.finmath.exception.CalculationException;
import net.finmath.montecarlo.automaticdifferentiation.backward.RandomVariableDifferentiableAADFactory.DiracDeltaApproximationMethod;
import net.finmath.montecarlo.interestrate.LIBORMonteCarloSimulationFromLIBORModel;
import net.finmath.montecarlo.interestrate.models.covariance.LIBORCovarianceModel;
import net.finmath.montecarlo.interestrate.models.covariance.LIBORCovarianceModelFromVolatilityAndCorrelation;
import net.finmath.montecarlo.interestrate.models.covariance.LIBORVolatilityModel;
import net.finmath.montecarlo.interestrate.models.covariance.LIBORVolatilityModelFourParameterExponentialForm;
import net.finmath.montecarlo.interestrate.products.AbstractTermStructureMonteCarloProduct;
import net.finmath.montecarlo.interestrate.products.BermudanSwaption;
import net.finmath.montecarlo.interestrate.products.Caplet;
import net.finmath.montecarlo.interestrate.products.SimpleSwap;
import net.finmath.montecarlo.interestrate.products.Swaption;
import net.finmath.montecarlo.conditionalexpectation.RegressionBasisFunctionsProvider;
import net.finmath.stochastic.RandomVariable;
import net.finmath.time.TimeDiscretization;

import static java.lang.foreign.ValueLayout.ADDRESS;
import static java.lang.foreign.ValueLayout.JAVA_DOUBLE;
import static java.lang.foreign.ValueLayout.JAVA_INT;

public class LIBORMonteCarloSimulationHIP extends LIBORMonteCarloSimulationFromLIBORModel {

	private final EulerSchemeFromProcessModelHIP process;
	// the AAD factory's diracDeltaApproximationMethod and diracDeltaApproximationWidthPerStdDev
	// (RandomVariableDifferentiableAADFactory.java:60-67): the reference's defaults
	private DiracDeltaApproximationMethod diracDeltaApproximationMethod = DiracDeltaApproximationMethod.DISCRETE_DELTA;
	private double diracDeltaApproximationWidthPerStdDev = 0.05;

	public LIBORMonteCarloSimulationHIP(EulerSchemeFromProcessModelHIP process) {
		super(process.liborModel(), process);
		this.process = process;
	}

	// ---- products with their constructor arguments ------------------------------------------------

	/** Caplet (products/Caplet.java:60-100) remembering its arguments. */
	public static class CapletHIP extends Caplet {
		private static final long serialVersionUID = 1L;
		final double maturity, periodLength, strike, daycountFraction;
		final boolean isFloorlet;
		public CapletHIP(double maturity, double periodLength, double strike, double daycountFraction, boolean isFloorlet) {
			super(maturity, periodLength, strike, daycountFraction, isFloorlet, Caplet.ValueUnit.VALUE);
			this.maturity = maturity; this.periodLength = periodLength; this.strike = strike;
			this.daycountFraction = daycountFraction; this.isFloorlet = isFloorlet;
		}
		public CapletHIP(double maturity, double periodLength, double strike) {
			this(maturity, periodLength, strike, periodLength, false);
		}
	}

	/** SimpleSwap (products/SimpleSwap.java:39-62) remembering its arguments. */
	public static class SimpleSwapHIP extends SimpleSwap {
		private static final long serialVersionUID = 1L;
		final double[] fixingDates, paymentDates, swaprates, notional;
		final boolean isPayFix;
		public SimpleSwapHIP(double[] fixingDates, double[] paymentDates, double[] swaprates, boolean isPayFix, double[] notional) {
			super(fixingDates, paymentDates, swaprates, isPayFix, notional);
			this.fixingDates = fixingDates; this.paymentDates = paymentDates; this.swaprates = swaprates;
			this.notional = notional; this.isPayFix = isPayFix;
		}
	}

	/** BermudanSwaption (products/BermudanSwaption.java:53-86) remembering its arguments. */
	public static class BermudanSwaptionHIP extends BermudanSwaption {
		private static final long serialVersionUID = 1L;
		final boolean[] exercise;
		final double[] fixingDates, periodLengths, paymentDates, notionals, swaprates;
		final boolean isCallable;
		final int basis;
		/** basis: FinmathHip.FM_BASIS_REFERENCE (the product's own basis functions) or FM_BASIS_SWAP_RATE
		 *  with the provider that computes 1, S, S^2, N (the same functions on the Java side). */
		public BermudanSwaptionHIP(boolean[] isPeriodStartDateExerciseDate, double[] fixingDates, double[] periodLength,
				double[] paymentDates, double[] periodNotionals, double[] swaprates, boolean isCallable,
				RegressionBasisFunctionsProvider provider, int basis) {
			super(isPeriodStartDateExerciseDate, fixingDates, periodLength, paymentDates, periodNotionals, swaprates,
					isCallable, provider);
			this.exercise = isPeriodStartDateExerciseDate; this.fixingDates = fixingDates;
			this.periodLengths = periodLength; this.paymentDates = paymentDates; this.notionals = periodNotionals;
			this.swaprates = swaprates; this.isCallable = isCallable; this.basis = basis;
		}
	}

	/** The fm_product of a product, or null when the fused kernel does not know it. */
	static MemorySegment lower(Arena arena, AbstractTermStructureMonteCarloProduct product) {
		final MemorySegment p = arena.allocate(FinmathHip.PRODUCT_SIZE, 8);
		if(product instanceof Swaption) {
			final Swaption s = (Swaption) product;
			p.set(JAVA_INT, FinmathHip.P_TYPE, FinmathHip.FM_PRODUCT_SWAPTION);
			p.set(JAVA_INT, FinmathHip.P_NUM_PERIODS, s.getFixingDates().length);
			p.set(JAVA_DOUBLE, FinmathHip.P_EXERCISE_DATE, s.getExerciseDate());
			p.set(ADDRESS, FinmathHip.P_FIXING_DATES, FinmathHip.doubles(arena, s.getFixingDates()));
			p.set(ADDRESS, FinmathHip.P_PAYMENT_DATES, FinmathHip.doubles(arena, s.getPaymentDates()));
			p.set(ADDRESS, FinmathHip.P_PERIOD_LENGTHS, FinmathHip.doubles(arena, s.getPeriodLengths()));
			p.set(ADDRESS, FinmathHip.P_SWAPRATES, FinmathHip.doubles(arena, s.getSwaprates()));
			p.set(JAVA_DOUBLE, FinmathHip.P_NOTIONAL, s.getNotional());
			return p;
		}
		if(product instanceof CapletHIP) {
			final CapletHIP c = (CapletHIP) product;
			p.set(JAVA_INT, FinmathHip.P_TYPE, FinmathHip.FM_PRODUCT_CAPLET);
			p.set(JAVA_INT, FinmathHip.P_NUM_PERIODS, 1);
			p.set(JAVA_DOUBLE, FinmathHip.P_EXERCISE_DATE, c.maturity);
			p.set(ADDRESS, FinmathHip.P_PERIOD_LENGTHS, FinmathHip.doubles(arena, new double[] { c.periodLength }));
			p.set(ADDRESS, FinmathHip.P_SWAPRATES, FinmathHip.doubles(arena, new double[] { c.strike }));
			p.set(JAVA_DOUBLE, FinmathHip.P_NOTIONAL, c.daycountFraction);
			p.set(JAVA_INT, FinmathHip.P_IS_FLOORLET, c.isFloorlet ? 1 : 0);
			return p;
		}
		if(product instanceof SimpleSwapHIP) {
			final SimpleSwapHIP s = (SimpleSwapHIP) product;
			p.set(JAVA_INT, FinmathHip.P_TYPE, FinmathHip.FM_PRODUCT_SWAP);
			p.set(JAVA_INT, FinmathHip.P_NUM_PERIODS, s.fixingDates.length);
			p.set(ADDRESS, FinmathHip.P_FIXING_DATES, FinmathHip.doubles(arena, s.fixingDates));
			p.set(ADDRESS, FinmathHip.P_PAYMENT_DATES, FinmathHip.doubles(arena, s.paymentDates));
			p.set(ADDRESS, FinmathHip.P_SWAPRATES, FinmathHip.doubles(arena, s.swaprates));
			p.set(ADDRESS, FinmathHip.P_NOTIONALS, FinmathHip.doubles(arena, s.notional));
			p.set(JAVA_INT, FinmathHip.P_IS_PAY_FIX, s.isPayFix ? 1 : 0);
			p.set(JAVA_DOUBLE, FinmathHip.P_NOTIONAL, 1.0);
			return p;
		}
		if(product instanceof BermudanSwaptionHIP) {
			final BermudanSwaptionHIP b = (BermudanSwaptionHIP) product;
			final int[] flags = new int[b.exercise.length];
			for(int i = 0; i < flags.length; i++) {
				flags[i] = b.exercise[i] ? 1 : 0;
			}
			p.set(JAVA_INT, FinmathHip.P_TYPE, FinmathHip.FM_PRODUCT_BERMUDAN_SWAPTION);
			p.set(JAVA_INT, FinmathHip.P_NUM_PERIODS, b.fixingDates.length);
			p.set(ADDRESS, FinmathHip.P_FIXING_DATES, FinmathHip.doubles(arena, b.fixingDates));
			p.set(ADDRESS, FinmathHip.P_PAYMENT_DATES, FinmathHip.doubles(arena, b.paymentDates));
			p.set(ADDRESS, FinmathHip.P_PERIOD_LENGTHS, FinmathHip.doubles(arena, b.periodLengths));
			p.set(ADDRESS, FinmathHip.P_SWAPRATES, FinmathHip.doubles(arena, b.swaprates));
			p.set(ADDRESS, FinmathHip.P_NOTIONALS, FinmathHip.doubles(arena, b.notionals));
			p.set(ADDRESS, FinmathHip.P_EXERCISE_FLAGS, FinmathHip.ints(arena, flags));
			p.set(JAVA_INT, FinmathHip.P_IS_CANCELABLE, b.isCallable ? 0 : 1);
			p.set(JAVA_INT, FinmathHip.P_BASIS, b.basis);
			p.set(JAVA_DOUBLE, FinmathHip.P_NOTIONAL, 1.0);
			return p;
		}
		return null;
	}

	/** Values of the products at time 0 (AbstractMonteCarloProduct.getValue(model)). */
	public double[] getValues(List<AbstractTermStructureMonteCarloProduct> products) throws CalculationException {
		final double[] values = new double[products.size()];
		try(Arena arena = Arena.ofConfined()) {
			final List<Integer> fused = new ArrayList<>();
			final List<MemorySegment> lowered = new ArrayList<>();
			for(int q = 0; q < products.size(); q++) {
				final MemorySegment p = lower(arena, products.get(q));
				if(p != null) {
					fused.add(q);
					lowered.add(p);
				}
				else {
					values[q] = products.get(q).getValue(this);
				}
			}
			if(!fused.isEmpty()) {
				final MemorySegment array = arena.allocate(FinmathHip.PRODUCT_SIZE * lowered.size(), 8);
				for(int i = 0; i < lowered.size(); i++) {
					MemorySegment.copy(lowered.get(i), 0, array, i * FinmathHip.PRODUCT_SIZE, FinmathHip.PRODUCT_SIZE);
				}
				final MemorySegment v = arena.allocate(JAVA_DOUBLE, lowered.size());
				final MemorySegment e = arena.allocate(JAVA_DOUBLE, lowered.size());
				FinmathHip.check((int) FinmathHip.call(FinmathHip.VALUE_PRODUCTS, process.context(), array,
						lowered.size(), v, e));
				for(int i = 0; i < fused.size(); i++) {
					values[fused.get(i)] = v.getAtIndex(JAVA_DOUBLE, i);
				}
			}
		}
		return values;
	}

	/**
	 * The derivative of a Bermudan's exercise indicator in its trigger within getGradient, as the
	 * properties of the reference's RandomVariableDifferentiableAADFactory choose it. The engine
	 * takes the enum's ordinals (fm_lmm_set_dirac_delta); DISCRETE_DELTA, ONE and ZERO are supported.
	 */
	public void setDiracDeltaApproximation(DiracDeltaApproximationMethod method, double widthPerStdDev) {
		diracDeltaApproximationMethod = method;
		diracDeltaApproximationWidthPerStdDev = widthPerStdDev;
	}

	/**
	 * Gradient of sum_q weights[q] value_q, keyed like getModelParameters(). Every product must be
	 * one the fused kernel knows; averages over paths are constants, as in
	 * RandomVariableDifferentiableAAD.getGradient (backward/RandomVariableDifferentiableAAD.java:165-219).
	 */
	public Map<String, Double> getGradient(List<AbstractTermStructureMonteCarloProduct> products, double[] weights)
			throws CalculationException {
		final double[] tenor = process.tenorDates();
		final int n = tenor.length - 1, nT = process.getTimeDiscretization().getNumberOfTimeSteps();
		final Map<String, Double> gradient = new HashMap<>();
		try(Arena arena = Arena.ofConfined()) {
			final MemorySegment array = arena.allocate(FinmathHip.PRODUCT_SIZE * products.size(), 8);
			for(int q = 0; q < products.size(); q++) {
				final MemorySegment p = lower(arena, products.get(q));
				if(p == null) {
					throw new UnsupportedOperationException("the adjoint differentiates the products the fused kernel knows");
				}
				MemorySegment.copy(p, 0, array, q * FinmathHip.PRODUCT_SIZE, FinmathHip.PRODUCT_SIZE);
			}
			final MemorySegment value = arena.allocate(JAVA_DOUBLE);
			final MemorySegment dL0 = arena.allocate(JAVA_DOUBLE, n);
			final MemorySegment dSigma = arena.allocate(JAVA_DOUBLE, (long) nT * n);
			FinmathHip.check((int) FinmathHip.call(FinmathHip.SET_DIRAC_DELTA, process.context(),
					diracDeltaApproximationMethod.ordinal(), diracDeltaApproximationWidthPerStdDev));
			FinmathHip.check((int) FinmathHip.call(FinmathHip.ADJOINT, process.context(), array, products.size(),
					FinmathHip.doubles(arena, weights), value, dL0, dSigma));
			for(int j = 0; j < n; j++) {
				gradient.put("FORWARD(" + tenor[j] + "," + tenor[j + 1] + ")", dL0.getAtIndex(JAVA_DOUBLE, j));
			}
			final double[] parameterVega = parameterVega(dSigma.toArray(JAVA_DOUBLE), nT, n);
			if(parameterVega != null) {
				for(int i = 0; i < parameterVega.length; i++) {
					gradient.put("COVARIANCEMODELPARAMETER(" + i + ")", parameterVega[i]);
				}
			}
		}
		return gradient;
	}

	// dV/d(a, b, c, d) = sum_{i,j} dV/dsigma_j(t_i) dsigma/d(a, b, c, d) at tau = T_j - t_ci > 0
	private double[] parameterVega(double[] dSigma, int nT, int n) {
		final LIBORCovarianceModel covariance = process.liborModel().getCovarianceModel();
		if(!(covariance instanceof LIBORCovarianceModelFromVolatilityAndCorrelation)) {
			return null;
		}
		final LIBORVolatilityModel volatility = ((LIBORCovarianceModelFromVolatilityAndCorrelation) covariance).getVolatilityModel();
		if(!(volatility instanceof LIBORVolatilityModelFourParameterExponentialForm) || volatility.getParameter() == null) {
			return null;
		}
		final RandomVariable[] abcd = volatility.getParameter();
		final double a = abcd[0].doubleValue(), b = abcd[1].doubleValue(), c = abcd[2].doubleValue();
		final TimeDiscretization vt = volatility.getTimeDiscretization();
		final TimeDiscretization vtenor = volatility.getLiborPeriodDiscretization();
		final TimeDiscretization ct = covariance.getTimeDiscretization();
		final double[] pv = new double[4];
		for(int i = 0; i < nT; i++) {
			int ci = ct.getTimeIndex(process.getTime(i));
			if(ci < 0) {
				ci = -ci - 2;
			}
			for(int j = 0; j < n; j++) {
				final double tau = vtenor.getTime(j) - vt.getTime(ci);
				if(tau <= 0) {
					continue;
				}
				final double e = Math.exp(-c * tau), g = dSigma[i * n + j];
				pv[0] += g * e;
				pv[1] += g * tau * e;
				pv[2] += g * (-tau) * (a + b * tau) * e;
				pv[3] += g;
			}
		}
		return pv;
	}
}

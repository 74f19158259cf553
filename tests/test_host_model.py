"""Host-side model set-up of the mirror (time grids, curves, volatility,
correlation PCA, covariance tables) against the oracle and closed forms.
These run on the CPU: nothing here touches the device."""
import math

import numpy as np
import pytest

import _oracle as O
from finmath_hip import _finmath_host as H
from finmath_hip import workloads as W


def build_model(wl):
    td = H.TimeDiscretizationFromArray(0.0, wl.num_time_steps, wl.period)
    tenor = H.TimeDiscretizationFromArray(0.0, wl.num_libors, wl.period)
    fc = H.ForwardCurveInterpolation.createForwardCurveFromForwards(
        "forwardCurve", list(tenor.getAsDoubleArray()[:-1]), list(wl.forwards), wl.period)
    a, b, c, d = wl.vol_params
    vol = H.LIBORVolatilityModelFourParameterExponentialForm(td, tenor, a, b, c, d, False)
    corr = H.LIBORCorrelationModelExponentialDecay(td, tenor, wl.num_factors, wl.corr_decay)
    cov = H.LIBORCovarianceModelFromVolatilityAndCorrelation(td, tenor, vol, corr)
    model = H.LIBORMarketModelFromCovarianceModel.of(tenor, None, fc, H.DiscountCurveFromForwardCurve(fc), None, cov, [],
                                                     {"measure": "SPOT", "stateSpace": "LOGNORMAL"})
    return td, tenor, fc, vol, corr, cov, model


@pytest.mark.parametrize("name", ["c0", "c1", "c4"])
def test_tables_match_oracle(name):
    wl = W.ALL[name]()
    td, tenor, fc, vol, corr, cov, model = build_model(wl)
    t = O.ModelTables(wl)
    lam, var = cov.tables(td)
    assert np.max(np.abs(lam - t.lam)) < 1e-12          # PCA by QL vs the oracle's Jacobi
    assert np.max(np.abs(var - t.var)) <= 1e-16
    assert np.array_equal(model.getInitialForwards(), t.L0)
    dc = model.getDiscountCurve()
    pd = np.array([dc.getDiscountFactor(T) for T in tenor.getAsDoubleArray()])
    assert np.max(np.abs(pd - t.pd)) <= 1e-16


def test_factor_reduction_matches_numpy_pca():
    T = np.arange(30) * 0.5
    rho = np.exp(-0.07 * np.abs(T[:, None] - T[None, :]))
    f = H.factorReduction(rho, 4)
    g = O.factor_reduction(rho, 4)
    assert np.max(np.abs(f - g)) < 1e-12


def test_time_discretization_semantics():
    td = H.TimeDiscretizationFromArray(0.0, 40, 0.5)
    assert td.getNumberOfTimeSteps() == 40
    assert td.getTime(7) == 3.5 and td.getTimeStep(3) == 0.5
    assert td.getTimeIndex(3.5) == 7
    assert td.getTimeIndex(3.6) == -(8 + 1)           # Java binarySearch insertion point
    assert td.getTimeIndexNearestLessOrEqual(3.6) == 7
    assert td.getTimeIndex(3.5 + 1e-6) == 7           # rounding to the 1/8760 tick


def test_forward_curve_interpolation_and_offset_conversion():
    fc = H.ForwardCurveInterpolation.createForwardCurveFromForwards("f", [0.5, 1.0, 2.0], [0.02, 0.03, 0.05], 0.5)
    assert fc.getForward(0.0) == 0.02                   # constant extrapolation
    assert fc.getForward(1.0) == 0.03                   # node value
    assert abs(fc.getForward(1.5) - 0.04) < 1e-15       # linear
    assert fc.getForward(3.0) == 0.05
    f1 = fc.getForward(1.0, 1.0)                        # period conversion for another payment offset
    assert abs(f1 - ((1 + 0.03 * 0.5) ** 2 - 1.0)) < 1e-15


def test_swap_rate_flat_curve_is_the_curve_rate():
    fc = H.ForwardCurveInterpolation.createForwardCurveFromForwards("f", [0.0, 30.0], [0.03, 0.03], 0.5)
    tenor = H.TimeDiscretizationFromArray(5.0, 10, 0.5)
    assert abs(H.getForwardSwapRate(tenor, tenor, fc) - 0.03) < 1e-15


def test_volatility_four_parameter_form():
    td = H.TimeDiscretizationFromArray(0.0, 4, 0.5)
    tenor = H.TimeDiscretizationFromArray(0.0, 4, 0.5)
    v = H.LIBORVolatilityModelFourParameterExponentialForm(td, tenor, 0.2, 0.1, 0.25, 0.1, True)
    assert v.getVolatility(2, 1) == 0.0                 # already fixed
    tau = 1.0
    assert abs(v.getVolatility(0, 2) - ((0.1 * tau + 0.2) * math.exp(-0.25 * tau) + 0.1)) < 1e-16
    assert list(v.getParameter()) == [0.2, 0.1, 0.25, 0.1]


def test_brownian_motion_validation_message():
    td = H.TimeDiscretizationFromArray(0.0, 4, 0.5)
    with pytest.raises(ValueError, match="Number of paths must be greater or equal 1"):
        H.BrownianMotionFromMersenneRandomNumbers(td, 2, 0, 3141)


def test_model_rejects_calibration_and_unknown_properties():
    wl = W.c0()
    td, tenor, fc, vol, corr, cov, model = build_model(wl)
    with pytest.raises(NotImplementedError):
        H.LIBORMarketModelFromCovarianceModel.of(tenor, None, fc, None, None, cov, [object()], {})
    with pytest.raises(ValueError):
        H.LIBORMarketModelFromCovarianceModel.of(tenor, None, fc, None, None, cov, [], {"measure": "FOO"})


def test_local_volatility_wrappers_describe_their_factor():
    # DisplacedLocalVolatilityModel: L + d everywhere; BlendedLocalVolatilityModel:
    # (L - L a) + a f0(max(T_j - t_i, 0)), 1 in place of f0 without a curve
    wl = W.c0()
    td, tenor, fc, vol, corr, cov, model = build_model(wl)
    assert cov.localVolatility(td) is None
    alpha, b = H.DisplacedLocalVolatilityModel(cov, 4.0, False).localVolatility(td)
    assert alpha == 0.0 and b.shape == (wl.num_time_steps, wl.num_libors) and np.all(b == 4.0)
    t = O.ModelTables(wl)
    alpha, b = H.BlendedLocalVolatilityModel(cov, fc, 0.3, False).localVolatility(td)
    times, T = td.getAsDoubleArray(), tenor.getAsDoubleArray()
    expect = np.array([[0.3 * O.forward_curve(t.node_t, t.node_f, max(T[j] - times[i], 0.0))
                        for j in range(wl.num_libors)] for i in range(wl.num_time_steps)])
    assert alpha == 0.3 and np.array_equal(b, expect)
    # components past their fixing take f0(0)
    assert b[-1, 0] == 0.3 * O.forward_curve(t.node_t, t.node_f, 0.0)
    alpha, b = H.BlendedLocalVolatilityModel(cov, 0.3, False).localVolatility(td)
    assert np.all(b == 0.3)
    # the wrapper keeps the base model's tables
    lam0, var0 = cov.tables(td)
    lam1, var1 = H.DisplacedLocalVolatilityModel(cov, 4.0, False).tables(td)
    assert np.array_equal(lam0, lam1) and np.array_equal(var0, var1)


def test_local_volatility_wrapper_parameters_and_clones():
    wl = W.c0()
    td, tenor, fc, _, corr, _, model = build_model(wl)
    vol = H.LIBORVolatilityModelFourParameterExponentialForm(td, tenor, 0.2, 0.0, 0.25, 0.1, True)
    cov = H.LIBORCovarianceModelFromVolatilityAndCorrelation(td, tenor, vol, corr)
    fixed = H.DisplacedLocalVolatilityModel(cov, 4.0, False)
    assert list(fixed.getParameter()) == [0.2, 0.0, 0.25, 0.1]
    free = H.DisplacedLocalVolatilityModel(cov, 4.0, True)
    assert list(free.getParameter()) == [0.2, 0.0, 0.25, 0.1, 4.0]
    c = free.getCloneWithModifiedParameters([0.3, 0.01, 0.2, 0.05, 2.5])
    assert isinstance(c, H.DisplacedLocalVolatilityModel)
    assert c.getDisplacement() == 2.5 and list(c.getBaseCovarianceModel().getParameter()) == [0.3, 0.01, 0.2, 0.05]
    assert list(free.getParameter()) == [0.2, 0.0, 0.25, 0.1, 4.0]          # the original is unchanged
    c = fixed.getCloneWithModifiedParameters([0.3, 0.01, 0.2, 0.05])
    assert isinstance(c, H.DisplacedLocalVolatilityModel) and c.getDisplacement() == 4.0
    blended = H.BlendedLocalVolatilityModel(cov, fc, 0.5, True)
    c = blended.getCloneWithModifiedParameters([0.3, 0.01, 0.2, 0.05, 0.7])
    assert isinstance(c, H.BlendedLocalVolatilityModel) and c.getDisplacement() == 0.7
    assert blended.getDisplacement() == 0.5
    with pytest.raises(ValueError):
        free.getCloneWithModifiedParameters([0.3, 0.01, 0.2, 0.05])         # base model takes four
    with pytest.raises(Exception):
        H.DisplacedLocalVolatilityModel(free, 1.0, False)                  # wrappers do not nest


def test_stochastic_volatility_wrapper_parameters_and_nesting():
    wl = W.c0()
    td, tenor, fc, _, corr, _, model = build_model(wl)
    vol = H.LIBORVolatilityModelFourParameterExponentialForm(td, tenor, 0.2, 0.0, 0.25, 0.1, True)
    cov = H.LIBORCovarianceModelFromVolatilityAndCorrelation(td, tenor, vol, corr)
    bm = H.BrownianMotionFromMersenneRandomNumbers(td, 2, 1000, 7)
    blended = H.BlendedLocalVolatilityModel(cov, 0.0, False)
    sv = H.LIBORCovarianceModelStochasticVolatility(blended, bm, 0.01, -0.3, True)
    assert list(sv.getParameter()) == [0.2, 0.0, 0.25, 0.1, 0.01, -0.3]
    alpha, b = sv.localVolatility(td)                                       # the wrapped blend
    assert alpha == 0.0 and np.all(b == 0.0)
    c = sv.getCloneWithModifiedParameters([0.3, 0.01, 0.2, 0.05, 0.02, -0.5])
    assert isinstance(c, H.LIBORCovarianceModelStochasticVolatility)
    assert (c.getNu(), c.getRho()) == (0.02, -0.5)
    assert isinstance(c.getBaseCovarianceModel(), H.BlendedLocalVolatilityModel)
    fixed = H.LIBORCovarianceModelStochasticVolatility(cov, bm, 0.01, -0.3, False)
    assert list(fixed.getParameter()) == [0.2, 0.0, 0.25, 0.1]
    with pytest.raises(ValueError):
        H.LIBORCovarianceModelStochasticVolatility(cov, bm, 0.01, 1.5, False)          # rho outside [-1, 1]
    with pytest.raises(ValueError):
        H.LIBORCovarianceModelStochasticVolatility(cov, H.BrownianMotionFromMersenneRandomNumbers(td, 1, 1000, 7),
                                                   0.01, 0.0, False)                  # one factor
    with pytest.raises(Exception):
        H.LIBORCovarianceModelStochasticVolatility(sv, bm, 0.01, 0.0, False)          # no nesting
    with pytest.raises(Exception):
        H.DisplacedLocalVolatilityModel(sv, 1.0, False)                               # local vol goes inside


def test_heston_wrapper_parameters():
    wl = W.c0()
    td, tenor, fc, _, corr, _, model = build_model(wl)
    vol = H.LIBORVolatilityModelFourParameterExponentialForm(td, tenor, 0.2, 0.0, 0.25, 0.1, True)
    cov = H.LIBORCovarianceModelFromVolatilityAndCorrelation(td, tenor, vol, corr)
    bm = H.BrownianMotionFromMersenneRandomNumbers(td, 1, 1000, 7)
    hv = H.LIBORCovarianceModelStochasticHestonVolatility(cov, bm, 0.5, 0.8, 0.3, True)
    assert list(hv.getParameter()) == [0.2, 0.0, 0.25, 0.1, 0.5, 0.8, 0.3]
    c = hv.getCloneWithModifiedParameters([0.3, 0.01, 0.2, 0.05, 1.0, 0.9, 0.4])
    assert isinstance(c, H.LIBORCovarianceModelStochasticHestonVolatility)
    assert (c.getKappa(), c.getTheta(), c.getXi()) == (1.0, 0.9, 0.4)
    assert list(H.LIBORCovarianceModelStochasticHestonVolatility(cov, bm, 0.5, 0.8, 0.3, False).getParameter()) == \
        [0.2, 0.0, 0.25, 0.1]
    with pytest.raises(Exception):
        H.LIBORCovarianceModelStochasticVolatility(hv, H.BrownianMotionFromMersenneRandomNumbers(td, 2, 1000, 7),
                                                   0.1, 0.0, False)                    # no nesting


@pytest.mark.parametrize("F,K,T,vol", [(0.03, 0.03, 1.0, 0.2), (0.03, 0.04, 5.0, 0.35), (0.05, 0.02, 10.0, 0.1),
                                       (0.02, 0.021, 0.5, 0.6)])
def test_black_value_and_implied_volatility(F, K, T, vol):
    # AnalyticFormulas.blackScholesGeneralizedOptionValue against the closed form written with
    # scipy's normal distribution, and the implied volatility as its inverse
    from scipy.stats import norm
    pu = 3.7
    sd = vol * math.sqrt(T)
    d1 = (math.log(F / K) + 0.5 * sd * sd) / sd
    expect = (F * norm.cdf(d1) - K * norm.cdf(d1 - sd)) * pu
    v = H.blackScholesGeneralizedOptionValue(F, vol, T, K, pu)
    assert v == pytest.approx(expect, rel=1e-13)
    assert H.blackScholesOptionImpliedVolatility(F, T, K, pu, v) == pytest.approx(vol, rel=1e-10)
    assert H.blackScholesOptionImpliedVolatility(F, T, -0.01, pu, v) == 0.0        # strike <= 0: not an option


@pytest.mark.parametrize("F,K,T,vol", [(0.03, 0.03, 1.0, 0.006), (0.03, 0.045, 5.0, 0.01), (0.01, -0.005, 10.0, 0.004)])
def test_bachelier_value_and_implied_volatility(F, K, T, vol):
    from scipy.stats import norm
    pu = 2.5
    sd = vol * math.sqrt(T)
    d = (F - K) / sd
    expect = ((F - K) * norm.cdf(d) + sd * norm.pdf(d)) * pu
    v = H.bachelierOptionValue(F, vol, T, K, pu)
    assert v == pytest.approx(expect, rel=1e-13)
    assert H.bachelierOptionImpliedVolatility(F, T, K, pu, v) == pytest.approx(vol, rel=1e-10)


def test_swap_annuity_and_swaption_simple_units():
    wl = W.c0()
    td, tenor, fc, vol, corr, cov, model = build_model(wl)
    dc = H.DiscountCurveFromForwardCurve(fc)
    t = [2.0, 2.5, 3.0, 3.5, 4.0]
    expect = sum((t[i + 1] - t[i]) * dc.getDiscountFactor(t[i + 1]) for i in range(4)) / dc.getDiscountFactor(0.0)
    assert H.getSwapAnnuity(t, fc) == pytest.approx(expect, rel=1e-15)
    for unit in ("VALUE", "VOLATILITY", "VOLATILITYLOGNORMAL", "VOLATILITYNORMAL", "INTEGRATEDVARIANCELOGNORMAL",
                 "INTEGRATEDNORMALVARIANCE"):
        H.SwaptionSimple(0.03, t, unit)
    with pytest.raises(Exception):
        H.SwaptionSimple(0.03, t, "PRICE")


def test_piecewise_constant_volatility_buckets():
    # LIBORVolatilityModelPiecewiseConstant: one parameter per (simulation time, time to
    # maturity) bucket pair starting within the last tenor date, looked up at the nearest grid
    # points at or below t_i and T_j - t_i; 0 once fixed
    td = H.TimeDiscretizationFromArray(0.0, 20, 0.5)
    tenor = H.TimeDiscretizationFromArray(0.0, 20, 0.5)
    sg = [0.0, 1.0, 2.0, 5.0, 10.0]
    mg = [0.0, 1.0, 2.0, 5.0, 10.0]
    pairs = [(a, b) for a in range(5) for b in range(5) if sg[a] + mg[b] <= 10.0]
    assert len(pairs) == 18
    params = [0.004 + 0.0001 * k for k in range(len(pairs))]
    vol = H.LIBORVolatilityModelPiecewiseConstant(td, tenor, H.TimeDiscretizationFromArray(sg),
                                                  H.TimeDiscretizationFromArray(mg), params, True)
    assert list(vol.getParameter()) == params
    times, T = td.getAsDoubleArray(), tenor.getAsDoubleArray()
    for i in range(20):
        for j in range(20):
            tau = T[j] - times[i]
            if tau <= 0:
                assert vol.getVolatility(i, j) == 0.0
                continue
            a = max(k for k in range(5) if sg[k] <= times[i])
            b = max(k for k in range(5) if mg[k] <= tau)
            assert vol.getVolatility(i, j) == params[pairs.index((a, b))], (i, j)
    flat = H.LIBORVolatilityModelPiecewiseConstant(td, tenor, H.TimeDiscretizationFromArray(sg),
                                                   H.TimeDiscretizationFromArray(mg), 0.005, False)
    assert flat.getVolatility(3, 10) == 0.005 and list(flat.getParameter()) == []
    with pytest.raises(ValueError):
        H.LIBORVolatilityModelPiecewiseConstant(td, tenor, H.TimeDiscretizationFromArray(sg),
                                                H.TimeDiscretizationFromArray(mg), [0.01, 0.02], True)
    corr = H.LIBORCorrelationModelExponentialDecay(td, tenor, 3, 0.05)
    cov = H.LIBORCovarianceModelFromVolatilityAndCorrelation(td, tenor, vol, corr)
    c = cov.getCloneWithModifiedParameters([0.006] * 18)
    assert list(c.getParameter()) == [0.006] * 18

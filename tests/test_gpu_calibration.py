"""Calibration of the covariance model on the device (AbstractLIBORCovarianceModelParametric.
getCloneCalibrated, covariance/AbstractLIBORCovarianceModelParametric.java:156-276), through
the host mirror.

Targets are the prices of a swaption strip under known volatility parameters, simulated
on the calibration's own Brownian motion (the reference's defaults: 2,000 paths, seed
31415), so the calibration error has an exact zero at those parameters: Levenberg-
Marquardt started away from them must find them, and the calibrated clone must reprice
the targets.
"""
import numpy as np
import pytest

from finmath_hip import _finmath_host as H
from finmath_hip import workloads as W
from test_gpu_parity import say

pytestmark = pytest.mark.gpu

TRUE = [0.2, 0.05, 0.3, 0.1]


def build(vol_params, n=20, factors=3, calibrate_corr=False, corr=0.1):
    td = H.TimeDiscretizationFromArray(0.0, n, 0.5)
    tenor = H.TimeDiscretizationFromArray(0.0, n, 0.5)
    fc = H.ForwardCurveInterpolation.createForwardCurveFromForwards(
        "forwardCurve", list(tenor.getAsDoubleArray()[:-1]), list(0.02 + 0.001 * np.arange(n)), 0.5)
    vol = H.LIBORVolatilityModelFourParameterExponentialForm(td, tenor, *vol_params, True)
    cm = H.LIBORCorrelationModelExponentialDecay(td, tenor, factors, corr, calibrate_corr)
    cov = H.LIBORCovarianceModelFromVolatilityAndCorrelation(td, tenor, vol, cm)
    model = H.LIBORMarketModelFromCovarianceModel.of(tenor, None, fc, H.DiscountCurveFromForwardCurve(fc), None, cov,
                                                     [], {"measure": "SPOT", "stateSpace": "LOGNORMAL"})
    return td, tenor, fc, cov, model


def strip(fc):
    products = []
    for e in (1.0, 2.0, 3.0, 4.0, 5.0):
        fix = [e + 0.5 * i for i in range(10)]
        pay = [f + 0.5 for f in fix]
        t = H.TimeDiscretizationFromArray(fix + [pay[-1]])
        atm = H.getForwardSwapRate(t, t, fc)
        for k in (-0.01, 0.0, 0.01):
            products.append(H.Swaption(e, fix, pay, [atm + k] * 10))
    return products


def prices(model, td, factors, products, paths=2000, seed=31415):
    bm = H.BrownianMotionFromMersenneRandomNumbers(td, factors, paths, seed)
    sim = H.LIBORMonteCarloSimulationFromLIBORModel(H.EulerSchemeFromProcessModel(model, bm))
    return [r["value"] for r in sim.getValues(products)]


def test_calibration_recovers_the_parameters_that_generated_the_targets():
    say("case calibration four-parameter volatility, 15 swaptions, 2000 paths")
    td, tenor, fc, cov_true, model_true = build(TRUE)
    products = strip(fc)
    targets = prices(model_true, td, 3, products)
    td, tenor, fc, cov0, model0 = build([0.15, 0.0, 0.25, 0.15])
    calib = [H.CalibrationProduct(p, v, 1.0) for p, v in zip(products, targets)]
    cov_cal, info = cov0.getCloneCalibrated(model0, calib, {"accuracy": 1e-12, "maxIterations": 200})
    say(f"  parameters {info['parameters']} iterations {info['iterations']} valuations {info['valuations']} "
        f"rms {info['rootMeanSquaredError']:.3e}")
    assert np.allclose(info["parameters"], TRUE, rtol=0, atol=2e-4)
    assert info["rootMeanSquaredError"] < 1e-7
    # the calibrated clone reprices the targets on the same Brownian motion
    model_cal = H.LIBORMarketModelFromCovarianceModel.of(tenor, None, fc, H.DiscountCurveFromForwardCurve(fc), None,
                                                         cov_cal, [], {"measure": "SPOT", "stateSpace": "LOGNORMAL"})
    rep = prices(model_cal, td, 3, products)
    assert np.max(np.abs(np.array(rep) - np.array(targets))) < 1e-6


def test_calibration_with_the_correlation_parameter_reduces_the_error():
    say("case calibration volatility + correlation decay")
    td, tenor, fc, _, model_true = build(TRUE, corr=0.2)
    products = strip(fc)
    targets = prices(model_true, td, 3, products)
    td, tenor, fc, cov0, model0 = build([0.15, 0.0, 0.25, 0.15], calibrate_corr=True, corr=0.05)
    calib = [H.CalibrationProduct(p, v, 1.0) for p, v in zip(products, targets)]
    start_rms = float(np.sqrt(np.mean((np.array(prices(model0, td, 3, products)) - np.array(targets)) ** 2)))
    cov_cal, info = cov0.getCloneCalibrated(model0, calib, {"maxIterations": 60})
    say(f"  parameters {info['parameters']} rms {start_rms:.3e} -> {info['rootMeanSquaredError']:.3e}")
    assert len(info["parameters"]) == 5
    assert info["rootMeanSquaredError"] < 0.05 * start_rms


def test_calibration_without_calibrateable_parameters_is_refused():
    td, tenor, fc, cov, model = build(TRUE)
    fixed = H.LIBORCovarianceModelFromVolatilityAndCorrelation(
        td, tenor, H.LIBORVolatilityModelFourParameterExponentialForm(td, tenor, *TRUE, False),
        H.LIBORCorrelationModelExponentialDecay(td, tenor, 3, 0.1, False))
    with pytest.raises(ValueError):
        fixed.getCloneCalibrated(model, [H.CalibrationProduct(strip(fc)[0], 0.01, 1.0)], {})


def test_calibration_under_a_displaced_wrapper_in_the_normal_state_space():
    # the reference's LIBORMarketModelCalibrationTest DISPLACED set-up: normal state space,
    # DisplacedLocalVolatilityModel(base, 1/0.25, not calibrateable) over a calibrateable base
    say("case calibration displaced wrapper d=4, normal state space, 15 swaptions, 2000 paths")
    true_vol = [0.004, 0.001, 0.3, 0.002]

    def displaced(vol_params):
        td, tenor, fc, cov, _ = build(vol_params)
        wrapped = H.DisplacedLocalVolatilityModel(cov, 1.0 / 0.25, False)
        model = H.LIBORMarketModelFromCovarianceModel.of(tenor, None, fc, H.DiscountCurveFromForwardCurve(fc), None,
                                                         wrapped, [], {"measure": "SPOT", "stateSpace": "NORMAL"})
        return td, tenor, fc, wrapped, model

    td, tenor, fc, _, model_true = displaced(true_vol)
    products = strip(fc)
    targets = prices(model_true, td, 3, products)
    td, tenor, fc, cov0, model0 = displaced([0.003, 0.0, 0.25, 0.003])
    assert list(cov0.getParameter()) == [0.003, 0.0, 0.25, 0.003]
    calib = [H.CalibrationProduct(p, v, 1.0) for p, v in zip(products, targets)]
    cov_cal, info = cov0.getCloneCalibrated(model0, calib, {"accuracy": 1e-12, "maxIterations": 200})
    say(f"  parameters {info['parameters']} iterations {info['iterations']} rms {info['rootMeanSquaredError']:.3e}")
    assert isinstance(cov_cal, H.DisplacedLocalVolatilityModel) and cov_cal.getDisplacement() == 4.0
    assert np.allclose(info["parameters"], true_vol, rtol=0, atol=2e-5)
    assert info["rootMeanSquaredError"] < 1e-8


def simple_strip(fc, unit):
    # ATM and out-of-the-money payers: an in-the-money Monte Carlo price can fall below the
    # intrinsic value, where no implied volatility exists
    products = []
    for e in (1.0, 2.0, 3.0, 4.0, 5.0):
        tenor = [e + 0.5 * i for i in range(11)]
        t = H.TimeDiscretizationFromArray(tenor)
        atm = H.getForwardSwapRate(t, t, fc)
        for k in (0.0, 0.005, 0.01):
            products.append(H.SwaptionSimple(atm + k, tenor, unit))
    return products


def test_swaption_simple_value_units():
    # SwaptionSimple (SwaptionSimple.java:68-104): VALUE is the Swaption's price; the volatility
    # units invert Black's and Bachelier's formulas at the curve's par rate and annuity
    from scipy.optimize import brentq
    from scipy.stats import norm
    say("case SwaptionSimple value units")
    td, tenor, fc, cov, model = build(TRUE)
    bm = H.BrownianMotionFromMersenneRandomNumbers(td, 3, 5000, 31415)
    sim = H.LIBORMonteCarloSimulationFromLIBORModel(H.EulerSchemeFromProcessModel(model, bm))
    tenor_s = [2.0 + 0.5 * i for i in range(11)]
    t = H.TimeDiscretizationFromArray(tenor_s)
    par = H.getForwardSwapRate(t, t, fc)
    K = par + 0.005
    ref = sim.getValues([H.Swaption(2.0, tenor_s[:-1], tenor_s[1:], [K] * 10)])[0]["value"]
    vals = sim.getValues([H.SwaptionSimple(K, tenor_s, u) for u in ("VALUE", "VOLATILITYLOGNORMAL", "VOLATILITYNORMAL",
                                                                      "INTEGRATEDVARIANCELOGNORMAL")])
    assert vals[0]["value"] == ref
    A = H.getSwapAnnuity(tenor_s, fc)

    def black(v):
        sd = v * math.sqrt(2.0)
        d1 = (math.log(par / K) + 0.5 * sd * sd) / sd
        return (par * norm.cdf(d1) - K * norm.cdf(d1 - sd)) * A - ref

    def bachelier(v):
        sd = v * math.sqrt(2.0)
        d = (par - K) / sd
        return ((par - K) * norm.cdf(d) + sd * norm.pdf(d)) * A - ref

    lv, nv = brentq(black, 1e-4, 5.0, xtol=1e-15), brentq(bachelier, 1e-6, 1.0, xtol=1e-16)
    say(f"  price {ref:.6e} lognormal vol {vals[1]['value']:.10f} normal vol {vals[2]['value']:.10f}")
    assert vals[1]["value"] == pytest.approx(lv, rel=1e-9) and vals[1]["error"] == 0.0
    assert vals[2]["value"] == pytest.approx(nv, rel=1e-9)
    assert vals[3]["value"] == pytest.approx(lv * lv * 2.0, rel=1e-9)


import math  # noqa: E402


def test_calibration_to_lognormal_volatilities():
    # AbstractLIBORCovarianceModelParametric.getCloneCalibrated with SwaptionSimple products in
    # VOLATILITYLOGNORMAL (the reference's calibration tests quote volatilities): targets are
    # the implied volatilities of the known parameters on the same Brownian motion
    say("case calibration to lognormal volatilities, 15 swaptions, 2000 paths")
    td, tenor, fc, cov_true, model_true = build(TRUE)
    products = simple_strip(fc, "VOLATILITYLOGNORMAL")
    targets = prices(model_true, td, 3, products)
    say(f"  target volatilities {np.round(targets, 4)}")
    assert all(0.05 < v < 1.0 for v in targets)
    td, tenor, fc, cov0, model0 = build([0.15, 0.0, 0.25, 0.15])
    calib = [H.CalibrationProduct(p, v, 1.0) for p, v in zip(products, targets)]
    cov_cal, info = cov0.getCloneCalibrated(model0, calib, {"accuracy": 1e-12, "maxIterations": 200})
    say(f"  parameters {info['parameters']} iterations {info['iterations']} rms {info['rootMeanSquaredError']:.3e}")
    assert np.allclose(info["parameters"], TRUE, rtol=0, atol=2e-4)
    assert info["rootMeanSquaredError"] < 1e-7


def test_concurrent_jacobian_passes_give_the_same_calibration():
    # numberOfThreads (LevenbergMarquardt.java's parallel Jacobian): the bumped parameter vectors
    # are priced on separate contexts at once; the iterates are the sequential run's bit for bit
    say("case calibration numberOfThreads 1 and 4")
    td, tenor, fc, _, model_true = build(TRUE)
    products = strip(fc)
    targets = prices(model_true, td, 3, products)
    td, tenor, fc, cov0, model0 = build([0.15, 0.0, 0.25, 0.15])
    calib = [H.CalibrationProduct(p, v, 1.0) for p, v in zip(products, targets)]
    runs = [cov0.getCloneCalibrated(model0, calib, {"maxIterations": 30, "numberOfThreads": k})[1] for k in (1, 4)]
    say(f"  iterations {runs[0]['iterations']} / {runs[1]['iterations']}")
    assert runs[0]["parameters"] == runs[1]["parameters"]
    assert runs[0]["iterations"] == runs[1]["iterations"] and runs[0]["valuations"] == runs[1]["valuations"]


def test_reference_displaced_calibration_setup():
    # LIBORMarketModelCalibrationTest's DISPLACED model (src/test/java/net/finmath/montecarlo/
    # interestrate/LIBORMarketModelCalibrationTest.java:270-300): piecewise-constant volatility on
    # simulation-time x time-to-maturity buckets, exponential-decay correlation 0.05,
    # DisplacedLocalVolatilityModel(1/0.25, not calibrateable), spot measure, normal state space;
    # here on 20 semiannual forwards with buckets {0, 1, 2, 5, 10} (18 parameters) and targets
    # from a known bucket pattern
    say("case reference displaced calibration set-up, 18 piecewise-constant volatilities")
    n = 20
    td = H.TimeDiscretizationFromArray(0.0, n, 0.5)
    tenor = H.TimeDiscretizationFromArray(0.0, n, 0.5)
    fc = H.ForwardCurveInterpolation.createForwardCurveFromForwards(
        "forwardCurve", list(tenor.getAsDoubleArray()[:-1]), list(0.02 + 0.001 * np.arange(n)), 0.5)
    grid = H.TimeDiscretizationFromArray([0.0, 1.0, 2.0, 5.0, 10.0])

    def displaced(params):
        vol = H.LIBORVolatilityModelPiecewiseConstant(td, tenor, grid, grid, params, True)
        corr = H.LIBORCorrelationModelExponentialDecay(td, tenor, 3, 0.05, False)
        cov = H.DisplacedLocalVolatilityModel(H.LIBORCovarianceModelFromVolatilityAndCorrelation(td, tenor, vol, corr),
                                              1.0 / 0.25, False)
        model = H.LIBORMarketModelFromCovarianceModel.of(tenor, None, fc, H.DiscountCurveFromForwardCurve(fc), None,
                                                         cov, [], {"measure": "SPOT", "stateSpace": "NORMAL"})
        return cov, model

    true = [0.005 + 0.0002 * ((7 * k) % 5) for k in range(18)]
    _, model_true = displaced(true)
    products = strip(fc)
    targets = prices(model_true, td, 3, products)
    cov0, model0 = displaced(0.005)
    assert len(cov0.getParameter()) == 18
    start = float(np.sqrt(np.mean((np.array(prices(model0, td, 3, products)) - np.array(targets)) ** 2)))
    cov_cal, info = cov0.getCloneCalibrated(model0, [H.CalibrationProduct(p, v, 1.0) for p, v in zip(products, targets)],
                                            {"maxIterations": 100})
    say(f"  rms {start:.3e} -> {info['rootMeanSquaredError']:.3e} in {info['iterations']} iterations, "
        f"{info['valuations']} passes")
    assert isinstance(cov_cal, H.DisplacedLocalVolatilityModel)
    assert info["rootMeanSquaredError"] < 0.02 * start

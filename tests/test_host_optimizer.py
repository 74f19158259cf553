"""The host mirror's Levenberg-Marquardt (optimizer/LevenbergMarquardt.java) and the
covariance model's parameter handling, on the CPU (no device).

Pins: the least-squares minimum found by scipy.optimize.least_squares (an independent
trust-region implementation) on a noisy curve fit; exact recovery of exactly representable
data; a rank-deficient problem (the pseudo-inverse step of solveLinearEquationSVD keeps it
finite); points with NaN values are rejected; the covariance parameters round trip.
"""
import math

import numpy as np
import pytest
from scipy.optimize import least_squares

from finmath_hip import _finmath_host as H

X = np.linspace(0.0, 4.0, 25)


def curve(p):
    return p[0] * np.exp(-p[1] * X) + p[2]


def test_exactly_representable_data_is_recovered():
    truth = [2.0, 0.7, 0.3]
    lm = H.LevenbergMarquardt([1.0, 0.2, 0.0], list(curve(truth)), 200)
    lm.setErrorTolerance(0.0)
    lm.run(lambda p: list(curve(p)))
    assert np.allclose(lm.getBestFitParameters(), truth, rtol=0, atol=1e-9)
    assert lm.getRootMeanSquaredError() < 1e-12


def test_noisy_fit_matches_an_independent_least_squares_solver():
    rng = np.random.default_rng(4)
    y = curve([2.0, 0.7, 0.3]) + rng.normal(0, 0.02, X.size)
    lm = H.LevenbergMarquardt([1.0, 0.2, 0.0], list(y), 500)
    lm.setErrorTolerance(1e-15).setParameterSteps([1e-7, 1e-7, 1e-7])
    lm.run(lambda p: list(curve(p)))
    ref = least_squares(lambda p: curve(p) - y, [1.0, 0.2, 0.0], xtol=1e-15, ftol=1e-15, gtol=1e-15)
    assert np.allclose(lm.getBestFitParameters(), ref.x, rtol=0, atol=1e-6)
    assert lm.getRootMeanSquaredError() == pytest.approx(math.sqrt(np.mean(ref.fun ** 2)), rel=1e-8)


def test_weights_enter_the_error():
    y = curve([2.0, 0.7, 0.3])
    y2 = y.copy()
    y2[0] += 1.0                                   # an outlier with zero weight is ignored
    w = np.ones(X.size)
    w[0] = 0.0
    lm = H.LevenbergMarquardt([1.0, 0.2, 0.0], list(y2), 300)
    lm.setWeights(list(w)).setErrorTolerance(0.0)
    lm.run(lambda p: list(curve(p)))
    assert np.allclose(lm.getBestFitParameters(), [2.0, 0.7, 0.3], atol=1e-8)


def test_rank_deficient_problem_stays_finite():
    # only p0 + p1 is identified: the normal equations are singular
    y = 1.5 * X
    lm = H.LevenbergMarquardt([0.1, 0.2], list(y), 100)
    lm.setErrorTolerance(0.0)
    lm.run(lambda p: list((p[0] + p[1]) * X))
    p = lm.getBestFitParameters()
    assert all(math.isfinite(v) for v in p)
    assert p[0] + p[1] == pytest.approx(1.5, abs=1e-10)


def test_points_with_nan_values_are_rejected():
    y = curve([2.0, 0.7, 0.3])
    def f(p):
        if p[1] < 0.5:
            return [math.nan] * X.size            # invalid region
        return list(curve(p))
    lm = H.LevenbergMarquardt([1.0, 1.5, 0.0], list(y), 300)
    lm.setErrorTolerance(0.0)
    lm.run(f)
    assert lm.getBestFitParameters()[1] >= 0.5
    assert np.allclose(lm.getBestFitParameters(), [2.0, 0.7, 0.3], atol=1e-7)


def test_covariance_parameters_round_trip():
    td = H.TimeDiscretizationFromArray(0.0, 20, 0.5)
    tenor = H.TimeDiscretizationFromArray(0.0, 20, 0.5)
    vol = H.LIBORVolatilityModelFourParameterExponentialForm(td, tenor, 0.2, 0.05, 0.3, 0.1, True)
    corr = H.LIBORCorrelationModelExponentialDecay(td, tenor, 3, 0.1, True)
    cov = H.LIBORCovarianceModelFromVolatilityAndCorrelation(td, tenor, vol, corr)
    assert cov.getParameter() == [0.2, 0.05, 0.3, 0.1, 0.1]
    c2 = cov.getCloneWithModifiedParameters([0.25, 0.0, 0.2, 0.12, 0.07])
    assert c2.getParameter() == [0.25, 0.0, 0.2, 0.12, 0.07]
    tau = 3.5 - 1.0
    sigma = (0.0 * tau + 0.25) * math.exp(0.2 * (-tau)) + 0.12
    assert c2.getFactorLoading(2, 7) == pytest.approx(
        list(sigma * np.asarray(H.LIBORCorrelationModelExponentialDecay(td, tenor, 3, 0.07, True).factorMatrix()[7])),
        rel=1e-15)
    fixed = H.LIBORCovarianceModelFromVolatilityAndCorrelation(
        td, tenor, H.LIBORVolatilityModelFourParameterExponentialForm(td, tenor, 0.2, 0.0, 0.25, 0.1, False),
        H.LIBORCorrelationModelExponentialDecay(td, tenor, 3, 0.1, False))
    assert fixed.getParameter() == []


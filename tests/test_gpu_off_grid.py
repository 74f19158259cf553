"""Forwards and numeraires off the tenor grid on the device random variables
(randomvariable.LIBORMonteCarloSimulation): LOG_LINEAR_UNCORRECTED interpolation of forwards
(LIBORMarketModelFromCovarianceModel.java:1236-1381), the numeraire between tenor dates
(:968-1015) and the log-linear discount-curve adjustment (:883-906).

Bars: the oracle's off-grid forwards and numeraires within 1e-12 relative (both measures,
with and without discount curve, LINEAR, LOG_LINEAR_UNCORRECTED and LOG_LINEAR_CORRECTED, the
last also under a displaced local-volatility wrapper); the reference's LIBORMarketModelValuationTest.testBond
(|bond - P^d(T)| < 8e-3 for maturities every 0.125 to 20y) and testFRA (|par FRA| < 2e-3 for
starts every 0.125) on the reference's model, 20,000 paths, PREDICTOR_CORRECTOR.
"""
import numpy as np
import pytest

import _oracle as O
from finmath_hip import _abi
from finmath_hip import _finmath_host as H
from finmath_hip import randomvariable as RV
from finmath_hip import workloads as W
from test_gpu_parity import say
from test_oracle_lmm import FineGrid

pytestmark = pytest.mark.gpu


INTERPOLATION = ["LINEAR", "LOG_LINEAR_UNCORRECTED", "LOG_LINEAR_CORRECTED"]


def context(wl, t, P, measure, scheme="EULER_FUNCTIONAL", with_discount=True, local_vol=None, state_space="LOGNORMAL"):
    ctx = _abi.LmmContext(t.times, t.tenor, wl.num_factors, P, wl.seed, measure=measure, scheme=scheme,
                          store_paths=True, state_space=state_space)
    ctx.set_initial_forwards(t.L0)
    ctx.set_covariance(t.lam, t.var)
    ctx.set_discount_curve(t.pd if with_discount else None)
    if local_vol is not None:
        ctx.set_local_volatility(*local_vol)
    ctx.simulate()
    return ctx


def curve(t):
    return H.ForwardCurveInterpolation.createForwardCurveFromForwards("forwardCurve", list(t.node_t), list(t.node_f),
                                                                      t.wl.period)


def rel(g, r):
    return float(np.max(np.abs(g - r) / np.maximum(np.abs(r), 1e-3)))


@pytest.mark.parametrize("interpolation", INTERPOLATION)
@pytest.mark.parametrize("with_discount", [True, False])
@pytest.mark.parametrize("measure", ["SPOT", "TERMINAL"])
def test_off_grid_numeraire_and_forwards_match_oracle(measure, with_discount, interpolation):
    # terminal: the simulation grid every 0.125 holds the off-grid times; spot: the 0.5 grid
    forwards = np.array([0.02 + 0.0015 * j for j in range(12)])
    if measure == "TERMINAL":
        wl = FineGrid("fine", 12, 0.5, 3, 3001, 0.1, forwards)
    else:
        wl = W.Workload("coarse", 12, 0.5, 3, 3001, 0.1, forwards)
    t = O.ModelTables(wl)
    P = wl.num_paths
    say(f"case off-grid numeraire/forwards {measure} discount={with_discount} {interpolation} P={P}")
    ctx = context(wl, t, P, measure, with_discount=with_discount)
    sim = RV.LIBORMonteCarloSimulation(ctx, curve(t), interpolation)
    om = O.OracleModel(t, P, measure=1 if measure == "TERMINAL" else 0, scheme=2, with_discount=with_discount,
                       interpolation=INTERPOLATION.index(interpolation))
    worst = 0.0
    for T in np.arange(0, 12 * 4 + 1) * 0.125:
        worst = max(worst, rel(sim.getNumeraire(T).getRealizations(), om.numeraire(T)))
    say(f"  worst relative numeraire difference {worst:.3e}")
    assert worst < 1e-12
    worst = 0.0
    cases = [(S, S, S + 0.5) for S in np.arange(1, 40) * 0.125] + \
            [(0.0, 0.875, 1.375), (1.0, 0.75, 2.0), (0.3, 1.125, 2.625), (5.0, 2.375, 3.0), (0.0, 1.0, 2.5)]
    for time, S, E in cases:
        worst = max(worst, rel(sim.getForwardRate(time, S, E).getRealizations(), om.forward_rate(time, S, E)))
    say(f"  worst relative forward difference {worst:.3e}")
    assert worst < 1e-12


@pytest.mark.parametrize("state_space", ["NORMAL", "LOGNORMAL"])
def test_corrected_interpolation_under_a_displaced_wrapper(state_space):
    # the drift adjustment integrates the wrapped loadings lambda (L + d)
    wl = FineGrid("fine", 10, 0.5, 3, 2001, 0.1, np.array([0.02 + 0.002 * j for j in range(10)]))
    t = O.ModelTables(wl, vol_params=(0.005, 0.0, 0.25, 0.002) if state_space == "NORMAL" else (0.05, 0.0, 0.25, 0.02))
    lv = (0.0, np.full((wl.num_time_steps, wl.num_libors), 4.0))
    P = wl.num_paths
    say(f"case off-grid corrected forwards, displaced wrapper, {state_space} P={P}")
    ctx = context(wl, t, P, "SPOT", local_vol=lv, state_space=state_space)
    sim = RV.LIBORMonteCarloSimulation(ctx, curve(t), "LOG_LINEAR_CORRECTED")
    om = O.OracleModel(t, P, scheme=2, state_space=1 if state_space == "LOGNORMAL" else 0, local_vol=lv,
                       interpolation=2)
    worst = 0.0
    for time, S, E in [(S, S, S + 0.5) for S in np.arange(1, 30) * 0.125] + [(0.5, 1.875, 3.125)]:
        worst = max(worst, rel(sim.getForwardRate(time, S, E).getRealizations(), om.forward_rate(time, S, E)))
    for T in np.arange(1, 40) * 0.125:
        worst = max(worst, rel(sim.getNumeraire(T).getRealizations(), om.numeraire(T)))
    say(f"  worst relative difference {worst:.3e}")
    assert worst < 1e-12


def test_on_grid_forwards_are_the_process_values():
    wl = W.c0().with_paths(2000)
    t = O.ModelTables(wl)
    ctx = context(wl, t, wl.num_paths, "SPOT")
    sim = RV.LIBORMonteCarloSimulation(ctx, curve(t))
    for ti, j in ((0, 3), (4, 4), (6, 9)):
        T = float(t.tenor[j])
        a = sim.getForwardRate(float(t.times[ti]), T, T + 0.5).getRealizations(ctx.P)
        assert np.array_equal(a, ctx.process_value(min(ti, j), j))


def test_reference_bond_and_fra_assertions():
    # LIBORMarketModelValuationTest (:84-161, :164-256): 40 semiannual forwards, flat 5 % curve
    # on nodes {0.5, 1, 2, 5, 40}, volatility (0.2, 0, 0.25, 0.3), 6 factors, correlation decay
    # 0.1, 20,000 paths, seed 3141, PREDICTOR_CORRECTOR, simulation every 0.5
    from test_gpu_reference_valuation import create_libor_market_model, par_swaprate
    _, td, tenor, fc, dc, cov = create_libor_market_model()
    say("case reference testBond / testFRA, 20000 paths, off-grid every 0.125")
    lam, var = cov.tables(td)
    times, T = td.getAsDoubleArray(), tenor.getAsDoubleArray()
    ctx = _abi.LmmContext(times, T, 6, 20_000, 3141, scheme="PREDICTOR_CORRECTOR", store_paths=True)
    ctx.set_initial_forwards([fc.getForward(x) for x in T[:-1]])
    ctx.set_covariance(lam, var)
    ctx.set_discount_curve([dc.getDiscountFactor(x) for x in T])
    ctx.simulate()
    sim = RV.LIBORMonteCarloSimulation(ctx, fc)
    N0, w0 = sim.getNumeraire(0.0), sim.getMonteCarloWeights(0.0)
    worst = 0.0
    for maturity in np.arange(0, 161) * 0.125:
        # Bond.getValue (products/Bond.java:59-91)
        v = RV.RandomVariable.constant(1.0).div(sim.getNumeraire(maturity)).mult(sim.getMonteCarloWeights(maturity))
        bond = v.mult(N0).div(w0).getAverage()
        worst = max(worst, abs(bond - dc.getDiscountFactor(maturity)))
    say(f"  testBond maximum abs deviation {worst:.3e}")
    assert worst < 8e-3
    worst = 0.0
    for start in np.arange(0, 157) * 0.125:
        end = start + 0.5
        K = par_swaprate([start, end], fc)
        # SimpleSwap.getValue (products/SimpleSwap.java:137-170), one period, payer
        libor = sim.getForwardRate(start, start, end)
        payoff = libor.sub(K).mult(end - start).div(sim.getNumeraire(end)).mult(sim.getMonteCarloWeights(end))
        value = payoff.mult(N0).div(w0).getAverage()
        worst = max(worst, abs(value))
    say(f"  testFRA maximum abs deviation {worst:.3e}")
    assert worst < 2e-3


def test_corrected_interpolation_refuses_a_stochastic_volatility_scaling():
    wl = W.c0().with_paths(512)
    t = O.ModelTables(wl)
    ctx = context(wl, t, wl.num_paths, "SPOT")
    driver = _abi.LmmContext(t.times, t.tenor, 2, wl.num_paths, 7)
    ctx.set_stochastic_volatility(0.2, -0.3, driver)
    RV.LIBORMonteCarloSimulation(ctx, curve(t), "LOG_LINEAR_UNCORRECTED")
    with pytest.raises(_abi.UnsupportedOperationException):
        RV.LIBORMonteCarloSimulation(ctx, curve(t), "LOG_LINEAR_CORRECTED")

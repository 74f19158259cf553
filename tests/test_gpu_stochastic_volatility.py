"""GPU parity of the stochastic volatility covariance wrapper against the oracle, through the
C ABI: LIBORCovarianceModelStochasticVolatility (covariance/LIBORCovarianceModelStochasticVolatility.java:
174-256) multiplies every loading at t_i by S(t_i), the exponential of a functional Euler
process with drift -nu^2/2 and loadings (rho nu, sqrt(1 - rho^2) nu) on two factors of a
Brownian motion: a separate one, or two factors of the forwards' own (a BrownianMotionView).

Bars: process values and numeraires in FP64 within 1e-12 relative of the oracle, prices within
1e-10 relative, FP32-mixed prices within 0.1 bp of the FP64 oracle; bonds reprice the initial
curve within 5 Monte Carlo errors at 200,000 paths; removing the scaling restores the plain
model bit for bit; serialisation and the adjoint refuse the scaling.
"""
import math

import numpy as np
import pytest

import _oracle as O
from finmath_hip import _abi
from finmath_hip import workloads as W
from test_gpu_parity import say, to_products
from test_gpu_variants import MEASURES, SCHEMES

pytestmark = pytest.mark.gpu

NU, RHO, DRIVER_SEED = 0.3, -0.3, 31415


def sv_context(wl, t, P, measure="SPOT", scheme="EULER_FUNCTIONAL", state_space="NORMAL", *, blend=False,
               shared=None, precision="FP64", store_paths=False, with_discount=True, nu=NU):
    """A context with the scaling, and its oracle description. shared = (a, b): the scaling is
    driven by factors a, b of the context's own Brownian motion, else by a separate 2-factor one."""
    ctx = _abi.LmmContext(t.times, t.tenor, wl.num_factors, P, wl.seed, measure=measure, scheme=scheme,
                          state_space=state_space, precision=precision, store_paths=store_paths)
    ctx.set_initial_forwards(t.L0)
    ctx.set_covariance(t.lam, t.var)
    ctx.set_discount_curve(t.pd if with_discount else None)
    lv = None
    if blend:
        # BlendedLocalVolatilityModel(cov, 0.0): the factor L - L 0 + 0 f0 = L
        lv = (0.0, np.zeros((wl.num_time_steps, wl.num_libors)))
        ctx.set_local_volatility(*lv)
    if shared is None:
        driver = _abi.LmmContext(t.times, t.tenor, 2, P, DRIVER_SEED)
        ctx.set_stochastic_volatility(nu, RHO, driver, 0, 1)
        sv = (nu, RHO, DRIVER_SEED, 2, 0, 1)
    else:
        ctx.set_stochastic_volatility(nu, RHO, ctx, *shared)
        sv = (nu, RHO, wl.seed, wl.num_factors, *shared)
    return ctx, lv, sv


def tables(wl, state_space, blend):
    if blend or state_space == "LOGNORMAL":
        return O.ModelTables(wl, vol_params=(0.2, 0.0, 0.25, 0.1))
    return O.ModelTables(wl, vol_params=(0.006, 0.0, 0.25, 0.002))


def oracle(t, P, measure, scheme, state_space, lv, sv, **kw):
    return O.OracleModel(t, P, measure=MEASURES.index(measure), scheme=SCHEMES.index(scheme),
                         state_space=1 if state_space == "LOGNORMAL" else 0, local_vol=lv, stochastic_vol=sv, **kw)


@pytest.mark.parametrize("measure,scheme,state_space,blend,shared", [
    ("SPOT", "EULER_FUNCTIONAL", "NORMAL", True, None),
    ("SPOT", "EULER_FUNCTIONAL", "LOGNORMAL", False, None),
    ("TERMINAL", "PREDICTOR_CORRECTOR", "NORMAL", True, None),
    ("TERMINAL", "EULER", "LOGNORMAL", False, None),
    ("SPOT", "PREDICTOR_CORRECTOR", "LOGNORMAL", False, (0, 2)),
    ("SPOT", "EULER_FUNCTIONAL", "NORMAL", True, (1, 2)),
])
def test_process_and_numeraire_match_oracle(measure, scheme, state_space, blend, shared):
    wl = W.c0()
    t = tables(wl, state_space, blend)
    P = 4001
    say(f"case stochastic-vol process c0 P={P} {measure} {scheme} {state_space} blend={blend} shared={shared}")
    ctx, lv, sv = sv_context(wl, t, P, measure, scheme, state_space, blend=blend, shared=shared, store_paths=True)
    ctx.simulate()
    om = oracle(t, P, measure, scheme, state_space, lv, sv)
    worst = 0.0
    for ti in range(wl.num_time_steps + 1):
        for j in range(wl.num_libors):
            g, r = ctx.process_value(ti, j), om.process_value(ti, j)
            worst = max(worst, float(np.max(np.abs(g - r) / np.maximum(np.abs(r), 1e-3))))
    say(f"  worst relative process difference {worst:.3e}")
    assert worst < 1e-12
    for k in range(wl.num_libors + 1):
        g, r = ctx.numeraire(t.tenor[k]), om.numeraire(t.tenor[k])
        assert np.max(np.abs(g - r) / np.abs(r)) < 1e-12, k


@pytest.mark.parametrize("measure,scheme", [("SPOT", "EULER_FUNCTIONAL"), ("TERMINAL", "PREDICTOR_CORRECTOR")])
def test_prices_match_oracle(measure, scheme):
    # the reference's smile set-up: normal state space, blended (log-normal) local volatility,
    # stochastic volatility scaling
    wl = W.c1()
    t = tables(wl, "NORMAL", True)
    P = 10_000
    say(f"case stochastic-vol prices c1 P={P} {measure} {scheme}")
    ctx, lv, sv = sv_context(wl, t, P, measure, scheme, "NORMAL", blend=True)
    vals, errs = ctx.value_products(to_products(wl.products))
    om = oracle(t, P, measure, scheme, "NORMAL", lv, sv)
    worst = 0.0
    for q, s in enumerate(wl.products):
        rv, re_ = om.value(s)
        worst = max(worst, abs(vals[q] - rv) / abs(rv))
        assert abs(errs[q] - re_) <= 1e-8 * abs(re_)
    say(f"  worst relative price difference {worst:.3e}")
    assert worst <= 1e-10
    fctx, _, _ = sv_context(wl, t, P, measure, scheme, "NORMAL", blend=True, precision="FP32_MIXED")
    fvals, _ = fctx.value_products(to_products(wl.products))
    worst = max(abs(fvals[q] - om.value(s)[0]) * 1e4 for q, s in enumerate(wl.products))
    say(f"  worst FP32-mixed price error {worst:.4f} bp")
    assert worst < 0.1


def test_bonds_reprice_the_initial_curve_at_full_size():
    # vol of vol 0.1: with 0.3 the scaling's log-volatility reaches 1.3 at 20y and the Euler
    # scheme's discretisation bias of the 20y bond (measured 0.7 %, reproduced by the oracle)
    # exceeds the Monte Carlo error
    wl = W.c1()
    t = tables(wl, "NORMAL", True)
    P = 200_000
    say(f"case stochastic-vol bonds c1 P={P}")
    ctx, _, _ = sv_context(wl, t, P, "SPOT", "EULER_FUNCTIONAL", "NORMAL", blend=True, store_paths=True,
                           with_discount=False, nu=0.1)
    ctx.simulate()
    N0 = ctx.numeraire(0.0)
    worst = 0.0
    for k in range(wl.num_libors + 1):
        x = N0 / ctx.numeraire(t.tenor[k])
        m, se = float(np.mean(x)), float(np.std(x) / math.sqrt(P))
        exact = float(np.prod(1.0 / (1.0 + 0.5 * t.L0[:k])))
        worst = max(worst, abs(m - exact) / (se + 1e-15))
        assert abs(m - exact) < 5.0 * se + 1e-12, (k, m, exact, se)
    say(f"  worst bond deviation {worst:.2f} standard errors")


def test_removing_the_scaling_restores_the_plain_model_and_refusals():
    wl = W.c0()
    t = tables(wl, "NORMAL", False)
    P = 3000
    say(f"case stochastic-vol removal and refusals c0 P={P}")
    ctx, _, _ = sv_context(wl, t, P)
    prods = to_products(wl.products)
    scaled, _ = ctx.value_products(prods)
    with pytest.raises(_abi.UnsupportedOperationException):
        ctx.serialize()
    ctx.set_volatility_factorization(np.asarray(t.vol), np.asarray(t.factors))
    with pytest.raises(_abi.UnsupportedOperationException):
        ctx.adjoint(prods[:1])
    ctx.set_stochastic_volatility(NU, RHO, None)
    a, _ = ctx.value_products(prods)
    ref = _abi.LmmContext(t.times, t.tenor, wl.num_factors, P, wl.seed, state_space="NORMAL")
    ref.set_initial_forwards(t.L0)
    ref.set_covariance(t.lam, t.var)
    ref.set_discount_curve(t.pd)
    b, _ = ref.value_products(prods)
    assert np.array_equal(a, b)
    assert np.all(np.abs(scaled - b) > 1e-7)
    with pytest.raises(_abi.IllegalArgumentException):
        ctx.set_stochastic_volatility(NU, 1.5, ref)                          # rho outside [-1, 1]
    other = _abi.LmmContext(t.times, t.tenor, 2, P + 1, DRIVER_SEED)
    with pytest.raises(_abi.IllegalArgumentException):
        ctx.set_stochastic_volatility(NU, RHO, other)                        # other paths
    with pytest.raises(_abi.IllegalArgumentException):
        ctx.set_stochastic_volatility(NU, RHO, ref, 0, 5)                    # factor out of range


def test_host_mirror_wrapper_runs_the_same_simulation():
    # LIBORCovarianceModelStochasticVolatility(BlendedLocalVolatilityModel(cov, 0.0), bm2, nu, rho)
    # through the host mirror gives the C ABI context's price bit for bit
    from finmath_hip import _finmath_host as H
    wl = W.c0()
    P = 3000
    say(f"case stochastic-vol host mirror c0 P={P}")
    td = H.TimeDiscretizationFromArray(0.0, wl.num_time_steps, wl.period)
    tenor = H.TimeDiscretizationFromArray(0.0, wl.num_libors, wl.period)
    fc = H.ForwardCurveInterpolation.createForwardCurveFromForwards(
        "forwardCurve", list(tenor.getAsDoubleArray()[:-1]), list(wl.forwards), wl.period)
    vol = H.LIBORVolatilityModelFourParameterExponentialForm(td, tenor, 0.2, 0.0, 0.25, 0.1, False)
    corr = H.LIBORCorrelationModelExponentialDecay(td, tenor, wl.num_factors, wl.corr_decay)
    cov = H.LIBORCovarianceModelFromVolatilityAndCorrelation(td, tenor, vol, corr)
    bm2 = H.BrownianMotionFromMersenneRandomNumbers(td, 2, P, DRIVER_SEED)
    wrapped = H.LIBORCovarianceModelStochasticVolatility(H.BlendedLocalVolatilityModel(cov, 0.0, False), bm2, NU, RHO,
                                                         False)
    model = H.LIBORMarketModelFromCovarianceModel.of(tenor, None, fc, H.DiscountCurveFromForwardCurve(fc), None,
                                                     wrapped, [], {"measure": "SPOT", "stateSpace": "NORMAL"})
    bm = H.BrownianMotionFromMersenneRandomNumbers(td, wl.num_factors, P, wl.seed)
    sim = H.LIBORMonteCarloSimulationFromLIBORModel(H.EulerSchemeFromProcessModel(model, bm))
    s = wl.products[0]
    v = sim.getValues([H.Swaption(s["exercise"], list(s["fixing"]), list(s["payment"]), list(s["strike"]))])[0]["value"]
    lam, var = cov.tables(td)
    ctx = _abi.LmmContext(td.getAsDoubleArray(), tenor.getAsDoubleArray(), wl.num_factors, P, wl.seed,
                          state_space="NORMAL")
    ctx.set_initial_forwards(model.getInitialForwards())
    ctx.set_covariance(lam, var)
    dc = model.getDiscountCurve()
    ctx.set_discount_curve([dc.getDiscountFactor(T) for T in tenor.getAsDoubleArray()])
    ctx.set_local_volatility(*wrapped.localVolatility(td))
    driver = _abi.LmmContext(td.getAsDoubleArray(), tenor.getAsDoubleArray(), 2, P, DRIVER_SEED)
    ctx.set_stochastic_volatility(NU, RHO, driver, 0, 1)
    ref, _ = ctx.value_products([_abi.Product.swaption(s["exercise"], s["fixing"], s["payment"], s["strike"])])
    assert v == ref[0]


def test_calibration_with_a_stochastic_volatility_wrapper():
    # the reference's smile set-up (LIBORMarketModelCalibrationSmileTest.java:216-227): normal
    # state space, blended local volatility 0.0, a calibrateable stochastic volatility scaling;
    # targets from known parameters on the same Brownian motions across three strikes
    from finmath_hip import _finmath_host as H
    from test_gpu_calibration import strip, prices
    say("case calibration with stochastic volatility wrapper, 15 swaptions, 2000 paths")
    n = 20
    td = H.TimeDiscretizationFromArray(0.0, n, 0.5)
    tenor = H.TimeDiscretizationFromArray(0.0, n, 0.5)
    fc = H.ForwardCurveInterpolation.createForwardCurveFromForwards(
        "forwardCurve", list(tenor.getAsDoubleArray()[:-1]), list(0.02 + 0.001 * np.arange(n)), 0.5)
    bm2 = H.BrownianMotionFromMersenneRandomNumbers(td, 2, 2000, DRIVER_SEED)

    def build(p):
        vol = H.LIBORVolatilityModelFourParameterExponentialForm(td, tenor, *p[:4], True)
        corr = H.LIBORCorrelationModelExponentialDecay(td, tenor, 3, 0.1, False)
        cov = H.LIBORCovarianceModelFromVolatilityAndCorrelation(td, tenor, vol, corr)
        sv = H.LIBORCovarianceModelStochasticVolatility(H.BlendedLocalVolatilityModel(cov, 0.0, False), bm2, p[4], p[5],
                                                        True)
        model = H.LIBORMarketModelFromCovarianceModel.of(tenor, None, fc, H.DiscountCurveFromForwardCurve(fc), None,
                                                         sv, [], {"measure": "SPOT", "stateSpace": "NORMAL"})
        return sv, model

    true = [0.2, 0.05, 0.3, 0.1, 0.3, -0.4]
    _, model_true = build(true)
    products = strip(fc)
    targets = prices(model_true, td, 3, products)
    sv0, model0 = build([0.15, 0.0, 0.25, 0.15, 0.2, -0.2])
    calib = [H.CalibrationProduct(p, v, 1.0) for p, v in zip(products, targets)]
    start = float(np.sqrt(np.mean((np.array(prices(model0, td, 3, products)) - np.array(targets)) ** 2)))
    cov_cal, info = sv0.getCloneCalibrated(model0, calib, {"maxIterations": 100})
    say(f"  parameters {info['parameters']} rms {start:.3e} -> {info['rootMeanSquaredError']:.3e}")
    assert isinstance(cov_cal, H.LIBORCovarianceModelStochasticVolatility)
    assert len(info["parameters"]) == 6
    assert info["rootMeanSquaredError"] < 0.01 * start


@pytest.mark.parametrize("measure,scheme,state_space,blend,shared", [
    ("SPOT", "EULER_FUNCTIONAL", "NORMAL", True, None),
    ("TERMINAL", "PREDICTOR_CORRECTOR", "LOGNORMAL", False, None),
    ("SPOT", "EULER", "LOGNORMAL", False, 2),
])
def test_heston_process_and_numeraire_match_oracle(measure, scheme, state_space, blend, shared):
    # LIBORCovarianceModelStochasticHestonVolatility; xi = 1.2 lets V reach zero on many paths
    wl = W.c0()
    t = tables(wl, state_space, blend)
    P = 4001
    kappa, theta, xi = 0.8, 0.9, 1.2
    say(f"case heston process c0 P={P} {measure} {scheme} {state_space} blend={blend} shared={shared}")
    ctx = _abi.LmmContext(t.times, t.tenor, wl.num_factors, P, wl.seed, measure=measure, scheme=scheme,
                          state_space=state_space, store_paths=True)
    ctx.set_initial_forwards(t.L0)
    ctx.set_covariance(t.lam, t.var)
    ctx.set_discount_curve(t.pd)
    lv = None
    if blend:
        lv = (0.0, np.zeros((wl.num_time_steps, wl.num_libors)))
        ctx.set_local_volatility(*lv)
    if shared is None:
        driver = _abi.LmmContext(t.times, t.tenor, 1, P, DRIVER_SEED)
        ctx.set_heston_volatility(kappa, theta, xi, driver, 0)
        hv = (kappa, theta, xi, DRIVER_SEED, 1, 0)
    else:
        ctx.set_heston_volatility(kappa, theta, xi, ctx, shared)
        hv = (kappa, theta, xi, wl.seed, wl.num_factors, shared)
    ctx.simulate()
    om = O.OracleModel(t, P, measure=MEASURES.index(measure), scheme=SCHEMES.index(scheme),
                       state_space=1 if state_space == "LOGNORMAL" else 0, local_vol=lv, heston_vol=hv)
    worst = 0.0
    for ti in range(wl.num_time_steps + 1):
        for j in range(wl.num_libors):
            g, r = ctx.process_value(ti, j), om.process_value(ti, j)
            worst = max(worst, float(np.max(np.abs(g - r) / np.maximum(np.abs(r), 1e-3))))
    say(f"  worst relative process difference {worst:.3e}")
    assert worst < 1e-12
    for k in range(wl.num_libors + 1):
        g, r = ctx.numeraire(t.tenor[k]), om.numeraire(t.tenor[k])
        assert np.max(np.abs(g - r) / np.abs(r)) < 1e-12, k
    prods = to_products(wl.products)
    vals, _ = ctx.value_products(prods)
    for q, s in enumerate(wl.products):
        rv, _ = om.value(s)
        assert abs(vals[q] - rv) <= 1e-10 * abs(rv)


def test_heston_host_mirror_runs_the_same_simulation():
    from finmath_hip import _finmath_host as H
    wl = W.c0()
    P = 3000
    say(f"case heston host mirror c0 P={P}")
    td = H.TimeDiscretizationFromArray(0.0, wl.num_time_steps, wl.period)
    tenor = H.TimeDiscretizationFromArray(0.0, wl.num_libors, wl.period)
    fc = H.ForwardCurveInterpolation.createForwardCurveFromForwards(
        "forwardCurve", list(tenor.getAsDoubleArray()[:-1]), list(wl.forwards), wl.period)
    vol = H.LIBORVolatilityModelFourParameterExponentialForm(td, tenor, 0.2, 0.0, 0.25, 0.1, False)
    corr = H.LIBORCorrelationModelExponentialDecay(td, tenor, wl.num_factors, wl.corr_decay)
    cov = H.LIBORCovarianceModelFromVolatilityAndCorrelation(td, tenor, vol, corr)
    bm1 = H.BrownianMotionFromMersenneRandomNumbers(td, 1, P, DRIVER_SEED)
    wrapped = H.LIBORCovarianceModelStochasticHestonVolatility(cov, bm1, 0.8, 0.9, 0.5, False)
    model = H.LIBORMarketModelFromCovarianceModel.of(tenor, None, fc, H.DiscountCurveFromForwardCurve(fc), None,
                                                     wrapped, [], {"measure": "SPOT", "stateSpace": "LOGNORMAL"})
    bm = H.BrownianMotionFromMersenneRandomNumbers(td, wl.num_factors, P, wl.seed)
    sim = H.LIBORMonteCarloSimulationFromLIBORModel(H.EulerSchemeFromProcessModel(model, bm))
    s = wl.products[0]
    v = sim.getValues([H.Swaption(s["exercise"], list(s["fixing"]), list(s["payment"]), list(s["strike"]))])[0]["value"]
    lam, var = cov.tables(td)
    ctx = _abi.LmmContext(td.getAsDoubleArray(), tenor.getAsDoubleArray(), wl.num_factors, P, wl.seed)
    ctx.set_initial_forwards(model.getInitialForwards())
    ctx.set_covariance(lam, var)
    dc = model.getDiscountCurve()
    ctx.set_discount_curve([dc.getDiscountFactor(T) for T in tenor.getAsDoubleArray()])
    driver = _abi.LmmContext(td.getAsDoubleArray(), tenor.getAsDoubleArray(), 1, P, DRIVER_SEED)
    ctx.set_heston_volatility(0.8, 0.9, 0.5, driver, 0)
    ref, _ = ctx.value_products([_abi.Product.swaption(s["exercise"], s["fixing"], s["payment"], s["strike"])])
    assert v == ref[0]

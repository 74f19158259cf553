"""GPU parity of the local-volatility covariance wrappers against the oracle, through the
C ABI: DisplacedLocalVolatilityModel (loadings times L + d,
covariance/DisplacedLocalVolatilityModel.java:175-186) and BlendedLocalVolatilityModel
(loadings times (L - L a) + a f0(T_j - t_i), covariance/BlendedLocalVolatilityModel.java:265-283),
on every measure, both state spaces, the Euler and predictor-corrector schemes.

Bars: process values and numeraires in FP64 within 1e-12 relative of the oracle, prices
within 1e-10 relative, FP32-mixed prices within 0.1 bp of the FP64 oracle; the bonds
N(0) E[1/N(T_k)] reprice the initial curve within the Monte Carlo error at 200,000 paths
(the drift is the no-arbitrage drift of whatever loadings the wrapper produces); removing
the wrapper restores the plain model bit for bit; the adjoint refuses the wrapper.
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

DISPLACEMENT = 1.0 / 0.25          # the reference's calibration tests: DisplacedLocalVolatilityModel(cov, 1/0.25)
BLEND = 0.5


def setup(wl, kind, state_space):
    """Model tables and the wrapper (alpha, offsets[nT][n]) of one case."""
    nT, n = wl.num_time_steps, wl.num_libors
    if kind == "displaced":
        # absolute volatilities times (L + 4): 2-3 % normal, 20-30 % log-normal
        vol = (0.005, 0.0, 0.25, 0.002) if state_space == "NORMAL" else (0.05, 0.0, 0.25, 0.02)
        t = O.ModelTables(wl, vol_params=vol)
        return t, (0.0, np.full((nT, n), DISPLACEMENT))
    # blended: log-normal-like volatilities times (L - L a) + a f0(max(T_j - t_i, 0))
    t = O.ModelTables(wl, vol_params=(0.2, 0.0, 0.25, 0.1) if state_space == "NORMAL" else (2.0, 0.0, 0.25, 1.0))
    times, tenor = t.times, t.tenor
    b = np.array([[BLEND * O.forward_curve(t.node_t, t.node_f, max(tenor[j] - times[i], 0.0)) for j in range(n)]
                  for i in range(nT)])
    return t, (BLEND, b)


def lv_context(wl, t, lv, P, measure, scheme, state_space, *, precision="FP64", store_paths=False,
               with_discount=True):
    ctx = _abi.LmmContext(t.times, t.tenor, wl.num_factors, P, wl.seed, measure=measure, scheme=scheme,
                          state_space=state_space, precision=precision, store_paths=store_paths)
    ctx.set_initial_forwards(t.L0)
    ctx.set_covariance(t.lam, t.var)
    ctx.set_discount_curve(t.pd if with_discount else None)
    ctx.set_local_volatility(*lv)
    return ctx


def lv_oracle(t, lv, P, measure, scheme, state_space, **kw):
    return O.OracleModel(t, P, measure=MEASURES.index(measure), scheme=SCHEMES.index(scheme),
                         state_space=1 if state_space == "LOGNORMAL" else 0, local_vol=lv, **kw)


@pytest.mark.parametrize("kind", ["displaced", "blended"])
@pytest.mark.parametrize("state_space", ["NORMAL", "LOGNORMAL"])
@pytest.mark.parametrize("scheme", ["EULER_FUNCTIONAL", "PREDICTOR_CORRECTOR"])
@pytest.mark.parametrize("measure", MEASURES)
def test_process_and_numeraire_match_oracle(measure, scheme, state_space, kind):
    wl = W.c0()
    t, lv = setup(wl, kind, state_space)
    P = 4001
    say(f"case local-vol process c0 P={P} {kind} {measure} {scheme} {state_space}")
    ctx = lv_context(wl, t, lv, P, measure, scheme, state_space, store_paths=True)
    ctx.simulate()
    om = lv_oracle(t, lv, P, measure, scheme, state_space)
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


@pytest.mark.parametrize("kind,measure,scheme", [("displaced", "SPOT", "EULER_FUNCTIONAL"),
                                                 ("blended", "TERMINAL", "PREDICTOR_CORRECTOR"),
                                                 ("displaced", "SPOT", "EULER")])
def test_prices_match_oracle(kind, measure, scheme):
    wl = W.c1()
    t, lv = setup(wl, kind, "NORMAL")
    P = 10_000
    say(f"case local-vol prices c1 P={P} {kind} {measure} {scheme} NORMAL")
    ctx = lv_context(wl, t, lv, P, measure, scheme, "NORMAL")
    vals, errs = ctx.value_products(to_products(wl.products))
    om = lv_oracle(t, lv, P, measure, scheme, "NORMAL")
    worst = 0.0
    for q, s in enumerate(wl.products):
        rv, re_ = om.value(s)
        worst = max(worst, abs(vals[q] - rv) / abs(rv))
        assert abs(errs[q] - re_) <= 1e-8 * abs(re_)
    say(f"  worst relative price difference {worst:.3e}")
    assert worst <= 1e-10


@pytest.mark.parametrize("measure,scheme", [("SPOT", "EULER_FUNCTIONAL"), ("TERMINAL", "PREDICTOR_CORRECTOR")])
def test_fp32_mixed_within_a_tenth_of_a_basis_point(measure, scheme):
    wl = W.c1()
    t, lv = setup(wl, "displaced", "NORMAL")
    P = 20_000
    say(f"case local-vol fp32 mixed c1 P={P} {measure} {scheme}")
    ctx = lv_context(wl, t, lv, P, measure, scheme, "NORMAL", precision="FP32_MIXED")
    vals, _ = ctx.value_products(to_products(wl.products))
    om = lv_oracle(t, lv, P, measure, scheme, "NORMAL")
    worst = max(abs(vals[q] - om.value(s)[0]) * 1e4 for q, s in enumerate(wl.products))
    say(f"  worst FP32-mixed price error {worst:.4f} bp")
    assert worst < 0.1


@pytest.mark.parametrize("kind", ["displaced", "blended"])
@pytest.mark.parametrize("measure", MEASURES)
def test_bonds_reprice_the_initial_curve_at_full_size(kind, measure):
    # c1 shape (40 forwards, 5 factors) at 200,000 paths; without discount curve
    # N(0) E[1/N(T_k)] = prod_{l<k} 1/(1 + d L_l(0)) up to the Monte Carlo error
    wl = W.c1()
    t, lv = setup(wl, kind, "NORMAL")
    P = 200_000
    say(f"case local-vol bonds c1 P={P} {kind} {measure}")
    ctx = lv_context(wl, t, lv, P, measure, "EULER_FUNCTIONAL", "NORMAL", store_paths=True, with_discount=False)
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


def test_removing_the_wrapper_restores_the_plain_model():
    wl = W.c0()
    t, lv = setup(wl, "displaced", "NORMAL")
    P = 3000
    say(f"case local-vol removal c0 P={P}")
    plain = lv_context(wl, t, lv, P, "SPOT", "EULER_FUNCTIONAL", "NORMAL")
    plain.set_local_volatility(0.0, None)
    prods = to_products(wl.products)
    a, _ = plain.value_products(prods)
    ref = _abi.LmmContext(t.times, t.tenor, wl.num_factors, P, wl.seed, state_space="NORMAL")
    ref.set_initial_forwards(t.L0)
    ref.set_covariance(t.lam, t.var)
    ref.set_discount_curve(t.pd)
    b, _ = ref.value_products(prods)
    assert np.array_equal(a, b)
    wrapped, _ = lv_context(wl, t, lv, P, "SPOT", "EULER_FUNCTIONAL", "NORMAL").value_products(prods)
    assert np.all(np.abs(wrapped - b) > 1e-6)


def test_serialised_context_keeps_the_wrapper():
    wl = W.c0()
    t, lv = setup(wl, "blended", "NORMAL")
    P = 2001
    say(f"case local-vol serialisation c0 P={P}")
    ctx = lv_context(wl, t, lv, P, "SPOT", "PREDICTOR_CORRECTOR", "NORMAL")
    blob = ctx.serialize()
    again = _abi.LmmContext.deserialize(blob)
    assert again.serialize() == blob
    prods = to_products(wl.products)
    a, _ = ctx.value_products(prods)
    b, _ = again.value_products(prods)
    assert np.array_equal(a, b)


def test_adjoint_refuses_the_wrapper():
    wl = W.c0()
    t, lv = setup(wl, "displaced", "LOGNORMAL")
    say("case local-vol adjoint refusal")
    ctx = lv_context(wl, t, lv, 64, "SPOT", "EULER_FUNCTIONAL", "LOGNORMAL")
    ctx.set_volatility_factorization(np.asarray(t.vol), np.asarray(t.factors))
    with pytest.raises(_abi.UnsupportedOperationException):
        ctx.adjoint(to_products(wl.products[:1]))


def test_offsets_must_be_finite():
    wl = W.c0()
    t, lv = setup(wl, "displaced", "NORMAL")
    ctx = lv_context(wl, t, lv, 64, "SPOT", "EULER_FUNCTIONAL", "NORMAL")
    bad = np.array(lv[1])
    bad[3, 4] = np.nan
    with pytest.raises(_abi.IllegalArgumentException):
        ctx.set_local_volatility(0.0, bad)


def test_host_mirror_wrapper_runs_the_same_simulation():
    # DisplacedLocalVolatilityModel / BlendedLocalVolatilityModel through the host mirror
    # give the C ABI context's prices bit for bit
    from finmath_hip import _finmath_host as H
    wl = W.c0()
    P = 3000
    say(f"case local-vol host mirror c0 P={P}")
    td = H.TimeDiscretizationFromArray(0.0, wl.num_time_steps, wl.period)
    tenor = H.TimeDiscretizationFromArray(0.0, wl.num_libors, wl.period)
    fc = H.ForwardCurveInterpolation.createForwardCurveFromForwards(
        "forwardCurve", list(tenor.getAsDoubleArray()[:-1]), list(wl.forwards), wl.period)
    vol = H.LIBORVolatilityModelFourParameterExponentialForm(td, tenor, 0.2, 0.0, 0.25, 0.1, False)
    corr = H.LIBORCorrelationModelExponentialDecay(td, tenor, wl.num_factors, wl.corr_decay)
    cov = H.LIBORCovarianceModelFromVolatilityAndCorrelation(td, tenor, vol, corr)
    swaption = wl.products[0]
    prod = H.Swaption(swaption["exercise"], list(swaption["fixing"]), list(swaption["payment"]),
                      list(swaption["strike"]))
    for wrapped in (H.DisplacedLocalVolatilityModel(cov, 0.05, False), H.BlendedLocalVolatilityModel(cov, fc, 0.5, False)):
        model = H.LIBORMarketModelFromCovarianceModel.of(tenor, None, fc, H.DiscountCurveFromForwardCurve(fc), None,
                                                         wrapped, [], {"measure": "SPOT", "stateSpace": "NORMAL"})
        bm = H.BrownianMotionFromMersenneRandomNumbers(td, wl.num_factors, P, wl.seed)
        sim = H.LIBORMonteCarloSimulationFromLIBORModel(H.EulerSchemeFromProcessModel(model, bm))
        v = sim.getValues([prod])[0]["value"]
        lam, var = cov.tables(td)
        ctx = _abi.LmmContext(td.getAsDoubleArray(), tenor.getAsDoubleArray(), wl.num_factors, P, wl.seed,
                              state_space="NORMAL")
        ctx.set_initial_forwards(model.getInitialForwards())
        ctx.set_covariance(lam, var)
        dc = model.getDiscountCurve()
        ctx.set_discount_curve([dc.getDiscountFactor(T) for T in tenor.getAsDoubleArray()])
        ctx.set_local_volatility(*wrapped.localVolatility(td))
        ref, _ = ctx.value_products([_abi.Product.swaption(swaption["exercise"], swaption["fixing"],
                                                           swaption["payment"], swaption["strike"])])
        assert v == ref[0]

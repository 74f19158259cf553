"""GPU parity of the model variants against the oracle, through the C ABI: the terminal
measure (LIBORMarketModelFromCovarianceModel.java:1162-1179,1020-1042), the carried-state
EULER and PREDICTOR_CORRECTOR schemes and their functional forms
(EulerSchemeFromProcessModel.java:232-314), both state spaces.

Bars: process values and numeraires in FP64 within 1e-12 relative of the oracle, prices
within 1e-10 relative (Bermudans 1e-9), FP32-mixed prices within 0.1 bp of the FP64
oracle; zero-volatility closed forms; the adjoint refuses the variants it does not
differentiate.
"""
import numpy as np
import pytest

import _oracle as O
from finmath_hip import _abi
from finmath_hip import workloads as W
from test_gpu_parity import say, to_products
from test_gpu_bermudan import berm_product

pytestmark = pytest.mark.gpu

SCHEMES = ["EULER", "PREDICTOR_CORRECTOR", "EULER_FUNCTIONAL", "PREDICTOR_CORRECTOR_FUNCTIONAL"]
MEASURES = ["SPOT", "TERMINAL"]
NORMAL_VOL = (0.006, 0.0, 0.25, 0.002)      # absolute volatilities for the normal state space


def variant_context(wl, t, P, measure, scheme, state_space="LOGNORMAL", *, precision="FP64", store_paths=False,
                    with_discount=True, cap=1e5):
    ctx = _abi.LmmContext(t.times, t.tenor, wl.num_factors, P, wl.seed, measure=measure, scheme=scheme,
                          state_space=state_space, precision=precision, store_paths=store_paths, libor_cap=cap)
    ctx.set_initial_forwards(t.L0)
    ctx.set_covariance(t.lam, t.var)
    ctx.set_discount_curve(t.pd if with_discount else None)
    return ctx


def oracle(t, P, measure, scheme, state_space="LOGNORMAL", **kw):
    return O.OracleModel(t, P, measure=MEASURES.index(measure), scheme=SCHEMES.index(scheme),
                         state_space=1 if state_space == "LOGNORMAL" else 0, **kw)


@pytest.mark.parametrize("state_space", ["LOGNORMAL", "NORMAL"])
@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("measure", MEASURES)
def test_process_and_numeraire_match_oracle(measure, scheme, state_space):
    wl = W.c0()
    t = O.ModelTables(wl, vol_params=None if state_space == "LOGNORMAL" else NORMAL_VOL)
    P = 4001
    say(f"case process c0 P={P} {measure} {scheme} {state_space}")
    ctx = variant_context(wl, t, P, measure, scheme, state_space, store_paths=True)
    ctx.simulate()
    om = oracle(t, P, measure, scheme, state_space)
    worst = 0.0
    for ti in range(wl.num_time_steps + 1):
        for j in range(wl.num_libors):
            g, r = ctx.process_value(ti, j), om.process_value(ti, j)
            # normal-state forwards may come close to zero: relative to max(|L|, 1e-3)
            worst = max(worst, float(np.max(np.abs(g - r) / np.maximum(np.abs(r), 1e-3))))
    say(f"  worst relative process difference {worst:.3e}")
    assert worst < 1e-12
    for k in range(wl.num_libors + 1):
        g, r = ctx.numeraire(t.tenor[k]), om.numeraire(t.tenor[k])
        assert np.max(np.abs(g - r) / np.abs(r)) < 1e-12, k


@pytest.mark.parametrize("precision", ["FP64", "FP64_LOGSTATE"])
@pytest.mark.parametrize("scheme", ["EULER", "EULER_FUNCTIONAL", "PREDICTOR_CORRECTOR"])
def test_binding_cap_matches_oracle(scheme, precision):
    # EULER caps the value and carries the uncapped state; the functional schemes
    # re-derive the state from the capped value (FP64_LOGSTATE caps the log-state, which
    # is EULER_FUNCTIONAL up to rounding, and leaves EULER to the general kernel)
    wl = W.c0()
    t = O.ModelTables(wl, vol_params=(0.5, 0.0, 0.25, 0.2))
    P, cap = 3001, 0.04
    say(f"case binding cap c0 P={P} {scheme} {precision}")
    ctx = variant_context(wl, t, P, "SPOT", scheme, precision=precision, store_paths=True, cap=cap)
    ctx.simulate()
    om = oracle(t, P, "SPOT", scheme, cap=cap)
    worst, capped = 0.0, 0
    for ti in range(1, wl.num_time_steps + 1):
        for j in range(wl.num_libors):
            g, r = ctx.process_value(ti, j), om.process_value(ti, j)
            worst = max(worst, float(np.max(np.abs(g - r) / np.abs(r))))
            capped += int(np.sum(r == cap))
    say(f"  worst relative process difference {worst:.3e}, capped values {capped}")
    assert capped > 10_000
    assert worst < 1e-12


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("measure", MEASURES)
def test_prices_match_oracle(measure, scheme):
    wl = W.c1()
    t = O.ModelTables(wl)
    P = 10_000
    say(f"case prices c1 P={P} {measure} {scheme}")
    ctx = variant_context(wl, t, P, measure, scheme)
    vals, errs = ctx.value_products(to_products(wl.products))
    om = oracle(t, P, measure, scheme)
    worst = 0.0
    for q, s in enumerate(wl.products):
        rv, re_ = om.value(s)
        worst = max(worst, abs(vals[q] - rv) / abs(rv))
        assert abs(errs[q] - re_) <= 1e-8 * abs(re_)
    say(f"  worst relative price difference {worst:.3e}")
    assert worst <= 1e-10


@pytest.mark.parametrize("measure", MEASURES)
def test_swaps_and_caplets_without_discount_curve(measure):
    # under the terminal measure the period payoffs are evaluated at their payment dates
    wl = W.c0()
    t = O.ModelTables(wl)
    P = 5000
    say(f"case swaps+caplets c0 P={P} {measure} no discount curve")
    ctx = variant_context(wl, t, P, measure, "PREDICTOR_CORRECTOR", with_discount=False)
    om = oracle(t, P, measure, "PREDICTOR_CORRECTOR", with_discount=False)
    sw = wl.products[1]
    prods = [_abi.Product.swap(sw["fixing"], sw["payment"], sw["strike"]),
             _abi.Product.caplet(2.0, 0.5, 0.03), _abi.Product.caplet(4.5, 0.5, 0.035, is_floorlet=True)]
    vals, _ = ctx.value_products(prods)
    refs = [om.swap(sw)[0], om.caplet(2.0, 0.5, 0.03)[0], om.caplet(4.5, 0.5, 0.035, floorlet=True)[0]]
    for g, r in zip(vals, refs):
        assert abs(g - r) <= 1e-10 * abs(r), (g, r)


@pytest.mark.parametrize("measure,scheme", [("TERMINAL", "EULER_FUNCTIONAL"), ("SPOT", "PREDICTOR_CORRECTOR"),
                                            ("TERMINAL", "PREDICTOR_CORRECTOR")])
def test_swaptions_sharing_their_periods_match_oracle(measure, scheme):
    # the FP64 terminal and predictor-corrector kernels value swaptions over the same periods in
    # one walk, at most four at a time: six such swaptions (flat and per-period strikes, other
    # notionals) interleaved with a caplet and a swaption over fewer periods
    wl = W.c0()
    t = O.ModelTables(wl)
    P = 6000
    say(f"case swaptions sharing periods c0 P={P} {measure} {scheme}")
    fix, pay = [2.0, 2.5, 3.0, 3.5], [2.5, 3.0, 3.5, 4.0]
    sw = [dict(type="swaption", exercise=2.0, fixing=fix, payment=pay, strike=[0.03] * 4),
          dict(type="swaption", exercise=2.0, fixing=fix, payment=pay, strike=[0.02, 0.025, 0.03, 0.035], notional=2.0),
          dict(type="swaption", exercise=2.0, fixing=fix[:3], payment=pay[:3], strike=[0.03] * 3),
          dict(type="swaption", exercise=2.0, fixing=fix, payment=pay, strike=[0.04] * 4, notional=0.25),
          dict(type="swaption", exercise=2.0, fixing=fix, payment=pay, strike=[0.05] * 4),
          dict(type="swaption", exercise=2.0, fixing=fix, payment=pay, strike=[0.01] * 4, notional=0.5),
          dict(type="swaption", exercise=2.0, fixing=fix, payment=pay, strike=[0.035] * 4, notional=3.0)]
    prods = [_abi.Product.swaption(s["exercise"], s["fixing"], s["payment"], s["strike"],
                                   notional=s.get("notional", 1.0)) for s in sw]
    prods.insert(2, _abi.Product.caplet(2.0, 0.5, 0.03))
    ctx = variant_context(wl, t, P, measure, scheme)
    vals, errs = ctx.value_products(prods)
    om = oracle(t, P, measure, scheme)
    ref = [om.swaption(s) for s in sw]
    ref.insert(2, om.caplet(2.0, 0.5, 0.03))
    for q, (rv, re_) in enumerate(ref):
        say(f"  q={q}: gpu {vals[q]:.15e} oracle {rv:.15e}")
        assert abs(vals[q] - rv) <= 1e-10 * abs(rv)
        assert abs(errs[q] - re_) <= 1e-8 * abs(re_)


@pytest.mark.parametrize("measure,scheme", [("TERMINAL", "EULER_FUNCTIONAL"), ("SPOT", "PREDICTOR_CORRECTOR"),
                                            ("TERMINAL", "PREDICTOR_CORRECTOR")])
@pytest.mark.parametrize("basis", [0, 1])
def test_bermudan_matches_oracle(measure, scheme, basis):
    wl = W.c0()
    t = O.ModelTables(wl)
    P = 20_000
    spec = W.bermudan_spec(2.0, 10.0, 0.5, 0.03, 1.0, True, basis=basis)
    say(f"case bermudan c0 P={P} {measure} {scheme} basis={basis}")
    ctx = variant_context(wl, t, P, measure, scheme)
    vals, _ = ctx.value_products([berm_product(spec)])
    rv, _ = oracle(t, P, measure, scheme).bermudan(spec)
    say(f"  gpu {vals[0]:.15e} oracle {rv:.15e}")
    assert abs(vals[0] - rv) <= 1e-9 * abs(rv)


@pytest.mark.parametrize("scheme", ["EULER_FUNCTIONAL", "PREDICTOR_CORRECTOR"])
def test_terminal_zero_volatility_closed_form(scheme):
    wl = W.c0()
    t = O.ModelTables(wl, vol_params=(0.0, 0.0, 0.0, 0.0))
    P = 777
    say(f"case terminal zero-vol c0 P={P} {scheme}")
    ctx = variant_context(wl, t, P, "TERMINAL", scheme, with_discount=False)
    s = wl.products[0]
    vals, errs = ctx.value_products(to_products([s]))
    L = t.L0
    e = int(round(s["exercise"] / 0.5))
    K = s["strike"][0]
    annuity_swap, df = 0.0, 1.0
    for j in range(e, e + len(s["fixing"])):
        df /= 1 + L[j] * 0.5
        annuity_swap += (L[j] - K) * 0.5 * df
    P_e = np.prod(1.0 / (1.0 + L[:e] * 0.5))
    assert vals[0] == pytest.approx(max(annuity_swap, 0.0) * P_e, rel=1e-13)


@pytest.mark.parametrize("measure,scheme", [("TERMINAL", "EULER_FUNCTIONAL"), ("SPOT", "PREDICTOR_CORRECTOR"),
                                            ("TERMINAL", "EULER")])
def test_fp32_mixed_within_a_tenth_of_a_basis_point(measure, scheme):
    wl = W.c1()
    t = O.ModelTables(wl)
    P = 20_000
    say(f"case fp32 mixed c1 P={P} {measure} {scheme}")
    ctx = variant_context(wl, t, P, measure, scheme, precision="FP32_MIXED")
    vals, _ = ctx.value_products(to_products(wl.products))
    om = oracle(t, P, measure, scheme)
    worst = max(abs(vals[q] - om.value(s)[0]) * 1e4 for q, s in enumerate(wl.products))
    say(f"  worst FP32-mixed price error {worst:.4f} bp")
    assert worst < 0.1


def test_fp64_logstate_request_on_the_terminal_measure_runs_the_fp64_general_kernel():
    wl = W.c0()
    t = O.ModelTables(wl)
    P = 3000
    say(f"case FP64_LOGSTATE terminal c0 P={P}")
    ctx = variant_context(wl, t, P, "TERMINAL", "EULER", precision="FP64_LOGSTATE")
    vals, _ = ctx.value_products(to_products(wl.products))
    om = oracle(t, P, "TERMINAL", "EULER")
    for q, s in enumerate(wl.products):
        rv, _ = om.value(s)
        assert abs(vals[q] - rv) <= 1e-10 * abs(rv)


@pytest.mark.parametrize("measure,scheme", [("TERMINAL", "EULER_FUNCTIONAL"), ("SPOT", "PREDICTOR_CORRECTOR"),
                                            ("SPOT", "EULER")])
def test_adjoint_refuses_variants_it_does_not_differentiate(measure, scheme):
    wl = W.c0()
    t = O.ModelTables(wl)
    say(f"case adjoint refusal {measure} {scheme}")
    ctx = variant_context(wl, t, 64, measure, scheme)
    ctx.set_volatility_factorization(np.asarray(t.vol), np.asarray(t.factors))
    with pytest.raises(_abi.UnsupportedOperationException):
        ctx.adjoint(to_products(wl.products[:1]))

"""GPU Bermudan swaption (regression exercise) against the oracle, through the C ABI.

Bars: FP64 within 1e-9 relative of the oracle (the regression solves the same normal
equations; different summation order moves the coefficients by ~cond * 1e-16, which
flips no exercise decision at these path counts), the zero-volatility closed form to
1e-12, FP32-mixed within 0.1 bp of the FP64 oracle price. The pathwise adjoint (exercise
decisions held fixed) against central finite differences with common random numbers.
"""
import numpy as np
import pytest

import _oracle as O
from finmath_hip import _abi
from finmath_hip import workloads as W
from test_gpu_parity import make_context, say

pytestmark = pytest.mark.gpu


def berm_product(spec):
    return _abi.Product.bermudan(spec["fixing"], spec["payment"], spec["period_lengths"], spec["strike"],
                                 spec["exercise_flags"], is_callable=spec["callable"], basis=spec["basis"])


@pytest.mark.parametrize("basis", [0, 1])
@pytest.mark.parametrize("callable_", [True, False])
def test_bermudan_fp64_matches_oracle(basis, callable_):
    wl = W.c0()
    t = O.ModelTables(wl)
    P = 20_000
    spec = W.bermudan_spec(2.0, 10.0, 0.5, 0.03, 1.0, callable_, basis=basis)
    say(f"case bermudan c0 P={P} basis={basis} callable={callable_}")
    ctx = make_context(wl, t, P)
    vals, errs = ctx.value_products([berm_product(spec)])
    rv, re_ = O.OracleModel(t, P).bermudan(spec)
    say(f"  gpu {vals[0]:.15e} oracle {rv:.15e} err {errs[0]:.3e}/{re_:.3e}")
    assert abs(vals[0] - rv) <= 1e-9 * abs(rv)
    assert abs(errs[0] - re_) <= 1e-7 * abs(re_)


def test_bermudan_priced_with_europeans_in_one_pass():
    wl = W.c0()
    t = O.ModelTables(wl)
    P = 10_000
    say(f"case bermudan + strip c0 P={P}")
    spec = W.bermudan_spec(2.0, 10.0, 0.5, 0.03, 1.0, True, basis=0)
    sw = wl.products[0]
    ctx = make_context(wl, t, P)
    prods = [_abi.Product.swaption(sw["exercise"], sw["fixing"], sw["payment"], sw["strike"]), berm_product(spec),
             _abi.Product.swap(wl.products[1]["fixing"], wl.products[1]["payment"], wl.products[1]["strike"])]
    vals, _ = ctx.value_products(prods)
    om = O.OracleModel(t, P)
    assert abs(vals[0] - om.swaption(sw)[0]) <= 1e-10 * abs(vals[0])
    assert abs(vals[1] - om.bermudan(spec)[0]) <= 1e-9 * abs(vals[1])
    assert abs(vals[2] - om.swap(wl.products[1])[0]) <= 1e-10 * abs(vals[2])


@pytest.mark.parametrize("basis", [0, 1])
def test_bermudan_zero_volatility_closed_form(basis):
    wl = W.c0()
    wl.forwards = 0.02 + 0.001 * np.arange(wl.num_libors)
    t = O.ModelTables(wl, vol_params=(0.0, 0.0, 0.0, 0.0))
    P = 1000
    strike = 0.03
    say(f"case bermudan zero-vol c0 P={P} basis={basis}")
    ctx = make_context(wl, t, P)
    spec = W.bermudan_spec(2.0, 10.0, 0.5, strike, 1.0, True, basis=basis)
    vals, errs = ctx.value_products([berm_product(spec)])
    fix, pay, pl = spec["fixing"], spec["payment"], spec["period_lengths"]
    flows = np.array([(t.L0[int(round(f / 0.5))] - strike) * pl[p] * t.pd[int(round(pay[p] / 0.5))]
                      for p, f in enumerate(fix)])
    U = [flows[p:].sum() for p in range(fix.size) if spec["exercise_flags"][p]]
    assert vals[0] == pytest.approx(max(0.0, max(U)), rel=1e-12)


def test_bermudan_fp32_mixed_within_a_tenth_of_a_basis_point():
    wl = W.c0()
    t = O.ModelTables(wl)
    P = 20_000
    spec = W.bermudan_spec(2.0, 10.0, 0.5, 0.03, 1.0, True, basis=1)
    say(f"case bermudan fp32 c0 P={P}")
    ctx = make_context(wl, t, P, precision="FP32_MIXED")
    vals, _ = ctx.value_products([berm_product(spec)])
    rv, _ = O.OracleModel(t, P).bermudan(spec)
    say(f"  fp32 {vals[0]:.12e} oracle {rv:.12e} diff {(vals[0] - rv) * 1e4:.5f} bp")
    assert abs(vals[0] - rv) * 1e4 < 0.1


def test_bermudan_80_forwards_10_factors():
    # the 80 x 10 model of the Bermudan workload at a path count the oracle finishes in seconds
    wl = W.c4()
    t = O.ModelTables(wl)
    P = 8192
    spec = wl.products[0]
    say(f"case bermudan c4 model P={P} FP64 and FP32_MIXED")
    rv, re_ = O.OracleModel(t, P).bermudan(spec)
    v64, _ = make_context(wl, t, P).value_products([berm_product(spec)])
    v32, _ = make_context(wl, t, P, precision="FP32_MIXED").value_products([berm_product(spec)])
    say(f"  fp64 {v64[0]:.15e} fp32 {v32[0]:.12e} oracle {rv:.15e} +- {re_:.2e}")
    assert abs(v64[0] - rv) <= 1e-9 * abs(rv)
    assert abs(v32[0] - rv) * 1e4 < 0.1


def test_bermudan_is_deterministic():
    wl = W.c0()
    t = O.ModelTables(wl)
    say("case bermudan determinism c0 P=30000")
    ctx = make_context(wl, t, 30_000)
    spec = W.bermudan_spec(2.0, 10.0, 0.5, 0.03, 1.0, True, basis=0)
    a, ea = ctx.value_products([berm_product(spec)])
    b, eb = ctx.value_products([berm_product(spec)])
    assert np.array_equal(a, b) and np.array_equal(ea, eb)


def _adjoint_context(wl, t, P, precision="FP64", with_discount=True, dirac="ZERO"):
    # ZERO holds the exercise decisions fixed, which the finite differences below reproduce
    # (the DISCRETE_DELTA default is compared with the oracle's tape in test_gpu_aad_oracle.py)
    ctx = make_context(wl, t, P, precision=precision, with_discount=with_discount)
    ctx.set_volatility_factorization(t.vol, t.factors)
    ctx.set_dirac_delta(dirac)
    return ctx


def test_bermudan_adjoint_value_equals_pricing():
    wl = W.c0()
    t = O.ModelTables(wl)
    P = 20_000
    spec = W.bermudan_spec(2.0, 10.0, 0.5, 0.03, 1.0, True, basis=0)
    say(f"case bermudan adjoint value c0 P={P}")
    v, _ = make_context(wl, t, P).value_products([berm_product(spec)])
    va, dl0, dsig = _adjoint_context(wl, t, P).adjoint([berm_product(spec)])
    assert va == pytest.approx(v[0], rel=1e-12)
    assert np.all(np.isfinite(dl0)) and np.all(np.isfinite(dsig))


def _fd_delta(wl, t, P, spec, j, h):
    vals = []
    for sgn in (1.0, -1.0):
        L0 = t.L0.copy()
        L0[j] += sgn * h
        ctx = make_context(wl, t, P, with_discount=False)
        ctx.set_initial_forwards(L0)
        vals.append(ctx.value_products([berm_product(spec)])[0][0])
    return (vals[0] - vals[1]) / (2 * h)


@pytest.mark.parametrize("callable_", [True, False])
def test_bermudan_deltas_zero_volatility_exact(callable_):
    # every path is the same and a small bump moves no exercise decision: the pathwise
    # derivative with fixed decisions is the derivative of the price
    wl = W.c0()
    wl.forwards = 0.02 + 0.001 * np.arange(wl.num_libors)
    t = O.ModelTables(wl, vol_params=(0.0, 0.0, 0.0, 0.0))
    P = 256
    spec = W.bermudan_spec(2.0, 10.0, 0.5, 0.03, 1.0, callable_, basis=0)
    say(f"case bermudan adjoint zero-vol deltas callable={callable_}")
    va, dl0, _ = _adjoint_context(wl, t, P, with_discount=False).adjoint([berm_product(spec)])
    assert abs(va) > 1e-3
    for j in range(0, wl.num_libors, 2):
        fd = _fd_delta(wl, t, P, spec, j, 1e-6)
        say(f"  delta[{j}] adjoint {dl0[j]:.10e} fd {fd:.10e}")
        assert abs(dl0[j] - fd) <= 1e-6 * max(1.0, abs(fd))


def _fixed_decision_value(wl, t, P, spec, L0, ex_times):
    # the Bermudan's value with the exercise time of every path given: the callable swap
    # pays the periods fixing at or after the exercise time, from the materialised process
    ctx = make_context(wl, t, P, store_paths=True, with_discount=False)
    ctx.set_initial_forwards(L0)
    ctx.simulate()
    v = np.zeros(P)
    for f, pd, pl, K in zip(spec["fixing"], spec["payment"], spec["period_lengths"], spec["strike"]):
        ti = int(round(f / wl.period))
        L = ctx.process_value(ti, ti)
        pays = ex_times <= f
        v += np.where(pays, (L - K) * pl / ctx.numeraire(pd), 0.0)
    return v.mean()


def test_bermudan_deltas_against_fixed_decision_differences():
    # stochastic paths, no discount curve: central differences of the price with the
    # exercise time of every path held at the one the induction chose (exerciseTime of
    # getValues), which is the derivative the pathwise adjoint computes
    wl = W.c0()
    t = O.ModelTables(wl)
    P = 20_000
    spec = W.bermudan_spec(2.0, 10.0, 0.5, 0.03, 1.0, True, basis=0)
    say(f"case bermudan adjoint deltas c0 P={P}")
    ctx = _adjoint_context(wl, t, P, with_discount=False)
    v, dl0, _ = ctx.adjoint([berm_product(spec)])
    ex = ctx.bermudan_exercise_times(0)
    assert np.any(np.isfinite(ex)) and np.any(np.isinf(ex))
    assert _fixed_decision_value(wl, t, P, spec, t.L0, ex) == pytest.approx(v, rel=1e-12)
    h = 1e-6
    for j in np.argsort(-np.abs(dl0))[:6]:
        up, dn = t.L0.copy(), t.L0.copy()
        up[j] += h
        dn[j] -= h
        fd = (_fixed_decision_value(wl, t, P, spec, up, ex) - _fixed_decision_value(wl, t, P, spec, dn, ex)) / (2 * h)
        say(f"  delta[{j}] adjoint {dl0[j]:.10e} fd {fd:.10e}")
        assert abs(dl0[j] - fd) <= 1e-5 * np.max(np.abs(dl0))


def test_bermudan_forward_vegas_against_central_differences():
    wl = W.c0()
    t = O.ModelTables(wl)
    P = 20_000
    spec = W.bermudan_spec(2.0, 10.0, 0.5, 0.03, 1.0, True, basis=1)
    say(f"case bermudan adjoint vegas c0 P={P}")
    _, _, dsig = _adjoint_context(wl, t, P, with_discount=False).adjoint([berm_product(spec)])
    vega = dsig.sum(axis=0)                  # parallel bump of sigma_j over all times
    h = 1e-5
    for j in np.argsort(-np.abs(vega))[:4]:
        vals = []
        for sgn in (1.0, -1.0):
            vol = t.vol.copy()
            vol[:, j] += sgn * h
            lam = vol[:, :, None] * t.factors[None, :, :]
            ctx = make_context(wl, t, P, with_discount=False)
            ctx.set_covariance(lam, vol * vol * 1.0)
            vals.append(ctx.value_products([berm_product(spec)])[0][0])
        fd = (vals[0] - vals[1]) / (2 * h)
        say(f"  vega[{j}] adjoint {vega[j]:.8e} fd {fd:.8e}")
        assert abs(vega[j] - fd) < 5e-2 * np.max(np.abs(vega))


def test_bermudan_adjoint_fp32_close_to_fp64():
    wl = W.c0()
    t = O.ModelTables(wl)
    P = 20_000
    spec = W.bermudan_spec(2.0, 10.0, 0.5, 0.03, 1.0, True, basis=1)
    say(f"case bermudan adjoint fp32 vs fp64 c0 P={P}")
    g64 = _adjoint_context(wl, t, P).adjoint([berm_product(spec)])
    g32 = _adjoint_context(wl, t, P, "FP32_MIXED").adjoint([berm_product(spec)])
    assert abs(g32[0] - g64[0]) < 1e-5
    assert np.max(np.abs(g32[1] - g64[1])) < 1e-2 * np.max(np.abs(g64[1]))
    assert np.max(np.abs(g32[2] - g64[2])) < 1e-2 * np.max(np.abs(g64[2]))

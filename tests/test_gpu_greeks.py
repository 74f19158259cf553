"""Adjoint Greeks of the c3 workload: the 30-swaption strip on the 40-forward, 5-factor model
with the model's discount curve (DiscountCurveFromForwardCurve), FP32 paths with FP64
gradient sums, checkpoints every 4 steps.

The reference's AAD treats every path average as a constant
(RandomVariableDifferentiableAAD.java:165-219; getAverage returns a double), so the
numeraire adjustment avg(N(0)/N(T)) of getNumeraire
(LIBORMarketModelFromCovarianceModel.java:860-876) is held at its base value. The finite
differences here do the same: a price with the discount curve is (adj_k / avg_k) times the
average of payoff/N(T_k) over paths, and the average of payoff/N(T_k) is the price of the
same product without a discount curve. So the frozen-adjustment difference quotient is
r_q (V_q^nodisc(theta + h) - V_q^nodisc(theta - h)) / 2h with r_q = V_q^disc / V_q^nodisc at
the base parameters, all on the same Brownian paths (common random numbers, the reference's
own check: LIBORMarketModelNormalAADSensitivitiesTest.java:553,711).
"""
import numpy as np
import pytest

import _oracle as O
from finmath_hip import _abi
from finmath_hip import workloads as W

pytestmark = pytest.mark.gpu

P = 20_000


def say(*a):
    print(*a, flush=True)


def strip(wl):
    return [_abi.Product.swaption(s["exercise"], s["fixing"], s["payment"], s["strike"]) for s in wl.products]


def context(wl, t, *, precision, discount, L0=None, vol=None, stride=4):
    vol = t.vol if vol is None else vol
    ctx = _abi.LmmContext(t.times, t.tenor, wl.num_factors, P, wl.seed, precision=precision, checkpoint_stride=stride)
    ctx.set_initial_forwards(t.L0 if L0 is None else L0)
    ctx.set_covariance(vol[:, :, None] * t.factors[None, :, :], vol * vol * 1.0)
    ctx.set_discount_curve(t.pd if discount else None)
    ctx.set_volatility_factorization(vol, t.factors)
    return ctx


def prices(wl, t, *, precision, discount, **kw):
    ctx = context(wl, t, precision=precision, discount=discount, **kw)
    v, _ = ctx.value_products(strip(wl))
    ctx.close()
    return v


@pytest.fixture(scope="module")
def c3():
    wl = W.c3().with_paths(P)
    t = O.ModelTables(wl)
    out = {}
    for prec in ("FP32_MIXED", "FP64"):
        ctx = context(wl, t, precision=prec, discount=True)
        v, dL, dsig = ctx.adjoint(strip(wl))
        vd, _ = ctx.value_products(strip(wl))
        ctx.close()
        vn = prices(wl, t, precision=prec, discount=False)
        out[prec] = dict(value=v, dL=dL, dsig=dsig, disc=vd, ratio=vd / vn)
    return wl, t, out


def frozen_fd(wl, t, prec, ratio, h, **bump):
    """Central difference of sum_q r_q V_q^nodisc in the bumped input (common random numbers)."""
    up = prices(wl, t, precision=prec, discount=False, **{k: v[0] for k, v in bump.items()})
    dn = prices(wl, t, precision=prec, discount=False, **{k: v[1] for k, v in bump.items()})
    return float(np.sum(ratio * (up - dn))) / (2 * h)


def test_c3_value_is_the_price_with_discount_curve(c3):
    wl, t, out = c3
    for prec, g in out.items():
        say(f"case c3 adjoint value {prec} P={P}: {g['value']:.15e}")
        # the adjoint's forward pass is the pricing pass (the same kernel, with checkpoints)
        assert g["value"] == pytest.approx(float(np.sum(g["disc"])), rel=1e-12)
        assert np.all(np.isfinite(g["dL"])) and np.all(np.isfinite(g["dsig"]))
    # the strip is worth more than its MC error and the discount curve moves it
    assert np.all(out["FP64"]["disc"] > 0) and np.max(np.abs(out["FP64"]["ratio"] - 1.0)) > 1e-6


DELTA_J = tuple(range(0, 40, 3))
VEGA_J = (1, 5, 12, 19, 26, 33, 39)


@pytest.fixture(scope="module")
def c3_differences(c3):
    """FP64 frozen-adjustment central differences (h = 1e-5) of the deltas and forward vegas.
    The FP32-mixed adjoint is checked against these too: central differences of FP32 paths
    at a small bump are dominated by the paths' rounding (about 1e-7 of the price over 2h)."""
    wl, t, out = c3
    ratio, h = out["FP64"]["ratio"], 1e-5
    deltas, vegas = {}, {}
    for j in DELTA_J:
        up, dn = t.L0.copy(), t.L0.copy()
        up[j] += h
        dn[j] -= h
        deltas[j] = frozen_fd(wl, t, "FP64", ratio, h, L0=(up, dn))
    for j in VEGA_J:
        up, dn = t.vol.copy(), t.vol.copy()
        alive = t.vol[:, j] > 0
        up[alive, j] += h
        dn[alive, j] -= h
        vegas[j] = frozen_fd(wl, t, "FP64", ratio, h, vol=(up, dn))
    return deltas, vegas


@pytest.mark.parametrize("prec,tol", [("FP64", 2e-4), ("FP32_MIXED", 2e-3)])
def test_c3_deltas_against_frozen_adjustment_differences(c3, c3_differences, prec, tol):
    wl, t, out = c3
    g = out[prec]
    say(f"case c3 deltas {prec} P={P}")
    worst = 0.0
    for j, fd in c3_differences[0].items():
        err = abs(g["dL"][j] - fd) / np.max(np.abs(g["dL"]))
        worst = max(worst, err)
        say(f"  delta[{j}] adjoint {g['dL'][j]:.10e} fd {fd:.10e} err {err:.2e}")
    assert worst < tol


@pytest.mark.parametrize("prec,tol", [("FP64", 2e-4), ("FP32_MIXED", 2e-3)])
def test_c3_forward_vegas_against_frozen_adjustment_differences(c3, c3_differences, prec, tol):
    wl, t, out = c3
    g = out[prec]
    vega = g["dsig"].sum(axis=0)
    say(f"case c3 forward vegas {prec} P={P}")
    worst = 0.0
    for j, fd in c3_differences[1].items():
        err = abs(vega[j] - fd) / np.max(np.abs(vega))
        worst = max(worst, err)
        say(f"  vega[{j}] adjoint {vega[j]:.10e} fd {fd:.10e} err {err:.2e}")
    assert worst < tol


def test_c3_fp32_greeks_close_to_fp64(c3):
    _, _, out = c3
    a, b = out["FP32_MIXED"], out["FP64"]
    say("case c3 fp32 vs fp64 gradients")
    assert np.max(np.abs(a["dL"] - b["dL"])) < 1e-3 * np.max(np.abs(b["dL"]))
    assert np.max(np.abs(a["dsig"] - b["dsig"])) < 1e-3 * np.max(np.abs(b["dsig"]))


# ---- the four volatility parameters through the host mirror's getGradient -------------------

def host_model(wl, params, *, discount, sim_dt=None):
    """The workload's model through the host mirror; the covariance model lives on the tenor
    grid and the simulation on sim_dt (default: the tenor grid)."""
    from finmath_hip import _finmath_host as H
    tenor = H.TimeDiscretizationFromArray(0.0, wl.num_libors, wl.period)
    dt = wl.period if sim_dt is None else sim_dt
    td = H.TimeDiscretizationFromArray(0.0, int(round(wl.num_libors * wl.period / dt)), dt)
    fc = H.ForwardCurveInterpolation.createForwardCurveFromForwards(
        "forwardCurve", list(tenor.getAsDoubleArray()[:-1]), list(wl.forwards), wl.period)
    vol = H.LIBORVolatilityModelFourParameterExponentialForm(tenor, tenor, *params, True)
    corr = H.LIBORCorrelationModelExponentialDecay(tenor, tenor, wl.num_factors, wl.corr_decay)
    cov = H.LIBORCovarianceModelFromVolatilityAndCorrelation(tenor, tenor, vol, corr)
    dc = H.DiscountCurveFromForwardCurve(fc) if discount else None
    model = H.LIBORMarketModelFromCovarianceModel.of(tenor, None, fc, dc, None, cov, [],
                                                     {"measure": "SPOT", "stateSpace": "LOGNORMAL"})
    return H, model, td


def host_process(wl, params, *, discount, precision, sim_dt=None, paths=P):
    H, model, td = host_model(wl, params, discount=discount, sim_dt=sim_dt)
    bm = H.BrownianMotionFromMersenneRandomNumbers(td, wl.num_factors, paths, wl.seed)
    process = H.EulerSchemeFromProcessModel(model, bm, "EULER_FUNCTIONAL", precision)
    return H, process


def host_strip(H, wl):
    return [H.Swaption(s["exercise"], list(s["fixing"]), list(s["payment"]), list(s["strike"])) for s in wl.products]


def host_prices(wl, params, *, discount, precision, sim_dt=None):
    H, process = host_process(wl, params, discount=discount, precision=precision, sim_dt=sim_dt)
    sim = H.LIBORMonteCarloSimulationFromLIBORModel(process)
    return np.array([r["value"] for r in sim.getValues(host_strip(H, wl))])


@pytest.mark.parametrize("wl_name,prec,sim_dt,tol", [
    ("c3", "FP64", None, 5e-4),
    ("c3", "FP32_MIXED", None, 3e-3),
    # the covariance model on the tenor grid (0.5), the simulation on a finer grid (0.25):
    # each volatility entry drives two simulation steps and its tau is the entry's own time
    ("c0", "FP64", 0.25, 5e-4),
])
def test_parameter_vegas_against_frozen_adjustment_differences(wl_name, prec, sim_dt, tol):
    """The adjoint in the path precision against FP64 central differences (h = 1e-6)."""
    wl = (W.c3() if wl_name == "c3" else W.c0()).with_paths(P)
    if wl_name == "c0":
        wl.products = [s for s in wl.products if s["type"] == "swaption"]
    base = tuple(wl.vol_params)
    h = 1e-6
    say(f"case parameter vegas {wl_name} {prec} sim_dt={sim_dt} P={P}")
    H, process = host_process(wl, base, discount=True, precision=prec, sim_dt=sim_dt)
    g = process.getGradient(host_strip(H, wl), [], 4)
    ratio = host_prices(wl, base, discount=True, precision="FP64", sim_dt=sim_dt) / \
        host_prices(wl, base, discount=False, precision="FP64", sim_dt=sim_dt)
    pv = np.asarray(g["parameterVega"])
    assert pv.size == 4
    for k in range(4):
        up, dn = list(base), list(base)
        up[k] += h
        dn[k] -= h
        fd = float(np.sum(ratio * (host_prices(wl, up, discount=False, precision="FP64", sim_dt=sim_dt) -
                                   host_prices(wl, dn, discount=False, precision="FP64", sim_dt=sim_dt)))) / (2 * h)
        err = abs(pv[k] - fd) / max(abs(fd), 1e-2 * np.max(np.abs(pv)))
        say(f"  d/d{'abcd'[k]} adjoint {pv[k]:.10e} fd {fd:.10e} err {err:.2e}")
        assert err < tol
    # the volatility-model entries are the simulation-step entries summed per entry
    vm = np.asarray(g["volatilityModel"])
    assert vm.sum() == pytest.approx(np.asarray(g["volatility"]).sum(), rel=1e-12)


def test_volatility_model_entries_on_a_coarser_grid_against_differences():
    """LIBORVolatilityModelFromGivenMatrix leaves on the tenor grid with the simulation on a
    finer grid: the gradient of one matrix entry against central differences of that entry."""
    from finmath_hip import _finmath_host as H
    wl = W.c0().with_paths(P)
    tenor = H.TimeDiscretizationFromArray(0.0, wl.num_libors, wl.period)
    td = H.TimeDiscretizationFromArray(0.0, 2 * wl.num_libors, wl.period / 2)
    fc = H.ForwardCurveInterpolation.createForwardCurveFromForwards(
        "forwardCurve", list(tenor.getAsDoubleArray()[:-1]), list(wl.forwards), wl.period)
    n = wl.num_libors
    # [numberOfTimeSteps][numberOfLibors] of the tenor grid
    base = np.array([[0.0 if j <= i else 0.15 + 0.002 * j - 0.001 * i for j in range(n)] for i in range(n)])

    def proc(volm, precision="FP64"):
        vm = H.LIBORVolatilityModelFromGivenMatrix(tenor, tenor, volm)
        corr = H.LIBORCorrelationModelExponentialDecay(tenor, tenor, wl.num_factors, wl.corr_decay)
        cov = H.LIBORCovarianceModelFromVolatilityAndCorrelation(tenor, tenor, vm, corr)
        model = H.LIBORMarketModelFromCovarianceModel.of(tenor, None, fc, None, None, cov, [],
                                                         {"measure": "SPOT", "stateSpace": "LOGNORMAL"})
        bm = H.BrownianMotionFromMersenneRandomNumbers(td, wl.num_factors, P, wl.seed)
        return H.EulerSchemeFromProcessModel(model, bm, "EULER_FUNCTIONAL", precision)

    spec = W.c0().products[0]
    prod = [H.Swaption(spec["exercise"], list(spec["fixing"]), list(spec["payment"]), list(spec["strike"]))]
    say(f"case volatility model entries coarse grid P={P}")
    g = proc(base).getGradient(prod, [], 4)
    vm = np.asarray(g["volatilityModel"])
    assert vm.shape == (n + 1, n)
    h = 1e-6
    for (i, j) in ((0, 12), (4, 12), (8, 15), (9, 10)):
        vals = []
        for sgn in (1.0, -1.0):
            m = base.copy()
            m[i, j] += sgn * h
            sim = H.LIBORMonteCarloSimulationFromLIBORModel(proc(m))
            vals.append(sim.getValues(prod)[0]["value"])
        fd = (vals[0] - vals[1]) / (2 * h)
        say(f"  dV/dsigma[{i}][{j}] adjoint {vm[i, j]:.10e} fd {fd:.10e}")
        assert abs(vm[i, j] - fd) <= 5e-4 * np.max(np.abs(vm))


def test_inconsistent_factorization_is_refused():
    wl = W.c0().with_paths(2000)
    t = O.ModelTables(wl)
    say("case adjoint refuses a factorization of other tables")
    ctx = _abi.LmmContext(t.times, t.tenor, wl.num_factors, 2000, wl.seed)
    ctx.set_initial_forwards(t.L0)
    ctx.set_covariance(t.lam, t.var)
    ctx.set_volatility_factorization(t.vol * 1.01, t.factors)
    prods = [_abi.Product.swaption(s["exercise"], s["fixing"], s["payment"], s["strike"])
             for s in wl.products if s["type"] == "swaption"]
    with pytest.raises(_abi.IllegalStateException):
        ctx.adjoint(prods)
    # a covariance update drops the factorization set before it
    ctx.set_volatility_factorization(t.vol, t.factors)
    ctx.set_covariance(t.lam, t.var)
    with pytest.raises(_abi.IllegalStateException):
        ctx.adjoint(prods)
    ctx.set_volatility_factorization(t.vol, t.factors)
    v, dL, _ = ctx.adjoint(prods)
    assert np.isfinite(v) and np.all(np.isfinite(dL))
    ctx.close()

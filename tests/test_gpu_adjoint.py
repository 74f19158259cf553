"""Pathwise adjoint (reverse sweep) against central finite differences with common
random numbers, the reference's own check for its AAD gradients
(src/test/java/net/finmath/montecarlo/interestrate/LIBORMarketModelNormalAADSensitivitiesTest.java:553,711).

The model has no discount curve here, as in that reference test, so the price has
no path average (numeraire adjustment) whose dependence AAD would hold constant.
"""
import sys

import numpy as np
import pytest

import _oracle as O
from finmath_hip import _abi
from finmath_hip import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def context(wl, t, P, *, L0=None, vol=None, precision="FP64", stride=4, discount=False):
    vol = t.vol if vol is None else vol
    ctx = _abi.LmmContext(t.times, t.tenor, wl.num_factors, P, wl.seed, precision=precision, checkpoint_stride=stride)
    ctx.set_initial_forwards(t.L0 if L0 is None else L0)
    lam = vol[:, :, None] * t.factors[None, :, :]
    ctx.set_covariance(lam, vol * vol * 1.0)
    ctx.set_discount_curve(t.pd if discount else None)
    ctx.set_volatility_factorization(vol, t.factors)
    return ctx


def strip(wl):
    return [_abi.Product.swaption(s["exercise"], s["fixing"], s["payment"], s["strike"]) for s in wl.products]


def price(wl, t, P, **kw):
    ctx = context(wl, t, P, **kw)
    v, _ = ctx.value_products(strip(wl))
    ctx.close()
    return float(np.sum(v))


@pytest.fixture(scope="module")
def setup():
    wl = W.c1()
    t = O.ModelTables(wl)
    return wl, t, 20_000


def test_adjoint_value_equals_pricing(setup):
    wl, t, P = setup
    say(f"case adjoint value c1 P={P}")
    ctx = context(wl, t, P)
    v, dL, dsig = ctx.adjoint(strip(wl))
    vals, _ = ctx.value_products(strip(wl))
    assert abs(v - np.sum(vals)) <= 1e-13 * abs(v)
    assert np.all(np.isfinite(dL)) and np.all(np.isfinite(dsig))


def test_deltas_against_central_differences(setup):
    wl, t, P = setup
    say(f"case adjoint deltas c1 P={P}")
    ctx = context(wl, t, P)
    _, dL, _ = ctx.adjoint(strip(wl))
    h = 1e-5
    worst = 0.0
    for j in list(range(1, 40, 3)):
        up, dn = t.L0.copy(), t.L0.copy()
        up[j] += h
        dn[j] -= h
        fd = (price(wl, t, P, L0=up) - price(wl, t, P, L0=dn)) / (2 * h)
        worst = max(worst, abs(dL[j] - fd) / max(abs(fd), 1e-3 * np.max(np.abs(dL))))
        say(f"  delta[{j}] adjoint {dL[j]:.10e} fd {fd:.10e}")
    assert worst < 2e-4


def test_forward_vegas_against_central_differences(setup):
    wl, t, P = setup
    say(f"case adjoint forward vegas c1 P={P}")
    ctx = context(wl, t, P)
    _, _, dsig = ctx.adjoint(strip(wl))
    vega = dsig.sum(axis=0)
    h = 1e-5
    worst = 0.0
    for j in (2, 9, 17, 25, 33):
        up, dn = t.vol.copy(), t.vol.copy()
        alive = t.vol[:, j] > 0
        up[alive, j] += h
        dn[alive, j] -= h
        fd = (price(wl, t, P, vol=up) - price(wl, t, P, vol=dn)) / (2 * h)
        worst = max(worst, abs(vega[j] - fd) / max(abs(fd), 1e-3 * np.max(np.abs(vega))))
        say(f"  vega[{j}] adjoint {vega[j]:.10e} fd {fd:.10e}")
    assert worst < 2e-4


@pytest.mark.parametrize("stride", [1, 2, 5, 40])
def test_checkpoint_stride_does_not_change_the_gradient(setup, stride):
    wl, t, _ = setup
    P = 4000
    say(f"case adjoint stride={stride} P={P}")
    g4 = context(wl, t, P, stride=4).adjoint(strip(wl))
    gs = context(wl, t, P, stride=stride).adjoint(strip(wl))
    assert np.array_equal(g4[1], gs[1]) and np.array_equal(g4[2], gs[2])


def test_fp32_adjoint_close_to_fp64(setup):
    wl, t, P = setup
    say(f"case adjoint fp32 vs fp64 P={P}")
    g64 = context(wl, t, P).adjoint(strip(wl))
    g32 = context(wl, t, P, precision="FP32_MIXED").adjoint(strip(wl))
    assert abs(g32[0] - g64[0]) < 1e-5 * abs(g64[0])
    assert np.max(np.abs(g32[1] - g64[1])) < 1e-3 * np.max(np.abs(g64[1]))
    assert np.max(np.abs(g32[2] - g64[2])) < 1e-3 * np.max(np.abs(g64[2]))


def test_adjoint_against_oracle_finite_differences():
    # an independent check: finite differences of the CPU oracle's prices
    wl = W.c0()
    t = O.ModelTables(wl)
    P = 4000
    say(f"case adjoint vs oracle fd c0 P={P}")
    prods = [_abi.Product.swaption(s["exercise"], s["fixing"], s["payment"], s["strike"])
             for s in wl.products if s["type"] == "swaption"]
    ctx = context(wl, t, P)
    _, dL, _ = ctx.adjoint(prods)
    spec = wl.products[0]
    h = 1e-5
    for j in (3, 10, 15):
        up, dn = t.L0.copy(), t.L0.copy()
        up[j] += h
        dn[j] -= h
        vu = O.OracleModel(O.ModelTables(wl, L0=up), P, with_discount=False).swaption(spec)[0]
        vd = O.OracleModel(O.ModelTables(wl, L0=dn), P, with_discount=False).swaption(spec)[0]
        fd = (vu - vd) / (2 * h)
        say(f"  delta[{j}] adjoint {dL[j]:.10e} oracle fd {fd:.10e}")
        assert abs(dL[j] - fd) <= 2e-4 * max(abs(fd), 1e-3 * np.max(np.abs(dL)))

"""Products on device random variables through the model interface (finmath_hip.products over
randomvariable.LIBORMonteCarloSimulation): Bond, SimpleSwap, Caplet and Swaption as their
reference classes value them (montecarlo/interestrate/products/Bond.java:59-91,
SimpleSwap.java:137-170, Caplet.java:115-140, Swaption.java:137-199).

Bars: on the tenor grid the fused kernel's prices within 1e-12 relative; off the grid the
oracle's forwards and numeraires composed in numpy within 1e-12 relative, with the standard
error of the same per-path values.
"""
import math

import numpy as np
import pytest

import _oracle as O
from finmath_hip import _abi
from finmath_hip import _finmath_host as H
from finmath_hip import products as PR
from finmath_hip import randomvariable as RV
from finmath_hip import workloads as W
from test_gpu_parity import say

pytestmark = pytest.mark.gpu


def simulation(wl, t, P, with_discount=True):
    ctx = _abi.LmmContext(t.times, t.tenor, wl.num_factors, P, wl.seed, store_paths=True)
    ctx.set_initial_forwards(t.L0)
    ctx.set_covariance(t.lam, t.var)
    ctx.set_discount_curve(t.pd if with_discount else None)
    ctx.simulate()
    fc = H.ForwardCurveInterpolation.createForwardCurveFromForwards("forwardCurve", list(t.node_t), list(t.node_f),
                                                                  wl.period)
    return ctx, RV.LIBORMonteCarloSimulation(ctx, fc)


def test_products_on_the_grid_match_the_fused_kernel():
    wl = W.c0()
    t = O.ModelTables(wl)
    P = 6001
    say(f"case device-RV products on the grid c0 P={P}")
    ctx, sim = simulation(wl, t, P)
    fix = [5.0 + 0.5 * i for i in range(10)]
    pay = [f + 0.5 for f in fix]
    generic = [PR.Swaption(5.0, fix, pay, [0.03] * 10), PR.SimpleSwap(fix, pay, [0.031] * 10),
               PR.Caplet(3.0, 0.5, 0.029), PR.Caplet(4.5, 0.5, 0.035, isFloorlet=True)]
    fused = [_abi.Product.swaption(5.0, fix, pay, [0.03] * 10), _abi.Product.swap(fix, pay, [0.031] * 10),
             _abi.Product.caplet(3.0, 0.5, 0.029), _abi.Product.caplet(4.5, 0.5, 0.035, is_floorlet=True)]
    vals, errs = ctx.value_products(fused)
    for g, v, e in zip(generic, vals, errs):
        r = g.getValues(sim)
        say(f"  {type(g).__name__}: device RV {r['value']:.15e} fused {v:.15e}")
        assert r["value"] == pytest.approx(v, rel=1e-12, abs=1e-17)
        assert r["error"] == pytest.approx(e, rel=1e-9)


def test_products_off_the_grid_match_the_oracle():
    wl = W.c0()
    t = O.ModelTables(wl)
    P = 4001
    say(f"case device-RV products off the grid c0 P={P}")
    ctx, sim = simulation(wl, t, P)
    om = O.OracleModel(t, P, scheme=2)
    N0 = om.numeraire(0.0)

    def expect(per_path):
        x = per_path * N0
        return float(np.mean(x)), float(np.std(x) / math.sqrt(P))

    # a caplet between tenor dates, a swap on a quarter-shifted schedule, a bond between dates
    r = PR.Caplet(2.125, 0.5, 0.03).getValues(sim)
    v, e = expect(np.maximum(om.forward_rate(2.125, 2.125, 2.625) - 0.03, 0.0) * 0.5 / om.numeraire(2.625))
    assert r["value"] == pytest.approx(v, rel=1e-12) and r["error"] == pytest.approx(e, rel=1e-8)
    fix = [1.25 + 0.5 * i for i in range(6)]
    pay = [f + 0.5 for f in fix]
    r = PR.SimpleSwap(fix, pay, [0.03] * 6).getValues(sim)
    per = sum((om.forward_rate(f, f, p) - 0.03) * (p - f) / om.numeraire(p) for f, p in zip(fix, pay))
    v, e = expect(per)
    say(f"  off-grid swap {r['value']:.12e} oracle {v:.12e}")
    assert r["value"] == pytest.approx(v, rel=1e-12) and r["error"] == pytest.approx(e, rel=1e-8)
    r = PR.Bond(3.375).getValues(sim)
    v, _ = expect(1.0 / om.numeraire(3.375))
    assert r["value"] == pytest.approx(v, rel=1e-13)

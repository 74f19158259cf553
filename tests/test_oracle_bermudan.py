"""Pins of the oracle's Bermudan swaption (BermudanSwaption.java:91-157 with
MonteCarloConditionalExpectationRegression.java:79-150).

- brute force: the backward induction written from the regression equations with
  numpy (normal equations solved by numpy's pseudo-inverse) on the oracle's own,
  separately pinned, process values and numeraires;
- closed form: at zero volatility every path is the same, the Gram matrix has rank
  one, the pseudo-inverse regression returns the continuation value itself, and the
  Bermudan is worth max(0, max_e U_e) with U_e the discounted swap from exercise e;
- Monte Carlo sanity: the Bermudan is worth at least each of its Europeans, up to
  the Monte Carlo error of the difference.
"""
import numpy as np
import pytest

import _oracle as O
from finmath_hip import workloads as W


def numpy_bermudan(om, spec):
    fix, pay, K = np.asarray(spec["fixing"]), np.asarray(spec["payment"]), np.asarray(spec["strike"])
    pl = np.asarray(spec["period_lengths"])
    flags = np.asarray(spec["exercise_flags"])
    P, dt = om.P, om.t.wl.period
    n = om.n
    w = 1.0 / P
    values, under = np.zeros(P), np.zeros(P)
    for per in reversed(range(fix.size)):
        ti = int(round(fix[per] / dt))
        L = om.process_value(ti, ti)
        payoff = (L - K[per]) * pl[per] / om.numeraire(pay[per]) * w
        if spec["callable"]:
            under = under + payoff
        else:
            values = values + payoff
        if flags[per]:
            Ls = [om.process_value(ti, j) for j in range(n)]
            if spec["basis"] == 0:
                ds = 1.0 / (Ls[ti] * (pay[per] - fix[per]) + 1.0)
                acc = np.prod([1.0 + Ls[j] * dt for j in range(ti, int(round(pay[-1] / dt)))], axis=0)
                dl = 1.0 / (((acc - 1.0) / (pay[-1] - fix[per])) * (pay[-1] - fix[per]) + 1.0)
                X = np.stack([np.ones(P), ds, ds * ds, dl, dl * dl, 1.0 / om.numeraire(fix[per])], axis=1)
            else:
                disc = np.cumprod([1.0 / (1.0 + Ls[int(round(fix[p] / dt))] * pl[p]) for p in range(per, fix.size)],
                                  axis=0)
                S = (1.0 - disc[-1]) / np.sum(disc * pl[per:, None], axis=0)
                X = np.stack([np.ones(P), S, S * S, om.numeraire(fix[per])], axis=1)
            G = X.T @ X / P
            rhs = X.T @ (values - under) / P
            x = np.linalg.pinv(G, rcond=X.shape[1] * 2.0 ** -52, hermitian=True) @ rhs
            trig = X @ x
            values = np.where(trig >= 0.0, values, under)
    v = values * om.numeraire(0.0) / w
    return v.mean(), v.std() / np.sqrt(P)


@pytest.mark.parametrize("basis", [0, 1])
@pytest.mark.parametrize("callable_", [True, False])
def test_bermudan_against_numpy_backward_induction(basis, callable_):
    wl = W.c0()
    om = O.OracleModel(O.ModelTables(wl), 3000)
    spec = W.bermudan_spec(2.0, 10.0, 0.5, 0.03, 1.0, callable_, basis=basis)
    v, e = om.bermudan(spec)
    rv, re_ = numpy_bermudan(om, spec)
    assert v == pytest.approx(rv, rel=1e-9, abs=1e-12)
    assert e == pytest.approx(re_, rel=1e-6)


@pytest.mark.parametrize("basis", [0, 1])
@pytest.mark.parametrize("forwards,strike", [("flat", 0.025), ("upward", 0.03)])
def test_bermudan_zero_volatility_closed_form(basis, forwards, strike):
    wl = W.c0()
    if forwards == "upward":
        wl.forwards = 0.02 + 0.001 * np.arange(wl.num_libors)
    t = O.ModelTables(wl, vol_params=(0.0, 0.0, 0.0, 0.0))
    om = O.OracleModel(t, 64)
    spec = W.bermudan_spec(2.0, 10.0, 0.5, strike, 1.0, True, basis=basis)
    v, e = om.bermudan(spec)
    fix, pay, pl = spec["fixing"], spec["payment"], spec["period_lengths"]
    j = [int(round(f / 0.5)) for f in fix]
    flows = np.array([(t.L0[j[p]] - strike) * pl[p] * t.pd[int(round(pay[p] / 0.5))] for p in range(fix.size)])
    U = [flows[p:].sum() for p in range(fix.size) if spec["exercise_flags"][p]]
    expected = max(0.0, max(U))
    assert expected > 1e-3
    assert v == pytest.approx(expected, rel=1e-12)
    assert e < 1e-15


def test_bermudan_dominates_its_europeans():
    wl = W.c0()
    om = O.OracleModel(O.ModelTables(wl), 20_000)
    spec = W.bermudan_spec(2.0, 10.0, 0.5, 0.03, 1.0, True, basis=0)
    v, _, pv, _ = om.bermudan(spec, per_path=True)
    for ex in (2.0, 5.0, 8.0):
        _, _, pe = om.swaption(W.swaption_spec(ex, 10.0 - ex, 0.5, 0.03), per_path=True)
        d = pv - pe
        assert d.mean() > -3.0 * d.std() / np.sqrt(d.size)

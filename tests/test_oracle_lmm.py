"""Pins of the oracle's LMM path, numeraire and products.

- brute force: a per-path restatement written from the model equations (not from
  the reference's code structure) on tiny inputs;
- closed forms: zero volatility gives deterministic forwards and prices;
- invariants: the adjusted numeraire reprices the discount curve exactly
  (LIBORMarketModelFromCovarianceModel.java:869-874), par swaps are worth ~0,
  the Monte Carlo error shrinks like 1/sqrt(P).
"""
import math

import numpy as np
import pytest

import _oracle as O
from finmath_hip import workloads as W


def tiny_workload(n=4, F=2, P=6):
    w = W.Workload("tiny", n, 0.5, F, P, 0.2, np.array([0.03, 0.035, 0.04, 0.045][:n]))
    return w


def brute_force_paths(tables, P):
    """log-Euler under the spot measure, written from the SDE:
    dlog L_j = (sum_{l=m(t)+1..j} d L_l/(1+d L_l) lambda_j.lambda_l - var_j/2) dt + lambda_j . dW."""
    times, tenor = tables.times, tables.tenor
    nT, n = times.size - 1, tenor.size - 1
    F = tables.factors.shape[1]
    dW = O.brownian(times, F, P, tables.wl.seed)
    L = np.empty((nT + 1, n, P))
    L[0] = np.minimum(np.exp(np.log(tables.L0)), 1e5)[:, None]
    for i in range(nT):
        dt = times[i + 1] - times[i]
        first = int(np.searchsorted(tenor, times[i], side="right"))
        L[i + 1] = L[i]
        for p in range(P):
            lam = tables.lam[i]                                     # [n][F]
            for j in range(first, n):
                mu = 0.0
                for l in range(first, j + 1):
                    d = tenor[l + 1] - tenor[l]
                    mu += d * L[i, l, p] / (1 + d * L[i, l, p]) * float(lam[j] @ lam[l])
                mu -= 0.5 * tables.var[i, j]
                y = math.log(L[i, j, p]) + mu * dt + float(lam[j] @ dW[i, :, p])
                L[i + 1, j, p] = min(math.exp(y), 1e5)
    return L


def test_process_against_brute_force():
    wl = tiny_workload()
    t = O.ModelTables(wl)
    P = 6
    om = O.OracleModel(t, P)
    ref = brute_force_paths(t, P)
    for i in range(wl.num_time_steps + 1):
        for j in range(wl.num_libors):
            assert np.allclose(om.process_value(i, j), ref[i, j], rtol=1e-13, atol=0)


def test_brownian_increments_are_the_generator_stream():
    wl = tiny_workload()
    t = O.ModelTables(wl)
    om = O.OracleModel(t, 5)
    ref = O.brownian(t.times, wl.num_factors, 5, wl.seed)
    for i in range(wl.num_time_steps):
        for k in range(wl.num_factors):
            assert np.array_equal(om.brownian_increment(i, k), ref[i, k])


def test_zero_volatility_closed_form():
    wl = W.c0()
    t = O.ModelTables(wl, vol_params=(0.0, 0.0, 0.0, 0.0))
    om = O.OracleModel(t, 64)
    L = t.L0
    # forwards stay at their initial values (up to the exp(log(.)) round trip of each step)
    for i in range(wl.num_time_steps + 1):
        for j in range(wl.num_libors):
            assert np.allclose(om.process_value(i, j), L[j], rtol=1e-14, atol=0)
    # payer swaption: max(sum_j (L_j-K) d P(T_e, T_{j+1}), 0) / N(T_e), N(T_k) = prod (1 + L d)
    s = wl.products[0]
    e = int(round(s["exercise"] / 0.5))
    K = s["strike"][0]
    annuity_swap = 0.0
    df = 1.0
    for j in range(e, e + len(s["fixing"])):
        df /= 1 + L[j] * 0.5
        annuity_swap += (L[j] - K) * 0.5 * df
    N_e = np.prod(1 + L[:e] * 0.5)
    v, err = om.swaption(s)
    assert abs(v - max(annuity_swap, 0.0) / N_e) < 1e-14
    assert err == 0.0


def test_adjusted_numeraire_reprices_discount_curve():
    wl = W.c0()
    t = O.ModelTables(wl)
    om = O.OracleModel(t, 2000)
    for k in range(wl.num_libors + 1):
        N = om.numeraire(t.tenor[k])
        assert abs(np.mean(1.0 / N) - t.pd[k]) < 1e-14


def test_par_swap_is_worth_zero_within_mc_error():
    wl = W.c1()
    t = O.ModelTables(wl)
    om = O.OracleModel(t, 4000)
    node_t = t.tenor[:-1]
    par = W.forward_swap_rate(node_t, wl.forwards, 0.5, 2.0, 12.0)
    v, e = om.swap(W.swap_spec(2.0, 12.0, 0.5, par))
    assert abs(v) < 4 * e


def test_mc_error_scales_like_inverse_sqrt_paths():
    wl = W.c0()
    t = O.ModelTables(wl)
    s = wl.products[0]
    _, e1 = O.OracleModel(t, 2000).swaption(s)
    _, e4 = O.OracleModel(t, 8000).swaption(s)
    assert 0.4 < e4 / e1 < 0.6


def test_payer_swaption_dominates_intrinsic_value():
    # E[max(S,0)] >= max(E[S],0) (Jensen) for the same paths
    wl = W.c0()
    t = O.ModelTables(wl)
    om = O.OracleModel(t, 4000)
    s = wl.products[0]
    v, _, pp = om.swaption(s, per_path=True)
    assert v >= 0 and np.all(pp >= 0)

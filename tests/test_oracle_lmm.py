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


def brute_force_paths(tables, P, measure=0, scheme=2, state_space=1, cap=1e5, local_vol=None, scaling=None):
    """Euler scheme per path, written from the SDE of the forward rates:
      log-normal: d log L_j = (mu_j - var_j / 2) dt + lambda_j . dW
      normal:     d L_j     =  mu_j dt             + lambda_j . dW
    with w_l = d L_l / (1 + d L_l) (log-normal) or d / (1 + d L_l) (normal) and
      spot measure:     mu_j =  sum_{l = m(t)+1 .. j}  w_l lambda_j . lambda_l
      terminal measure: mu_j = -sum_{l = j+1 .. n-1}   w_l lambda_j . lambda_l.
    The predictor-corrector (scheme 1, 3) steps with the average of the drifts at the
    start values and at the predicted values; schemes 0, 1 carry the state, 2, 3
    re-derive it from the value (they differ only when the cap binds).
    local_vol = (alpha, b[nT][n]): the loadings of L_j are scaled by ((1 - alpha) L_j + b_ij),
    with L_j the value the drift or the diffusion is evaluated on, and the log-normal
    variance is the squared norm of the scaled loading. scaling[nT][P]: every loading of path p
    at step i is also multiplied by scaling[i, p]."""
    times, tenor = tables.times, tables.tenor
    nT, n = times.size - 1, tenor.size - 1
    F = tables.factors.shape[1]
    dW = O.brownian(times, F, P, tables.wl.seed)
    lognormal = state_space == 1
    to_value = (lambda y: min(math.exp(y), cap)) if lognormal else (lambda y: min(y, cap))
    L = np.empty((nT + 1, n, P))
    Y0 = np.log(np.maximum(tables.L0, 0.0)) if lognormal else tables.L0.copy()
    L[0] = np.array([to_value(y) for y in Y0])[:, None]
    state = np.repeat(Y0[:, None], P, axis=1)

    def loading(i, Lv, p):
        lam = tables.lam[i]                                         # [n][F]
        if local_vol is not None:
            alpha, b = local_vol[0], np.broadcast_to(local_vol[1], (nT, n))
            lam = lam * ((1.0 - alpha) * Lv + b[i])[:, None]
        if scaling is not None:
            lam = lam * scaling[i, p]
        return lam

    def drift(i, Lv, first, p):
        lam = loading(i, Lv, p)
        mu = np.zeros(n)
        for j in range(first, n):
            ls = range(first, j + 1) if measure == 0 else range(j + 1, n)
            sgn = 1.0 if measure == 0 else -1.0
            for l in ls:
                d = tenor[l + 1] - tenor[l]
                w = d * Lv[l] / (1 + d * Lv[l]) if lognormal else d / (1 + d * Lv[l])
                mu[j] += sgn * w * float(lam[j] @ lam[l])
            if lognormal:
                plain = local_vol is None and scaling is None
                mu[j] -= 0.5 * (tables.var[i, j] if plain else float(lam[j] @ lam[j]))
        return mu

    for i in range(nT):
        dt = times[i + 1] - times[i]
        first = int(np.searchsorted(tenor, times[i], side="right"))
        L[i + 1] = L[i]
        for p in range(P):
            y0 = state[:, p] if scheme in (0, 1) else (np.log(L[i, :, p]) if lognormal else L[i, :, p].copy())
            mu = drift(i, L[i, :, p], first, p)
            diffusion = loading(i, L[i, :, p], p) @ dW[i, :, p]
            y = y0 + mu * dt + diffusion
            if scheme in (1, 3):
                mup = drift(i, np.array([to_value(v) for v in y]), first, p)
                y = y0 + 0.5 * (mu + mup) * dt + diffusion
            for j in range(first, n):
                state[j, p] = y[j]
                L[i + 1, j, p] = to_value(y[j])
    return L


@pytest.mark.parametrize("measure", [0, 1])
@pytest.mark.parametrize("scheme", [0, 1, 2, 3])
@pytest.mark.parametrize("state_space", [1, 0])
def test_process_against_brute_force(measure, scheme, state_space):
    wl = tiny_workload()
    # the normal state space takes absolute volatilities
    t = O.ModelTables(wl, vol_params=(0.2, 0.0, 0.25, 0.1) if state_space == 1 else (0.006, 0.0, 0.25, 0.002))
    P = 6
    om = O.OracleModel(t, P, measure=measure, scheme=scheme, state_space=state_space)
    ref = brute_force_paths(t, P, measure, scheme, state_space)
    for i in range(wl.num_time_steps + 1):
        for j in range(wl.num_libors):
            assert np.allclose(om.process_value(i, j), ref[i, j], rtol=1e-12, atol=1e-16), (i, j)


@pytest.mark.parametrize("measure", [0, 1])
@pytest.mark.parametrize("scheme", [0, 1, 2, 3])
def test_binding_cap_against_brute_force(measure, scheme):
    # a cap that binds on many path-steps: the carried schemes keep the uncapped state,
    # the functional ones restart from log(cap)
    wl = tiny_workload(P=8)
    t = O.ModelTables(wl, vol_params=(0.6, 0.0, 0.25, 0.3))
    P, cap = 8, 0.04
    om = O.OracleModel(t, P, measure=measure, scheme=scheme, cap=cap)
    ref = brute_force_paths(t, P, measure, scheme, 1, cap=cap)
    assert np.mean(ref[1:] == cap) > 0.05
    for i in range(wl.num_time_steps + 1):
        for j in range(wl.num_libors):
            assert np.allclose(om.process_value(i, j), ref[i, j], rtol=1e-12, atol=0), (i, j)


@pytest.mark.parametrize("measure", [0, 1])
@pytest.mark.parametrize("scheme", [0, 1, 2, 3])
@pytest.mark.parametrize("state_space", [0, 1])
@pytest.mark.parametrize("kind", ["displaced", "blended"])
def test_local_volatility_against_brute_force(measure, scheme, state_space, kind):
    # DisplacedLocalVolatilityModel (L + d) and BlendedLocalVolatilityModel ((L - L a) + a f0)
    # on normal-state-space volatilities, as the reference's calibration tests set them up;
    # the blend's offset varies by time index and component
    wl = tiny_workload()
    t = O.ModelTables(wl, vol_params=(0.06, 0.0, 0.25, 0.02))
    nT, n = wl.num_time_steps, wl.num_libors
    if kind == "displaced":
        lv = (0.0, np.full((nT, n), 0.25))
    else:
        a = 0.4
        lv = (a, a * (0.03 + 0.002 * np.add.outer(np.arange(nT), np.arange(n))))
    P = 6
    om = O.OracleModel(t, P, measure=measure, scheme=scheme, state_space=state_space, local_vol=lv)
    ref = brute_force_paths(t, P, measure, scheme, state_space, local_vol=lv)
    plain = brute_force_paths(t, P, measure, scheme, state_space)
    assert np.max(np.abs(ref - plain)) > 1e-4          # the wrapper changes the paths
    for i in range(wl.num_time_steps + 1):
        for j in range(wl.num_libors):
            assert np.allclose(om.process_value(i, j), ref[i, j], rtol=1e-12, atol=1e-16), (i, j)


def sv_scaling(times, P, nu, rho, seed, F=2, a=0, b=1):
    """exp of the Euler scheme of dX = -nu^2/2 dt + nu (rho dB_a + sqrt(1 - rho^2) dB_b), X(0) = 0,
    on the increments of an F-factor Brownian motion, [nT][P]."""
    dB = O.brownian(times, F, P, seed)
    X = np.zeros(P)
    out = np.empty((times.size - 1, P))
    for i in range(times.size - 1):
        out[i] = np.exp(X)
        X = X - 0.5 * nu * nu * (times[i + 1] - times[i]) + nu * (rho * dB[i, a] + math.sqrt(1 - rho * rho) * dB[i, b])
    return out


@pytest.mark.parametrize("measure,scheme,state_space,lv", [(0, 2, 1, False), (1, 1, 0, True), (0, 3, 0, True),
                                                           (1, 0, 1, False)])
def test_stochastic_volatility_against_brute_force(measure, scheme, state_space, lv):
    # LIBORCovarianceModelStochasticVolatility: loadings times S(t_i) of a log-Euler process on
    # its own Brownian motion, over a plain or a blended local-volatility model
    wl = tiny_workload()
    t = O.ModelTables(wl, vol_params=(0.2, 0.0, 0.25, 0.1) if state_space == 1 else (0.006, 0.0, 0.25, 0.002))
    nT, n = wl.num_time_steps, wl.num_libors
    P = 6
    local = (0.5, 0.5 * 0.03) if lv else None
    if lv:
        t = O.ModelTables(wl, vol_params=(0.2, 0.0, 0.25, 0.1))
    sv = (0.4, -0.3, 777, 2, 0, 1)
    scaling = sv_scaling(t.times, P, 0.4, -0.3, 777)
    om = O.OracleModel(t, P, measure=measure, scheme=scheme, state_space=state_space,
                       local_vol=None if local is None else (local[0], np.full((nT, n), local[1])), stochastic_vol=sv)
    ref = brute_force_paths(t, P, measure, scheme, state_space, scaling=scaling,
                            local_vol=None if local is None else (local[0], np.full((nT, n), local[1])))
    assert np.ptp(scaling[1:]) > 0.1                         # the scaling moves the paths
    for i in range(nT + 1):
        for j in range(n):
            assert np.allclose(om.process_value(i, j), ref[i, j], rtol=1e-12, atol=1e-16), (i, j)


def test_stochastic_volatility_scaling_is_a_martingale_and_vanishes_without_vol_of_vol():
    # E[S(t)] = 1 (log-normal martingale); nu = 0 gives S = 1 and the plain model bit for bit
    wl = W.c0()
    t = O.ModelTables(wl)
    P = 40_000
    S = sv_scaling(t.times, P, 0.3, -0.5, 99)
    for i in range(1, wl.num_time_steps):
        assert abs(np.mean(S[i]) - 1.0) < 4 * np.std(S[i]) / math.sqrt(P)
    small = tiny_workload()
    ts = O.ModelTables(small)
    a = O.OracleModel(ts, 6, stochastic_vol=(0.0, -0.3, 5, 2, 0, 1))
    b = O.OracleModel(ts, 6)
    # the log-normal drift sums the squared loadings instead of sigma^2: equal up to rounding
    for i in range(small.num_time_steps + 1):
        for j in range(small.num_libors):
            assert np.allclose(a.process_value(i, j), b.process_value(i, j), rtol=1e-14, atol=0)


def heston_scaling(times, P, kappa, theta, xi, seed, F=1, a=0):
    """sqrt(max(V, 0)) of the Euler scheme of dV = kappa (theta - V) dt + xi sqrt(max(V, 0)) dB,
    V(0) = 1, on factor a of an F-factor Brownian motion, [nT][P]."""
    dB = O.brownian(times, F, P, seed)
    V = np.ones(P)
    out = np.empty((times.size - 1, P))
    for i in range(times.size - 1):
        root = np.sqrt(np.maximum(V, 0.0))
        out[i] = root
        V = V + kappa * (theta - V) * (times[i + 1] - times[i]) + xi * root * dB[i, a]
    return out


@pytest.mark.parametrize("measure,scheme,state_space", [(0, 2, 0), (1, 1, 1)])
def test_heston_volatility_against_brute_force(measure, scheme, state_space):
    # LIBORCovarianceModelStochasticHestonVolatility: loadings times sqrt(max(V, 0)) of a
    # square-root variance process on its own Brownian motion; xi = 1.5 makes V hit zero
    wl = tiny_workload()
    t = O.ModelTables(wl, vol_params=(0.2, 0.0, 0.25, 0.1) if state_space == 1 else (0.006, 0.0, 0.25, 0.002))
    P = 8
    hv = (0.8, 0.9, 1.5, 4242, 2, 1)
    scaling = heston_scaling(t.times, P, 0.8, 0.9, 1.5, 4242, F=2, a=1)
    assert np.min(scaling) == 0.0 and np.ptp(scaling) > 0.5
    om = O.OracleModel(t, P, measure=measure, scheme=scheme, state_space=state_space, heston_vol=hv)
    ref = brute_force_paths(t, P, measure, scheme, state_space, scaling=scaling)
    for i in range(wl.num_time_steps + 1):
        for j in range(wl.num_libors):
            assert np.allclose(om.process_value(i, j), ref[i, j], rtol=1e-12, atol=1e-16), (i, j)


def test_heston_variance_mean_reverts_like_its_expectation():
    # E[V(t_{i+1})] = E[V(t_i)] + kappa (theta - E[V(t_i)]) dt for the Euler scheme (the increment
    # has mean zero), so E[V] = theta + (1 - theta)(1 - kappa dt)^i
    wl = W.c0()
    t = O.ModelTables(wl)
    P = 40_000
    S = heston_scaling(t.times, P, 0.5, 0.6, 0.3, 17)
    for i in (4, 10, 19):
        expect = 0.6 + 0.4 * (1 - 0.5 * 0.5) ** i
        V = S[i] ** 2
        assert abs(np.mean(V) - expect) < 4 * np.std(V) / math.sqrt(P) + 1e-3, i


def test_full_blend_with_unit_offset_is_the_plain_model():
    # blend a = 1 with offset 1: f = (L - L) + 1 = 1 exactly, so in the normal state space
    # (no variance term, which the wrapper sums from the loadings) the paths are the plain
    # model's bit for bit
    wl = tiny_workload()
    t = O.ModelTables(wl, vol_params=(0.006, 0.0, 0.25, 0.002))
    P = 6
    for scheme in (0, 2, 3):
        om = O.OracleModel(t, P, scheme=scheme, state_space=0, local_vol=(1.0, 1.0))
        plain = O.OracleModel(t, P, scheme=scheme, state_space=0)
        for i in range(wl.num_time_steps + 1):
            for j in range(wl.num_libors):
                assert np.array_equal(om.process_value(i, j), plain.process_value(i, j))


def test_terminal_numeraire_against_brute_force():
    # N(T_k) = P(T_k; T_n) = prod_{l >= k} 1 / (1 + d L_l(T_k))
    wl = tiny_workload()
    t = O.ModelTables(wl)
    P = 6
    om = O.OracleModel(t, P, measure=1, with_discount=False)
    ref = brute_force_paths(t, P, measure=1)
    n = wl.num_libors
    for k in range(n + 1):
        ti = k               # simulation grid = tenor grid
        expect = np.ones(P)
        for l in range(k, n):
            expect = expect / (1 + 0.5 * ref[ti, l])
        assert np.allclose(om.numeraire(t.tenor[k]), expect, rtol=1e-13, atol=0), k


@pytest.mark.parametrize("measure,scheme", [(0, 2), (0, 1), (1, 2), (1, 1), (1, 0)])
def test_bonds_are_martingale_prices_without_discount_curve(measure, scheme):
    # N(0) E[1 / N(T_k)] = P(0; T_k) = prod_{l < k} 1 / (1 + d L_l(0)) up to the Monte Carlo
    # error (the Euler bias of the drift is far below it at these sizes); a wrong drift
    # sign or summation range in either measure moves the bonds by several errors
    wl = W.c0()
    t = O.ModelTables(wl)
    P = 20000
    om = O.OracleModel(t, P, measure=measure, scheme=scheme, with_discount=False)
    N0 = om.numeraire(0.0)
    assert np.ptp(N0) == 0.0
    for k in range(wl.num_libors + 1):
        x = N0[0] / om.numeraire(t.tenor[k])
        m, se = float(np.mean(x)), float(np.std(x) / math.sqrt(P))
        exact = float(np.prod(1.0 / (1.0 + 0.5 * t.L0[:k])))
        assert abs(m - exact) < 4.0 * se + 1e-12, (k, m, exact, se)


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


@pytest.mark.parametrize("scheme", [1, 2])
def test_terminal_zero_volatility_closed_form(scheme):
    # zero volatility: deterministic forwards, the same prices in both measures
    # (N(0) E[payoff / N(T_e)] with N(T) = P(T; T_n) under the terminal measure)
    wl = W.c0()
    t = O.ModelTables(wl, vol_params=(0.0, 0.0, 0.0, 0.0))
    om = O.OracleModel(t, 64, measure=1, scheme=scheme, with_discount=False)
    L = t.L0
    n = wl.num_libors
    for k in range(n + 1):
        assert np.allclose(om.numeraire(t.tenor[k]), np.prod(1.0 / (1.0 + 0.5 * L[k:])), rtol=1e-14, atol=0)
    s = wl.products[0]
    e = int(round(s["exercise"] / 0.5))
    K = s["strike"][0]
    annuity_swap, df = 0.0, 1.0
    for j in range(e, e + len(s["fixing"])):
        df /= 1 + L[j] * 0.5
        annuity_swap += (L[j] - K) * 0.5 * df
    P_e = np.prod(1.0 / (1.0 + L[:e] * 0.5))
    v, err = om.swaption(s)
    assert abs(v - max(annuity_swap, 0.0) * P_e) < 1e-14
    assert err < 1e-15


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


# ---- off-grid forwards and numeraires (LOG_LINEAR_UNCORRECTED) --------------------------

class FineGrid(W.Workload):
    """Tenor every `period`, simulation every period / 4 (the terminal numeraire off the tenor
    grid needs the time on the simulation grid)."""

    @property
    def num_time_steps(self) -> int:
        return self.num_libors * 4

    def times(self):
        return W.round_ticks(np.arange(self.num_time_steps + 1) * self.period / 4)


def short_forward(t, T, period):
    """ForwardCurveInterpolation.getForward(model, T, dt) of the tables' curve: the node forward
    interpolated at T, converted from the curve's payment offset to dt."""
    f = O.forward_curve(t.node_t, t.node_f, T)
    return f if period == t.wl.period else (math.exp(math.log(1 + f * t.wl.period) * period / t.wl.period) - 1.0) / period


def zero_vol_bond(t, T):
    """P(0; T) at zero volatility without discount curve: the tenor bonds up to the next tenor
    date T_u, times the curve's forward bond from T to T_u (the interpolated long forward and
    the analytic interpolation cancel when the forwards are the curve's)."""
    tenor = t.tenor
    u = int(np.searchsorted(tenor, T - 1e-12))
    if abs(tenor[u] - T) < 1e-9:
        return float(np.prod(1.0 / (1.0 + 0.5 * t.L0[:u])))
    short = tenor[u] - T
    return float(np.prod(1.0 / (1.0 + 0.5 * t.L0[:u]))) * (1 + short_forward(t, T, short) * short)


@pytest.mark.parametrize("interpolation", [0, 1, 2])
@pytest.mark.parametrize("measure", [0, 1])
def test_off_grid_bonds_zero_volatility_closed_form(measure, interpolation):
    # every interpolation method (LINEAR, LOG_LINEAR_UNCORRECTED, LOG_LINEAR_CORRECTED, whose
    # drift adjustment vanishes without volatility) gives the curve's short forward bond
    wl = FineGrid("fine", 8, 0.5, 2, 5, 0.2, np.array([0.02, 0.025, 0.03, 0.034, 0.037, 0.039, 0.04, 0.041]))
    t = O.ModelTables(wl, vol_params=(0.0, 0.0, 0.0, 0.0))
    om = O.OracleModel(t, 5, measure=measure, with_discount=False, interpolation=interpolation)
    N0 = om.numeraire(0.0)
    for T in np.arange(1, 31) * 0.125:
        bond = N0 / om.numeraire(T)
        assert np.allclose(bond, zero_vol_bond(t, T), rtol=2e-14, atol=0), T
    # with the discount curve the bond is its log-linear interpolation between tenor dates
    om = O.OracleModel(t, 5, measure=measure)
    N0 = om.numeraire(0.0)
    for T in np.arange(1, 31) * 0.125:
        k = int(np.floor(T / 0.5 + 1e-9))
        k = min(k, wl.num_libors - 1)
        frac = (T - t.tenor[k]) / 0.5
        expect = t.pd[k] * (t.pd[k + 1] / t.pd[k]) ** frac
        assert np.allclose(N0 / om.numeraire(T), expect, rtol=2e-14, atol=0), T


def test_off_grid_forwards_zero_volatility_are_bond_ratios():
    # at zero volatility every (interpolated) forward is the forward of the zero-vol bonds
    wl = W.Workload("fwd", 8, 0.5, 2, 5, 0.2, np.array([0.02, 0.025, 0.03, 0.034, 0.037, 0.039, 0.04, 0.041]))
    t = O.ModelTables(wl, vol_params=(0.0, 0.0, 0.0, 0.0))
    om = O.OracleModel(t, 5, with_discount=False)
    cases = [(S, S + 0.5) for S in np.arange(1, 25) * 0.125] + [(0.75, 2.0), (1.0, 2.625), (0.375, 1.0), (1.125, 1.375)]
    for S, E in cases:
        for time in (0.0, S, S + 0.3):
            F = om.forward_rate(time, S, E)
            expect = (zero_vol_bond(t, S) / zero_vol_bond(t, E) - 1.0) / (E - S)
            assert np.allclose(F, expect, rtol=1e-12, atol=0), (time, S, E)


def test_interpolated_forward_reads_the_long_forward_at_the_right_time_index():
    # S = 0.875 inside the tenor period [0.5, 1.0): the evaluation time is rounded to the
    # nearest simulation time (LMM:1244-1252) and the long forward L_1 read there, at most at its
    # fixing T_1 (LMM:1328-1334; L_1 is frozen afterwards); its one-plus is raised to
    # alpha = (1.0 - S) / 0.5 and scaled by the curve's short / interpolated one-plus (LMM:1341-1380)
    wl = tiny_workload(n=4, P=64)
    t = O.ModelTables(wl)
    om = O.OracleModel(t, 64, with_discount=False)
    ratio = (1 + short_forward(t, 0.875, 0.125) * 0.125) / math.exp(math.log(1 + short_forward(t, 0.5, 0.5) * 0.5) * 0.25)

    def expect(L):
        return (np.exp(np.log(L * 0.5 + 1.0) * 0.25) * ratio - 1.0) / 0.125

    F = om.forward_rate(0.875, 0.875, 1.0)          # nearest time 1.0, after T_1: read at index 1
    assert np.allclose(F, expect(om.process_value(1, 1)), rtol=1e-13, atol=0)
    assert np.ptp(F) > 0
    assert np.array_equal(om.forward_rate(0.375, 0.875, 1.0), F)   # nearest time 0.5 = T_1 (not 0.0)
    F0 = om.forward_rate(0.125, 0.875, 1.0)         # nearest time 0.0: the time-0 forward
    assert np.ptp(F0) == 0 and np.allclose(F0, expect(t.L0[1]), rtol=1e-13)


@pytest.mark.parametrize("state_space", [1, 0])
def test_log_linear_corrected_forward_against_the_integral(state_space):
    # LOG_LINEAR_CORRECTED (LMM:1355-1360,1432-1499): the log-linear one-plus of the long forward
    # L_j is corrected by exp(-A dt_s (T_j - S) / 2), A the trapezoid integral over the simulation
    # times up to the fixing of sigma_j(t)^2 / (1 + L_j(t) dT)^2, times L_j(t)^2 in the log-normal
    # state space; written here from that formula on the process values
    wl = FineGrid("fine", 6, 0.5, 2, 16, 0.2, np.array([0.02, 0.025, 0.03, 0.034, 0.037, 0.039]))
    t = O.ModelTables(wl, vol_params=(0.3, 0.0, 0.25, 0.1) if state_space == 1 else (0.008, 0.0, 0.25, 0.003))
    P = 16
    corr = O.OracleModel(t, P, with_discount=False, state_space=state_space, interpolation=2)
    unc = O.OracleModel(t, P, with_discount=False, state_space=state_space, interpolation=1)
    times = t.times
    for j, S in ((2, 1.125), (3, 1.875), (4, 2.25)):
        Tj, Te = t.tenor[j], t.tenor[j + 1]
        fix = int(np.searchsorted(times, Tj - 1e-12))
        L = [corr.process_value(min(s, fix), j) for s in range(fix + 1)]
        g = []
        for s in range(fix + 1):
            x = t.var[s, j] / (1 + L[s] * 0.5) ** 2
            g.append(x * L[s] ** 2 if state_space == 1 else x)
        A = sum(0.5 * (times[s] - times[s - 1]) * (g[s] + g[s - 1]) for s in range(1, fix + 1))
        small = Te - S
        ratio = (1 + short_forward(t, S, small) * small) / math.exp(math.log(1 + short_forward(t, Tj, 0.5) * 0.5) * small / 0.5)
        one_plus = np.exp(np.log(L[fix] * 0.5 + 1) * (small / 0.5) - A * (0.5 * small * (Tj - S))) * ratio
        expect = (one_plus - 1) / small
        got = corr.forward_rate(S, S, Te)
        assert np.allclose(got, expect, rtol=1e-12, atol=0), (j, S)
        # the correction is small but present
        assert 0 < np.max(np.abs(got - unc.forward_rate(S, S, Te))) < 1e-3


def test_reference_bond_and_fra_assertions_on_the_oracle():
    # LIBORMarketModelValuationTest.testBond (|bond - P^d(T)| < 8e-3 at 0.125 steps) and
    # testFRA (|par FRA| < 2e-3), on a smaller model: 16 forwards, 3 factors, 5,000 paths
    from finmath_hip import _finmath_host as H
    wl = W.Workload("val", 16, 0.5, 3, 5000, 0.1, np.full(16, 0.05), vol_params=(0.2, 0.0, 0.25, 0.3))
    t = O.ModelTables(wl)
    om = O.OracleModel(t, wl.num_paths, scheme=1)
    N0 = om.numeraire(0.0)
    worst = 0.0
    for T in np.arange(0, 8 * 8 + 1) * 0.125:
        bond = float(np.mean(N0 / om.numeraire(T)))
        k = min(int(np.floor(T / 0.5 + 1e-9)), wl.num_libors - 1)
        pd = t.pd[k] * (t.pd[k + 1] / t.pd[k]) ** ((T - t.tenor[k]) / 0.5)
        worst = max(worst, abs(bond - pd))
    assert worst < 8e-3
    fc = H.ForwardCurveInterpolation.createForwardCurveFromForwards("fc", list(t.node_t), list(t.node_f), 0.5)
    worst = 0.0
    for S in np.arange(0, 8 * 8 - 3) * 0.125:
        E = S + 0.5
        par = H.getForwardSwapRate(H.TimeDiscretizationFromArray([S, E]), H.TimeDiscretizationFromArray([S, E]), fc)
        v = (om.forward_rate(S, S, E) - par) * 0.5 / om.numeraire(E) * N0
        worst = max(worst, abs(float(np.mean(v))))
    assert worst < 2e-3

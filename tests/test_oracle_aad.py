"""Pins of the oracle's tape AAD (oracle/lmm_aad.cpp), a restatement of the reference's
RandomVariableDifferentiableAAD.

- Each operator's derivative against its closed form at 1e-15, as the reference's
  RandomVariableDifferentiableAADTest.java:61-261 asserts (add, sub, mult, div with a
  repeated argument and with two, exp, log), and the other operators the LMM path uses.
- The conditional expectation is self-adjoint: the derivative through it is the least-squares
  projection (numpy lstsq) of the node's derivative on the basis.
- choose under DISCRETE_DELTA is the derivative of choose with the trigger's indicator
  replaced by a ramp of width eps = 0.05 sd(trigger) (central differences of the ramp form).
- The whole tape against central differences of the plain oracle's prices (common random
  numbers): deltas, volatility entries and the four volatility parameters, with and without
  discount curve (the numeraire-adjustment average held at its base value, as the tape
  holds it), both state spaces, and the Bermudan with fixed exercise decisions (ZERO).
"""
import math

import numpy as np
import pytest

import _oracle as O
from finmath_hip import workloads as W


# ---- operators ---------------------------------------------------------------------------

def scalar_tape():
    return O.AadTape(1)


def d(t, r, x):
    return t.derivative(r, x)[0]


def test_reference_operator_cases():
    # RandomVariableDifferentiableAADTest.java:65-262, same inputs and bounds
    t = scalar_tape()
    x = t.input(2.0)
    assert d(t, t.op("add", x, x), x) == pytest.approx(2.0, abs=1e-15)
    assert d(t, t.op("sub", x, x), x) == pytest.approx(0.0, abs=1e-15)
    assert d(t, t.op("mult", x, x), x) == pytest.approx(2.0 * 2.0, abs=1e-15)
    assert d(t, t.op("div", x, x), x) == pytest.approx(0.0, abs=1e-15)
    assert d(t, t.op("exp", x), x) == pytest.approx(math.exp(2.0), abs=1e-15)
    assert d(t, t.op("log", x), x) == pytest.approx(0.5, abs=1e-15)
    x0, x1 = t.input(6.0), t.input(2.0)
    y = t.op("add", x0, x1)
    assert (d(t, y, x0), d(t, y, x1)) == (1.0, 1.0)
    y = t.op("sub", x0, x1)
    assert (d(t, y, x0), d(t, y, x1)) == (1.0, -1.0)
    y = t.op("mult", x0, x1)
    assert (d(t, y, x0), d(t, y, x1)) == (2.0, 6.0)
    y = t.op("div", x0, x1)
    assert d(t, y, x0) == pytest.approx(0.5, abs=1e-15)
    assert d(t, y, x1) == pytest.approx(-6.0 / 4.0, abs=1e-15)


def test_operators_of_the_lmm_path():
    rng = np.random.default_rng(7)
    P = 64
    t = O.AadTape(P)
    xv, rv = rng.uniform(0.5, 2.0, P), rng.uniform(0.01, 0.08, P)
    x, r = t.input(xv), t.input(rv)
    dt = 0.5
    inv = t.op("invert", x)
    assert np.allclose(t.derivative(inv, x), -1.0 / xv**2, rtol=1e-15, atol=0)
    y = t.op("discount", x, r, imm=dt)
    assert np.allclose(t.value(y), xv / (1.0 + rv * dt), rtol=1e-15)
    assert np.allclose(t.derivative(y, x), 1.0 / (1.0 + rv * dt), rtol=1e-15, atol=0)
    assert np.allclose(t.derivative(y, r), -xv * dt / (1.0 + rv * dt) ** 2, rtol=1e-14, atol=0)
    y = t.op("accrue", x, r, imm=dt)
    assert np.allclose(t.derivative(y, x), 1.0 + rv * dt, rtol=1e-15, atol=0)
    assert np.allclose(t.derivative(y, r), xv * dt, rtol=1e-15, atol=0)
    y = t.op("addProduct", x, r, x)
    assert np.allclose(t.derivative(y, x), 1.0 + rv, rtol=1e-15, atol=0)
    assert np.allclose(t.derivative(y, r), xv, rtol=1e-15, atol=0)
    y = t.op("addProduct", x, r, imm=-0.5)
    assert np.allclose(t.derivative(y, r), -0.5, rtol=0, atol=0)
    # cap: derivative 1 strictly below, 0 at and above; floor: 1 at and above, 0 below
    c = float(np.median(xv))
    y = t.op("cap", x, imm=c)
    assert np.array_equal(t.derivative(y, x), np.where(xv - c >= 0, 0.0, 1.0))
    y = t.op("floor", x, imm=c)
    assert np.array_equal(t.derivative(y, x), np.where(xv - c >= 0, 1.0, 0.0))
    t1 = O.AadTape(2)
    z = t1.input([1.5, 0.5])
    assert np.array_equal(t1.derivative(t1.op("cap", z, imm=1.5), z), [0.0, 1.0])
    assert np.array_equal(t1.derivative(t1.op("floor", z, imm=0.5), z), [1.0, 1.0])


@pytest.mark.parametrize("method", ["ONE", "ZERO", "DISCRETE_DELTA"])
def test_choose_trigger_derivative(method):
    rng = np.random.default_rng(11)
    P = 4000
    t = O.AadTape(P, dirac=method)
    tv = rng.standard_normal(P)
    av, bv = rng.uniform(1, 2, P), rng.uniform(-1, 0, P)
    trig, a, b = t.input(tv), t.input(av), t.input(bv)
    y = t.op("choose", trig, a, b)
    assert np.array_equal(t.value(y), np.where(tv >= 0, av, bv))
    assert np.array_equal(t.derivative(y, a), np.where(tv >= 0, 1.0, 0.0))
    assert np.array_equal(t.derivative(y, b), np.where(tv >= 0, 0.0, 1.0))
    g = t.derivative(y, trig)
    if method == "ONE":
        assert np.array_equal(g, av - bv)
    elif method == "ZERO":
        assert np.array_equal(g, np.zeros(P))
    else:
        # the derivative of choose with the indicator replaced by the ramp
        # H(x) = clip((x + eps/2) / eps, 0, 1), eps = 0.05 sd(trigger): central differences
        # of H(x) a + (1 - H(x)) b per path (eps held fixed, as the tape holds the average)
        eps = 0.05 * np.sqrt(np.mean((tv - tv.mean()) ** 2))
        inside = (tv + eps / 2 >= 0) & (tv - eps / 2 < 0)
        assert inside.sum() > 20
        h = 1e-9 * eps

        def ramp_value(x):
            H = np.clip((x + eps / 2) / eps, 0.0, 1.0)
            return H * av + (1 - H) * bv
        fd = (ramp_value(tv + h) - ramp_value(tv - h)) / (2 * h)
        far = np.abs(np.abs(tv) - eps / 2) > 1e3 * h          # away from the ramp's corners
        assert np.allclose(g[far], fd[far], rtol=1e-5, atol=1e-6 * np.max(np.abs(fd)))
        assert np.array_equal(g != 0, inside)


def test_conditional_expectation_is_self_adjoint():
    rng = np.random.default_rng(3)
    P = 500
    t = O.AadTape(P)
    s = rng.standard_normal(P)
    basis = np.stack([np.ones(P), s, s * s])
    ce = t.estimator(basis)
    yv = np.exp(s) + 0.1 * rng.standard_normal(P)
    uv = rng.uniform(-1, 1, P)
    y, u = t.input(yv), t.input(uv, differentiable=False)
    c = t.op("condexp", y, imm=ce)
    X = basis.T

    def project(v):
        return X @ np.linalg.lstsq(X, v, rcond=None)[0]
    assert np.allclose(t.value(c), project(yv), rtol=1e-12, atol=1e-12)
    r = t.op("mult", c, u)
    # d(sum_p u_p CE(y)_p)/dy = CE(u): the projection is symmetric
    assert np.allclose(t.derivative(r, y), project(uv), rtol=1e-10, atol=1e-12)


def test_choose_of_a_conditional_expectation_against_the_ramp_form():
    # the Bermudan's exercise step values = choose(CE(values - under), values, under): the
    # trigger's derivative is regressed on the basis (self-adjoint CE) and reaches both
    # arguments; against central differences of the ramp form with eps and the regression
    # coefficients' Gram matrix held fixed
    rng = np.random.default_rng(5)
    P = 3000
    t = O.AadTape(P, dirac="DISCRETE_DELTA")
    s = rng.standard_normal(P)
    basis = np.stack([np.ones(P), s, s * s])
    ce = t.estimator(basis)
    vv = 0.3 * s + 0.05 * rng.standard_normal(P)
    uv = 0.1 + 0.02 * rng.standard_normal(P)
    v, u = t.input(vv), t.input(uv)
    trig = t.op("condexp", t.op("sub", v, u), imm=ce)
    y = t.op("choose", trig, v, u)
    X = basis.T
    tv = t.value(trig)
    eps = 0.05 * np.sqrt(np.mean((tv - tv.mean()) ** 2))

    def total(vx, ux):
        tr = X @ np.linalg.lstsq(X, vx - ux, rcond=None)[0]
        H = np.clip((tr + eps / 2) / eps, 0.0, 1.0)
        return np.sum(H * vx + (1 - H) * ux)
    gv, gu = t.derivative(y, v), t.derivative(y, u)
    h = 1e-7
    for p in rng.choice(P, 12, replace=False):
        e = np.zeros(P)
        e[p] = h
        fdv = (total(vv + e, uv) - total(vv - e, uv)) / (2 * h)
        fdu = (total(vv, uv + e) - total(vv, uv - e)) / (2 * h)
        assert gv[p] == pytest.approx(fdv, rel=1e-4, abs=1e-6)
        assert gu[p] == pytest.approx(fdu, rel=1e-4, abs=1e-6)


# ---- the LMM tape against the plain oracle's central differences --------------------------

P = 2000


def plain(t, *, L0=None, vol=None, with_discount, state_space=1, scheme=2, cap=1e5):
    tt = O.ModelTables(t.wl, L0=t.L0 if L0 is None else L0)
    tt.vol = t.vol if vol is None else vol
    return O.OracleModel(tt, P, with_discount=with_discount, state_space=state_space, scheme=scheme, cap=cap,
                         threads=4)


def strip_value(m, specs, weights):
    return sum(w * m.swaption(s)[0] for s, w in zip(specs, weights))


def frozen_ratio(t, specs, **kw):
    """r_q = V_q(with discount curve) / V_q(without): the numeraire adjustment of the base run."""
    md, mn = plain(t, with_discount=True, **kw), plain(t, with_discount=False, **kw)
    r = np.array([md.swaption(s)[0] / mn.swaption(s)[0] for s in specs])
    md.close(), mn.close()
    return r


def model_c0():
    wl = W.c0().with_paths(P)
    wl.forwards = 0.02 + 0.001 * np.arange(wl.num_libors)
    specs = [W.swaption_spec(e, 4.0, 0.5, k) for e in (2.0, 4.0) for k in (0.025, 0.035)]
    return wl, O.ModelTables(wl), specs


@pytest.mark.parametrize("with_discount", [False, True])
def test_tape_deltas_and_vegas_against_plain_differences(with_discount):
    wl, t, specs = model_c0()
    w = np.array([1.0, 0.5, -0.7, 2.0])
    v, dY, dvol, _ = O.aad_swaptions(t, P, specs, w, with_discount=with_discount)
    m = plain(t, with_discount=with_discount)
    assert v == pytest.approx(strip_value(m, specs, w), rel=1e-14)
    m.close()
    ratio = frozen_ratio(t, specs) if with_discount else np.ones(len(specs))
    wr = w * ratio

    def fd_of(**bump):
        up = strip_value(plain(t, with_discount=False, **{k: b[0] for k, b in bump.items()}), specs, wr)
        dn = strip_value(plain(t, with_discount=False, **{k: b[1] for k, b in bump.items()}), specs, wr)
        return up, dn
    h = 1e-6
    for j in (0, 3, 7, 12):
        # the leaf is the log-forward Y_j(0): bump L_j(0) by the factor exp(+-h)
        up, dn = fd_of(L0=(t.L0 * np.where(np.arange(wl.num_libors) == j, math.exp(h), 1.0),
                           t.L0 * np.where(np.arange(wl.num_libors) == j, math.exp(-h), 1.0)))
        fd = (up - dn) / (2 * h)
        assert dY[j] == pytest.approx(fd, rel=2e-6, abs=1e-9 * np.max(np.abs(dY)))
    for (i, j) in ((0, 5), (3, 6), (6, 9), (8, 11)):
        vu, vd = t.vol.copy(), t.vol.copy()
        vu[i, j] += h
        vd[i, j] -= h
        up, dn = fd_of(vol=(vu, vd))
        fd = (up - dn) / (2 * h)
        assert dvol[i, j] == pytest.approx(fd, rel=2e-5, abs=2e-6 * np.max(np.abs(dvol)))
    # entries of fixed forwards (T_j <= t_i) get nothing
    fixed = t.times[:-1, None] >= t.tenor[None, :-1]
    assert np.all(dvol[fixed] == 0.0)


def test_tape_four_parameter_vegas_against_plain_differences():
    wl, t, specs = model_c0()
    w = np.ones(len(specs))
    base = tuple(wl.vol_params)
    v, _, _, dabcd = O.aad_swaptions(t, P, specs, w, with_discount=True, vol_mode=1, abcd=base)
    ratio = frozen_ratio(t, specs)
    # a bump small enough that no path's swaption crosses its floor (the payoff's kink)
    h = 1e-7
    for q in range(4):
        vals = []
        for sgn in (1.0, -1.0):
            pr = list(base)
            pr[q] += sgn * h
            vol = O.vol_fourparam(t.times, t.tenor[:-1], *pr)
            vals.append(strip_value(plain(t, vol=vol, with_discount=False), specs, w * ratio))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert dabcd[q] == pytest.approx(fd, rel=1e-6, abs=1e-8 * np.max(np.abs(dabcd)))


def test_tape_normal_state_space_euler_without_cap():
    # the reference's own AAD test set-up: NORMAL, EULER, liborCap infinite, no discount curve
    # (LIBORMarketModelNormalAADSensitivitiesTest.java:288-402)
    wl, t, specs = model_c0()
    t.vol = t.vol * 0.02           # normal volatilities
    w = np.ones(len(specs))
    v, dY, dvol, _ = O.aad_swaptions(t, P, specs, w, with_discount=False, state_space=0, scheme=0, cap=None)
    h = 1e-7
    for j in (2, 6, 10):
        e = np.where(np.arange(wl.num_libors) == j, h, 0.0)
        up = strip_value(plain(t, L0=t.L0 + e, with_discount=False, state_space=0, scheme=0, cap=None), specs, w)
        dn = strip_value(plain(t, L0=t.L0 - e, with_discount=False, state_space=0, scheme=0, cap=None), specs, w)
        assert dY[j] == pytest.approx((up - dn) / (2 * h), rel=1e-5, abs=1e-8)
    j = 8
    vu, vd = t.vol.copy(), t.vol.copy()
    vu[:, j] += h
    vd[:, j] -= h
    up = strip_value(plain(t, vol=vu, with_discount=False, state_space=0, scheme=0, cap=None), specs, w)
    dn = strip_value(plain(t, vol=vd, with_discount=False, state_space=0, scheme=0, cap=None), specs, w)
    assert dvol[:, j].sum() == pytest.approx((up - dn) / (2 * h), rel=1e-5, abs=1e-8)


def test_tape_bermudan_zero_method_is_the_fixed_decision_derivative():
    # ZERO: the exercise decisions do not move; central differences of the plain oracle's
    # value with every path's exercise time held at the base run's
    wl, t, _ = model_c0()
    spec = W.bermudan_spec(2.0, 8.0, 0.5, 0.03, 1.0, True, basis=1)
    v, dY, _, _ = O.aad_bermudan(t, P, spec, with_discount=False, dirac="ZERO")
    m = plain(t, with_discount=False)
    vb, _, _, ex = m.bermudan(spec, per_path=True)
    assert v == pytest.approx(vb, rel=1e-14)
    m.close()

    def fixed_value(L0):
        mm = plain(t, L0=L0, with_discount=False)
        val = np.zeros(P)
        for f, pd, pl, K in zip(spec["fixing"], spec["payment"], spec["period_lengths"], spec["strike"]):
            ti = int(round(f / wl.period))
            L = mm.process_value(ti, ti)
            val += np.where(ex <= f, (L - K) * pl / mm.numeraire(pd), 0.0)
        mm.close()
        return val.mean()
    h = 1e-6
    for j in np.argsort(-np.abs(dY))[:4]:
        up = t.L0 * np.where(np.arange(wl.num_libors) == j, math.exp(h), 1.0)
        dn = t.L0 * np.where(np.arange(wl.num_libors) == j, math.exp(-h), 1.0)
        fd = (fixed_value(up) - fixed_value(dn)) / (2 * h)
        assert dY[j] == pytest.approx(fd, rel=1e-6, abs=1e-9)


def test_tape_bermudan_discrete_delta_adds_the_exercise_boundary_term():
    wl, t, _ = model_c0()
    spec = W.bermudan_spec(2.0, 8.0, 0.5, 0.03, 1.0, True, basis=1)
    _, dz, _, _ = O.aad_bermudan(t, P, spec, with_discount=True, dirac="ZERO")
    _, dd, _, _ = O.aad_bermudan(t, P, spec, with_discount=True, dirac="DISCRETE_DELTA")
    # the boundary term is not zero, and it is small against the delta (the exercise
    # boundary is near-optimal)
    diff = np.max(np.abs(dd - dz))
    assert diff > 0 and diff < 0.2 * np.max(np.abs(dz))

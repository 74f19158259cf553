"""GPU parity of the HIP path against the CPU oracle, through the C ABI.

Bars: the MT19937 word stream and uniforms bit-exact; Brownian increments equal
up to the last-ulp differences of the device log in the AS241 tail branch;
process values and prices in FP64 within 1e-10 relative of the oracle
(the plain-form tolerance), at sizes the oracle finishes in seconds.
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
    sys.stdout.flush()


def make_context(wl, tables, num_paths, *, precision="FP64", store_paths=False, with_discount=True, cap=1e5):
    ctx = _abi.LmmContext(tables.times, tables.tenor, wl.num_factors, num_paths, wl.seed,
                          precision=precision, store_paths=store_paths, libor_cap=cap)
    ctx.set_initial_forwards(tables.L0)
    ctx.set_covariance(tables.lam, tables.var)
    ctx.set_discount_curve(tables.pd if with_discount else None)
    return ctx


def to_products(specs):
    out = []
    for s in specs:
        if s["type"] == "swaption":
            out.append(_abi.Product.swaption(s["exercise"], s["fixing"], s["payment"], s["strike"]))
        else:
            out.append(_abi.Product.swap(s["fixing"], s["payment"], s["strike"], is_pay_fix=s.get("pay_fix", True)))
    return out


@pytest.mark.parametrize("seed,offset,count", [(3141, 0, 3_000_000), (3141, 123_457, 1000), (-5, 0, 70_000),
                                               (1352, 4095, 4099)])
def test_mt_words_bit_exact(seed, offset, count):
    say(f"case mt_words seed={seed} offset={offset} count={count}")
    gpu = _abi.mt19937_words(seed, offset, count)
    ref = O.mt_words(seed, offset, count)
    assert np.array_equal(gpu, ref)


def test_mt_doubles_bit_exact():
    say("case mt_doubles seed=3141 count=200000")
    assert np.array_equal(_abi.mt19937_doubles(3141, 0, 200_000), O.mt_doubles(3141, 0, 200_000))


def test_brownian_increments_match_oracle():
    wl = W.c0()
    t = O.ModelTables(wl)
    P = 10_000
    say(f"case brownian c0 P={P}")
    ctx = make_context(wl, t, P)
    ref = O.brownian(t.times, wl.num_factors, P, wl.seed)
    exact = 0
    total = 0
    for ti in range(wl.num_time_steps):
        for k in range(wl.num_factors):
            g = ctx.brownian_increment(ti, k)
            r = ref[ti, k]
            assert np.max(np.abs(g - r) / np.maximum(np.abs(r), 1e-300)) < 8e-16
            exact += int(np.sum(g == r))
            total += P
    say(f"bit-exact increments: {exact}/{total}")
    assert exact / total > 0.999


@pytest.mark.parametrize("precision", ["FP64", "FP32_MIXED"])
def test_brownian_increments_on_a_non_uniform_grid(precision):
    # steps of different lengths: every increment takes the sqrt(dt) of its own step
    times = W.round_ticks(np.array([0.0, 0.1, 0.5, 0.75, 1.0, 1.5, 2.0, 3.0]))
    tenor = W.round_ticks(np.array([0.0, 0.5, 1.0, 1.5, 2.0]))
    F, P, n, nT = 2, 3001, 4, 7
    say(f"case brownian non-uniform grid {precision} P={P}")
    ctx = _abi.LmmContext(times, tenor, F, P, 3141, precision=precision)
    ctx.set_initial_forwards(np.full(n, 0.03))
    ctx.set_covariance(np.full((nT, n, F), 0.1), np.full((nT, n), 0.02))
    ref = O.brownian(times, F, P, 3141)
    tol = 8e-16 if precision == "FP64" else 3e-6
    for ti in range(nT):
        for k in range(F):
            g = ctx.brownian_increment(ti, k)
            r = ref[ti, k]
            assert np.max(np.abs(g - r) / np.maximum(np.abs(r), 1e-3)) < tol


@pytest.mark.parametrize("precision", ["FP64", "FP64_LOGSTATE"])
@pytest.mark.parametrize("name,P", [("c0", 10_000), ("c1", 4_000)])
def test_process_values_match_oracle(name, P, precision):
    wl = W.ALL[name]()
    t = O.ModelTables(wl)
    say(f"case process {name} P={P} {precision}")
    ctx = make_context(wl, t, P, store_paths=True, precision=precision)
    ctx.simulate()
    om = O.OracleModel(t, P)
    worst = 0.0
    for ti in range(wl.num_time_steps + 1):
        for j in range(wl.num_libors):
            g, r = ctx.process_value(ti, j), om.process_value(ti, j)
            worst = max(worst, float(np.max(np.abs(g - r) / np.abs(r))))
    say(f"worst relative process difference {worst:.3e}")
    assert worst < 1e-12
    for k in range(wl.num_libors + 1):
        T = t.tenor[k]
        g, r = ctx.numeraire(T), om.numeraire(T)
        assert np.max(np.abs(g - r) / np.abs(r)) < 1e-12


@pytest.mark.parametrize("precision", ["FP64", "FP64_LOGSTATE"])
@pytest.mark.parametrize("name,P", [("c0", 10_000), ("c1", 20_000)])
def test_prices_match_oracle_fp64(name, P, precision):
    wl = W.ALL[name]()
    t = O.ModelTables(wl)
    say(f"case prices {name} P={P} {precision}")
    ctx = make_context(wl, t, P, precision=precision)
    vals, errs = ctx.value_products(to_products(wl.products))
    om = O.OracleModel(t, P)
    for q, s in enumerate(wl.products):
        rv, re_ = om.value(s)
        say(f"  {s['type']} q={q}: gpu {vals[q]:.15e} oracle {rv:.15e} err {errs[q]:.3e}/{re_:.3e}")
        assert abs(vals[q] - rv) <= 1e-10 * abs(rv)
        assert abs(errs[q] - re_) <= 1e-8 * abs(re_)


def test_prices_without_discount_curve_and_without_cap():
    wl = W.c0()
    t = O.ModelTables(wl)
    P = 5000
    say(f"case prices c0 no-discount no-cap P={P}")
    ctx = make_context(wl, t, P, with_discount=False, cap=None)
    vals, _ = ctx.value_products(to_products(wl.products))
    om = O.OracleModel(t, P, with_discount=False, cap=None)
    for q, s in enumerate(wl.products):
        rv, _ = om.value(s)
        assert abs(vals[q] - rv) <= 1e-10 * abs(rv)


def test_products_sharing_an_exercise_date_match_oracle():
    # several products evaluated at one time index: five swaptions over the same periods (flat and
    # per-period strikes, other notionals), interleaved with a caplet, a swaption over fewer
    # periods and a swap
    wl = W.c0()
    t = O.ModelTables(wl)
    P = 6000
    say(f"case products at one exercise date c0 P={P}")
    fix, pay = [2.0, 2.5, 3.0, 3.5], [2.5, 3.0, 3.5, 4.0]
    sw = [dict(type="swaption", exercise=2.0, fixing=fix, payment=pay, strike=[0.03] * 4),
          dict(type="swaption", exercise=2.0, fixing=fix, payment=pay, strike=[0.02, 0.025, 0.03, 0.035], notional=2.0),
          dict(type="swaption", exercise=2.0, fixing=fix[:3], payment=pay[:3], strike=[0.03] * 3),
          dict(type="swaption", exercise=2.0, fixing=fix, payment=pay, strike=[0.04] * 4),
          dict(type="swaption", exercise=2.0, fixing=fix, payment=pay, strike=[0.05] * 4),
          dict(type="swaption", exercise=2.0, fixing=fix, payment=pay, strike=[0.01] * 4, notional=0.5)]
    swap = dict(type="swap", fixing=[1.0, 1.5, 2.0], payment=[1.5, 2.0, 2.5], strike=[0.03] * 3)
    ctx = make_context(wl, t, P)
    prods = []
    for s in sw[:2]:
        prods.append(_abi.Product.swaption(s["exercise"], s["fixing"], s["payment"], s["strike"],
                                           notional=s.get("notional", 1.0)))
    prods.append(_abi.Product.caplet(2.0, 0.5, 0.03))
    for s in sw[2:]:
        prods.append(_abi.Product.swaption(s["exercise"], s["fixing"], s["payment"], s["strike"],
                                           notional=s.get("notional", 1.0)))
    prods.append(_abi.Product.swap(swap["fixing"], swap["payment"], swap["strike"]))
    vals, errs = ctx.value_products(prods)
    om = O.OracleModel(t, P)
    ref = [om.swaption(s) for s in sw[:2]] + [om.caplet(2.0, 0.5, 0.03)] + [om.swaption(s) for s in sw[2:]] + \
          [om.swap(swap)]
    for q, (rv, re_) in enumerate(ref):
        say(f"  q={q}: gpu {vals[q]:.15e} oracle {rv:.15e}")
        assert abs(vals[q] - rv) <= 1e-10 * abs(rv)
        assert abs(errs[q] - re_) <= 1e-8 * abs(re_)


def test_repeated_pricing_is_deterministic():
    wl = W.c0()
    t = O.ModelTables(wl)
    say("case determinism c0 P=30000")
    ctx = make_context(wl, t, 30_000)
    a, ea = ctx.value_products(to_products(wl.products))
    b, eb = ctx.value_products(to_products(wl.products))
    assert np.array_equal(a, b) and np.array_equal(ea, eb)


def test_ragged_path_counts():
    # path counts that leave partial wavefronts and partial workgroups
    wl = W.c0()
    t = O.ModelTables(wl)
    for P in (1, 63, 65, 257):
        say(f"case ragged c0 P={P}")
        ctx = make_context(wl, t, P)
        vals, _ = ctx.value_products(to_products(wl.products))
        om = O.OracleModel(t, P)
        for q, s in enumerate(wl.products):
            rv, _ = om.value(s)
            assert abs(vals[q] - rv) <= 1e-10 * max(abs(rv), 1e-12)


def test_fp32_mixed_prices_within_a_tenth_of_a_basis_point():
    wl = W.c1()
    t = O.ModelTables(wl)
    P = 20_000
    say(f"case fp32 mixed vs oracle c1 P={P}")
    ctx = make_context(wl, t, P, precision="FP32_MIXED")
    vals, errs = ctx.value_products(to_products(wl.products))
    om = O.OracleModel(t, P)
    worst = 0.0
    for q, s in enumerate(wl.products):
        rv, _ = om.value(s)
        worst = max(worst, abs(vals[q] - rv) * 1e4)
    say(f"worst FP32-mixed price error {worst:.4f} bp")
    assert worst < 0.1


def test_materialised_process_agrees_with_fused_prices():
    # the swaption recomputed from getProcessValue/getNumeraire equals the fused pass
    wl = W.c0()
    t = O.ModelTables(wl)
    P = 5000
    say(f"case materialised vs fused c0 P={P}")
    ctx = make_context(wl, t, P, store_paths=True)
    ctx.simulate()
    s = wl.products[0]
    e = int(round(s["exercise"] / 0.5))
    V = np.zeros(P)
    for j in reversed(range(e, e + len(s["fixing"]))):
        L = ctx.process_value(e, j)
        V = (V + (L - s["strike"][0]) * 0.5 * 1.0) / (1.0 + L * 0.5) * 1.0
    x = np.maximum(V, 0.0) / ctx.numeraire(s["exercise"])
    fused, _ = ctx.value_products(to_products([s]))
    assert abs(np.mean(x) - fused[0]) < 1e-12 * abs(fused[0])


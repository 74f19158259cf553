"""Device random variables (include/finmath_rv.h) against numpy and the reference's
formulas, through the C ABI.

Bars: operations that IEEE arithmetic rounds correctly (+, -, *, /, sqrt, min, max,
choose and the compound forms of RandomVariableFromDoubleArray, evaluated without fused
multiply-add) bit-exact against numpy; exp, log, pow, expm1, sin, cos within 2 ulp;
Java's Math.min/max rules for NaN and signed zeros; averages within 2 ulp of the
reference's sequential Kahan sum (RandomVariableFromDoubleArray.java:287-307) and
variances to 1e-14; a lazily fused expression bit-exact against the same expression
materialised one operation at a time; products valued with the algebra on the process
of a context equal the fused path kernel's prices to 1e-12.
"""
import math

import numpy as np
import pytest

import _oracle as O
from finmath_hip import _abi
from finmath_hip import randomvariable as R
from finmath_hip import workloads as W
from test_gpu_parity import make_context, say

pytestmark = pytest.mark.gpu

N = 100_003


def kahan_average(x, p=None):
    """RandomVariableFromDoubleArray.getAverage (Kahan summation, :287-331), sequentially."""
    s, err = 0.0, 0.0
    for i in range(len(x)):
        v = (x[i] if p is None else x[i] * p[i]) - err
        t = s + v
        err = (t - s) - v
        s = t
    return s / len(x)


def ulps(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b) / np.maximum(np.spacing(np.abs(b)), 5e-324)


@pytest.fixture(scope="module")
def data():
    rng = np.random.default_rng(2024)
    x = rng.standard_normal(N)
    y = rng.uniform(0.5, 2.0, N)
    z = rng.uniform(-1.0, 1.0, N)
    return x, y, z, R.RandomVariable.from_array(x, 1.0), R.RandomVariable.from_array(y, 2.0), \
        R.RandomVariable.from_array(z, 0.5)


def test_correctly_rounded_operations_are_bit_exact(data):
    say("case rv elementwise bit-exact")
    x, y, z, X, Y, Z = data
    with np.errstate(all="ignore"):
        cases = [
            (X.add(Y), x + y), (X.sub(Y), x - y), (X.bus(Y), y - x), (X.mult(Y), x * y), (X.div(Y), x / y),
            (X.vid(Y), y / x), (X.add(0.25), x + 0.25), (X.mult(-3.0), x * -3.0), (Y.sqrt(), np.sqrt(y)),
            (X.abs(), np.abs(x)), (Y.invert(), 1.0 / y), (X.squared(), x * x),
            (X.addProduct(Y, Z), x + y * z), (X.addProduct(Y, 0.5), x + y * 0.5),
            (X.addRatio(Z, Y), x + z / y), (X.subRatio(Z, Y), x - z / y),
            (Y.accrue(Z, 0.5), y * (1 + z * 0.5)), (Y.discount(Z, 0.5), y / (1.0 + z * 0.5)),
            (X.choose(Y, Z), np.where(x >= 0.0, y, z)), (X.cap(Z), np.minimum(x, z)), (X.floor(0.0), np.maximum(x, 0.0)),
            (X.addSumProduct([Y, Z], [Z, Y]), (x + y * z) + z * y),
        ]
    for k, (rv, ref) in enumerate(cases):
        got = rv.getRealizations()
        assert np.array_equal(got, ref), k


def test_transcendental_operations_within_two_ulp(data):
    say("case rv transcendental")
    x, y, z, X, Y, Z = data
    for rv, ref in [(X.exp(), np.exp(x)), (Y.log(), np.log(y)), (Y.pow(1.7), np.power(y, 1.7)),
                    (Z.expm1(), np.expm1(z)), (X.sin(), np.sin(x)), (X.cos(), np.cos(x))]:
        assert np.max(ulps(rv.getRealizations(), ref)) <= 2.0


def test_java_min_max_and_choose_special_values():
    say("case rv special values")
    inf, nan = math.inf, math.nan
    a = np.array([nan, 1.0, -0.0, 0.0, -inf, inf, 2.0, 0.0])
    b = np.array([1.0, nan, 0.0, -0.0, 3.0, 3.0, 2.0, 0.0])
    A, B = R.RandomVariable.from_array(a), R.RandomVariable.from_array(b)
    mn, mx = A.cap(B).getRealizations(), A.floor(B).getRealizations()
    # Math.min/max: NaN wins, -0.0 < +0.0
    assert np.isnan(mn[0]) and np.isnan(mn[1]) and np.isnan(mx[0]) and np.isnan(mx[1])
    assert math.copysign(1, mn[2]) < 0 and math.copysign(1, mn[3]) < 0
    assert math.copysign(1, mx[2]) > 0 and math.copysign(1, mx[3]) > 0
    assert mn[4] == -inf and mx[5] == inf and mn[6] == 2.0
    # choose: NaN and -0.0 triggers (NaN >= 0 is false, -0.0 >= 0 is true)
    ch = A.choose(R.RandomVariable.constant(1.0), R.RandomVariable.constant(-1.0)).getRealizations()
    assert list(ch[:4]) == [-1.0, 1.0, 1.0, 1.0]
    assert list(A.isNaN().getRealizations()[:2]) == [1.0, 0.0]


def test_average_and_variance_match_the_reference_formulas(data):
    say("case rv average/variance")
    x, y, z, X, Y, Z = data
    ref = kahan_average(x)
    assert abs(X.getAverage() - ref) <= 2 * np.spacing(abs(ref))
    var = kahan_average((x - ref) * (x - ref))
    assert abs(X.getVariance() - var) <= 1e-14 * var
    assert X.getStandardError() == pytest.approx(math.sqrt(var) / math.sqrt(N), rel=1e-14)
    # getAverage(probabilities) = sum(x p) / n
    w = R.RandomVariable.from_array(y)
    refw = kahan_average(x, y)
    assert abs(X.getAverage(w) - refw) <= 4 * np.spacing(abs(refw)) + 1e-18
    # fused: the average of an expression without materialising it
    e = X.mult(Y).add(Z).exp()
    assert e.getAverage() == pytest.approx(kahan_average(np.exp(x * y + z)), rel=2e-15)


def test_average_is_reproducible_and_size_independent_in_accuracy():
    say("case rv reproducible sums")
    rng = np.random.default_rng(5)
    for n in (1, 2, 511, 513, 1_000_001):
        x = rng.uniform(-1, 1, n) * 1e3 + 1e6
        X = R.RandomVariable.from_array(x)
        a1, a2 = X.getAverage(), X.getAverage()
        assert a1 == a2
        exact = float(np.sum(x.astype(np.longdouble)) / n)
        assert abs(a1 - exact) <= 2 * np.spacing(exact)


def test_lazy_fusion_equals_step_by_step_evaluation(data):
    say("case rv fusion vs eager")
    x, y, z, X, Y, Z = data
    lazy = X
    eager = X
    for k in range(12):
        lazy = lazy.mult(Y).add(Z).discount(Y, 0.5).cap(5.0)
        eager = eager.mult(Y).cache().add(Z).cache().discount(Y, 0.5).cache().cap(5.0).cache()
    assert np.array_equal(lazy.getRealizations(), eager.getRealizations())


def test_expressions_beyond_one_program_are_split(data):
    say("case rv long expression")
    x, y, z, X, Y, Z = data
    e, ref = X, x.copy()
    for k in range(120):                       # about 360 instructions: several programs
        e = e.mult(0.999).add(Y).sub(Z.mult(float(k % 7)))
        ref = ref * 0.999 + y - z * float(k % 7)
    assert np.array_equal(e.getRealizations(), ref)
    # many distinct inputs (more than 16)
    rng = np.random.default_rng(9)
    arrays = [rng.uniform(0, 1, 1000) for _ in range(40)]
    s, ref = R.RandomVariable.from_array(arrays[0]), arrays[0].copy()
    for a in arrays[1:]:
        s = s.add(R.RandomVariable.from_array(a))
        ref = ref + a
    assert np.array_equal(s.getRealizations(), ref)


def test_constants_stay_on_the_host():
    c = R.RandomVariable.constant(2.0).mult(3.0).add(1.0).exp()
    assert c.isDeterministic() and c.devicePointer() == 0
    assert c.doubleValue() == math.exp(7.0)
    assert c.getAverage() == math.exp(7.0) and c.getVariance() == 0.0


def test_mismatched_sizes_raise():
    with pytest.raises(_abi.IllegalArgumentException):
        R.RandomVariable.from_array(np.ones(10)).add(R.RandomVariable.from_array(np.ones(11)))


def test_swaption_valued_with_the_algebra_equals_the_fused_price():
    # Swaption.getValue (products/Swaption.java:137-199) written with random-variable
    # operations on the materialised process
    wl = W.c0()
    t = O.ModelTables(wl)
    P = 20_000
    say(f"case rv swaption c0 P={P}")
    ctx = make_context(wl, t, P, store_paths=True)
    ctx.simulate()
    s = wl.products[0]
    fused, _ = ctx.value_products([_abi.Product.swaption(s["exercise"], s["fixing"], s["payment"], s["strike"])])
    ex = s["exercise"]
    e = int(round(ex / 0.5))
    v = R.RandomVariable.constant(0.0)
    for p in reversed(range(len(s["fixing"]))):
        fix, pay, K = s["fixing"][p], s["payment"][p], s["strike"][p]
        libor = R.process_value(ctx, e, int(round(fix / 0.5)))
        payoff = libor.sub(K).mult(pay - fix).mult(1.0)
        v = v.add(payoff).discount(libor, pay - max(fix, ex)).mult(1.0)
    values = v.floor(0.0)
    values = values.div(R.numeraire(ctx, ex)).mult(R.monte_carlo_weights(ctx))
    values = values.mult(R.numeraire(ctx, 0.0)).div(R.monte_carlo_weights(ctx))
    value = values.getAverage()
    say(f"  algebra {value:.15e} fused {fused[0]:.15e}")
    assert abs(value - fused[0]) <= 1e-12 * abs(fused[0])


def test_digital_caplet_from_process_values_matches_numpy():
    # a product the fused kernel does not know: payoff 1{L > K} d / N(T_pay) (the
    # reference's DigitalCaplet through the RandomVariable interface)
    wl = W.c0()
    t = O.ModelTables(wl)
    P = 30_001
    say(f"case rv digital caplet c0 P={P}")
    ctx = make_context(wl, t, P, store_paths=True)
    ctx.simulate()
    K, T, dt = 0.03, 3.0, 0.5
    L = R.process_value(ctx, int(T / 0.5), int(T / 0.5))
    N = R.numeraire(ctx, T + dt)
    payoff = L.sub(K).choose(R.RandomVariable.constant(dt), R.RandomVariable.constant(0.0)).div(N)
    value = payoff.getAverage()
    l_host = ctx.process_value(int(T / 0.5), int(T / 0.5))
    n_host = ctx.numeraire(T + dt)
    ref = kahan_average(np.where(l_host - K >= 0.0, dt, 0.0) / n_host)
    assert abs(value - ref) <= 4 * np.spacing(ref)
    assert np.array_equal(L.getRealizations(), l_host)
    assert np.array_equal(N.getRealizations(), n_host)

"""Path ranges and serialisation through the C ABI.

A context with path_offset p0 simulates paths [p0, p0 + P) of the seed's stream: its
increments are the oracle's increments of those paths (the generator jumps the seed state
by 2 p0 nT F words), its process values are the full run's bit for bit, and prices without
discount curve combine linearly. A serialised context regenerates the same increments,
process and prices bit for bit (the reference serialises parameters and seed and
regenerates transient state: LIBORMarketModelFromCovarianceModel.java:1734-1745,
BrownianMotionFromMersenneRandomNumbers.java:253-257).
"""
import numpy as np
import pytest

import _oracle as O
from finmath_hip import _abi
from finmath_hip import workloads as W
from test_gpu_parity import say, to_products

pytestmark = pytest.mark.gpu


def context(wl, t, P, offset=0, with_discount=True, store_paths=True, precision="FP64"):
    ctx = _abi.LmmContext(t.times, t.tenor, wl.num_factors, P, wl.seed, precision=precision,
                          store_paths=store_paths, path_offset=offset)
    ctx.set_initial_forwards(t.L0)
    ctx.set_covariance(t.lam, t.var)
    ctx.set_discount_curve(t.pd if with_discount else None)
    return ctx


def test_offset_increments_are_the_oracle_stream_of_those_paths():
    wl = W.c0()
    t = O.ModelTables(wl)
    p0, P = 2_345, 1_111
    say(f"case path offset increments c0 offset={p0} P={P}")
    ctx = context(wl, t, P, offset=p0, store_paths=False)
    ref = O.brownian(t.times, wl.num_factors, p0 + P, wl.seed)[:, :, p0:]
    exact = 0
    for ti in range(wl.num_time_steps):
        for k in range(wl.num_factors):
            g = ctx.brownian_increment(ti, k)
            assert np.max(np.abs(g - ref[ti, k]) / np.maximum(np.abs(ref[ti, k]), 1e-300)) < 8e-16
            exact += int(np.sum(g == ref[ti, k]))
    assert exact / (P * wl.num_time_steps * wl.num_factors) > 0.999


@pytest.mark.parametrize("precision", ["FP64", "FP32_MIXED"])
def test_path_ranges_reproduce_the_full_run(precision):
    wl = W.c0()
    t = O.ModelTables(wl)
    parts = [(0, 1000), (1000, 1237), (2237, 763)]
    P = sum(n for _, n in parts)
    say(f"case path ranges c0 P={P} parts={parts} {precision}")
    full = context(wl, t, P, precision=precision)
    full.simulate()
    for p0, n in parts:
        part = context(wl, t, n, offset=p0, precision=precision)
        part.simulate()
        for ti in (1, 7, wl.num_time_steps):
            for j in (ti, wl.num_libors - 1):
                if j >= wl.num_libors:
                    continue
                assert np.array_equal(part.process_value(ti, j), full.process_value(ti, j)[p0:p0 + n])
        part.close()


def test_prices_of_path_ranges_combine_to_the_full_price():
    # without discount curve every price is a plain average, so the ranges combine linearly
    wl = W.c0()
    t = O.ModelTables(wl)
    parts = [(0, 4096), (4096, 3000), (7096, 2904)]
    P = sum(n for _, n in parts)
    say(f"case path-range prices c0 P={P}")
    prods = to_products(wl.products)
    full, _ = context(wl, t, P, with_discount=False, store_paths=False).value_products(prods)
    acc = np.zeros(len(prods))
    for p0, n in parts:
        v, _ = context(wl, t, n, offset=p0, with_discount=False, store_paths=False).value_products(prods)
        acc += v * n
    assert np.allclose(acc / P, full, rtol=1e-13, atol=1e-16)


def test_serialised_context_regenerates_the_same_paths_and_prices():
    wl = W.c0()
    t = O.ModelTables(wl)
    P = 3_001
    say(f"case serialisation c0 P={P}")
    ctx = context(wl, t, P, offset=17)
    ctx.set_volatility_factorization(np.asarray(t.vol), np.asarray(t.factors))
    blob = ctx.serialize()
    again = _abi.LmmContext.deserialize(blob)
    assert again.serialize() == blob
    prods = to_products(wl.products)
    a, ea = ctx.value_products(prods)
    b, eb = again.value_products(prods)
    assert np.array_equal(a, b) and np.array_equal(ea, eb)
    ctx.simulate()
    again.simulate()
    for ti in (0, 5, wl.num_time_steps):
        for j in (5, wl.num_libors - 1):
            assert np.array_equal(ctx.process_value(ti, j), again.process_value(ti, j))
    va, da, sa = ctx.adjoint(prods[:1])
    vb, db, sb = again.adjoint(prods[:1])
    assert va == vb and np.array_equal(da, db) and np.array_equal(sa, sb)
    with pytest.raises(_abi.IllegalArgumentException):
        _abi.LmmContext.deserialize(blob[:-8])


def test_path_range_beyond_the_jump_table_is_refused():
    wl = W.c0()
    t = O.ModelTables(wl)
    with pytest.raises(_abi.UnsupportedOperationException):
        context(wl, t, 10, offset=2**41 // (2 * wl.num_time_steps * wl.num_factors))

"""fm_lmm_value_products_noalloc (no device allocation) against fm_lmm_value_products, through
the C ABI. Both run the same generator, path kernel and reduction order on the same
increments, so values and standard errors must be identical bit for bit. Two contexts in a row,
the second smaller, so stale scratch contents of the first call would show in the second."""
import numpy as np
import pytest

import _oracle as O
from finmath_hip import _abi
from finmath_hip import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def context(wl, t, P, precision="FP64", regenerate=False):
    ctx = _abi.LmmContext(t.times, t.tenor, wl.num_factors, P, wl.seed, precision=precision,
                          regenerate_increments=regenerate)
    ctx.set_initial_forwards(t.L0)
    ctx.set_covariance(t.lam, t.var)
    ctx.set_discount_curve(t.pd)
    return ctx


def swaptions(specs):
    return [_abi.Product.swaption(s["exercise"], s["fixing"], s["payment"], s["strike"]) for s in specs]


@pytest.fixture(scope="module")
def c1():
    wl = W.c1()
    return wl, O.ModelTables(wl)


@pytest.mark.parametrize("precision", ["FP64", "FP32_MIXED"])
def test_noalloc_equals_value_products_two_calls(c1, precision):
    wl, t = c1
    strip = swaptions(wl.products)
    # first call: 200,000 paths, the 30-swaption strip
    say(f"case noalloc {precision} P=200000 products=30")
    ctx = context(wl, t, 200_000, precision)
    v_ref, e_ref = ctx.value_products(strip)
    v, e = ctx.value_products_noalloc(strip)
    assert np.array_equal(v, v_ref) and np.array_equal(e, e_ref)
    ctx.close()
    # second call, smaller: 50,003 paths (a ragged last wavefront), 7 swaptions and 3 caplets
    say(f"case noalloc {precision} P=50003 products=10")
    prods = swaptions(wl.products[::4]) + [_abi.Product.caplet(m, 0.5, 0.025) for m in (1.0, 4.5, 9.5)]
    ctx = context(wl, t, 50_003, precision)
    v_ref, e_ref = ctx.value_products(prods)
    v, e = ctx.value_products_noalloc(prods)
    assert np.array_equal(v, v_ref) and np.array_equal(e, e_ref)
    # repeated calls on the same context give the same values
    v2, e2 = ctx.value_products_noalloc(prods)
    assert np.array_equal(v2, v_ref) and np.array_equal(e2, e_ref)
    ctx.close()


def test_noalloc_regenerating_increments(c1):
    wl, t = c1
    strip = swaptions(wl.products[:5])
    say("case noalloc regenerate_increments P=30000")
    ctx = context(wl, t, 30_000, regenerate=True)
    v_ref, e_ref = ctx.value_products(strip)
    v, e = ctx.value_products_noalloc(strip)
    assert np.array_equal(v, v_ref) and np.array_equal(e, e_ref)
    ctx.close()


def test_noalloc_refusals(c1):
    wl, t = c1
    strip = swaptions(wl.products[:3])
    say("case noalloc refusals")
    # a fresh context has no device buffers yet: the call would have to allocate
    ctx = context(wl, t, 1000)
    with pytest.raises(_abi.IllegalStateException):
        ctx.value_products_noalloc(strip)
    ctx.value_products(strip)
    # swaps need per-path columns
    swap = _abi.Product.swap(np.arange(1.0, 5.0, 0.5), np.arange(1.5, 5.5, 0.5), np.full(8, 0.03))
    with pytest.raises(_abi.UnsupportedOperationException):
        ctx.value_products_noalloc([swap])
    ctx.close()


def test_noalloc_rejects_a_pass_beyond_its_scratch(c1):
    # 4,000,000 paths x (30 + 40 + 1) partial columns x 16 B / 64 = 71 MB > 64 MiB
    wl, t = c1
    strip = swaptions(wl.products)
    say("case noalloc beyond scratch P=4000000 FP32_MIXED")
    ctx = context(wl, t, 4_000_000, "FP32_MIXED")
    ctx.value_products(strip[:1])
    with pytest.raises(_abi.IllegalArgumentException, match="bytes of scratch"):
        ctx.value_products_noalloc(strip)
    # within the limit on the same context
    v_ref, _ = ctx.value_products(strip[:10])
    v, _ = ctx.value_products_noalloc(strip[:10])
    assert np.array_equal(v, v_ref)
    ctx.close()

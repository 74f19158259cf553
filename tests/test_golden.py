"""Golden vectors (tests/golden/lmm_golden.json, written by tests/golden/make_golden.py).

CPU: the oracle still reproduces its committed vectors (words exactly, reals to 1e-15).
GPU: the engine reproduces them through the C ABI (words and doubles exactly, process
values and numeraires to 1e-12, prices to 1e-10 relative, the FP64 bar).
"""
import json
import os

import numpy as np
import pytest

import _oracle as O
from finmath_hip import workloads as W

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lmm_golden.json")


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def test_oracle_reproduces_generator_vectors(golden):
    for rec in golden["mt"]:
        assert O.mt_words(golden["seed"], rec["offset"], len(rec["words"])).tolist() == rec["words"]
    for rec in golden["doubles"]:
        assert O.mt_doubles(golden["seed"], rec["offset"], len(rec["values"])).tolist() == rec["values"]


def test_oracle_reproduces_model_vectors(golden):
    wl = W.ALL[golden["model"]]()
    om = O.OracleModel(O.ModelTables(wl), golden["paths"])
    for rec in golden["brownian"]:
        np.testing.assert_allclose(om.brownian_increment(rec["t"], rec["k"])[rec["paths"]], rec["values"], rtol=1e-15)
    for rec in golden["process"]:
        np.testing.assert_allclose(om.process_value(rec["t"], rec["j"])[rec["paths"]], rec["values"], rtol=1e-15)
    for rec in golden["numeraire"]:
        np.testing.assert_allclose(om.numeraire(rec["T"])[rec["paths"]], rec["values"], rtol=1e-15)
    for rec in golden["products"]:
        v, e = om.value(rec["spec"])
        assert v == pytest.approx(rec["value"], rel=1e-14, abs=1e-18)
        assert e == pytest.approx(rec["error"], rel=1e-12)
    om.close()


@pytest.mark.gpu
def test_engine_reproduces_golden_vectors(golden):
    from finmath_hip import _abi
    from test_gpu_parity import make_context, say, to_products
    for rec in golden["mt"]:
        say(f"case golden mt offset={rec['offset']}")
        assert _abi.mt19937_words(golden["seed"], rec["offset"], len(rec["words"])).tolist() == rec["words"]
    for rec in golden["doubles"]:
        assert _abi.mt19937_doubles(golden["seed"], rec["offset"], len(rec["values"])).tolist() == rec["values"]
    wl = W.ALL[golden["model"]]()
    t = O.ModelTables(wl)
    say(f"case golden model {golden['model']} P={golden['paths']}")
    ctx = make_context(wl, t, golden["paths"], store_paths=True)
    ctx.simulate()
    for rec in golden["brownian"]:
        np.testing.assert_allclose(ctx.brownian_increment(rec["t"], rec["k"])[rec["paths"]], rec["values"], rtol=8e-16)
    for rec in golden["process"]:
        np.testing.assert_allclose(ctx.process_value(rec["t"], rec["j"])[rec["paths"]], rec["values"], rtol=1e-12)
    for rec in golden["numeraire"]:
        np.testing.assert_allclose(ctx.numeraire(rec["T"])[rec["paths"]], rec["values"], rtol=1e-12)
    vals, errs = ctx.value_products(to_products([r["spec"] for r in golden["products"]]))
    for q, rec in enumerate(golden["products"]):
        assert abs(vals[q] - rec["value"]) <= 1e-10 * abs(rec["value"])
        assert abs(errs[q] - rec["error"]) <= 1e-8 * abs(rec["error"])

"""Writes the golden vectors of tests/golden/lmm_golden.json from the CPU oracle.

    python tests/golden/make_golden.py

The vectors are small (a few hundred numbers): MT19937 words and nextDouble values of
the reference's seed at stream offsets inside and across path blocks, Brownian
increments, process values, numeraires and product prices of the 20-forward,
3-factor model at 2,000 paths. Tests compare the oracle (regression of the
restatement) and the GPU engine (through the C ABI) against them.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "finmath-lib_amd"))

import _oracle as O  # noqa: E402
from finmath_hip import workloads as W  # noqa: E402

PATHS = 2000
SAMPLE_PATHS = [0, 1, 2, 997, 1999]
PROCESS_CELLS = [(0, 0), (1, 1), (5, 7), (10, 10), (10, 19), (19, 19), (20, 19)]


def main():
    out = {"seed": 3141, "mt": [], "doubles": [], "model": "c0", "paths": PATHS}
    for offset in (0, 623, 624, 120 * 1000 + 5, 1_000_003):
        out["mt"].append({"offset": offset, "words": [int(x) for x in O.mt_words(3141, offset, 8)]})
    for offset in (0, 60, 500_001):
        out["doubles"].append({"offset": offset, "values": [float(x) for x in O.mt_doubles(3141, offset, 6)]})
    wl = W.c0()
    t = O.ModelTables(wl)
    om = O.OracleModel(t, PATHS)
    out["brownian"] = [{"t": ti, "k": k, "paths": SAMPLE_PATHS,
                        "values": [float(x) for x in om.brownian_increment(ti, k)[SAMPLE_PATHS]]}
                       for ti, k in [(0, 0), (3, 2), (19, 1)]]
    out["process"] = [{"t": ti, "j": j, "paths": SAMPLE_PATHS,
                       "values": [float(x) for x in om.process_value(ti, j)[SAMPLE_PATHS]]}
                      for ti, j in PROCESS_CELLS]
    out["numeraire"] = [{"T": float(T), "paths": SAMPLE_PATHS,
                         "values": [float(x) for x in om.numeraire(float(T))[SAMPLE_PATHS]]}
                        for T in (0.5, 5.0, 10.0)]
    prods = []
    for s in wl.products:
        v, e = om.value(s)
        prods.append({"spec": {k: (list(map(float, x)) if isinstance(x, (list, tuple, np.ndarray)) else x)
                               for k, x in s.items()}, "value": v, "error": e})
    out["products"] = prods
    om.close()
    path = os.path.join(HERE, "lmm_golden.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()

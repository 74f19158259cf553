"""Workload definitions of the LMM benchmarks (shapes, market data, products).

c0: 20 semiannual forwards, 3 factors, flat 3 % curve, 10,000 paths, seed 3141:
    5y-into-5y payer swaption at 3 % and a 10y payer swap at 3 %.
c1: 40 semiannual forwards, 5 factors, f_i = 0.02 + 0.001 i, 1,000,000 paths:
    payer swaptions with expiries 1..10y into 10y swaps at ATM and +-100bp.
c2: the c1 model with 4,000,000 paths in FP32-mixed arithmetic.
c3: the c1 model with 1,000,000 paths, adjoint Greeks of the c1 strip.
c4: 80 forwards, 10 factors, flat 3.5 %, FP32-mixed, Bermudan payer swaption 10y no-call,
    then callable annually into the swap to 30y at 3.5 %, regression basis 1, S, S^2, N.

Volatility sigma_j(t) = (a + b (T_j - t)) exp(-c (T_j - t)) + d with
(a, b, c, d) = (0.2, 0.0, 0.25, 0.1); correlation exp(-beta |T_i - T_j|).
Seed 3141 is the seed of the reference's LMM valuation tests
(src/test/java/net/finmath/montecarlo/interestrate/LIBORMarketModelValuationTest.java:156).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Workload:
    name: str
    num_libors: int
    period: float
    num_factors: int
    num_paths: int
    corr_decay: float
    forwards: np.ndarray                 # forward curve node values at T_0..T_{n-1}
    vol_params: tuple = (0.2, 0.0, 0.25, 0.1)
    seed: int = 3141
    precision: str = "FP64"
    products: list = field(default_factory=list)   # list of dicts (type, ...)

    @property
    def num_time_steps(self) -> int:
        return self.num_libors

    def times(self) -> np.ndarray:
        return round_ticks(np.arange(self.num_time_steps + 1) * self.period)

    def tenor(self) -> np.ndarray:
        return round_ticks(np.arange(self.num_libors + 1) * self.period)

    def with_paths(self, num_paths: int, name: str | None = None) -> "Workload":
        w = Workload(**{**self.__dict__})
        w.num_paths = int(num_paths)
        if name:
            w.name = name
        return w

    @property
    def path_steps(self) -> int:
        return self.num_paths * self.num_time_steps


def round_ticks(t) -> np.ndarray:
    """TimeDiscretizationFromArray rounding to the 1/8760 tick (time/TimeDiscretizationFromArray.java:387-389)."""
    tick = 1.0 / (365.0 * 24.0)
    return np.rint(np.asarray(t, dtype=np.float64) / tick) * tick


def _curve_forward(node_t, node_f, t):
    """Linear interpolation of forwards with constant extrapolation, node values exact."""
    node_t = np.asarray(node_t)
    idx = np.searchsorted(node_t, t)
    if idx < len(node_t) and node_t[idx] == t:
        return float(node_f[idx])
    if idx == 0:
        return float(node_f[0])
    if idx >= len(node_t):
        return float(node_f[-1])
    i = idx - 1
    slope = (node_f[i + 1] - node_f[i]) / (node_t[i + 1] - node_t[i])
    return float(node_f[i] * 1.0 + slope * (t - node_t[i]))


def discount_factor(node_t, node_f, payment_offset, maturity) -> float:
    """DiscountCurveFromForwardCurve.getDiscountFactor (marketdata/model/curves/DiscountCurveFromForwardCurve.java:115-143)."""
    time, df = 0.0, 1.0
    while time < maturity:
        df /= 1.0 + _curve_forward(node_t, node_f, time) * min(payment_offset, maturity - time) * 1.0
        time += payment_offset
    return df


def forward_swap_rate(node_t, node_f, period, start, end) -> float:
    """Single-curve par swap rate (marketdata/products/Swap.java:146-174)."""
    fix = np.arange(start, end - 1e-12, period)
    pay = fix + period
    ev = fix[0]
    p_ev = discount_factor(node_t, node_f, period, ev)
    annuity = sum(period * discount_factor(node_t, node_f, period, p) for p in pay) / p_ev
    floatleg = sum(_curve_forward(node_t, node_f, f) * period * discount_factor(node_t, node_f, period, p)
                   for f, p in zip(fix, pay)) / p_ev
    return floatleg / annuity


def swaption_spec(expiry, length, period, strike):
    fix = round_ticks(np.arange(expiry, expiry + length - 1e-12, period))
    pay = round_ticks(fix + period)
    return {"type": "swaption", "exercise": float(round_ticks(expiry)), "fixing": fix, "payment": pay,
            "strike": np.full(fix.size, strike)}


def swap_spec(start, end, period, strike, pay_fix=True):
    fix = round_ticks(np.arange(start, end - 1e-12, period))
    pay = round_ticks(fix + period)
    return {"type": "swap", "fixing": fix, "payment": pay, "strike": np.full(fix.size, strike), "pay_fix": pay_fix}


def bermudan_spec(first_exercise, end, period, strike, exercise_every=1.0, callable_=True, basis=0):
    """Bermudan payer swaption: swap periods from first_exercise to end, exercisable at the
    period starts first_exercise, first_exercise + exercise_every, ... (before the last period)."""
    fix = round_ticks(np.arange(first_exercise, end - 1e-12, period))
    pay = round_ticks(fix + period)
    steps = int(round(exercise_every / period))
    flags = np.array([1 if (i % steps == 0) else 0 for i in range(fix.size)], dtype=np.int32)
    return {"type": "bermudan", "fixing": fix, "payment": pay, "period_lengths": pay - fix,
            "strike": np.full(fix.size, strike), "exercise_flags": flags, "callable": callable_, "basis": basis}


def c0() -> Workload:
    n = 20
    w = Workload("c0", n, 0.5, 3, 10_000, 0.1, np.full(n, 0.03))
    w.products = [swaption_spec(5.0, 5.0, 0.5, 0.03), swap_spec(0.0, 10.0, 0.5, 0.03)]
    return w


def swaption_strip(w: Workload, expiries=range(1, 11), length=10.0, shifts=(-0.01, 0.0, 0.01)):
    node_t = w.tenor()[:-1]
    out = []
    for e in expiries:
        atm = forward_swap_rate(node_t, w.forwards, w.period, float(e), float(e) + length)
        for s in shifts:
            out.append(swaption_spec(float(e), length, w.period, atm + s))
    return out


def c1() -> Workload:
    n = 40
    w = Workload("c1", n, 0.5, 5, 1_000_000, 0.05, 0.02 + 0.001 * np.arange(n))
    w.products = swaption_strip(w)
    return w


def c2() -> Workload:
    w = c1().with_paths(4_000_000, "c2")
    w.precision = "FP32_MIXED"
    return w


def c3() -> Workload:
    w = c1().with_paths(1_000_000, "c3")
    w.precision = "FP32_MIXED"
    return w


def c4() -> Workload:
    n = 80
    w = Workload("c4", n, 0.5, 10, 1_000_000, 0.05, np.full(n, 0.035))
    w.precision = "FP32_MIXED"
    # 10y no-call, then callable annually into the swap to 30y, strike 3.5 %, basis 1, S, S^2, N
    w.products = [bermudan_spec(10.0, 30.0, 0.5, 0.035, 1.0, True, basis=1)]
    return w


ALL = {"c0": c0, "c1": c1, "c2": c2, "c3": c3, "c4": c4}

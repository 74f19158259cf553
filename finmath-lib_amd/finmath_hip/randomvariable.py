"""Device random variables with the reference's RandomVariable interface.

`RandomVariable` follows net.finmath.stochastic.RandomVariable as implemented by
RandomVariableFromDoubleArray (montecarlo/RandomVariableFromDoubleArray.java): immutable,
element-wise operations, compensated averages. Operations are recorded lazily, as
RandomVariableLazyEvaluation (montecarlo/RandomVariableLazyEvaluation.java) does, and
compiled into one register program of the C ABI (include/finmath_rv.h) when a value is
needed: materialising an expression, or averaging it without materialising, is one pass
over device memory. Constants behave like RandomVariableFromDoubleArray's deterministic
values (not like Scalar, whose discount formula differs).
"""
from __future__ import annotations

import ctypes
import math
from ctypes import POINTER, c_double, c_int32, c_int64, c_void_p

import numpy as np

from . import _abi

OP = dict(LOAD=0, CONST=1, ADD=2, SUB=3, MULT=4, DIV=5, CAP=6, FLOOR=7, POW=8, EXP=9, LOG=10, SQRT=11, ABS=12,
          INVERT=13, SQUARED=14, ADD_PRODUCT=15, ADD_RATIO=16, SUB_RATIO=17, ACCRUE=18, DISCOUNT=19, CHOOSE=20,
          ISNAN=21, EXPM1=22, SIN=23, COS=24)
MAX_INSTRUCTIONS, MAX_INPUTS, REGISTERS = 64, 16, 16


class _Handle:
    """Owns one fm_rv* of the library."""

    def __init__(self, ptr: int):
        self.ptr = ptr

    def __del__(self):
        if self.ptr:
            try:
                _abi.load().fm_rv_release(c_void_p(self.ptr))
            except Exception:
                pass
            self.ptr = None


class _Leaf:
    __slots__ = ("handle", "time", "det", "value", "size")

    def __init__(self, handle, time, det, value, size):
        self.handle, self.time, self.det, self.value, self.size = handle, time, det, value, size


class _Op:
    __slots__ = ("op", "args", "imm", "time", "size", "det")

    def __init__(self, op, args, imm=0.0):
        self.op, self.args, self.imm = op, tuple(args), float(imm)
        self.time = max(a.time for a in self.args)
        sizes = {a.size for a in self.args if not a.det}
        if len(sizes) > 1:
            raise _abi.IllegalArgumentException(f"random variables of different sizes {sorted(sizes)}")
        self.size = sizes.pop() if sizes else 1
        self.det = all(a.det for a in self.args)


def _leaf_from_ptr(ptr: int) -> _Leaf:
    lib = _abi.load()
    t, v = c_double(), c_double()
    n, d = c_int64(), c_int32()
    _abi.check(lib.fm_rv_info(c_void_p(ptr), ctypes.byref(t), ctypes.byref(n), ctypes.byref(d), ctypes.byref(v)))
    return _Leaf(_Handle(ptr), t.value, bool(d.value), v.value, n.value)


class _Overflow(Exception):
    pass


def _compile(root):
    """Register program for the DAG under `root`: (inputs, instructions, result register).
    Shared sub-expressions are computed once; a register is freed after its last use."""
    order, seen = [], set()
    stack = [(root, False)]                 # iterative post-order (expressions can be deep)
    while stack:
        node, expanded = stack.pop()
        if expanded:
            order.append(node)
            continue
        if id(node) in seen:
            continue
        seen.add(id(node))
        stack.append((node, True))
        if isinstance(node, _Op):
            for a in reversed(node.args):
                if id(a) not in seen:
                    stack.append((a, False))
    uses = {}
    for node in order:
        if isinstance(node, _Op):
            for a in node.args:
                uses[id(a)] = uses.get(id(a), 0) + 1
    inputs, input_of, reg_of, free = [], {}, {}, list(range(REGISTERS - 1, -1, -1))
    prog = []
    for node in order:
        if isinstance(node, _Leaf):
            if node.det:
                ins = (OP["CONST"], 0, 0, 0, node.value)
            else:
                key = id(node.handle)
                if key not in input_of:
                    if len(inputs) == MAX_INPUTS:
                        raise _Overflow()
                    input_of[key] = len(inputs)
                    inputs.append(node.handle)
                ins = (OP["LOAD"], input_of[key], 0, 0, 0.0)
        else:
            regs = [reg_of[id(a)] for a in node.args] + [0] * (3 - len(node.args))
            ins = (node.op, regs[0], regs[1], regs[2], node.imm)
            for a in node.args:                       # free the registers whose last use this is
                uses[id(a)] -= 1
                if uses[id(a)] == 0 and a is not root:
                    free.append(reg_of[id(a)])
        if not free:
            raise _Overflow()
        dst = free.pop()
        reg_of[id(node)] = dst
        prog.append((ins[0], dst, ins[1], ins[2], ins[3], ins[4]))
        if len(prog) > MAX_INSTRUCTIONS:
            raise _Overflow()
    return inputs, prog, reg_of[id(root)]


def _program_args(inputs, prog):
    arr = (c_void_p * max(len(inputs), 1))(*[h.ptr for h in inputs])
    ins = (_abi.FmRvInstr * len(prog))()
    for k, (op, dst, a, b, c, imm) in enumerate(prog):
        ins[k].op, ins[k].dst, ins[k].a, ins[k].b, ins[k].c, ins[k].imm = op, dst, a, b, c, imm
    return arr, ins


def _subtree_size(node):
    count, stack, seen = 0, [node], set()
    while stack:
        n = stack.pop()
        if id(n) in seen:
            continue
        seen.add(id(n))
        count += 1
        if isinstance(n, _Op):
            stack.extend(n.args)
    return count


def _materialize(node) -> _Leaf:
    if isinstance(node, _Leaf):
        return node
    while True:
        try:
            inputs, prog, res = _compile(node)
            break
        except _Overflow:
            # evaluate the largest operand on its own first, then retry
            big = max((a for a in node.args if isinstance(a, _Op)), key=_subtree_size)
            node = _Op(node.op, [(_materialize(a) if a is big else a) for a in node.args], node.imm)
    arr, ins = _program_args(inputs, prog)
    out = c_void_p()
    _abi.check(_abi.load().fm_rv_evaluate(len(inputs), arr, len(prog), ins, res, node.time, ctypes.byref(out)))
    return _leaf_from_ptr(out.value)


def _as_node(x):
    if isinstance(x, RandomVariable):
        return x._node
    return _Leaf(None, 0.0, True, float(x), 1)


class RandomVariable:
    """A random variable on the device (RandomVariableFromDoubleArray semantics)."""

    __slots__ = ("_node",)

    def __init__(self, node):
        self._node = node

    # ---- construction (RandomVariableFactory.createRandomVariable, RandomVariableFactory.java:22-95)
    @staticmethod
    def from_array(values, time: float = 0.0) -> "RandomVariable":
        v = np.ascontiguousarray(np.asarray(values, dtype=np.float64))
        out = c_void_p()
        _abi.check(_abi.load().fm_rv_from_host(time, v.size, v.ctypes.data_as(POINTER(c_double)), ctypes.byref(out)))
        return RandomVariable(_leaf_from_ptr(out.value))

    @staticmethod
    def constant(value: float, time: float = 0.0) -> "RandomVariable":
        return RandomVariable(_Leaf(None, time, True, float(value), 1))

    @staticmethod
    def _from_ptr(ptr: int) -> "RandomVariable":
        return RandomVariable(_leaf_from_ptr(ptr))

    # ---- properties
    def getFiltrationTime(self) -> float:
        return self._node.time

    def isDeterministic(self) -> bool:
        return self._node.det

    def size(self) -> int:
        return self._node.size

    def _leaf(self) -> _Leaf:
        if not isinstance(self._node, _Leaf):
            self._node = _materialize(self._node)
        return self._node

    def cache(self) -> "RandomVariable":
        self._leaf()
        return self

    def doubleValue(self) -> float:
        leaf = self._leaf()
        if leaf.det:
            return leaf.value
        raise _abi.UnsupportedOperationException("doubleValue of a stochastic random variable")

    def getRealizations(self, size: int | None = None) -> np.ndarray:
        leaf = self._leaf()
        n = leaf.size if size is None else int(size)
        out = np.empty(n)
        if leaf.det:
            out[:] = leaf.value
            return out
        _abi.check(_abi.load().fm_rv_to_host(c_void_p(leaf.handle.ptr), n, out.ctypes.data_as(POINTER(c_double))))
        return out

    def get(self, i: int) -> float:
        leaf = self._leaf()
        return leaf.value if leaf.det else float(self.getRealizations()[i])

    def devicePointer(self) -> int:
        leaf = self._leaf()
        return 0 if leaf.det else leaf.handle.ptr

    # ---- reductions (RandomVariableFromDoubleArray.java:287-428)
    def _reduce(self, probabilities, want_variance):
        node = self._node
        if isinstance(node, _Op) and node.det:
            node = self._leaf()
        if isinstance(node, _Leaf) and node.det and probabilities is None:
            return node.value, 0.0
        while True:
            try:
                inputs, prog, res = _compile(node)
                break
            except _Overflow:
                node = self._leaf()
        prob = None
        if probabilities is not None:
            prob = probabilities if isinstance(probabilities, RandomVariable) else RandomVariable.constant(probabilities)
            pl = prob._leaf()
            if pl.det:
                cp = c_void_p()
                _abi.check(_abi.load().fm_rv_constant(pl.time, pl.value, ctypes.byref(cp)))
                prob_handle = _Handle(cp.value)
            else:
                prob_handle = pl.handle
        arr, ins = _program_args(inputs, prog)
        avg, var = c_double(), c_double()
        _abi.check(_abi.load().fm_rv_evaluate_average(
            len(inputs), arr, len(prog), ins, res, c_void_p(prob_handle.ptr) if prob is not None else None,
            ctypes.byref(avg), ctypes.byref(var) if want_variance else None))
        return avg.value, var.value

    def getAverage(self, probabilities=None) -> float:
        return self._reduce(probabilities, False)[0]

    def getVariance(self, probabilities=None) -> float:
        if self.isDeterministic():
            return 0.0
        return self._reduce(probabilities, True)[1]

    def getStandardDeviation(self, probabilities=None) -> float:
        return 0.0 if self.isDeterministic() else math.sqrt(self.getVariance(probabilities))

    def getStandardError(self) -> float:
        if self.isDeterministic():
            return 0.0
        return self.getStandardDeviation() / math.sqrt(self.size())

    def average(self) -> "RandomVariable":
        return RandomVariable.constant(self.getAverage())

    # ---- element operations
    def _op(self, name, *args, imm=0.0):
        return RandomVariable(_Op(OP[name], [self._node] + [_as_node(a) for a in args], imm))

    def add(self, x): return self._op("ADD", x)
    def sub(self, x): return self._op("SUB", x)
    def bus(self, x): return RandomVariable(_Op(OP["SUB"], [_as_node(x), self._node]))
    def mult(self, x): return self._op("MULT", x)
    def div(self, x): return self._op("DIV", x)
    def vid(self, x): return RandomVariable(_Op(OP["DIV"], [_as_node(x), self._node]))
    def cap(self, x): return self._op("CAP", x)
    def floor(self, x): return self._op("FLOOR", x)
    def pow(self, exponent): return self._op("POW", exponent)
    def exp(self): return self._op("EXP")
    def expm1(self): return self._op("EXPM1")
    def log(self): return self._op("LOG")
    def sqrt(self): return self._op("SQRT")
    def abs(self): return self._op("ABS")
    def invert(self): return self._op("INVERT")
    def squared(self): return self._op("SQUARED")
    def sin(self): return self._op("SIN")
    def cos(self): return self._op("COS")
    def isNaN(self): return self._op("ISNAN")

    def accrue(self, rate, periodLength: float):
        """this * (1 + rate * periodLength) (RandomVariableFromDoubleArray.java:1279-1305)."""
        return self._op("ACCRUE", rate, imm=periodLength)

    def discount(self, rate, periodLength: float):
        """this / (1.0 + rate * periodLength) (RandomVariableFromDoubleArray.java:1308-1333)."""
        return self._op("DISCOUNT", rate, imm=periodLength)

    def choose(self, valueIfTriggerNonNegative, valueIfTriggerNegative):
        """this >= 0 ? a : b (RandomVariableFromDoubleArray.java:1342-1363)."""
        return self._op("CHOOSE", valueIfTriggerNonNegative, valueIfTriggerNegative)

    def addProduct(self, factor1, factor2):
        return self._op("ADD_PRODUCT", factor1, factor2)

    def addSumProduct(self, factor1, factor2):
        """Sequential addProduct over the lists (RandomVariableFromDoubleArray.java:1430-1437)."""
        r = self
        for f1, f2 in zip(factor1, factor2):
            r = r.addProduct(f1, f2)
        return r

    def addRatio(self, numerator, denominator): return self._op("ADD_RATIO", numerator, denominator)
    def subRatio(self, numerator, denominator): return self._op("SUB_RATIO", numerator, denominator)

    __add__ = add
    __sub__ = sub
    __mul__ = mult
    __truediv__ = div

    def __radd__(self, x): return RandomVariable(_Op(OP["ADD"], [_as_node(x), self._node]))
    def __rsub__(self, x): return self.bus(x)
    def __rmul__(self, x): return RandomVariable(_Op(OP["MULT"], [_as_node(x), self._node]))
    def __rtruediv__(self, x): return self.vid(x)
    def __neg__(self): return self.mult(-1.0)


def process_value(ctx, time_index: int, component: int) -> RandomVariable:
    """getProcessValue(timeIndex, component) of a simulated context (fm_lmm_process_value_rv)."""
    out = c_void_p()
    _abi.check(_abi.load().fm_lmm_process_value_rv(ctx.h, int(time_index), int(component), ctypes.byref(out)))
    return RandomVariable._from_ptr(out.value)


def numeraire(ctx, time: float) -> RandomVariable:
    """getNumeraire(time) of a simulated context (fm_lmm_numeraire_rv)."""
    out = c_void_p()
    _abi.check(_abi.load().fm_lmm_numeraire_rv(ctx.h, float(time), ctypes.byref(out)))
    return RandomVariable._from_ptr(out.value)


def monte_carlo_weights(ctx) -> RandomVariable:
    out = c_void_p()
    _abi.check(_abi.load().fm_lmm_weights_rv(ctx.h, ctypes.byref(out)))
    return RandomVariable._from_ptr(out.value)


# ---- the model's getters on device random variables, with the off-grid rules ------------------

_TICK = 1.0 / (365.0 * 24.0)


def _round_tick(t: float) -> float:
    """TimeDiscretizationFromArray tick rounding (time/TimeDiscretizationFromArray.java:387-389)."""
    return float(np.rint(t / _TICK) * _TICK)


def _time_index(grid: np.ndarray, t: float) -> int:
    """TimeDiscretization.getTimeIndex: the index of t on the grid, else -(insertion point) - 1."""
    t = _round_tick(t)
    k = int(np.searchsorted(grid, t))
    if k < grid.size and grid[k] == t:
        return k
    return -k - 1


class LIBORMonteCarloSimulation:
    """getLIBOR / getForwardRate / getNumeraire / getMonteCarloWeights of a simulated context
    (store_paths = 1) as device random variables, for any time: the rules of
    LIBORMarketModelFromCovarianceModel for forwards and numeraires off the tenor grid
    (LIBORMarketModelFromCovarianceModel.java:883-906,968-1015,1236-1381) composed from the
    context's process values, with the interpolation LOG_LINEAR_UNCORRECTED (the reference's
    default), LOG_LINEAR_CORRECTED or LINEAR and simulation times rounded ROUND_NEAREST.

    forward_curve: the model's forward curve (getForward(fixingTime, paymentOffset), as the host
    mirror's ForwardCurveInterpolation), needed for interpolated forwards. On-grid numeraires
    are the engine's own (fm_lmm_numeraire_rv); off the grid the numeraire is composed here, in
    RandomVariableFromDoubleArray arithmetic, so it can differ from a Scalar chain in the last
    place.
    """

    def __init__(self, ctx, forward_curve=None, interpolation: str = "LOG_LINEAR_UNCORRECTED"):
        if ctx.times is None or ctx.cfg is None:
            raise _abi.IllegalArgumentException("the context must be created with its time grids")
        if interpolation not in ("LOG_LINEAR_UNCORRECTED", "LOG_LINEAR_CORRECTED", "LINEAR"):
            raise _abi.UnsupportedOperationException(f"interpolation {interpolation} is not supported")
        self.ctx = ctx
        self.forward_curve = forward_curve
        self.interpolation = interpolation
        self.times = np.array([_round_tick(t) for t in ctx.times])
        self.tenor = np.array([_round_tick(t) for t in ctx.tenor])
        self.terminal = ctx.cfg.measure == _abi.MEASURE["TERMINAL"]
        self.lognormal = ctx.cfg.state_space == _abi.STATE_SPACE["LOGNORMAL"]
        if interpolation == "LOG_LINEAR_CORRECTED" and ctx.factor_loading is None:
            raise _abi.IllegalArgumentException("LOG_LINEAR_CORRECTED needs the context's factor loadings")
        if interpolation == "LOG_LINEAR_CORRECTED" and getattr(ctx, "_sv_driver", None) is not None:
            # the drift adjustment would integrate the per-path scaling, which stays on the device
            raise _abi.UnsupportedOperationException("LOG_LINEAR_CORRECTED under a stochastic volatility scaling")
        self._numeraires = {}
        self._adjustments = None

    # -- model primitives --
    def getLIBOR(self, time_index: int, libor_index: int) -> RandomVariable:
        return process_value(self.ctx, time_index, libor_index)

    def getMonteCarloWeights(self, time: float = 0.0) -> RandomVariable:
        return monte_carlo_weights(self.ctx)

    def _process_time_index(self, time: float) -> int:
        """Nearest simulation time index (LMM:1244-1252, ROUND_NEAREST)."""
        ti = _time_index(self.times, time)
        if ti < 0:
            ti = -ti - 2
            if time - self.times[ti] > self.times[ti + 1] - time:
                ti += 1
        return ti

    def _curve_forward(self, fixing: float, period: float) -> float:
        if self.forward_curve is None:
            raise _abi.IllegalArgumentException("interpolated forwards need the model's forward curve")
        fc = self.forward_curve
        return float(fc(fixing, period) if callable(fc) else fc.getForward(fixing, period))

    def _interpolation_drift_adjustment(self, e: int, j: int) -> RandomVariable:
        """getInterpolationDriftAdjustmentEvaluated (LMM:1432-1499): the trapezoid integral over
        the simulation times up to index e of |lambda_j(t_s)|^2 / (1 + L_j(t_s) dT)^2 (times
        L_j(t_s)^2 in the log-normal state space), L_j read at most at its fixing index."""
        lam = self.ctx.factor_loading
        lv = self.ctx.local_volatility
        dT = float(self.tenor[j + 1] - self.tenor[j])
        fix = _time_index(self.times, self.tenor[j])
        fix = -fix - 2 if fix < 0 else fix
        adj, prev = None, None
        for s in range(e + 1):
            if s >= lam.shape[0]:
                raise _abi.IllegalArgumentException("factor loading beyond the simulation grid")
            L = self.getLIBOR(min(s, fix), j)
            if lv is None:
                g = 0.0
                for x in lam[s, j]:
                    g = g + float(x) * float(x)
                g = RandomVariable.constant(g)
            else:
                alpha, b = lv
                f = L if alpha == 0.0 else L.sub(L.mult(alpha))
                f = f.add(float(b[s, j]))
                g = RandomVariable.constant(0.0)
                for x in lam[s, j]:
                    g = g.add(f.mult(float(x)).squared())
            g = g.div(L.mult(dT).add(1.0).squared())
            if self.lognormal:
                g = g.mult(L.squared())
            g = g.cache()
            if s > 0:
                step = 0.5 * float(self.times[s] - self.times[s - 1])
                term = g.add(prev).mult(step)
                adj = term if adj is None else adj.add(term)
                adj = adj.cache()
            prev = g
        return RandomVariable.constant(0.0) if adj is None else adj

    def _one_plus_interpolated(self, ti: int, start: float, j: int) -> RandomVariable:
        """getOnePlusInterpolatedLIBORDt (LMM:1324-1381)."""
        ts, te = self.tenor[j], self.tenor[j + 1]
        tdt = te - ts
        if ts < self.times[ti]:
            fi = _time_index(self.times, ts)
            if fi < 0:
                raise _abi.IllegalArgumentException("tenor discretization not part of time discretization")
            ti = min(ti, fi)
        long_ = self.getLIBOR(ti, j).mult(tdt).add(1.0)
        small = te - start
        alpha = small / tdt
        if self.interpolation == "LINEAR":
            interp = long_.mult(alpha).add(1 - alpha)
        elif self.interpolation == "LOG_LINEAR_UNCORRECTED":
            interp = long_.log().mult(alpha).exp()
        else:
            coef = 0.5 * small * (ts - start)
            a = self._interpolation_drift_adjustment(ti, j).mult(coef)
            interp = long_.log().mult(alpha).sub(a).exp()
        a_long = 1 + self._curve_forward(ts, tdt) * tdt
        a_short = 1 + self._curve_forward(start, small) * small
        if self.interpolation == "LINEAR":
            a_interp = a_long * alpha + (1 - alpha)
        else:
            a_interp = math.exp(math.log(a_long) * alpha)
        return interp.mult(a_short / a_interp)

    def getForwardRate(self, time: float, start: float, end: float) -> RandomVariable:
        """getForwardRate (LMM:1236-1305)."""
        si, ei = _time_index(self.tenor, start), _time_index(self.tenor, end)
        time = min(time, start)
        ti = self._process_time_index(time)
        if ei < 0:
            pe = (-ei - 1) - 1
            next_end = float(self.tenor[pe + 1])
            long_ = self.getForwardRate(time, start, next_end).mult(next_end - start).add(1.0)
            return long_.div(self._one_plus_interpolated(ti, end, pe)).sub(1.0).div(end - start)
        if si < 0:
            ps = (-si - 1) - 1
            next_start = float(self.tenor[ps + 1])
            if next_start > end:
                raise _abi.IllegalArgumentException("interpolation not possible")
            if next_start == end:
                return self._one_plus_interpolated(ti, start, ps).sub(1.0).div(end - start)
            long_ = self.getForwardRate(time, next_start, end).mult(end - next_start).add(1.0)
            return long_.mult(self._one_plus_interpolated(ti, start, ps)).sub(1.0).div(end - start)
        if si + 1 == ei:
            return self.getLIBOR(ti, si)
        acc = None
        for j in range(si, ei):
            d = float(self.tenor[j + 1] - self.tenor[j])
            L = self.getLIBOR(ti, j)
            acc = L.mult(d).add(1.0) if acc is None else acc.accrue(L, d)
        return acc.sub(1.0).div(end - start)

    # -- numeraire --
    def _unadjusted_at(self, k: int) -> RandomVariable:
        """getNumerairetUnAdjustedAtLIBORIndex (LMM:1017-1074) from the process values."""
        if k in self._numeraires:
            return self._numeraires[k]
        n = self.tenor.size - 1
        if self.terminal:
            ti = _time_index(self.times, self.tenor[k])
            ti = -ti - 1 if ti < 0 else ti
            N = RandomVariable.constant(1.0)
            for l in range(k, n):
                N = N.discount(self.getLIBOR(ti, l), float(self.tenor[l + 1] - self.tenor[l]))
        elif k == 0:
            N = RandomVariable.constant(1.0)
        else:
            ti = _time_index(self.times, self.tenor[k - 1])
            ti = -ti - 1 if ti < 0 else ti
            N = self._unadjusted_at(k - 1).accrue(self.getLIBOR(ti, k - 1), float(self.tenor[k] - self.tenor[k - 1]))
        N = N.cache()
        self._numeraires[k] = N
        return N

    def _unadjusted(self, time: float) -> RandomVariable:
        """getNumerairetUnAdjusted (LMM:968-1015)."""
        k = _time_index(self.tenor, time)
        if k >= 0:
            return self._unadjusted_at(k)
        upper = -k - 1
        if upper - 1 < 0:
            raise _abi.IllegalArgumentException(f"numeraire requested for time {time}: unsupported")
        if self.terminal:
            N = RandomVariable.constant(1.0)
            for l in range(upper, self.tenor.size - 1):
                ti = _time_index(self.times, min(time, self.tenor[l]))
                if ti < 0:
                    raise _abi.IllegalArgumentException("terminal numeraire off the simulation grid")
                N = N.discount(self.getLIBOR(ti, l), float(self.tenor[l + 1] - self.tenor[l]))
        else:
            N = self._unadjusted_at(upper)
        T_up = float(self.tenor[upper])
        return N.discount(self.getForwardRate(time, time, T_up), T_up - time)

    def _adjustment(self, time: float) -> float:
        """getNumeraireDefaultableZeroBondAsOfTimeZero (LMM:883-944): the discount curve's chain
        on the tenor grid (Scalar.discount), log-linear in between."""
        pd = self.ctx.discount_factors
        if self._adjustments is None:
            a = [float(pd[0])]
            for i in range(self.tenor.size - 1):
                ts = float(self.tenor[i + 1] - self.tenor[i])
                fr = (pd[i] / pd[i + 1] - 1.0) / ts
                a.append(1.0 / (fr * (ts / a[-1]) + 1.0 / a[-1]))
            self._adjustments = a
        k = _time_index(self.tenor, time)
        if k >= 0:
            return self._adjustments[k]
        prev = min(-k - 2, self.tenor.size - 2)
        tp, tn = float(self.tenor[prev]), float(self.tenor[prev + 1])
        ap, an = self._adjustments[prev], self._adjustments[prev + 1]
        return ap * math.pow(an / ap, (time - tp) / (tn - tp))

    def getNumeraire(self, time: float) -> RandomVariable:
        """getNumeraire (LMM:860-876): the engine's numeraire on the tenor grid, the composed
        one (adjusted by the discount curve, when the context has one) off it."""
        if _time_index(self.tenor, time) >= 0:
            return numeraire(self.ctx, time)
        N = self._unadjusted(time)
        if self.ctx.discount_factors is not None:
            zb = N.invert().mult(self._unadjusted(0.0)).getAverage()
            N = N.mult(zb).div(self._adjustment(time))
        return N.cache()

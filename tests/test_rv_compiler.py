"""The random-variable program compiler (finmath_hip.randomvariable) on the CPU: the
register programs it emits, run by a numpy register machine written here, equal the
direct evaluation of the expression graph; shared sub-expressions are computed once;
register and instruction limits hold. Nothing here touches the device."""
import numpy as np
import pytest

from finmath_hip import randomvariable as R

SIZE = 257


class FakeHandle:
    def __init__(self, values):
        self.values = values
        self.ptr = 0


def leaf(values):
    return R.RandomVariable(R._Leaf(FakeHandle(np.asarray(values, dtype=float)), 0.0, False, 0.0, len(values)))


def direct(node):
    """Evaluate the graph recursively with numpy (the element formulas of the ABI)."""
    if isinstance(node, R._Leaf):
        return np.full(SIZE, node.value) if node.det else node.handle.values
    a = [direct(x) for x in node.args] + [None] * (3 - len(node.args))
    return run_op(node.op, a[0], a[1], a[2], node.imm)


def run_op(op, a, b, c, imm):
    O = R.OP
    with np.errstate(all="ignore"):
        return {
            O["ADD"]: lambda: a + b, O["SUB"]: lambda: a - b, O["MULT"]: lambda: a * b, O["DIV"]: lambda: a / b,
            O["CAP"]: lambda: np.minimum(a, b), O["FLOOR"]: lambda: np.maximum(a, b), O["EXP"]: lambda: np.exp(a),
            O["LOG"]: lambda: np.log(a), O["SQRT"]: lambda: np.sqrt(a), O["ABS"]: lambda: np.abs(a),
            O["INVERT"]: lambda: 1.0 / a, O["SQUARED"]: lambda: a * a, O["ADD_PRODUCT"]: lambda: a + b * c,
            O["ADD_RATIO"]: lambda: a + b / c, O["SUB_RATIO"]: lambda: a - b / c,
            O["ACCRUE"]: lambda: a * (1 + b * imm), O["DISCOUNT"]: lambda: a / (1.0 + b * imm),
            O["CHOOSE"]: lambda: np.where(a >= 0.0, b, c), O["POW"]: lambda: np.power(a, b),
        }[op]()


def machine(inputs, prog, result):
    """The register machine of include/finmath_rv.h on numpy vectors."""
    reg = [np.zeros(SIZE) for _ in range(R.REGISTERS)]
    for op, dst, a, b, c, imm in prog:
        if op == R.OP["LOAD"]:
            v = inputs[a].values
        elif op == R.OP["CONST"]:
            v = np.full(SIZE, imm)
        else:
            v = run_op(op, reg[a], reg[b], reg[c], imm)
        reg[dst] = v
    return reg[result]


@pytest.fixture
def data():
    rng = np.random.default_rng(11)
    return [leaf(rng.uniform(0.5, 2.0, SIZE)) for _ in range(6)]


def check(rv):
    inputs, prog, res = R._compile(rv._node)
    assert len(prog) <= R.MAX_INSTRUCTIONS and len(inputs) <= R.MAX_INPUTS
    assert all(0 <= p[1] < R.REGISTERS for p in prog)
    got = machine(inputs, prog, res)
    want = direct(rv._node)
    assert np.array_equal(got, want)
    return prog


def test_swaption_like_expression(data):
    L = data
    v = R.RandomVariable.constant(0.0)
    for j in reversed(range(5)):
        payoff = L[j].sub(0.03).mult(0.5).mult(1.0)
        v = v.add(payoff).discount(L[j], 0.5).mult(1.0)
    v = v.floor(0.0).div(L[5]).mult(1e-3)
    check(v)


def test_shared_subexpressions_are_computed_once(data):
    x = data[0].exp()
    y = x.mult(x).add(x.sqrt()).choose(x, data[1])
    prog = check(y)
    assert sum(1 for p in prog if p[0] == R.OP["EXP"]) == 1


def test_every_operation(data):
    a, b, c = data[0], data[1], data[2]
    exprs = [a.add(b), a.sub(b), a.bus(b), a.mult(b), a.div(b), a.vid(b), a.cap(b), a.floor(1.0), a.exp(), a.log(),
             a.sqrt(), a.abs(), a.invert(), a.squared(), a.addProduct(b, c), a.addRatio(b, c), a.subRatio(b, c),
             a.accrue(b, 0.25), a.discount(b, 0.5), a.sub(1.0).choose(b, c), a.pow(1.5), 2.0 - a, 3.0 / a]
    for e in exprs:
        check(e)


def test_deep_chain_stays_within_the_register_file(data):
    # a chain needs a few registers however long it is (15 links: 52 instructions)
    x = data[0]
    for k in range(15):
        x = x.mult(1.0001).add(data[k % 6])
    prog = check(x)
    assert len({p[1] for p in prog}) <= 6 + 3         # the six reused inputs stay live, plus the chain


def test_limits_raise_overflow_for_the_caller_to_split():
    rng = np.random.default_rng(3)
    # more than 16 distinct inputs
    xs = [leaf(rng.uniform(1, 2, SIZE)) for _ in range(R.MAX_INPUTS + 1)]
    s = xs[0]
    for x in xs[1:]:
        s = s.add(x)
    with pytest.raises(R._Overflow):
        R._compile(s._node)
    # a balanced product tree of 32 computed values keeps more than 16 registers alive
    ys = [xs[k % 8].add(float(k)) for k in range(32)]
    while len(ys) > 1:
        ys = [ys[i].mult(ys[i + 1]) for i in range(0, len(ys), 2)]
    with pytest.raises(R._Overflow):
        R._compile(ys[0]._node)


def test_sizes_must_agree():
    a = leaf(np.ones(SIZE))
    b = R.RandomVariable(R._Leaf(FakeHandle(np.ones(5)), 0.0, False, 0.0, 5))
    with pytest.raises(ValueError):
        a.add(b)


def test_filtration_time_is_the_latest_operand_time():
    a = R.RandomVariable(R._Leaf(FakeHandle(np.ones(SIZE)), 1.5, False, 0.0, SIZE))
    b = R.RandomVariable(R._Leaf(FakeHandle(np.ones(SIZE)), 3.0, False, 0.0, SIZE))
    assert a.add(b).getFiltrationTime() == 3.0
    assert a.mult(2.0).getFiltrationTime() == 1.5

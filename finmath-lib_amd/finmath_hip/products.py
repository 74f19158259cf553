"""Interest-rate products valued on device random variables through the model interface of
`randomvariable.LIBORMonteCarloSimulation` (getForwardRate, getNumeraire, getMonteCarloWeights),
for any dates: on the tenor grid, between tenor dates, or between simulation times. Each product
follows its reference class's getValue (montecarlo/interestrate/products/): the payoff algebra
runs as fused element programs over the paths, the average and standard error as compensated
device reductions. Products on the tenor grid can also go through the fused kernel
(`_abi.Product`), which values them inside the path pass.
"""
from __future__ import annotations

from .randomvariable import RandomVariable


class AbstractMonteCarloProduct:
    """getValue(model) = getValue(0, model).getAverage(); getValues = {value, error}
    (montecarlo/AbstractMonteCarloProduct.java:82-104)."""

    def getValueRandomVariable(self, evaluationTime: float, model) -> RandomVariable:
        raise NotImplementedError

    def getValue(self, model) -> float:
        return self.getValueRandomVariable(0.0, model).getAverage()

    def getValues(self, model) -> dict:
        v = self.getValueRandomVariable(0.0, model)
        return {"value": v.getAverage(), "error": v.getStandardError()}


def _to_evaluation_time(values: RandomVariable, evaluationTime: float, model) -> RandomVariable:
    """values * N(t) / w(t): back from numeraire-relative values (the last lines of every product)."""
    return values.mult(model.getNumeraire(evaluationTime)).div(model.getMonteCarloWeights(evaluationTime))


class Bond(AbstractMonteCarloProduct):
    """Zero-coupon bond paying 1 at maturity (products/Bond.java:59-91)."""

    def __init__(self, maturity: float):
        self.maturity = float(maturity)

    def getValueRandomVariable(self, evaluationTime, model):
        if evaluationTime > self.maturity:
            return RandomVariable.constant(0.0)
        values = RandomVariable.constant(1.0).div(model.getNumeraire(self.maturity))
        values = values.mult(model.getMonteCarloWeights(self.maturity))
        return _to_evaluation_time(values, evaluationTime, model)


class SimpleSwap(AbstractMonteCarloProduct):
    """Swap paying (L(T_i; T_i, T_{i+1}) - K_i) (T_{i+1} - T_i) at T_{i+1}, payer when isPayFix
    (products/SimpleSwap.java:137-170)."""

    def __init__(self, fixingDates, paymentDates, swaprates, isPayFix: bool = True, notional=None):
        self.fixingDates = [float(x) for x in fixingDates]
        self.paymentDates = [float(x) for x in paymentDates]
        self.swaprates = [float(x) for x in swaprates]
        self.isPayFix = isPayFix
        self.notional = None if notional is None else [float(x) for x in notional]

    def getValueRandomVariable(self, evaluationTime, model):
        values = RandomVariable.constant(0.0)
        for i, (fix, pay, k) in enumerate(zip(self.fixingDates, self.paymentDates, self.swaprates)):
            if pay < evaluationTime:     # a period paying at the evaluation time still counts
                continue
            libor = model.getForwardRate(fix, fix, pay)
            payoff = libor.sub(k).mult(pay - fix)
            if self.notional is not None:
                payoff = payoff.mult(self.notional[i])
            if not self.isPayFix:
                payoff = payoff.mult(-1.0)
            payoff = payoff.div(model.getNumeraire(pay)).mult(model.getMonteCarloWeights(pay))
            values = values.add(payoff).cache()
        return _to_evaluation_time(values, evaluationTime, model)


class Caplet(AbstractMonteCarloProduct):
    """max(L - K, 0) dcf paid at maturity + periodLength, or the floorlet -min(L - K, 0) dcf
    (products/Caplet.java:115-140), value unit VALUE."""

    def __init__(self, maturity: float, periodLength: float, strike: float, daycountFraction=None,
                 isFloorlet: bool = False):
        self.maturity, self.periodLength, self.strike = float(maturity), float(periodLength), float(strike)
        self.daycountFraction = self.periodLength if daycountFraction is None else float(daycountFraction)
        self.isFloorlet = isFloorlet

    def getValueRandomVariable(self, evaluationTime, model):
        payment = self.maturity + self.periodLength
        values = model.getForwardRate(self.maturity, self.maturity, payment)
        if not self.isFloorlet:
            values = values.sub(self.strike).floor(0.0).mult(self.daycountFraction)
        else:
            values = values.sub(self.strike).cap(0.0).mult(-1.0 * self.daycountFraction)
        values = values.div(model.getNumeraire(payment)).mult(model.getMonteCarloWeights(payment))
        return _to_evaluation_time(values, evaluationTime, model)


class Swaption(AbstractMonteCarloProduct):
    """Payer swaption exercised at exerciseDate: the swap's periods valued backwards with the
    forwards at the exercise date, floored at 0 (products/Swaption.java:137-199).
    discountingAdjustment: per-period factor for a discount curve other than the forward curve's
    own (1 when the model discounts on its forward curve)."""

    def __init__(self, exerciseDate: float, fixingDates, paymentDates, swaprates, periodLengths=None,
                 notional: float = 1.0, discountingAdjustment=None):
        self.exerciseDate = float(exerciseDate)
        self.fixingDates = [float(x) for x in fixingDates]
        self.paymentDates = [float(x) for x in paymentDates]
        self.swaprates = [float(x) for x in swaprates]
        self.periodLengths = None if periodLengths is None else [float(x) for x in periodLengths]
        self.notional = float(notional)
        self.discountingAdjustment = discountingAdjustment

    def getValueRandomVariable(self, evaluationTime, model):
        value = RandomVariable.constant(0.0)
        for period in reversed(range(len(self.fixingDates))):
            fix, pay, k = self.fixingDates[period], self.paymentDates[period], self.swaprates[period]
            if pay <= evaluationTime:
                break
            pl = self.periodLengths[period] if self.periodLengths is not None else pay - fix
            libor = model.getForwardRate(self.exerciseDate, fix, pay)
            payoff = libor.sub(k).mult(pl).mult(self.notional)
            discounting_date = max(fix, self.exerciseDate)
            adj = 1.0 if self.discountingAdjustment is None else float(self.discountingAdjustment[period])
            value = value.add(payoff)
            value = value.discount(libor, pay - discounting_date).mult(adj).cache()
        values = value.floor(0.0)
        values = values.div(model.getNumeraire(self.exerciseDate)).mult(model.getMonteCarloWeights(self.exerciseDate))
        return _to_evaluation_time(values, evaluationTime, model)

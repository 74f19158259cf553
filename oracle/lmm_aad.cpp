// Tape automatic differentiation (reverse mode) of the LMM pricing, for the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY. Only tests/ may load this (through liblmm_oracle.so); the
// product path never links, loads or calls anything in this directory.
//
// A restatement of the reference's operator-tree AAD,
// RandomVariableDifferentiableAAD (montecarlo/automaticdifferentiation/backward/
// RandomVariableDifferentiableAAD.java, "AAD" below; paths relative to
// src/main/java/net/finmath/ of the reference):
//   - every operation on a differentiable random variable appends a node holding the
//     operator, its differentiable arguments and the argument values the partial
//     derivatives need (AAD:66-71 operator set, :97-163 node);
//   - getGradient seeds the result with 1 and walks the nodes in descending id order,
//     adding derivative * partial to each argument (AAD:165-219, :659-704);
//   - the partial derivatives of each operator are those of AAD:230-452, evaluated on
//     the retained argument values with the RandomVariable arithmetic of orc_rv.hpp;
//   - a path average (getAverage) is a double, so it is a constant of the tape; the
//     conditional expectation is self-adjoint: the argument's derivative is the
//     conditional expectation of the node's derivative under the same estimator
//     (AAD:188-192); its basis functions are not differentiated;
//   - choose (the indicator of the Bermudan exercise) differentiates its trigger by the
//     factory's DiracDeltaApproximationMethod (RandomVariableDifferentiableAADFactory.java:60-67:
//     DISCRETE_DELTA by default, width 0.05 standard deviations): ONE gives Y - Z, ZERO
//     gives 0, DISCRETE_DELTA (Y - Z) 1{-eps/2 <= X < eps/2} / eps with
//     eps = width * sd(X) (AAD:404-424).
// The LMM restatement below follows lmm_oracle.cpp's plain one operation by operation
// (same values), with the model's leaves as differentiable inputs: the initial states
// Y_j(0) (log-forwards in the log-normal state space, LIBORMarketModelFromCovarianceModel
// .java:1081-1093), and either the volatility entries sigma_j(t_i) (the leaves of a
// LIBORVolatilityModelFromGivenMatrix) or the four parameters a, b, c, d of
// LIBORVolatilityModelFourParameterExponentialForm.getVolatility (:168-195:
// (a.addProduct(b, tau)).mult(c.mult(-tau).exp()).add(d)).
//
// Pins (tests/test_oracle_aad.py): each operator's derivative against its closed form at
// 1e-15, as RandomVariableDifferentiableAADTest.java:61-261 asserts; the whole tape
// against central finite differences of the plain oracle's prices (common random
// numbers, numeraire-adjustment averages frozen as the tape holds them).

#include "orc_rv.hpp"

#include <cstring>
#include <map>
#include <stdexcept>
#include <string>

namespace aad {

using orc::RV;

enum Op : int {
    LEAF = 0, ADD = 1, SUB = 2, MULT = 3, DIV = 4, EXP = 5, LOG = 6, INVERT = 7, CAP = 8, FLOOR = 9,
    ADDPRODUCT = 10, DISCOUNT = 11, ACCRUE = 12, CHOOSE = 13, CONDEXP = 14
};
// DiracDeltaApproximationMethod ordinals (RandomVariableDifferentiableAADFactory.java:33-39)
enum { DIRAC_ONE = 0, DIRAC_ZERO = 1, DIRAC_DISCRETE_DELTA = 2 };

// MonteCarloConditionalExpectationRegression over fixed basis values
// (montecarlo/conditionalexpectation/MonteCarloConditionalExpectationRegression.java:79-150):
// E[y | basis] = sum_i b_i x_i with x the pseudo-inverse solution of the averaged normal
// equations, as the plain oracle's Bermudan computes it.
struct CondExp {
    std::vector<RV> b;
    std::vector<double> XTX;
    int K = 0;
    void init() {
        K = (int)b.size();
        XTX.assign((size_t)K * K, 0.0);
        for (int i = 0; i < K; i++)
            for (int j = i; j < K; j++) {
                XTX[(size_t)i * K + j] = orc::average(orc::mult(b[(size_t)i], b[(size_t)j]));
                XTX[(size_t)j * K + i] = XTX[(size_t)i * K + j];
            }
    }
    RV apply(const RV& y, std::vector<double>* coef = nullptr) const {
        std::vector<double> XTy((size_t)K);
        for (int i = 0; i < K; i++) XTy[(size_t)i] = orc::average(orc::mult(y, b[(size_t)i]));
        const std::vector<double> x = orc::pinv_solve(XTX, XTy, K);
        if (coef) *coef = x;
        RV r = orc::mult(b[0], x[0]);
        for (int i = 1; i < K; i++) r = orc::addProduct(r, b[(size_t)i], x[(size_t)i]);
        return r;
    }
};

struct Node {
    int op = LEAF;
    int arg[3] = {-1, -1, -1};    // differentiable arguments (node ids), -1: none
    RV x[3];                      // argument values (X, Y, Z of AAD:236-238)
    int nargs = 0;
    int ce = -1;                  // conditional expectation estimator (CONDEXP)
};

struct Tape {
    std::vector<Node> nodes;
    std::vector<RV> value;
    std::vector<CondExp> ces;
    int dirac = DIRAC_DISCRETE_DELTA;
    double width = 0.05;
};

thread_local Tape* g_tape = nullptr;

// A random variable of the tape: its value and its node (-1: not differentiable).
struct AV {
    RV v;
    int id = -1;
};

struct Arg { int id; RV x; };

AV record(int op, RV v, std::initializer_list<Arg> args, int ce = -1) {
    bool any = false;
    for (const Arg& a : args) any = any || a.id >= 0;
    if (!any) return AV{std::move(v), -1};
    Node n;
    n.op = op;
    n.ce = ce;
    for (const Arg& a : args) { n.arg[n.nargs] = a.id; n.x[n.nargs] = a.x; n.nargs++; }
    g_tape->nodes.push_back(n);
    g_tape->value.push_back(v);
    return AV{std::move(v), (int)g_tape->nodes.size() - 1};
}

AV leaf(RV v) {
    g_tape->nodes.push_back(Node{});
    g_tape->value.push_back(v);
    return AV{std::move(v), (int)g_tape->nodes.size() - 1};
}
AV cst(RV v) { return AV{std::move(v), -1}; }
AV cst(double v) { return AV{orc::scalar(v), -1}; }
bool is_scalar_receiver(const AV& a) { return a.id < 0 && a.v.prio == 0; }

// Operations. A differentiable receiver records one node of its operator (AAD:
// type priority 3 above every other random variable); a Scalar receiver with a
// differentiable argument takes the Scalar's own route (stochastic/Scalar.java:275-357),
// which hands the operation to the argument.
AV add(const AV& a, double b) { return record(ADD, orc::add(a.v, b), {{a.id, a.v}, {-1, orc::scalar(b)}}); }
AV add(const AV& a, const AV& b) {
    if (is_scalar_receiver(a) && b.id >= 0) return add(b, a.v.val);                 // Scalar.java:275-278
    return record(ADD, orc::add(a.v, b.v), {{a.id, a.v}, {b.id, b.v}});
}
AV sub(const AV& a, double b) { return record(SUB, orc::sub(a.v, b), {{a.id, a.v}, {-1, orc::scalar(b)}}); }
AV sub(const AV& a, const AV& b) { return record(SUB, orc::sub(a.v, b.v), {{a.id, a.v}, {b.id, b.v}}); }
AV mult(const AV& a, double b) { return record(MULT, orc::mult(a.v, b), {{a.id, a.v}, {-1, orc::scalar(b)}}); }
AV mult(const AV& a, const AV& b) {
    if (is_scalar_receiver(a) && b.id >= 0) return mult(b, a.v.val);                // Scalar.java:290-293
    return record(MULT, orc::mult(a.v, b.v), {{a.id, a.v}, {b.id, b.v}});
}
AV div(const AV& a, const AV& b) { return record(DIV, orc::div(a.v, b.v), {{a.id, a.v}, {b.id, b.v}}); }
AV div(const AV& a, double b) { return record(DIV, orc::divd(a.v, b), {{a.id, a.v}, {-1, orc::scalar(b)}}); }
AV expv(const AV& a) { return record(EXP, orc::expv(a.v), {{a.id, a.v}}); }
AV logv(const AV& a) { return record(LOG, orc::logv(a.v), {{a.id, a.v}}); }
AV invert(const AV& a) { return record(INVERT, orc::invert(a.v), {{a.id, a.v}}); }
AV cap(const AV& a, double c) { return record(CAP, orc::cap(a.v, c), {{a.id, a.v}, {-1, orc::scalar(c)}}); }
AV floor_(const AV& a, double f) { return record(FLOOR, orc::floor_(a.v, f), {{a.id, a.v}, {-1, orc::scalar(f)}}); }
AV addProduct(const AV& a, const AV& f1, double f2) {
    if (is_scalar_receiver(a) && f1.id >= 0) return add(mult(f1, f2), a.v.val);     // Scalar.java:349-352
    return record(ADDPRODUCT, orc::addProduct(a.v, f1.v, f2), {{a.id, a.v}, {f1.id, f1.v}, {-1, orc::scalar(f2)}});
}
AV addProduct(const AV& a, const AV& f1, const AV& f2) {
    if (is_scalar_receiver(a) && (f1.id >= 0 || f2.id >= 0)) return add(mult(f1, f2), a.v.val);   // Scalar.java:354-357
    return record(ADDPRODUCT, orc::addProduct(a.v, f1.v, f2.v), {{a.id, a.v}, {f1.id, f1.v}, {f2.id, f2.v}});
}
AV discount(const AV& a, const AV& rate, double dt) {
    if (is_scalar_receiver(a) && rate.id >= 0) {                                   // Scalar.java:320-328
        if (a.v.val == 0) return mult(rate, 0.0);
        return invert(add(mult(rate, dt / a.v.val), 1.0 / a.v.val));
    }
    return record(DISCOUNT, orc::discount(a.v, rate.v, dt), {{a.id, a.v}, {rate.id, rate.v}, {-1, orc::scalar(dt)}});
}
AV accrue(const AV& a, const AV& rate, double dt) {
    if (is_scalar_receiver(a) && rate.id >= 0) return add(mult(rate, dt * a.v.val), a.v.val);   // Scalar.java:315-318
    return record(ACCRUE, orc::accrue(a.v, rate.v, dt), {{a.id, a.v}, {rate.id, rate.v}, {-1, orc::scalar(dt)}});
}
AV choose(const AV& trig, const AV& a, const AV& b) {
    return record(CHOOSE, orc::choose(trig.v, a.v, b.v), {{trig.id, trig.v}, {a.id, a.v}, {b.id, b.v}});
}
AV condexp(const AV& y, int ce) {
    return record(CONDEXP, g_tape->ces[(size_t)ce].apply(y.v), {{y.id, y.v}}, ce);
}

// ---- reverse sweep ----------------------------------------------------------------------
// Partial derivative of node n with respect to argument i (AAD:230-452).
RV partial(const Tape& t, const Node& n, int i) {
    using namespace orc;
    const RV& X = n.x[0];
    const RV& Y = n.x[1];
    const RV& Z = n.x[2];
    const RV one = scalar(1.0), zero = scalar(0.0);
    switch (n.op) {
    case ADD: return one;
    case SUB: return i == 0 ? one : scalar(-1.0);
    case MULT: return i == 0 ? Y : X;
    case DIV: return i == 0 ? invert(Y) : mult(div(X, mult(Y, Y)), -1.0);
    case EXP: return expv(X);
    case LOG: return invert(X);
    case INVERT: { RV v = invert(X); return mult(mult(v, v), -1.0); }
    case CAP: return i == 0 ? choose(sub(X, Y), zero, one) : choose(sub(X, Y), one, zero);
    case FLOOR: return i == 0 ? choose(sub(X, Y), one, zero) : choose(sub(X, Y), zero, one);
    case ADDPRODUCT: return i == 0 ? one : (i == 1 ? Z : Y);
    case DISCOUNT: {
        const RV den = add(mult(Y, Z), 1.0);
        if (i == 0) return invert(den);
        if (i == 1) return mult(div(mult(X, Z), mult(den, den)), -1.0);
        return mult(div(mult(X, Y), mult(den, den)), -1.0);
    }
    case ACCRUE: return i == 0 ? add(mult(Y, Z), 1.0) : (i == 1 ? mult(X, Z) : mult(X, Y));
    case CHOOSE:
        if (i == 0) {
            if (t.dirac == DIRAC_ONE) return sub(Y, Z);
            if (t.dirac == DIRAC_ZERO) return zero;
            const double eps = t.width * std::sqrt(variance(X));                   // getStandardDeviation
            if (std::isinf(eps)) return sub(Y, Z);
            if (!(eps > 0)) return zero;
            RV d = sub(Y, Z);
            d = mult(d, choose(add(X, eps / 2), one, zero));
            d = mult(d, choose(sub(X, eps / 2), zero, one));
            return divd(d, eps);
        }
        return i == 1 ? choose(X, one, zero) : choose(X, zero, one);
    case CONDEXP: return one;
    default: throw std::runtime_error("operator without derivative");
    }
}

// Derivatives of `result` with respect to every node (AAD:165-219 propagation, :659-704
// order): the seed is the Scalar 1; entries without a derivative stay empty.
std::vector<RV> gradient(const Tape& t, int result, std::vector<bool>& has) {
    using namespace orc;
    const size_t N = t.nodes.size();
    std::vector<RV> d(N);
    has.assign(N, false);
    d[(size_t)result] = scalar(1.0);
    has[(size_t)result] = true;
    for (int m = result; m >= 0; m--) {
        if (!has[(size_t)m]) continue;
        const Node& n = t.nodes[(size_t)m];
        for (int i = 0; i < n.nargs; i++) {
            const int a = n.arg[i];
            if (a < 0) continue;
            RV der = d[(size_t)m];
            if (n.op == CONDEXP) der = t.ces[(size_t)n.ce].apply(der);             // AAD:188-192
            const RV p = partial(t, n, i);
            if (!has[(size_t)a]) { d[(size_t)a] = mult(der, p); has[(size_t)a] = true; }
            else d[(size_t)a] = addProduct(d[(size_t)a], p, der);
        }
        if (n.op != LEAF) d[(size_t)m] = RV();       // only leaves are kept (isGradientRetainsLeafNodesOnly)
    }
    return d;
}

double gradient_average(const std::vector<RV>& d, const std::vector<bool>& has, int id) {
    return (id >= 0 && has[(size_t)id]) ? orc::average(d[(size_t)id]) : 0.0;
}

// ---- the LMM on the tape ------------------------------------------------------------------
// Spot measure, EULER_FUNCTIONAL or EULER, either state space; the same operations as
// lmm_oracle.cpp's Model (LIBORMarketModelFromCovarianceModel.java:1125-1191,1194-1223,
// 1017-1074, 860-876; EulerSchemeFromProcessModel.java:170-326).
struct Model {
    int nT = 0, n = 0, F = 0, P = 0;
    std::vector<double> times, tenor, fac, pd;
    double capv = INFINITY;
    int ss = 1, scheme = 2;
    bool has_disc = false;
    std::vector<AV> y0;                          // leaves Y_j(0)
    std::vector<AV> abcd;                        // leaves a, b, c, d (four-parameter mode)
    std::vector<std::vector<AV>> sig;            // sigma_j(t_i) [nT][n]
    std::vector<std::vector<std::vector<AV>>> lam;   // lambda_jk(t_i), built on first use
    std::vector<std::vector<AV>> var;
    std::vector<std::vector<AV>> proc;           // [nT+1][n]
    std::vector<std::vector<RV>> dW;
    std::map<int, AV> numeraires;
    std::vector<double> adjchain;

    int libor_period_index(double t) const { return orc::java_binary_search(tenor.data(), n + 1, orc_round_time(t)); }
    int process_time_index(double t) const { return orc::java_binary_search(times.data(), nT + 1, orc_round_time(t)); }
    int first_alive(int ti) const {
        int f = libor_period_index(times[(size_t)ti]) + 1;
        if (f < 0) f = -f - 1 + 1;
        return f;
    }
    void tables(int i, int j) {
        if (!lam[(size_t)i][(size_t)j].empty()) return;
        for (int k = 0; k < F; k++) lam[(size_t)i][(size_t)j].push_back(mult(sig[(size_t)i][(size_t)j], fac[(size_t)j * F + k]));
        var[(size_t)i][(size_t)j] = mult(mult(sig[(size_t)i][(size_t)j], sig[(size_t)i][(size_t)j]), 1.0);
    }
    AV transform(const AV& y) const {
        AV v = y;
        if (ss == 1) v = expv(v);
        if (!std::isinf(capv)) v = cap(v, capv);
        return v;
    }

    std::vector<AV> drift(int ti, const std::vector<AV>& real) {
        std::vector<AV> d((size_t)n);
        const int first = first_alive(ti);
        const AV zero = cst(0.0);
        std::vector<AV> S((size_t)F, zero);
        for (int j = first; j < n; j++) {
            tables(ti, j);
            const double dl = tenor[(size_t)j + 1] - tenor[(size_t)j];
            AV w = discount(cst(dl), real[(size_t)j], dl);
            if (ss == 1) w = mult(w, real[(size_t)j]);
            for (int k = 0; k < F; k++) S[(size_t)k] = addProduct(S[(size_t)k], w, lam[(size_t)ti][(size_t)j][(size_t)k]);
            AV r = zero;
            for (int k = 0; k < F; k++) r = addProduct(r, S[(size_t)k], lam[(size_t)ti][(size_t)j][(size_t)k]);
            d[(size_t)j] = r;
        }
        if (ss == 1)
            for (int j = first; j < n; j++) d[(size_t)j] = addProduct(d[(size_t)j], var[(size_t)ti][(size_t)j], -0.5);
        return d;
    }

    void simulate(int32_t seed) {
        std::vector<double> bm((size_t)nT * F * P);
        orc_brownian(times.data(), nT, F, P, seed, bm.data());
        dW.assign((size_t)nT, std::vector<RV>((size_t)F));
        for (int t = 0; t < nT; t++)
            for (int k = 0; k < F; k++)
                dW[(size_t)t][(size_t)k] = orc::rvarr(std::vector<double>(bm.begin() + ((size_t)t * F + k) * P,
                                                                          bm.begin() + ((size_t)t * F + k + 1) * P));
        lam.assign((size_t)nT, std::vector<std::vector<AV>>((size_t)n));
        var.assign((size_t)nT, std::vector<AV>((size_t)n));
        proc.assign((size_t)nT + 1, std::vector<AV>((size_t)n));
        std::vector<AV> state((size_t)n);
        for (int j = 0; j < n; j++) {                                            // Euler:187-192
            state[(size_t)j] = y0[(size_t)j];
            proc[0][(size_t)j] = transform(state[(size_t)j]);
        }
        for (int ti = 1; ti <= nT; ti++) {                                       // Euler:202-290
            const double dt = times[(size_t)ti] - times[(size_t)ti - 1];
            const int first = first_alive(ti - 1);
            const std::vector<AV> d = drift(ti - 1, proc[(size_t)ti - 1]);
            for (int j = 0; j < n; j++) {
                if (j < first) { proc[(size_t)ti][(size_t)j] = proc[(size_t)ti - 1][(size_t)j]; continue; }
                AV cur = scheme == 2 ? (ss == 1 ? logv(proc[(size_t)ti - 1][(size_t)j]) : proc[(size_t)ti - 1][(size_t)j])
                                     : state[(size_t)j];
                cur = addProduct(cur, d[(size_t)j], dt);                                            // Euler:248
                for (int k = 0; k < F; k++)
                    cur = addProduct(cur, lam[(size_t)ti - 1][(size_t)j][(size_t)k], cst(dW[(size_t)ti - 1][(size_t)k]));   // :252
                state[(size_t)j] = cur;
                proc[(size_t)ti][(size_t)j] = transform(cur);                                       // :255
            }
        }
        if (has_disc) {                                                          // LMM:915-944
            adjchain.assign((size_t)n + 1, 0.0);
            RV a = orc::scalar(pd[0]);
            adjchain[0] = a.val;
            for (int i = 0; i < n; i++) {
                const double ts = tenor[(size_t)i + 1] - tenor[(size_t)i];
                a = orc::discount(a, orc::scalar((pd[(size_t)i] / pd[(size_t)i + 1] - 1.0) / ts), ts);
                adjchain[(size_t)i + 1] = a.val;
            }
        }
    }

    int time_index_rounded_up(double T) const {
        int ti = process_time_index(T);
        if (ti < 0) ti = -ti - 1;
        return ti;
    }
    AV numeraire_unadjusted_at(int k) {                                          // LMM:1017-1074 (spot)
        auto it = numeraires.find(k);
        if (it != numeraires.end()) return it->second;
        AV N = cst(1.0);
        if (k != 0) {
            const int ti = time_index_rounded_up(tenor[(size_t)k - 1]);
            N = accrue(numeraire_unadjusted_at(k - 1), proc[(size_t)ti][(size_t)k - 1], tenor[(size_t)k] - tenor[(size_t)k - 1]);
        }
        numeraires[k] = N;
        return N;
    }
    AV numeraire(double T) {                                                     // LMM:860-876, tenor grid
        const int k = libor_period_index(T);
        if (k < 0) throw std::runtime_error("tape oracle: numeraire off the tenor grid");
        AV N = numeraire_unadjusted_at(k);
        if (has_disc) {
            const double zb = orc::average(orc::mult(orc::invert(N.v), numeraire_unadjusted_at(0).v));   // a constant
            N = div(mult(N, zb), cst(orc::scalar(adjchain[(size_t)k])));
        }
        return N;
    }
    // getForwardRate on the grid (LMM:1237-1305): one period is the process value at the
    // time index of min(time, start); several accrue.
    AV forward_rate(double time, double start, double end) {
        const int si = libor_period_index(start), ei = libor_period_index(end);
        if (si < 0 || ei < 0) throw std::runtime_error("tape oracle: forward rate off the tenor grid");
        time = std::min(time, start);
        int ti = process_time_index(time);
        if (ti < 0) {
            ti = -ti - 2;
            if (time - times[(size_t)ti] > times[(size_t)ti + 1] - time) ti++;
        }
        if (si + 1 == ei) return proc[(size_t)ti][(size_t)si];
        AV acc;
        bool have = false;
        for (int pi = si; pi < ei; pi++) {
            const double sub_ = tenor[(size_t)pi + 1] - tenor[(size_t)pi];
            const AV l = proc[(size_t)ti][(size_t)pi];
            acc = have ? accrue(acc, l, sub_) : add(mult(l, sub_), 1.0);
            have = true;
        }
        return div(sub(acc, 1.0), end - start);
    }
    AV weights() const { return cst(1.0 / P); }
};

// Swaption.getValue (products/Swaption.java:137-199), evaluation time 0, the discount curve
// being the forward curve's own (adjustment 1 per period).
AV swaption(Model& m, double exercise, int nper, const double* fix, const double* pay, const double* K) {
    AV v = cst(0.0);
    for (int per = nper - 1; per >= 0; per--) {
        const double fd = fix[per], pd_ = pay[per];
        if (pd_ <= 0.0) break;
        const AV libor = m.forward_rate(exercise, fd, pd_);
        const AV payoff = mult(mult(sub(libor, K[per]), pd_ - fd), 1.0);
        v = add(v, payoff);
        v = mult(discount(v, libor, pd_ - std::max(fd, exercise)), 1.0);
    }
    AV vals = floor_(v, 0.0);
    vals = mult(div(vals, m.numeraire(exercise)), m.weights());
    vals = div(mult(vals, m.numeraire(0.0)), m.weights());
    return vals;
}

// BermudanSwaption.getValues (products/BermudanSwaption.java:91-157) with the regression
// basis of lmm_oracle.cpp (basis 0: the reference's default, 1: 1, S, S^2, N).
AV bermudan(Model& m, int nper, const int* isEx, const double* fix, const double* pl, const double* pay,
            const double* K, int callable, int basis) {
    using namespace orc;
    AV values = cst(0.0), under = cst(0.0);
    for (int per = nper - 1; per >= 0; per--) {
        const double fd = fix[per];
        const AV libor = m.forward_rate(fd, fd, fd + pl[per]);
        AV payoff = mult(mult(sub(libor, K[per]), pl[per]), 1.0);
        payoff = mult(div(payoff, m.numeraire(pay[per])), m.weights());
        if (callable) under = add(under, payoff);
        else values = add(values, payoff);
        if (isEx[per]) {
            // basis functions on the values only (the estimator does not differentiate them)
            CondExp ce;
            ce.b.push_back(rvdet(1.0));
            if (basis == 0) {
                int idx = java_binary_search(fix, nper, fd);
                if (idx < 0) idx = -idx;
                if (idx >= nper) idx = nper - 1;
                const RV rs = m.forward_rate(fd, fd, pay[idx]).v;
                const RV ds = invert(orc::add(orc::mult(rs, pay[idx] - fd), 1.0));
                ce.b.push_back(ds);
                ce.b.push_back(powv(ds, 2.0));
                const RV rl = m.forward_rate(fd, fix[idx], pay[nper - 1]).v;
                const RV dl = invert(orc::add(orc::mult(rl, pay[nper - 1] - fix[idx]), 1.0));
                ce.b.push_back(dl);
                ce.b.push_back(powv(dl, 2.0));
                ce.b.push_back(invert(m.numeraire(fd).v));
            } else {
                RV Pd, A;
                for (int p = per; p < nper; p++) {
                    const RV L = m.forward_rate(fd, fix[p], pay[p]).v;
                    if (p == per) { Pd = invert(orc::add(orc::mult(L, pl[p]), 1.0)); A = orc::mult(Pd, pl[p]); }
                    else { Pd = orc::div(Pd, orc::add(orc::mult(L, pl[p]), 1.0)); A = orc::add(A, orc::mult(Pd, pl[p])); }
                }
                const RV S = orc::div(orc::add(orc::mult(Pd, -1.0), 1.0), A);
                ce.b.push_back(S);
                ce.b.push_back(orc::mult(S, S));
                ce.b.push_back(m.numeraire(fd).v);
            }
            ce.init();
            g_tape->ces.push_back(std::move(ce));
            const AV trig = condexp(sub(values, under), (int)g_tape->ces.size() - 1);
            values = choose(trig, values, under);
        }
    }
    return div(mult(values, m.numeraire(0.0)), m.weights());
}

} // namespace aad

extern "C" {

typedef struct orc_aad_spec {
    int nT; const double* times;       // nT+1 simulation times
    int n; const double* tenor;        // n+1 tenor dates
    int F; int P; int32_t seed;
    const double* L0;                  // n initial forwards
    const double* factors;             // [n][F]
    double cap;
    int scheme;                        // EULER 0, EULER_FUNCTIONAL 2
    int state_space;                   // NORMAL 0, LOGNORMAL 1
    const double* pd;                  // n+1 discount-curve factors at the tenor dates, or NULL
    int vol_mode;                      // 0: the entries sigma_j(t_i) are the leaves, 1: a, b, c, d
    const double* vol;                 // [nT][n] (vol_mode 0)
    const double* abcd;                // 4 (vol_mode 1)
    const double* vol_time;            // [nT] the volatility model's time of each step (vol_mode 1)
    int dirac;                         // DiracDeltaApproximationMethod ordinal
    double dirac_width;
} orc_aad_spec;

static std::string g_aad_err;
const char* orc_aad_last_error(void) { return g_aad_err.c_str(); }

static aad::Model build(const orc_aad_spec* s) {
    using namespace aad;
    Model m;
    m.nT = s->nT; m.n = s->n; m.F = s->F; m.P = s->P;
    m.times.assign(s->times, s->times + s->nT + 1);
    m.tenor.assign(s->tenor, s->tenor + s->n + 1);
    m.fac.assign(s->factors, s->factors + (size_t)s->n * s->F);
    m.capv = s->cap; m.ss = s->state_space; m.scheme = s->scheme;
    if (m.scheme != 0 && m.scheme != 2) throw std::runtime_error("tape oracle: EULER or EULER_FUNCTIONAL only");
    m.has_disc = s->pd != nullptr;
    if (m.has_disc) m.pd.assign(s->pd, s->pd + s->n + 1);
    for (int j = 0; j < m.n; j++) {                                              // LMM:1081-1093
        const double y = m.ss == 1 ? std::log(std::max(s->L0[j], 0.0)) : s->L0[j];
        m.y0.push_back(leaf(orc::scalar(y)));
    }
    m.sig.assign((size_t)m.nT, std::vector<AV>((size_t)m.n));
    if (s->vol_mode == 0) {
        for (int i = 0; i < m.nT; i++)
            for (int j = 0; j < m.n; j++) m.sig[(size_t)i][(size_t)j] = leaf(orc::scalar(s->vol[(size_t)i * m.n + j]));
    } else {
        for (int q = 0; q < 4; q++) m.abcd.push_back(leaf(orc::scalar(s->abcd[q])));
        for (int i = 0; i < m.nT; i++)
            for (int j = 0; j < m.n; j++) {
                const double tau = m.tenor[(size_t)j] - s->vol_time[i];
                if (tau <= 0) { m.sig[(size_t)i][(size_t)j] = cst(0.0); continue; }
                // (a.addProduct(b, tau)).mult(c.mult(-tau).exp()).add(d)
                m.sig[(size_t)i][(size_t)j] =
                    add(mult(addProduct(m.abcd[0], m.abcd[1], tau), expv(mult(m.abcd[2], -tau))), m.abcd[3]);
            }
    }
    m.simulate(s->seed);
    return m;
}

static void outputs(const aad::Tape& t, const aad::Model& m, const aad::AV& result, const orc_aad_spec* s,
                    double* value, double* dY0, double* dvol, double* dabcd) {
    using namespace aad;
    *value = orc::average(result.v);
    std::vector<bool> has;
    const std::vector<RV> d = result.id >= 0 ? gradient(t, result.id, has) : std::vector<RV>();
    if (result.id < 0) has.assign(t.nodes.size(), false);
    if (dY0) for (int j = 0; j < m.n; j++) dY0[j] = gradient_average(d, has, m.y0[(size_t)j].id);
    if (dvol && s->vol_mode == 0)
        for (int i = 0; i < m.nT; i++)
            for (int j = 0; j < m.n; j++) dvol[(size_t)i * m.n + j] = gradient_average(d, has, m.sig[(size_t)i][(size_t)j].id);
    if (dabcd && s->vol_mode == 1) for (int q = 0; q < 4; q++) dabcd[q] = gradient_average(d, has, m.abcd[(size_t)q].id);
}

// Sum over products of weight_q * Swaption_q.getValue, and its gradient: dY0[n] (leaves
// Y_j(0)), dvol[nT][n] (vol_mode 0) or dabcd[4] (vol_mode 1), each the path average of the
// leaf's derivative. Swaption q has nper[q] periods: fix/pay/strike hold them back to back.
int orc_aad_swaptions(const orc_aad_spec* s, int nprod, const double* exercise, const int* nper, const double* fix,
                      const double* pay, const double* strike, const double* weights, double* value, double* dY0,
                      double* dvol, double* dabcd) {
    try {
        aad::Tape tape;
        tape.dirac = s->dirac; tape.width = s->dirac_width;
        aad::g_tape = &tape;
        aad::Model m = build(s);
        aad::AV total;
        int off = 0;
        for (int q = 0; q < nprod; q++) {
            aad::AV v = aad::mult(aad::swaption(m, exercise[q], nper[q], fix + off, pay + off, strike + off),
                                  weights ? weights[q] : 1.0);
            total = q == 0 ? v : aad::add(total, v);
            off += nper[q];
        }
        outputs(tape, m, total, s, value, dY0, dvol, dabcd);
        aad::g_tape = nullptr;
        return 0;
    } catch (const std::exception& e) { g_aad_err = e.what(); aad::g_tape = nullptr; return -1; }
}

// BermudanSwaption.getValues and its gradient (outputs as orc_aad_swaptions).
int orc_aad_bermudan(const orc_aad_spec* s, int nper, const int* isExercise, const double* fix,
                     const double* periodLength, const double* pay, const double* swaprates, int isCallable,
                     int basis, double* value, double* dY0, double* dvol, double* dabcd) {
    try {
        aad::Tape tape;
        tape.dirac = s->dirac; tape.width = s->dirac_width;
        aad::g_tape = &tape;
        aad::Model m = build(s);
        const aad::AV v = aad::bermudan(m, nper, isExercise, fix, periodLength, pay, swaprates, isCallable, basis);
        outputs(tape, m, v, s, value, dY0, dvol, dabcd);
        aad::g_tape = nullptr;
        return 0;
    } catch (const std::exception& e) { g_aad_err = e.what(); aad::g_tape = nullptr; return -1; }
}

// ---- a bare tape for the operator pins ----------------------------------------------------
// Handles index the tape's random variables; ops take handles and return a new one.
struct orc_aad_tape {
    aad::Tape tape;
    std::vector<aad::AV> h;
};

void* orc_aad_tape_new(int dirac, double width) {
    auto* t = new orc_aad_tape();
    t->tape.dirac = dirac;
    t->tape.width = width;
    return t;
}
void orc_aad_tape_free(void* p) { delete (orc_aad_tape*)p; }

static orc::RV rv_from(int size, const double* v) {
    if (size <= 1) return orc::rvdet(v[0]);
    return orc::rvarr(std::vector<double>(v, v + size));
}

// A differentiable input (leaf) or a constant random variable of `size` realisations
// (size 1: deterministic).
int orc_aad_input(void* p, int differentiable, int size, const double* values) {
    auto* t = (orc_aad_tape*)p;
    aad::g_tape = &t->tape;
    t->h.push_back(differentiable ? aad::leaf(rv_from(size, values)) : aad::cst(rv_from(size, values)));
    aad::g_tape = nullptr;
    return (int)t->h.size() - 1;
}

// Conditional expectation estimator over `K` basis functions of `size` realisations each
// (basis laid out back to back). Returns its index.
int orc_aad_estimator(void* p, int K, int size, const double* basis) {
    auto* t = (orc_aad_tape*)p;
    aad::CondExp ce;
    for (int i = 0; i < K; i++) ce.b.push_back(rv_from(size, basis + (size_t)i * size));
    ce.init();
    t->tape.ces.push_back(std::move(ce));
    return (int)t->tape.ces.size() - 1;
}

// op: the Op codes above; a, b, c handles (unused: -1); imm the double operand of the
// forms with one (ADD/SUB/MULT/DIV/CAP/FLOOR with b = -1, ADDPRODUCT with c = -1,
// DISCOUNT/ACCRUE: the period length), or the estimator index (CONDEXP).
int orc_aad_op(void* p, int op, int a, int b, int c, double imm) {
    auto* t = (orc_aad_tape*)p;
    aad::g_tape = &t->tape;
    const auto& H = t->h;
    aad::AV r;
    try {
        switch (op) {
        case aad::ADD: r = b >= 0 ? aad::add(H[a], H[b]) : aad::add(H[a], imm); break;
        case aad::SUB: r = b >= 0 ? aad::sub(H[a], H[b]) : aad::sub(H[a], imm); break;
        case aad::MULT: r = b >= 0 ? aad::mult(H[a], H[b]) : aad::mult(H[a], imm); break;
        case aad::DIV: r = b >= 0 ? aad::div(H[a], H[b]) : aad::div(H[a], imm); break;
        case aad::EXP: r = aad::expv(H[a]); break;
        case aad::LOG: r = aad::logv(H[a]); break;
        case aad::INVERT: r = aad::invert(H[a]); break;
        case aad::CAP: r = aad::cap(H[a], imm); break;
        case aad::FLOOR: r = aad::floor_(H[a], imm); break;
        case aad::ADDPRODUCT: r = c >= 0 ? aad::addProduct(H[a], H[b], H[c]) : aad::addProduct(H[a], H[b], imm); break;
        case aad::DISCOUNT: r = aad::discount(H[a], H[b], imm); break;
        case aad::ACCRUE: r = aad::accrue(H[a], H[b], imm); break;
        case aad::CHOOSE: r = aad::choose(H[a], H[b], H[c]); break;
        case aad::CONDEXP: r = aad::condexp(H[a], (int)imm); break;
        default: g_aad_err = "unknown operator"; aad::g_tape = nullptr; return -1;
        }
    } catch (const std::exception& e) { g_aad_err = e.what(); aad::g_tape = nullptr; return -1; }
    t->h.push_back(r);
    aad::g_tape = nullptr;
    return (int)t->h.size() - 1;
}

// Realisations of handle h (size of out: the tape's path count, deterministic values repeated).
int orc_aad_value(void* p, int h, int size, double* out) {
    auto* t = (orc_aad_tape*)p;
    const orc::RV& v = t->h[(size_t)h].v;
    for (int i = 0; i < size; i++) out[i] = v.get(v.det ? 0 : (size_t)i);
    return 0;
}

// Derivative of handle `result` with respect to the input handle `wrt`, per path (out of
// `size` entries; zero where the result does not depend on it).
int orc_aad_derivative(void* p, int result, int wrt, int size, double* out) {
    auto* t = (orc_aad_tape*)p;
    const int rid = t->h[(size_t)result].id, wid = t->h[(size_t)wrt].id;
    for (int i = 0; i < size; i++) out[i] = 0.0;
    if (rid < 0 || wid < 0) return 0;
    std::vector<bool> has;
    const std::vector<orc::RV> d = aad::gradient(t->tape, rid, has);
    if (!has[(size_t)wid]) return 0;
    const orc::RV& g = d[(size_t)wid];
    for (int i = 0; i < size; i++) out[i] = g.get(g.det ? 0 : (size_t)i);
    return 0;
}

} // extern "C"

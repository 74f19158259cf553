// Shared pieces of the CPU oracle (TEST INFRASTRUCTURE ONLY): the reference's
// RandomVariable semantics, Java's binary search and the pseudo-inverse solve of the
// regression's normal equations. Included by lmm_oracle.cpp and lmm_aad.cpp.
#pragma once

#include <cstdint>
#include <cmath>
#include <vector>
#include <memory>
#include <algorithm>

extern "C" {
double orc_round_time(double t);
int orc_eigen_sym(const double* a_in, int n, double* evals, double* evecs);
int orc_brownian(const double* times, int nT, int F, int P, int32_t seed, double* out);
}

namespace orc {

// Java Arrays.binarySearch on a sorted array: index, or -(insertion point)-1.
inline int java_binary_search(const double* a, int n, double key) {
    int lo = 0, hi = n - 1;
    while (lo <= hi) {
        int mid = (lo + hi) >> 1;
        double v = a[mid];
        if (v < key) lo = mid + 1;
        else if (v > key) hi = mid - 1;
        else return mid;
    }
    return -(lo + 1);
}

// ---------------------------------------------------------------------------
// RandomVariable semantics: a value that is either deterministic or an array of
// path realisations, with the type priority of Scalar (0) and
// RandomVariableFromDoubleArray (1). Each op restates the one of
// stochastic/Scalar.java:200-371 or montecarlo/RandomVariableFromDoubleArray.java
// (line numbers at each op), including which operand's implementation runs.
// ---------------------------------------------------------------------------

struct RV {
    int prio = 0;            // 0 Scalar, 1 RandomVariableFromDoubleArray
    bool det = true;
    double val = 0.0;
    std::shared_ptr<const std::vector<double>> r;
    size_t size() const { return det ? 1 : r->size(); }
    double get(size_t i) const { return det ? val : (*r)[i]; }
};

inline RV scalar(double v) { RV x; x.prio = 0; x.det = true; x.val = v; return x; }
inline RV rvdet(double v) { RV x; x.prio = 1; x.det = true; x.val = v; return x; }
inline RV rvarr(std::vector<double>&& v) {
    RV x; x.prio = 1; x.det = false; x.r = std::make_shared<const std::vector<double>>(std::move(v)); return x;
}
template <class Fn> inline RV map1(const RV& a, Fn fn) {
    if (a.det) return a.prio == 0 ? scalar(fn(a.val)) : rvdet(fn(a.val));
    std::vector<double> o(a.r->size());
    const double* s = a.r->data();
    for (size_t i = 0; i < o.size(); i++) o[i] = fn(s[i]);
    return rvarr(std::move(o));
}
template <class Fn> inline RV map2(const RV& a, const RV& b, size_t n, Fn fn) {
    std::vector<double> o(n);
    for (size_t i = 0; i < n; i++) o[i] = fn(a.get(i), b.get(i));
    return rvarr(std::move(o));
}
inline size_t nmax(const RV& a, const RV& b) { return std::max(a.size(), b.size()); }

// unary / with double: Scalar.java:201-262, RandomVariableFromDoubleArray.java:743-990
inline RV add(const RV& a, double v) { return map1(a, [v](double x) { return x + v; }); }
inline RV sub(const RV& a, double v) { return map1(a, [v](double x) { return x - v; }); }
inline RV mult(const RV& a, double v) { return map1(a, [v](double x) { return x * v; }); }
inline RV divd(const RV& a, double v) { return map1(a, [v](double x) { return x / v; }); }
inline RV cap(const RV& a, double v) { return map1(a, [v](double x) { return std::min(x, v); }); }
inline RV floor_(const RV& a, double v) { return map1(a, [v](double x) { return std::max(x, v); }); }
inline RV expv(const RV& a) { return map1(a, [](double x) { return std::exp(x); }); }
inline RV logv(const RV& a) { return map1(a, [](double x) { return std::log(x); }); }
inline RV invert(const RV& a) { return map1(a, [](double x) { return 1.0 / x; }); }
inline RV powv(const RV& a, double e) { return map1(a, [e](double x) { return std::pow(x, e); }); }
inline RV sqrtv(const RV& a) { return map1(a, [](double x) { return std::sqrt(x); }); }

// binary ops
inline RV add(const RV& a, const RV& b);
inline RV mult(const RV& a, const RV& b);

inline RV add(const RV& a, const RV& b) {
    if (a.prio == 0) return add(b, a.val);                          // Scalar.java:275-278
    if (b.prio > a.prio) return add(b, a);
    if (a.det && b.det) return rvdet(a.val + b.val);                // RandomVariableFromDoubleArray.java:1028-1058
    if (a.det) return map2(a, b, nmax(a, b), [](double x, double y) { return x + y; });
    if (b.det) return add(a, b.val);
    return map2(a, b, nmax(a, b), [](double x, double y) { return x + y; });
}
inline RV sub(const RV& a, const RV& b) {
    if (a.prio == 0) return mult(sub(b, a.val), -1.0);              // Scalar.java:280-283
    if (a.det && b.det) return rvdet(a.val - b.val);                // RandomVariableFromDoubleArray.java:1061-1091
    if (a.det) return map2(a, b, nmax(a, b), [](double x, double y) { return x - y; });
    if (b.det) return sub(a, b.val);
    return map2(a, b, nmax(a, b), [](double x, double y) { return x - y; });
}
inline RV mult(const RV& a, const RV& b) {
    if (a.prio == 0) return mult(b, a.val);                         // Scalar.java:290-293
    if (a.det && b.det) return rvdet(a.val * b.val);                // RandomVariableFromDoubleArray.java:1124-1154
    if (b.det) return mult(a, b.val);
    if (a.det) return map2(a, b, nmax(a, b), [](double x, double y) { return x * y; });
    return map2(a, b, nmax(a, b), [](double x, double y) { return x * y; });
}
inline RV div(const RV& a, const RV& b) {
    if (a.prio == 0) return mult(invert(b), a.val);                 // Scalar.java:295-298
    if (a.det && b.det) return rvdet(a.val / b.val);                // RandomVariableFromDoubleArray.java:1157-1184
    return map2(a, b, nmax(a, b), [](double x, double y) { return x / y; });
}
// accrue: Scalar.java:315-318, RandomVariableFromDoubleArray.java:1279-1305
inline RV accrue(const RV& a, const RV& rate, double dt) {
    if (a.prio == 0) return add(mult(rate, dt * a.val), a.val);
    if (rate.det) return mult(a, 1.0 + rate.val * dt);
    return map2(a, rate, nmax(a, rate), [dt](double x, double r) { return x * (1 + r * dt); });
}
// discount: Scalar.java:320-328, RandomVariableFromDoubleArray.java:1308-1335
inline RV discount(const RV& a, const RV& rate, double dt) {
    if (a.prio == 0) {
        if (a.val == 0) return mult(rate, 0.0);
        return invert(add(mult(rate, dt / a.val), 1.0 / a.val));
    }
    if (rate.det) return divd(a, 1.0 + rate.val * dt);
    return map2(a, rate, nmax(a, rate), [dt](double x, double r) { return x / (1.0 + r * dt); });
}
// addProduct(RV, double): Scalar.java:349-352, RandomVariableFromDoubleArray.java:1366-1391
inline RV addProduct(const RV& a, const RV& f1, double f2) {
    if (a.prio == 0) return add(mult(f1, f2), a.val);
    if (f1.det) return add(a, f1.val * f2);
    return map2(a, f1, nmax(a, f1), [f2](double x, double y) { return x + y * f2; });
}
// addProduct(RV, RV): Scalar.java:354-357, RandomVariableFromDoubleArray.java:1395-1427
inline RV addProduct(const RV& a, const RV& f1, const RV& f2) {
    if (a.prio == 0) return add(mult(f1, f2), a.val);
    if (a.det && f1.det && f2.det) return rvdet(a.val + (f1.val * f2.val));
    if (f1.det && f2.det) return add(a, f1.val * f2.val);
    if (f2.det) return addProduct(a, f1, f2.val);
    if (f1.det) return addProduct(a, f2, f1.val);
    if (!a.det) {
        std::vector<double> o(std::max(a.size(), std::max(f1.size(), f2.size())));
        for (size_t i = 0; i < o.size(); i++) o[i] = a.get(i) + f1.get(i) * f2.get(i);
        return rvarr(std::move(o));
    }
    return add(a, mult(f1, f2));
}
// choose: Scalar.java:330-337, RandomVariableFromDoubleArray.java:1342-1363
inline RV choose(const RV& trig, const RV& a, const RV& b) {
    if (trig.det) return trig.val >= 0 ? a : b;
    std::vector<double> o(trig.size());
    for (size_t i = 0; i < o.size(); i++) o[i] = trig.get(i) >= 0.0 ? a.get(i) : b.get(i);
    return rvarr(std::move(o));
}
// Kahan average: RandomVariableFromDoubleArray.java:287-307
inline double average(const RV& a) {
    if (a.det) return a.val;
    const auto& v = *a.r;
    if (v.empty()) return NAN;
    double sum = 0.0, err = 0.0;
    for (double x : v) { double y = x - err; double t = sum + y; err = (t - sum) - y; sum = t; }
    return sum / (double)v.size();
}
// getVariance / getStandardError: RandomVariableFromDoubleArray.java:333-355,419-428
inline double variance(const RV& a) {
    if (a.det || a.size() == 1) return 0.0;
    double avg = average(a);
    const auto& v = *a.r;
    double sum = 0.0, err = 0.0;
    for (double x : v) { double y = (x - avg) * (x - avg) - err; double t = sum + y; err = (t - sum) - y; sum = t; }
    return sum / (double)v.size();
}
inline double stderror(const RV& a) {
    if (a.det) return 0.0;
    return std::sqrt(variance(a)) / std::sqrt((double)a.size());
}

// Pseudo-inverse solve of the regression's normal equations
// (MonteCarloConditionalExpectationRegression.java:136-147 with commons-math3's
// SingularValueDecomposition solver): singular values at or below
// max(K sigma_max 2^-52, sqrt(2^-1022)) are dropped. X^T X is symmetric positive
// semi-definite, so its singular values are |eigenvalues|.
inline std::vector<double> pinv_solve(const std::vector<double>& A, const std::vector<double>& b, int K) {
    std::vector<double> ev((size_t)K), V((size_t)K * K);
    orc_eigen_sym(A.data(), K, ev.data(), V.data());
    double smax = 0.0;
    for (int i = 0; i < K; i++) smax = std::max(smax, std::fabs(ev[(size_t)i]));
    const double tol = std::max(K * smax * 0x1.0p-52, std::sqrt(0x1.0p-1022));
    std::vector<double> x((size_t)K, 0.0);
    for (int i = 0; i < K; i++) {
        if (!(std::fabs(ev[(size_t)i]) > tol)) continue;
        double d = 0.0;
        for (int r = 0; r < K; r++) d += V[(size_t)r * K + i] * b[(size_t)r];
        d /= ev[(size_t)i];
        for (int r = 0; r < K; r++) x[(size_t)r] += V[(size_t)r * K + i] * d;
    }
    return x;
}

} // namespace orc

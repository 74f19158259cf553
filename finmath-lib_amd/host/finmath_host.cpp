// Implementation of the host-side mirror of the finmath-lib API (see finmath_host.hpp).
#include "finmath_host.hpp"

#include <atomic>
#include <exception>
#include <thread>

#include <algorithm>
#include <cstring>

namespace finmath {

void check(int status) {
    if (status == FM_OK) return;
    const std::string msg = fm_last_error();
    switch (status) {
        case FM_ERR_INVALID_ARGUMENT: throw IllegalArgumentException(msg);
        case FM_ERR_STATE: throw IllegalStateException(msg);
        case FM_ERR_UNSUPPORTED: throw UnsupportedOperationException(msg);
        default: throw CalculationException(msg);
    }
}

namespace {
// java.util.Arrays.binarySearch semantics
int jsearch(const std::vector<double>& a, double key) {
    int lo = 0, hi = (int)a.size() - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        if (a[(size_t)mid] < key) lo = mid + 1;
        else if (a[(size_t)mid] > key) hi = mid - 1;
        else return mid;
    }
    return -(lo + 1);
}
} // namespace

// ---- TimeDiscretizationFromArray -------------------------------------------------
TimeDiscretizationFromArray::TimeDiscretizationFromArray(double initial, int numberOfTimeSteps, double deltaT) {
    if (numberOfTimeSteps < 0) throw IllegalArgumentException("number of time steps must be non-negative");
    t_.resize((size_t)numberOfTimeSteps + 1);
    for (int i = 0; i <= numberOfTimeSteps; i++) t_[(size_t)i] = roundToTimeTickSize(initial + i * deltaT);
}

TimeDiscretizationFromArray::TimeDiscretizationFromArray(std::vector<double> times) {
    for (double& x : times) x = roundToTimeTickSize(x);
    std::sort(times.begin(), times.end());
    times.erase(std::unique(times.begin(), times.end()), times.end());
    t_ = std::move(times);
}

int TimeDiscretizationFromArray::getTimeIndex(double time) const { return jsearch(t_, roundToTimeTickSize(time)); }

int TimeDiscretizationFromArray::getTimeIndexNearestLessOrEqual(double time) const {
    int i = jsearch(t_, roundToTimeTickSize(time));
    return i < 0 ? -i - 2 : i;
}

// ---- curves ---------------------------------------------------------------------------
std::shared_ptr<ForwardCurveInterpolation> ForwardCurveInterpolation::createForwardCurveFromForwards(
    const std::string& name, const std::vector<double>& times, const std::vector<double>& forwards,
    double paymentOffset) {
    if (times.size() != forwards.size() || times.empty())
        throw IllegalArgumentException("times and forwards must have the same, non-zero length");
    auto c = std::make_shared<ForwardCurveInterpolation>();
    c->name_ = name;
    std::vector<size_t> idx(times.size());
    for (size_t i = 0; i < idx.size(); i++) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return times[a] < times[b]; });
    for (size_t i : idx) { c->t_.push_back(times[i]); c->f_.push_back(forwards[i]); }
    c->paymentOffset_ = paymentOffset;
    return c;
}

double ForwardCurveInterpolation::getForward(double t) const {
    const int i = jsearch(t_, t);
    if (i >= 0) return f_[(size_t)i];                            // node value
    const int iv = -i - 2;
    if (iv < 0) return f_.front();                               // constant extrapolation
    if (iv > (int)t_.size() - 2) return f_.back();
    const double slope = (f_[(size_t)iv + 1] - f_[(size_t)iv]) / (t_[(size_t)iv + 1] - t_[(size_t)iv]);
    return f_[(size_t)iv] * 1.0 + slope * (t - t_[(size_t)iv]);
}

double ForwardCurveInterpolation::getForward(double fixingTime, double paymentOffset) const {
    double forward = getForward(fixingTime);
    const double cpo = getPaymentOffset(fixingTime);
    if (paymentOffset != cpo)                                    // ForwardCurveInterpolation.java:338-347
        forward = (std::exp(std::log(1 + forward * cpo) * paymentOffset / cpo) - 1.0) / paymentOffset;
    return forward;
}

double DiscountCurveFromForwardCurve::getDiscountFactor(double maturity) const {
    double time = 0.0, df = 1.0;
    while (time < maturity) {
        const double po = fc_->getPaymentOffset(time);
        if (po <= 0) throw CalculationException("non-positive payment offset");
        df /= 1.0 + fc_->getForward(time) * std::min(po, maturity - time) * 1.0;
        time += po;
    }
    return df;
}

double getForwardSwapRate(const TimeDiscretizationFromArray& fixTenor, const TimeDiscretizationFromArray& floatTenor,
                          const ForwardCurveInterpolation& fc) {
    auto fcp = std::shared_ptr<const ForwardCurveInterpolation>(&fc, [](const ForwardCurveInterpolation*) {});
    DiscountCurveFromForwardCurve dc(fcp);
    const double ev = fixTenor.getTime(0);
    const double pev = dc.getDiscountFactor(ev);
    double annuity = 0.0;
    for (int i = 0; i < fixTenor.getNumberOfTimeSteps(); i++)
        annuity += fixTenor.getTimeStep(i) * dc.getDiscountFactor(fixTenor.getTime(i + 1));
    annuity /= pev;
    double floatLeg = 0.0;
    for (int i = 0; i < floatTenor.getNumberOfTimeSteps(); i++) {
        const double fix = floatTenor.getTime(i), pay = floatTenor.getTime(i + 1);
        floatLeg += fc.getForward(fix, pay - fix) * (pay - fix) * dc.getDiscountFactor(pay);
    }
    return (floatLeg / pev) / annuity;
}

// ---- volatility / correlation / covariance ------------------------------------------------
LIBORVolatilityModelFourParameterExponentialForm::LIBORVolatilityModelFourParameterExponentialForm(
    TimeDiscretizationPtr td, TimeDiscretizationPtr tenor, double a_, double b_, double c_, double d_, bool cal)
    : LIBORVolatilityModel(std::move(td), std::move(tenor)), a(a_), b(b_), c(c_), d(d_), isCalibrateable(cal) {}

double LIBORVolatilityModelFourParameterExponentialForm::getVolatility(int timeIndex, int liborIndex) const {
    const double tau = tenor_->getTime(liborIndex) - td_->getTime(timeIndex);
    if (tau <= 0) return 0.0;                                    // already fixed
    return (b * tau + a) * std::exp(c * (-tau)) + d;             // Scalar arithmetic of the reference
}

std::vector<double> LIBORVolatilityModelFourParameterExponentialForm::getParameter() const {
    if (!isCalibrateable) return {};
    return {a, b, c, d};
}

LIBORVolatilityModelFromGivenMatrix::LIBORVolatilityModelFromGivenMatrix(TimeDiscretizationPtr td,
                                                                         TimeDiscretizationPtr tenor,
                                                                         std::vector<double> vol)
    : LIBORVolatilityModel(std::move(td), std::move(tenor)), vol_(std::move(vol)) {
    if (vol_.size() != (size_t)td_->getNumberOfTimeSteps() * tenor_->getNumberOfTimeSteps())
        throw IllegalArgumentException("volatility matrix must be [numberOfTimeSteps][numberOfLibors]");
}

double LIBORVolatilityModelFromGivenMatrix::getVolatility(int timeIndex, int liborIndex) const {
    return vol_[(size_t)timeIndex * tenor_->getNumberOfTimeSteps() + liborIndex];
}

namespace {

// Symmetric eigen decomposition: Householder reduction to tridiagonal form, then
// the implicit QL iteration with Wilkinson shifts (Golub & Van Loan 8.3), which is
// also the method of commons-math3's EigenDecomposition. a is overwritten by the
// eigenvectors (columns); d receives the eigenvalues.
void symmetric_eigen(std::vector<double>& a, int n, std::vector<double>& d) {
    std::vector<double> e((size_t)n, 0.0);
    d.assign((size_t)n, 0.0);
    auto A = [&](int i, int j) -> double& { return a[(size_t)i * n + j]; };
    for (int i = n - 1; i > 0; i--) {
        const int l = i - 1;
        double h = 0.0, scale = 0.0;
        if (l > 0) {
            for (int k = 0; k <= l; k++) scale += std::fabs(A(i, k));
            if (scale == 0.0) e[(size_t)i] = A(i, l);
            else {
                for (int k = 0; k <= l; k++) { A(i, k) /= scale; h += A(i, k) * A(i, k); }
                double f = A(i, l);
                double g = (f >= 0.0) ? -std::sqrt(h) : std::sqrt(h);
                e[(size_t)i] = scale * g;
                h -= f * g;
                A(i, l) = f - g;
                f = 0.0;
                for (int j = 0; j <= l; j++) {
                    A(j, i) = A(i, j) / h;
                    g = 0.0;
                    for (int k = 0; k <= j; k++) g += A(j, k) * A(i, k);
                    for (int k = j + 1; k <= l; k++) g += A(k, j) * A(i, k);
                    e[(size_t)j] = g / h;
                    f += e[(size_t)j] * A(i, j);
                }
                const double hh = f / (h + h);
                for (int j = 0; j <= l; j++) {
                    f = A(i, j);
                    e[(size_t)j] = g = e[(size_t)j] - hh * f;
                    for (int k = 0; k <= j; k++) A(j, k) -= (f * e[(size_t)k] + g * A(i, k));
                }
            }
        } else {
            e[(size_t)i] = A(i, l);
        }
        d[(size_t)i] = h;
    }
    d[0] = 0.0;
    e[0] = 0.0;
    for (int i = 0; i < n; i++) {
        const int l = i - 1;
        if (d[(size_t)i] != 0.0) {
            for (int j = 0; j <= l; j++) {
                double g = 0.0;
                for (int k = 0; k <= l; k++) g += A(i, k) * A(k, j);
                for (int k = 0; k <= l; k++) A(k, j) -= g * A(k, i);
            }
        }
        d[(size_t)i] = A(i, i);
        A(i, i) = 1.0;
        for (int j = 0; j <= l; j++) A(j, i) = A(i, j) = 0.0;
    }
    // implicit QL
    for (int i = 1; i < n; i++) e[(size_t)i - 1] = e[(size_t)i];
    e[(size_t)n - 1] = 0.0;
    for (int l = 0; l < n; l++) {
        int iter = 0, m;
        do {
            for (m = l; m < n - 1; m++) {
                const double dd = std::fabs(d[(size_t)m]) + std::fabs(d[(size_t)m + 1]);
                if (std::fabs(e[(size_t)m]) <= 1e-16 * dd) break;
            }
            if (m != l) {
                if (iter++ == 60) throw CalculationException("eigen decomposition did not converge");
                double g = (d[(size_t)l + 1] - d[(size_t)l]) / (2.0 * e[(size_t)l]);
                double r = std::hypot(g, 1.0);
                g = d[(size_t)m] - d[(size_t)l] + e[(size_t)l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
                double s = 1.0, c = 1.0, p = 0.0;
                int i;
                for (i = m - 1; i >= l; i--) {
                    double f = s * e[(size_t)i];
                    const double b = c * e[(size_t)i];
                    e[(size_t)i + 1] = (r = std::hypot(f, g));
                    if (r == 0.0) { d[(size_t)i + 1] -= p; e[(size_t)m] = 0.0; break; }
                    s = f / r; c = g / r;
                    g = d[(size_t)i + 1] - p;
                    r = (d[(size_t)i] - g) * s + 2.0 * c * b;
                    d[(size_t)i + 1] = g + (p = s * r);
                    g = c * r - b;
                    for (int k = 0; k < n; k++) {
                        f = A(k, i + 1);
                        A(k, i + 1) = s * A(k, i) + c * f;
                        A(k, i) = c * A(k, i) - s * f;
                    }
                }
                if (r == 0.0 && i >= l) continue;
                d[(size_t)l] -= p; e[(size_t)l] = g; e[(size_t)m] = 0.0;
            }
        } while (m != l);
    }
}

// LinearAlgebra.getFactorMatrixUsingCommonsMath (functions/LinearAlgebra.java:392-445)
std::vector<double> factor_matrix(const std::vector<double>& corr, int n, int m) {
    std::vector<double> V = corr, ev;
    symmetric_eigen(V, n, ev);
    std::vector<int> idx((size_t)n);
    for (int i = 0; i < n; i++) idx[(size_t)i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return ev[(size_t)x] > ev[(size_t)y]; });
    std::vector<double> out((size_t)n * m);
    for (int f = 0; f < m; f++) {
        const int c = idx[(size_t)f];
        const double sign = V[(size_t)c] > 0.0 ? 1.0 : -1.0;     // first entry positive
        double norm2 = 0.0;
        for (int r = 0; r < n; r++) norm2 += V[(size_t)r * n + c] * V[(size_t)r * n + c];
        const double lambda = std::max(ev[(size_t)c], 0.0);
        for (int r = 0; r < n; r++) out[(size_t)r * m + f] = sign * std::sqrt(lambda / norm2) * V[(size_t)r * n + c];
    }
    return out;
}

} // namespace

std::vector<double> factorReduction(const std::vector<double>& corr, int n, int m) {
    if (m < 1 || m > n) throw IllegalArgumentException("number of factors must be in [1, n]");
    std::vector<double> f = factor_matrix(corr, n, m);
    for (int r = 0; r < n; r++) {                                // renormalise rows (LinearAlgebra.java:461-475)
        double ss = 0.0;
        for (int k = 0; k < m; k++) ss += f[(size_t)r * m + k] * f[(size_t)r * m + k];
        if (ss != 0.0) for (int k = 0; k < m; k++) f[(size_t)r * m + k] = f[(size_t)r * m + k] / std::sqrt(ss);
        else for (int k = 0; k < m; k++) f[(size_t)r * m + k] = 1.0;
    }
    std::vector<double> red((size_t)n * n);
    for (int r = 0; r < n; r++)
        for (int c = 0; c < n; c++) {
            double s = 0.0;
            for (int k = 0; k < m; k++) s += f[(size_t)r * m + k] * f[(size_t)c * m + k];
            red[(size_t)r * n + c] = s;
        }
    return factor_matrix(red, n, m);                             // orthogonalised again (:478-481)
}

LIBORCorrelationModelExponentialDecay::LIBORCorrelationModelExponentialDecay(TimeDiscretizationPtr, TimeDiscretizationPtr tenor,
                                                                             int numberOfFactors, double a, bool calibrateable)
    : n_(tenor->getNumberOfTimeSteps()), m_(numberOfFactors), a_(a), calibrateable_(calibrateable) {
    a = std::max(a, 0.0);
    std::vector<double> rho((size_t)n_ * n_);
    for (int r = 0; r < n_; r++)
        for (int c = 0; c < n_; c++) rho[(size_t)r * n_ + c] = std::exp(-a * std::fabs(tenor->getTime(r) - tenor->getTime(c)));
    factor_ = factorReduction(rho, n_, m_);
    corr_.assign((size_t)n_ * n_, 0.0);
    for (int c1 = 0; c1 < n_; c1++) {
        for (int c2 = 0; c2 < c1; c2++) {
            double s = 0.0;
            for (int k = 0; k < m_; k++) s += factor_[(size_t)c1 * m_ + k] * factor_[(size_t)c2 * m_ + k];
            corr_[(size_t)c1 * n_ + c2] = corr_[(size_t)c2 * n_ + c1] = s;
        }
        corr_[(size_t)c1 * n_ + c1] = 1.0;
    }
}

std::vector<double> LIBORCovarianceModelFromVolatilityAndCorrelation::getFactorLoading(int timeIndex, int component) const {
    const int m = corr_->getNumberOfFactors();
    std::vector<double> out((size_t)m);
    const double v = vol_->getVolatility(timeIndex, component);
    for (int k = 0; k < m; k++) out[(size_t)k] = v * corr_->getFactorLoading(timeIndex, k, component);
    return out;
}

double LIBORCovarianceModelFromVolatilityAndCorrelation::getCovariance(int timeIndex, int c1, int c2) const {
    return (vol_->getVolatility(timeIndex, c1) * vol_->getVolatility(timeIndex, c2)) * corr_->getCorrelation(timeIndex, c1, c2);
}

int LIBORCovarianceModelFromVolatilityAndCorrelation::covarianceTimeIndex(double time) const {
    int ci = td_->getTimeIndex(time);
    return ci < 0 ? -ci - 2 : ci;
}

bool LIBORCovarianceModelFromVolatilityAndCorrelation::localVolatility(const TimeDiscretizationFromArray&, double& alpha,
                                                                       std::vector<double>& offsets) const {
    alpha = 0.0;
    offsets.clear();
    return false;
}

void LIBORCovarianceModelFromVolatilityAndCorrelation::tables(const TimeDiscretizationFromArray& sim,
                                                              std::vector<double>& lam, std::vector<double>& var) const {
    const int nT = sim.getNumberOfTimeSteps(), n = tenor_->getNumberOfTimeSteps(), m = getNumberOfFactors();
    lam.assign((size_t)nT * n * m, 0.0);
    var.assign((size_t)nT * n, 0.0);
    for (int i = 0; i < nT; i++) {
        // AbstractLIBORCovarianceModel.getFactorLoading(double time, ...): round-down time index (:54-60,69-76)
        const int ci = covarianceTimeIndex(sim.getTime(i));
        for (int j = 0; j < n; j++) {
            const std::vector<double> f = getFactorLoading(ci, j);
            for (int k = 0; k < m; k++) lam[((size_t)i * n + j) * m + k] = f[(size_t)k];
            var[(size_t)i * n + j] = getCovariance(ci, j, j);
        }
    }
}

// ---- piecewise-constant volatility ------------------------------------------------------------
LIBORVolatilityModelPiecewiseConstant::LIBORVolatilityModelPiecewiseConstant(
    TimeDiscretizationPtr td, TimeDiscretizationPtr tenor, TimeDiscretizationPtr simulationTimes,
    TimeDiscretizationPtr timesToMaturity, std::vector<double> volatility, bool isCalibrateable)
    : LIBORVolatilityModel(std::move(td), std::move(tenor)), sim_(std::move(simulationTimes)),
      ttm_(std::move(timesToMaturity)), calibrateable_(isCalibrateable) {
    const double maxMaturity = tenor_->getTime(tenor_->getNumberOfTimes() - 1);
    int count = 0;
    index_.assign((size_t)sim_->getNumberOfTimes(), std::vector<int>((size_t)ttm_->getNumberOfTimes(), -1));
    for (int i = 0; i < sim_->getNumberOfTimes(); i++)
        for (int k = 0; k < ttm_->getNumberOfTimes(); k++)
            if (!(sim_->getTime(i) + ttm_->getTime(k) > maxMaturity)) index_[(size_t)i][(size_t)k] = count++;
    if (volatility.size() == 1) vol_.assign((size_t)count, volatility[0]);
    else if ((int)volatility.size() == count) vol_ = std::move(volatility);
    else throw IllegalArgumentException("Volatility length does not match number of free parameters.");
}

double LIBORVolatilityModelPiecewiseConstant::getVolatility(int timeIndex, int liborIndex) const {
    const double time = td_->getTime(timeIndex);
    const double timeToMaturity = tenor_->getTime(liborIndex) - time;
    if (timeToMaturity <= 0) return 0.0;                          // already fixed
    int i = sim_->getTimeIndexNearestLessOrEqual(time);
    i = std::min(std::max(i, 0), sim_->getNumberOfTimes() - 1);
    int k = ttm_->getTimeIndexNearestLessOrEqual(timeToMaturity);
    k = std::min(std::max(k, 0), ttm_->getNumberOfTimes() - 1);
    const int p = index_[(size_t)i][(size_t)k];
    if (p < 0) throw IllegalStateException("no volatility parameter for this bucket");
    return vol_[(size_t)p];
}

// ---- local-volatility wrappers -------------------------------------------------------------
static bool isStochastic(const LIBORCovarianceModelFromVolatilityAndCorrelation& cov) {
    double a, b, c;
    BrownianMotionPtr bm;
    return cov.stochasticVolatility(a, b, bm) || cov.hestonVolatility(a, b, c, bm);
}

static const LIBORCovarianceModelFromVolatilityAndCorrelation& unwrapped(const CovarianceModelPtr& base) {
    if (!base) throw IllegalArgumentException("covariance model must be given");
    double a;
    std::vector<double> o;
    if (base->localVolatility(*base->getTimeDiscretization(), a, o))
        throw UnsupportedOperationException("local-volatility wrappers do not nest");
    if (isStochastic(*base))
        throw UnsupportedOperationException("a stochastic volatility wrapper goes outside a local-volatility wrapper");
    return *base;
}

// the base of a stochastic volatility wrapper: a plain model or one local-volatility wrapper
static const LIBORCovarianceModelFromVolatilityAndCorrelation& stochasticBase(const CovarianceModelPtr& base) {
    if (!base) throw IllegalArgumentException("covariance model must be given");
    if (isStochastic(*base)) throw UnsupportedOperationException("stochastic volatility wrappers do not nest");
    return *base;
}

LIBORCovarianceModelStochasticHestonVolatility::LIBORCovarianceModelStochasticHestonVolatility(
    CovarianceModelPtr covarianceModel, BrownianMotionPtr brownianMotion, double kappa, double theta, double xi,
    bool isCalibrateable)
    : LIBORCovarianceModelFromVolatilityAndCorrelation(stochasticBase(covarianceModel)),
      base_(std::move(covarianceModel)), bm_(std::move(brownianMotion)), kappa_(kappa), theta_(theta), xi_(xi),
      calibrateable_(isCalibrateable) {
    if (!bm_) throw IllegalArgumentException("the Heston volatility needs a Brownian motion");
    if (!std::isfinite(kappa) || !std::isfinite(theta) || !std::isfinite(xi))
        throw IllegalArgumentException("kappa, theta and xi must be finite");
}

LIBORCovarianceModelStochasticVolatility::LIBORCovarianceModelStochasticVolatility(CovarianceModelPtr covarianceModel,
                                                                                   BrownianMotionPtr brownianMotion,
                                                                                   double nu, double rho,
                                                                                   bool isCalibrateable)
    : LIBORCovarianceModelFromVolatilityAndCorrelation(stochasticBase(covarianceModel)),
      base_(std::move(covarianceModel)), bm_(std::move(brownianMotion)), nu_(nu), rho_(rho),
      calibrateable_(isCalibrateable) {
    if (!bm_ || bm_->getNumberOfFactors() < 2)
        throw IllegalArgumentException("the stochastic volatility needs a Brownian motion with at least two factors");
    if (!std::isfinite(nu) || !(rho >= -1.0 && rho <= 1.0))
        throw IllegalArgumentException("nu must be finite and rho in [-1, 1]");
}

// The scaling process of a stochastic volatility wrapper on the device: a context used only as
// the wrapper's Brownian motion drives it (factors 0 and 1); without a wrapper the scaling is
// removed.
static void applyStochasticVolatility(fm_lmm* c, const LIBORCovarianceModelFromVolatilityAndCorrelation& cov,
                                      const std::vector<double>& tenor, int paths) {
    double nu = 0.0, rho = 0.0, xi = 0.0;
    BrownianMotionPtr bm;
    const bool heston = cov.hestonVolatility(nu, rho, xi, bm);     // kappa, theta, xi
    if (!heston && !cov.stochasticVolatility(nu, rho, bm)) {
        check(fm_lmm_set_stochastic_volatility(c, 0.0, 0.0, nullptr, 0, 1));
        return;
    }
    if (bm->getNumberOfPaths() != paths)
        throw IllegalArgumentException("the stochastic volatility's Brownian motion has another number of paths");
    const auto& times = bm->getTimeDiscretization()->getAsDoubleArray();
    fm_lmm_config cfg{};
    cfg.num_time_steps = (int32_t)times.size() - 1;
    cfg.times = times.data();
    cfg.num_libors = (int32_t)tenor.size() - 1;
    cfg.tenor = tenor.data();
    cfg.num_factors = bm->getNumberOfFactors();
    cfg.num_paths = paths;
    cfg.seed = bm->getSeed();
    cfg.measure = FM_MEASURE_SPOT;
    cfg.state_space = FM_STATE_SPACE_LOGNORMAL;
    cfg.scheme = FM_SCHEME_EULER_FUNCTIONAL;
    cfg.precision = FM_PRECISION_FP64;
    cfg.libor_cap = 1e5;
    fm_lmm* driver = nullptr;
    check(fm_lmm_create(&cfg, &driver));
    std::unique_ptr<fm_lmm, void (*)(fm_lmm*)> guard(driver, fm_lmm_destroy);
    if (heston) check(fm_lmm_set_heston_volatility(c, nu, rho, xi, driver, 0));
    else check(fm_lmm_set_stochastic_volatility(c, nu, rho, driver, 0, 1));
}

DisplacedLocalVolatilityModel::DisplacedLocalVolatilityModel(CovarianceModelPtr covarianceModel, double displacement,
                                                             bool isCalibrateable)
    : LIBORCovarianceModelFromVolatilityAndCorrelation(unwrapped(covarianceModel)),
      base_(std::move(covarianceModel)), displacement_(displacement), calibrateable_(isCalibrateable) {
    if (!std::isfinite(displacement)) throw IllegalArgumentException("displacement must be finite");
}

bool DisplacedLocalVolatilityModel::localVolatility(const TimeDiscretizationFromArray& sim, double& alpha,
                                                    std::vector<double>& offsets) const {
    // L.add(d) (DisplacedLocalVolatilityModel.java:179)
    alpha = 0.0;
    offsets.assign((size_t)sim.getNumberOfTimeSteps() * getLiborPeriodDiscretization()->getNumberOfTimeSteps(),
                   displacement_);
    return true;
}

BlendedLocalVolatilityModel::BlendedLocalVolatilityModel(CovarianceModelPtr covarianceModel, ForwardCurvePtr forwardCurve,
                                                         double displacement, bool isCalibrateable)
    : LIBORCovarianceModelFromVolatilityAndCorrelation(unwrapped(covarianceModel)),
      base_(std::move(covarianceModel)), fc_(std::move(forwardCurve)), displacement_(displacement),
      calibrateable_(isCalibrateable) {
    if (!std::isfinite(displacement)) throw IllegalArgumentException("displacement must be finite");
}

bool BlendedLocalVolatilityModel::localVolatility(const TimeDiscretizationFromArray& sim, double& alpha,
                                                  std::vector<double>& offsets) const {
    // L.sub(L.mult(a)).add(a.mult(f0(max(T_j - t, 0)))) at the covariance model's round-down
    // time index t (BlendedLocalVolatilityModel.java:268-276)
    const int nT = sim.getNumberOfTimeSteps(), n = getLiborPeriodDiscretization()->getNumberOfTimeSteps();
    alpha = displacement_;
    offsets.assign((size_t)nT * n, 0.0);
    for (int i = 0; i < nT; i++) {
        const double t = getTimeDiscretization()->getTime(covarianceTimeIndex(sim.getTime(i)));
        for (int j = 0; j < n; j++) {
            double forward = 1.0;
            if (fc_) forward = fc_->getForward(std::max(getLiborPeriodDiscretization()->getTime(j) - t, 0.0));
            offsets[(size_t)i * n + j] = displacement_ * forward;
        }
    }
    return true;
}

// ---- LIBOR market model ------------------------------------------------------------------
std::shared_ptr<LIBORMarketModelFromCovarianceModel> LIBORMarketModelFromCovarianceModel::of(
    TimeDiscretizationPtr tenor, ForwardCurvePtr fc, DiscountCurvePtr dc, CovarianceModelPtr cov,
    const std::map<std::string, std::string>& props) {
    auto m = std::make_shared<LIBORMarketModelFromCovarianceModel>();
    m->tenor_ = std::move(tenor); m->fc_ = std::move(fc); m->dc_ = std::move(dc); m->cov_ = std::move(cov);
    auto up = [](std::string s) { for (auto& ch : s) ch = (char)std::toupper((unsigned char)ch); return s; };
    auto it = props.find("measure");
    if (it != props.end()) {
        const std::string v = up(it->second);
        if (v == "SPOT") m->measure = FM_MEASURE_SPOT;
        else if (v == "TERMINAL") m->measure = FM_MEASURE_TERMINAL;
        else throw IllegalArgumentException("unknown measure " + it->second);
    }
    it = props.find("stateSpace");
    if (it != props.end()) {
        const std::string v = up(it->second);
        if (v == "LOGNORMAL") m->stateSpace = FM_STATE_SPACE_LOGNORMAL;
        else if (v == "NORMAL") m->stateSpace = FM_STATE_SPACE_NORMAL;
        else throw IllegalArgumentException("unknown state space " + it->second);
    }
    it = props.find("liborCap");
    if (it != props.end()) m->liborCap = std::stod(it->second);
    return m;
}

std::vector<double> LIBORMarketModelFromCovarianceModel::getInitialForwards() const {
    const int n = getNumberOfLibors();
    std::vector<double> L0((size_t)n);
    for (int j = 0; j < n; j++) L0[(size_t)j] = fc_->getForward(tenor_->getTime(j), tenor_->getTimeStep(j));
    return L0;
}

BrownianMotionFromMersenneRandomNumbers::BrownianMotionFromMersenneRandomNumbers(TimeDiscretizationPtr td_, int F,
                                                                                 int P, int seed_)
    : td(std::move(td_)), numberOfFactors(F), numberOfPaths(P), seed(seed_) {
    // Validate.isTrue (montecarlo/BrownianMotionFromMersenneRandomNumbers.java:77-78)
    if (F <= 0) throw IllegalArgumentException("Number of factors must be greater or equal 1 (given " + std::to_string(F) + ").");
    if (P <= 0) throw IllegalArgumentException("Number of paths must be greater or equal 1 (given " + std::to_string(P) + ").");
}

// ---- RandomVariable ---------------------------------------------------------------------
double RandomVariable::getAverage() const {
    if (det_) return value_;
    if (r_.empty()) return NAN;
    double sum = 0.0, err = 0.0;
    for (double x : r_) { const double y = x - err; const double t = sum + y; err = (t - sum) - y; sum = t; }
    return sum / (double)r_.size();
}

double RandomVariable::getVariance() const {
    if (det_ || r_.size() == 1) return 0.0;
    const double avg = getAverage();
    double sum = 0.0, err = 0.0;
    for (double x : r_) {
        const double y = (x - avg) * (x - avg) - err;
        const double t = sum + y;
        err = (t - sum) - y;
        sum = t;
    }
    return sum / (double)r_.size();
}

double RandomVariable::getStandardError() const {
    if (det_) return 0.0;
    return std::sqrt(getVariance()) / std::sqrt((double)r_.size());
}

// ---- Euler scheme over the engine -------------------------------------------------------
EulerSchemeFromProcessModel::EulerSchemeFromProcessModel(LIBORMarketModelPtr model, BrownianMotionPtr driver, Scheme scheme,
                                                         int precision)
    : model_(std::move(model)), driver_(std::move(driver)), scheme_(scheme), precision_(precision) {
    if (!model_ || !driver_) throw IllegalArgumentException("model and driver must be given");
}

EulerSchemeFromProcessModel::~EulerSchemeFromProcessModel() { if (ctx_) fm_lmm_destroy(ctx_); }

void EulerSchemeFromProcessModel::ensureContext(bool storePaths) {
    if (ctx_ && (!storePaths || ctxStores_)) return;
    if (ctx_) { fm_lmm_destroy(ctx_); ctx_ = nullptr; simulated_ = false; }
    const auto& times = driver_->getTimeDiscretization()->getAsDoubleArray();
    const auto& tenor = model_->getLiborPeriodDiscretization()->getAsDoubleArray();
    fm_lmm_config cfg{};
    cfg.num_time_steps = (int32_t)times.size() - 1;
    cfg.times = times.data();
    cfg.num_libors = (int32_t)tenor.size() - 1;
    cfg.tenor = tenor.data();
    cfg.num_factors = driver_->getNumberOfFactors();
    cfg.num_paths = driver_->getNumberOfPaths();
    cfg.seed = driver_->getSeed();
    cfg.measure = model_->measure;
    cfg.state_space = model_->stateSpace;
    cfg.scheme = (int32_t)scheme_;
    cfg.precision = precision_;
    cfg.libor_cap = model_->liborCap;
    cfg.store_paths = storePaths ? 1 : 0;
    cfg.checkpoint_stride = ckptStride_;
    cfg.regenerate_increments = regenerate_ ? 1 : 0;
    cfg.path_offset = 0;
    if (driver_->getNumberOfFactors() != model_->getCovarianceModel()->getNumberOfFactors())
        throw IllegalArgumentException("number of factors of the driver and of the covariance model differ");
    fm_lmm* c = nullptr;
    check(fm_lmm_create(&cfg, &c));
    try {
        const std::vector<double> L0 = model_->getInitialForwards();
        check(fm_lmm_set_initial_forwards(c, L0.data()));
        std::vector<double> lam, var;
        model_->getCovarianceModel()->tables(*driver_->getTimeDiscretization(), lam, var);
        check(fm_lmm_set_covariance(c, lam.data(), var.data()));
        double lvAlpha = 0.0;
        std::vector<double> lvOffsets;
        if (model_->getCovarianceModel()->localVolatility(*driver_->getTimeDiscretization(), lvAlpha, lvOffsets))
            check(fm_lmm_set_local_volatility(c, lvAlpha, lvOffsets.data()));
        applyStochasticVolatility(c, *model_->getCovarianceModel(), tenor, driver_->getNumberOfPaths());
        if (model_->getDiscountCurve()) {
            std::vector<double> pd(tenor.size());
            for (size_t k = 0; k < tenor.size(); k++) pd[k] = model_->getDiscountCurve()->getDiscountFactor(tenor[k]);
            check(fm_lmm_set_discount_curve(c, pd.data()));
        }
        // volatility factorization (lambda = sigma f, variance = sigma^2 rho_jj) for the adjoint
        const auto cov = model_->getCovarianceModel();
        const auto& sim = *driver_->getTimeDiscretization();
        const int nT = sim.getNumberOfTimeSteps(), n = (int)tenor.size() - 1, F = cov->getNumberOfFactors();
        std::vector<double> sig((size_t)nT * n), fac((size_t)n * F), cd((size_t)n);
        for (int i = 0; i < nT; i++) {
            int ci = cov->getTimeDiscretization()->getTimeIndex(sim.getTime(i));
            if (ci < 0) ci = -ci - 2;
            for (int j = 0; j < n; j++) sig[(size_t)i * n + j] = cov->getVolatilityModel()->getVolatility(ci, j);
        }
        for (int j = 0; j < n; j++) {
            for (int k = 0; k < F; k++) fac[(size_t)j * F + k] = cov->getCorrelationModel()->getFactorLoading(0, k, j);
            cd[(size_t)j] = cov->getCorrelationModel()->getCorrelation(0, j, j);
        }
        check(fm_lmm_set_volatility_factorization(c, sig.data(), fac.data(), cd.data()));
    } catch (...) { fm_lmm_destroy(c); throw; }
    ctx_ = c;
    ctxStores_ = storePaths;
}

void EulerSchemeFromProcessModel::setDiracDeltaApproximation(const std::string& method, double widthPerStdDev) {
    if (!(widthPerStdDev > 0.0)) throw IllegalArgumentException("diracDeltaApproximationWidthPerStdDev must be positive");
    int m;
    if (method == "DISCRETE_DELTA") m = FM_DIRAC_DISCRETE_DELTA;
    else if (method == "ONE") m = FM_DIRAC_ONE;
    else if (method == "ZERO") m = FM_DIRAC_ZERO;
    else if (method == "REGRESSION_ON_DENSITY" || method == "REGRESSION_ON_DISTRIBUITON")
        throw UnsupportedOperationException("the regression Dirac delta approximations are not implemented");
    else throw IllegalArgumentException("unknown diracDeltaApproximationMethod " + method);
    diracMethod_ = m;
    diracWidth_ = widthPerStdDev;
}

Gradient EulerSchemeFromProcessModel::getGradient(const std::vector<ProductPtr>& products, const std::vector<double>& weights,
                                                  int checkpointStride) {
    if (ctx_ && checkpointStride != ckptStride_) { fm_lmm_destroy(ctx_); ctx_ = nullptr; simulated_ = false; }
    ckptStride_ = checkpointStride;
    ensureContext(false);
    std::deque<std::vector<double>> arena;
    std::vector<fm_product> lowered(products.size());
    for (size_t q = 0; q < products.size(); q++) products[q]->lower(arena, lowered[q]);
    if (!weights.empty() && weights.size() != products.size())
        throw IllegalArgumentException("one weight per product expected");
    const int n = getNumberOfComponents(), nT = getTimeDiscretization()->getNumberOfTimeSteps();
    Gradient g;
    g.delta.assign((size_t)n, 0.0);
    g.volatility.assign((size_t)nT * n, 0.0);
    check(fm_lmm_set_dirac_delta(ctx_, diracMethod_, diracWidth_));
    check(fm_lmm_adjoint(ctx_, lowered.data(), (int32_t)lowered.size(), weights.empty() ? nullptr : weights.data(),
                         &g.value, g.delta.data(), g.volatility.data()));
    g.forwardVega.assign((size_t)n, 0.0);
    for (int i = 0; i < nT; i++)
        for (int j = 0; j < n; j++) g.forwardVega[(size_t)j] += g.volatility[(size_t)i * n + j];
    // sigma_j(t_i) is the volatility model's entry at the covariance model's time index of t_i,
    // rounded down (AbstractLIBORCovarianceModel.java:45-60): the gradient with respect to the
    // model's own matrix entries (the leaves of LIBORVolatilityModelFromGivenMatrix) sums the
    // simulation steps that read each entry
    const auto cov = model_->getCovarianceModel();
    const auto vm = cov->getVolatilityModel();
    const auto& sim = *getTimeDiscretization();
    const int nTc = vm->getTimeDiscretization()->getNumberOfTimeSteps() + 1;
    g.volatilityModel.assign((size_t)nTc * n, 0.0);
    std::vector<int> ci_of((size_t)nT);
    for (int i = 0; i < nT; i++) {
        int ci = cov->getTimeDiscretization()->getTimeIndex(sim.getTime(i));
        if (ci < 0) ci = -ci - 2;
        ci_of[(size_t)i] = ci;
        for (int j = 0; j < n; j++) g.volatilityModel[(size_t)ci * n + j] += g.volatility[(size_t)i * n + j];
    }
    // chain rule to (a, b, c, d) of sigma = (a + b tau) exp(-c tau) + d with tau = T_j - t_ci > 0,
    // the volatility model's own time of the entry (LIBORVolatilityModelFourParameterExponentialForm.java:168-195)
    auto fp = std::dynamic_pointer_cast<const LIBORVolatilityModelFourParameterExponentialForm>(vm);
    if (fp) {
        g.parameterVega.assign(4, 0.0);
        const auto& vtd = *vm->getTimeDiscretization();
        const auto& vtenor = *vm->getLiborPeriodDiscretization();
        for (int i = 0; i < nT; i++)
            for (int j = 0; j < n; j++) {
                const double tau = vtenor.getTime(j) - vtd.getTime(ci_of[(size_t)i]);
                if (tau <= 0) continue;
                const double e = std::exp(-fp->c * tau), gs = g.volatility[(size_t)i * n + j];
                g.parameterVega[0] += gs * e;
                g.parameterVega[1] += gs * tau * e;
                g.parameterVega[2] += gs * (-tau) * (fp->a + fp->b * tau) * e;
                g.parameterVega[3] += gs;
            }
    }
    return g;
}

fm_lmm* EulerSchemeFromProcessModel::handle() { ensureContext(false); return ctx_; }

RandomVariable EulerSchemeFromProcessModel::getProcessValue(int timeIndex, int componentIndex) {
    ensureContext(true);
    if (!simulated_) { check(fm_lmm_simulate(ctx_)); simulated_ = true; }
    std::vector<double> out((size_t)getNumberOfPaths());
    check(fm_lmm_get_process_value(ctx_, timeIndex, componentIndex, out.data()));
    const double t = getTime(timeIndex);
    if (timeIndex == 0) return RandomVariable(t, out[0]);       // deterministic initial value
    return RandomVariable(t, std::move(out));
}

RandomVariable EulerSchemeFromProcessModel::getBrownianIncrement(int timeIndex, int factor) {
    ensureContext(false);
    std::vector<double> out((size_t)getNumberOfPaths());
    check(fm_lmm_get_brownian_increment(ctx_, timeIndex, factor, out.data()));
    return RandomVariable(getTime(timeIndex + 1), std::move(out));
}

RandomVariable EulerSchemeFromProcessModel::getNumeraire(double time) {
    ensureContext(true);
    if (!simulated_) { check(fm_lmm_simulate(ctx_)); simulated_ = true; }
    std::vector<double> out((size_t)getNumberOfPaths());
    check(fm_lmm_get_numeraire(ctx_, time, out.data()));
    const int k = model_->getLiborPeriodIndex(time);
    if (k <= 1) return RandomVariable(time, out[0]);           // N(T_0), N(T_1) are deterministic
    return RandomVariable(time, std::move(out));
}

void EulerSchemeFromProcessModel::valueProducts(const std::vector<ProductPtr>& products, std::vector<double>& values,
                                                std::vector<double>& errors) {
    ensureContext(false);
    std::deque<std::vector<double>> arena;
    std::vector<fm_product> lowered(products.size());
    for (size_t q = 0; q < products.size(); q++) products[q]->lower(arena, lowered[q]);
    values.assign(products.size(), 0.0);
    errors.assign(products.size(), 0.0);
    check(fm_lmm_value_products(ctx_, lowered.data(), (int32_t)lowered.size(), values.data(), errors.data()));
}

std::vector<double> EulerSchemeFromProcessModel::lastTimings() {
    std::vector<double> ms(5, 0.0);
    if (ctx_) check(fm_lmm_last_timings(ctx_, ms.data(), 5));
    return ms;
}

std::vector<std::pair<double, double>> LIBORMonteCarloSimulationFromLIBORModel::getValues(const std::vector<ProductPtr>& products) {
    std::vector<double> v, e;
    process_->valueProducts(products, v, e);
    std::vector<std::pair<double, double>> out;
    for (size_t q = 0; q < v.size(); q++) {
        if (products[q]->isValueUnitValue()) out.emplace_back(v[q], e[q]);
        else out.emplace_back(products[q]->toValueUnit(v[q], *process_->getModel()), 0.0);   // a constant
    }
    return out;
}

// ---- products ----------------------------------------------------------------------------
double Product::getValue(LIBORMonteCarloSimulationFromLIBORModel& model) const {
    return model.getValues({shared_from_this()})[0].first;
}

std::map<std::string, double> Product::getValues(LIBORMonteCarloSimulationFromLIBORModel& model) const {
    const auto r = model.getValues({shared_from_this()})[0];
    return {{"value", r.first}, {"error", r.second}};
}

namespace {
const double* keep(std::deque<std::vector<double>>& arena, const std::vector<double>& v) {
    if (v.empty()) return nullptr;
    arena.push_back(v);
    return arena.back().data();
}
} // namespace

Swaption::Swaption(double ex, std::vector<double> fix, std::vector<double> pay, std::vector<double> pl,
                   std::vector<double> k, double notional_)
    : exerciseDate(ex), notional(notional_), fixingDates(std::move(fix)), paymentDates(std::move(pay)),
      periodLengths(std::move(pl)), swaprates(std::move(k)) {
    if (fixingDates.size() != paymentDates.size() || fixingDates.size() != swaprates.size() ||
        (!periodLengths.empty() && periodLengths.size() != fixingDates.size()))
        throw IllegalArgumentException("swaption arrays must have the same length");
}

Swaption::Swaption(double ex, const TimeDiscretizationFromArray& swapTenor, double swaprate)
    : exerciseDate(ex), notional(1.0) {                           // Swaption.java:110-130
    for (int i = 0; i < swapTenor.getNumberOfTimeSteps(); i++) {
        fixingDates.push_back(swapTenor.getTime(i));
        paymentDates.push_back(swapTenor.getTime(i + 1));
        swaprates.push_back(swaprate);
    }
}

void Swaption::lower(std::deque<std::vector<double>>& arena, fm_product& o) const {
    std::memset(&o, 0, sizeof(o));
    o.type = FM_PRODUCT_SWAPTION;
    o.num_periods = (int32_t)fixingDates.size();
    o.exercise_date = exerciseDate;
    o.fixing_dates = keep(arena, fixingDates);
    o.payment_dates = keep(arena, paymentDates);
    o.period_lengths = keep(arena, periodLengths);
    o.swaprates = keep(arena, swaprates);
    o.notional = notional;
}

SimpleSwap::SimpleSwap(std::vector<double> fix, std::vector<double> pay, std::vector<double> k, bool payFix,
                       std::vector<double> notional_)
    : fixingDates(std::move(fix)), paymentDates(std::move(pay)), swaprates(std::move(k)), notional(std::move(notional_)),
      isPayFix(payFix) {
    if (fixingDates.size() != paymentDates.size() || fixingDates.size() != swaprates.size())
        throw IllegalArgumentException("swap arrays must have the same length");
}

void SimpleSwap::lower(std::deque<std::vector<double>>& arena, fm_product& o) const {
    std::memset(&o, 0, sizeof(o));
    o.type = FM_PRODUCT_SWAP;
    o.num_periods = (int32_t)fixingDates.size();
    o.fixing_dates = keep(arena, fixingDates);
    o.payment_dates = keep(arena, paymentDates);
    o.swaprates = keep(arena, swaprates);
    o.notionals = keep(arena, notional);
    o.is_pay_fix = isPayFix ? 1 : 0;
    o.notional = 1.0;
}

Caplet::Caplet(double m, double pl, double k, double dcf, bool floorlet)
    : maturity(m), periodLength(pl), strike(k), daycountFraction(dcf), isFloorlet(floorlet) {}

void Caplet::lower(std::deque<std::vector<double>>& arena, fm_product& o) const {
    std::memset(&o, 0, sizeof(o));
    o.type = FM_PRODUCT_CAPLET;
    o.num_periods = 1;
    o.exercise_date = maturity;
    o.period_lengths = keep(arena, {periodLength});
    o.swaprates = keep(arena, {strike});
    o.notional = daycountFraction;
    o.is_floorlet = isFloorlet ? 1 : 0;
}

// ---- SwaptionSimple and the analytic formulas -------------------------------------------------
SwaptionSimple::SwaptionSimple(double k, std::vector<double> tenor, const std::string& valueUnit)
    : swaprate(k), swapTenor(std::move(tenor)) {
    if (swapTenor.size() < 2) throw IllegalArgumentException("the swap tenor needs at least two dates");
    static const std::map<std::string, int> units = {
        {"VALUE", 0}, {"VOLATILITY", 1}, {"VOLATILITYLOGNORMAL", 1}, {"VOLATILITYNORMAL", 2},
        {"INTEGRATEDVARIANCE", 3}, {"INTEGRATEDVARIANCELOGNORMAL", 3}, {"INTEGRATEDLOGNORMALVARIANCE", 3},
        {"INTEGRATEDVARIANCENORMAL", 4}, {"INTEGRATEDNORMALVARIANCE", 4}};
    const auto it = units.find(valueUnit);
    if (it == units.end()) throw UnsupportedOperationException("Provided valueUnit not implemented.");
    unit_ = it->second;
}

void SwaptionSimple::lower(std::deque<std::vector<double>>& arena, fm_product& o) const {
    // the underlying Swaption(exercise T_0, fixings T_0..T_{n-1}, payments T_1..T_n, one rate) (:49-60)
    std::memset(&o, 0, sizeof(o));
    const size_t n = swapTenor.size() - 1;
    o.type = FM_PRODUCT_SWAPTION;
    o.num_periods = (int32_t)n;
    o.exercise_date = swapTenor[0];
    o.fixing_dates = keep(arena, std::vector<double>(swapTenor.begin(), swapTenor.end() - 1));
    o.payment_dates = keep(arena, std::vector<double>(swapTenor.begin() + 1, swapTenor.end()));
    o.swaprates = keep(arena, std::vector<double>(n, swaprate));
    o.notional = 1.0;
}

double SwaptionSimple::toValueUnit(double value, const LIBORMarketModelFromCovarianceModel& model) const {
    if (unit_ == 0) return value;
    const ForwardCurvePtr fc = model.getForwardRateCurve();
    if (!fc) throw IllegalArgumentException("volatility units need the model's forward curve");
    const TimeDiscretizationFromArray t(swapTenor);
    const double par = getForwardSwapRate(t, t, *fc);
    const double T = swapTenor[0];
    const double annuity = getSwapAnnuity(swapTenor, *fc);
    const bool lognormal = unit_ == 1 || unit_ == 3;
    const double vol = lognormal ? AnalyticFormulas::blackScholesOptionImpliedVolatility(par, T, swaprate, annuity, value)
                                 : AnalyticFormulas::bachelierOptionImpliedVolatility(par, T, swaprate, annuity, value);
    return unit_ <= 2 ? vol : vol * vol * T;
}

double getSwapAnnuity(const std::vector<double>& tenor, const ForwardCurveInterpolation& fc) {
    auto fcp = std::shared_ptr<const ForwardCurveInterpolation>(&fc, [](const ForwardCurveInterpolation*) {});
    DiscountCurveFromForwardCurve dc(fcp);
    double value = 0.0;
    for (size_t i = 0; i + 1 < tenor.size(); i++) {
        const double pay = tenor[i + 1];
        if (pay <= 0.0) continue;
        value += (tenor[i + 1] - tenor[i]) * dc.getDiscountFactor(pay);
    }
    return value / dc.getDiscountFactor(0.0);
}

namespace AnalyticFormulas {
namespace {
double Phi(double x) { return 0.5 * std::erfc(-x / std::sqrt(2.0)); }
double phi(double x) { return std::exp(-0.5 * x * x) / std::sqrt(2.0 * M_PI); }

// root of the increasing function value(vol) = target on [lo, hi]: Newton steps with the
// derivative, bisection whenever a step leaves the bracket, to a relative width of 1e-15
template <class F>
double invert(F value_and_vega, double target, double lo, double hi) {
    double v = 0.5 * (lo + hi);
    for (int it = 0; it < 500; it++) {
        double val, vega;
        value_and_vega(v, val, vega);
        const double err = val - target;
        if (err == 0.0) return v;
        if (err > 0) hi = v; else lo = v;
        double next = (vega > 0.0) ? v - err / vega : 0.5 * (lo + hi);
        if (!(next > lo && next < hi)) next = 0.5 * (lo + hi);
        if (std::fabs(next - v) <= 1e-15 * std::max(v, 1e-300) || hi - lo <= 1e-15 * std::max(hi, 1e-300)) return next;
        v = next;
    }
    return v;
}
} // namespace

double blackScholesGeneralizedOptionValue(double F, double vol, double T, double K, double pu) {
    if (T < 0) return 0.0;
    if (F <= 0 || K <= 0 || vol <= 0 || T == 0) return std::max(F - K, 0.0) * pu;
    const double sd = vol * std::sqrt(T);
    const double dp = (std::log(F / K) + 0.5 * vol * vol * T) / sd, dm = dp - sd;
    return (F * Phi(dp) - K * Phi(dm)) * pu;
}

double bachelierOptionValue(double F, double vol, double T, double K, double pu) {
    if (T < 0) return 0.0;
    if (vol <= 0 || T == 0) return std::max(F - K, 0.0) * pu;
    const double sd = vol * std::sqrt(T), d = (F - K) / sd;
    return ((F - K) * Phi(d) + sd * phi(d)) * pu;
}

double blackScholesOptionImpliedVolatility(double F, double T, double K, double pu, double value) {
    if (K <= 0.0) return 0.0;
    if (!(value > std::max(F - K, 0.0) * pu) || !(value < F * pu) || T <= 0) return NAN;
    double hi = 1.0;
    while (blackScholesGeneralizedOptionValue(F, hi, T, K, pu) < value && hi < 1e6) hi *= 2.0;
    return invert([&](double v, double& val, double& vega) {
        val = blackScholesGeneralizedOptionValue(F, v, T, K, pu);
        const double dp = (std::log(F / K) + 0.5 * v * v * T) / (v * std::sqrt(T));
        vega = F * std::sqrt(T) * phi(dp) * pu;
    }, value, 0.0, hi);
}

double bachelierOptionImpliedVolatility(double F, double T, double K, double pu, double value) {
    if (!(value > std::max(F - K, 0.0) * pu) || T <= 0) return NAN;
    double hi = 0.01;
    while (bachelierOptionValue(F, hi, T, K, pu) < value && hi < 1e6) hi *= 2.0;
    return invert([&](double v, double& val, double& vega) {
        val = bachelierOptionValue(F, v, T, K, pu);
        vega = std::sqrt(T) * phi((F - K) / (v * std::sqrt(T))) * pu;
    }, value, 0.0, hi);
}
} // namespace AnalyticFormulas

BermudanSwaption::BermudanSwaption(std::vector<bool> exercise, std::vector<double> fix, std::vector<double> pl,
                                   std::vector<double> pay, std::vector<double> notionals, std::vector<double> k,
                                   bool callable, RegressionBasisFunctionsProviderPtr prov)
    : fixingDates(std::move(fix)), periodLengths(std::move(pl)), paymentDates(std::move(pay)),
      periodNotionals(std::move(notionals)), swaprates(std::move(k)), isCallable(callable), provider(std::move(prov)) {
    const size_t m = fixingDates.size();
    if (exercise.size() != m || periodLengths.size() != m || paymentDates.size() != m || swaprates.size() != m ||
        (!periodNotionals.empty() && periodNotionals.size() != m))
        throw IllegalArgumentException("Bermudan swaption arrays must have the same length");
    for (bool e : exercise) exerciseFlags_.push_back(e ? 1 : 0);
}

std::vector<double> BermudanSwaption::getExerciseTimes() const {
    std::vector<double> t;
    for (size_t i = 0; i < fixingDates.size(); i++) if (exerciseFlags_[i]) t.push_back(fixingDates[i]);
    return t;
}

void BermudanSwaption::lower(std::deque<std::vector<double>>& arena, fm_product& o) const {
    std::memset(&o, 0, sizeof(o));
    o.type = FM_PRODUCT_BERMUDAN_SWAPTION;
    o.num_periods = (int32_t)fixingDates.size();
    o.fixing_dates = keep(arena, fixingDates);
    o.payment_dates = keep(arena, paymentDates);
    o.period_lengths = keep(arena, periodLengths);
    o.swaprates = keep(arena, swaprates);
    o.notionals = periodNotionals.empty() ? nullptr : keep(arena, periodNotionals);
    o.exercise_flags = exerciseFlags_.data();
    o.is_cancelable = isCallable ? 0 : 1;
    o.basis = provider ? provider->basisId() : FM_BASIS_REFERENCE;
    o.notional = 1.0;
}

// ---- optimizer/LevenbergMarquardt.java -------------------------------------------------------
LevenbergMarquardt::LevenbergMarquardt(std::vector<double> initialParameters, std::vector<double> targetValues,
                                       int maxIteration)
    : initial_(std::move(initialParameters)), targets_(std::move(targetValues)), maxIteration_(maxIteration) {
    weights_.assign(targets_.size(), 1.0);
}

LevenbergMarquardt& LevenbergMarquardt::setWeights(std::vector<double> weights) {
    if (weights.size() != targets_.size()) throw IllegalArgumentException("one weight per target value");
    weights_ = std::move(weights);
    return *this;
}

LevenbergMarquardt& LevenbergMarquardt::setParameterSteps(std::vector<double> parameterSteps) {
    if (parameterSteps.size() != initial_.size()) throw IllegalArgumentException("one step per parameter");
    steps_ = std::move(parameterSteps);
    return *this;
}

LevenbergMarquardt& LevenbergMarquardt::setErrorTolerance(double errorTolerance) {
    tolerance_ = errorTolerance;
    return *this;
}

double LevenbergMarquardt::meanSquaredError(const std::vector<double>& v) const {       // :733-742
    double e = 0.0;
    for (size_t k = 0; k < v.size(); k++) {
        const double dev = v[k] - targets_[k];
        e += weights_[k] * dev * dev;
    }
    return e / (double)v.size();
}

LevenbergMarquardt& LevenbergMarquardt::setNumberOfThreads(int numberOfThreads) {
    if (numberOfThreads < 1) throw IllegalArgumentException("numberOfThreads must be positive");
    threads_ = numberOfThreads;
    return *this;
}

void LevenbergMarquardt::run(const ObjectiveFunction& setValues) {
    const int threads = threads_;
    threads_ = 1;
    try {
        run(WorkerObjectiveFunction([&](const std::vector<double>& p, std::vector<double>& v, int) { setValues(p, v); }));
    } catch (...) { threads_ = threads; throw; }
    threads_ = threads;
}

void LevenbergMarquardt::run(const WorkerObjectiveFunction& evaluate) {                    // :640-731
    const ObjectiveFunction setValues = [&](const std::vector<double>& p, std::vector<double>& v) { evaluate(p, v, 0); };
    const int np = (int)initial_.size(), nv = (int)targets_.size();
    std::vector<double> parameterTest = initial_, valueTest((size_t)nv), valueCurrent((size_t)nv);
    std::vector<std::vector<double>> derivative((size_t)np, std::vector<double>((size_t)nv));
    parameterCurrent_ = initial_;
    errorMeanSquaredCurrent_ = INFINITY;
    double rmsChange = INFINITY;
    bool derivativeValid = false;
    iteration_ = 0;
    while (true) {
        iteration_++;
        setValues(parameterTest, valueTest);
        const double errorTest = meanSquaredError(valueTest);
        if (errorTest < errorMeanSquaredCurrent_) {                 // false for NaN: the point is rejected
            rmsChange = std::sqrt(errorMeanSquaredCurrent_) - std::sqrt(errorTest);
            parameterCurrent_ = parameterTest;
            valueCurrent = valueTest;
            errorMeanSquaredCurrent_ = errorTest;
            derivativeValid = false;
            lambda_ /= lambdaDivisor_;
        } else {
            rmsChange = std::sqrt(errorTest) - std::sqrt(errorMeanSquaredCurrent_);
            lambda_ *= lambdaMultiplicator_;
        }
        // done(): :623-637
        if (iteration_ > maxIteration_ || rmsChange <= tolerance_ || std::isinf(lambda_)) break;
        // updateParameterTest(): :749-815
        if (!derivativeValid) {                                     // forward differences (:550-620)
            auto column = [&](int i, int worker) {
                std::vector<double> p = parameterCurrent_;
                const double h = steps_.empty() ? (std::fabs(p[(size_t)i]) + 1) * 1e-8 : steps_[(size_t)i];
                p[(size_t)i] += h;
                std::vector<double>& d = derivative[(size_t)i];
                evaluate(p, d, worker);
                for (int k = 0; k < nv; k++) {
                    d[(size_t)k] = (d[(size_t)k] - valueCurrent[(size_t)k]) / h;
                    if (std::isnan(d[(size_t)k])) d[(size_t)k] = 0.0;
                }
            };
            const int workers = std::min(threads_, np);
            if (workers <= 1) {
                for (int i = 0; i < np; i++) column(i, 0);
            } else {
                // worker w takes the columns i = w, w + workers, ...; each column is independent
                std::vector<std::thread> pool;
                std::vector<std::exception_ptr> failed((size_t)workers);
                for (int w = 0; w < workers; w++)
                    pool.emplace_back([&, w] {
                        try {
                            for (int i = w; i < np; i += workers) column(i, w);
                        } catch (...) { failed[(size_t)w] = std::current_exception(); }
                    });
                for (auto& t : pool) t.join();
                for (auto& e : failed)
                    if (e) std::rethrow_exception(e);
            }
            derivativeValid = true;
        }
        std::vector<double> increment((size_t)np, 0.0);
        bool invalid = true;
        while (invalid && std::isfinite(lambda_)) {
            invalid = false;
            std::vector<double> H((size_t)np * np), beta((size_t)np);
            for (int i = 0; i < np; i++) {
                for (int j = i; j < np; j++) {
                    double h = 0.0;
                    for (int k = 0; k < nv; k++) h += weights_[(size_t)k] * derivative[(size_t)i][(size_t)k] * derivative[(size_t)j][(size_t)k];
                    if (i == j) h = (h == 0.0) ? lambda_ : h * (1 + lambda_);   // LEVENBERG_MARQUARDT regularisation
                    H[(size_t)i * np + j] = H[(size_t)j * np + i] = h;
                }
                double b = 0.0;
                for (int k = 0; k < nv; k++)
                    b += weights_[(size_t)k] * (targets_[(size_t)k] - valueCurrent[(size_t)k]) * derivative[(size_t)i][(size_t)k];
                beta[(size_t)i] = b;
            }
            // pseudo-inverse solve (LinearAlgebra.solveLinearEquationSVD): eigen-decomposition of the
            // symmetric H, singular values below max(n sigma_max 2^-52, sqrt(2^-1022)) dropped
            std::vector<double> V = H, ev;
            symmetric_eigen(V, np, ev);
            double smax = 0.0;
            for (double e : ev) smax = std::max(smax, std::fabs(e));
            const double tol = std::max(np * smax * 0x1.0p-52, std::sqrt(0x1.0p-1022));
            for (int j = 0; j < np; j++) {
                if (!(std::fabs(ev[(size_t)j]) > tol)) continue;
                double vb = 0.0;
                for (int i = 0; i < np; i++) vb += V[(size_t)i * np + j] * beta[(size_t)i];
                for (int i = 0; i < np; i++) increment[(size_t)i] += V[(size_t)i * np + j] * (vb / ev[(size_t)j]);
            }
            for (double x : increment)
                if (!std::isfinite(x)) invalid = true;
            if (invalid) {
                lambda_ *= 16;
                std::fill(increment.begin(), increment.end(), 0.0);
            }
        }
        for (int i = 0; i < np; i++) parameterTest[(size_t)i] = parameterCurrent_[(size_t)i] + increment[(size_t)i];
    }
}

// ---- calibration of the covariance model -----------------------------------------------------
std::vector<double> getParameter(const LIBORCovarianceModelFromVolatilityAndCorrelation& cov) {
    if (auto* w = dynamic_cast<const LIBORCovarianceModelStochasticHestonVolatility*>(&cov)) {   // Heston:84-104
        std::vector<double> p = getParameter(*w->getBaseCovarianceModel());
        if (w->isCalibrateable()) { p.push_back(w->getKappa()); p.push_back(w->getTheta()); p.push_back(w->getXi()); }
        return p;
    }
    if (auto* w = dynamic_cast<const LIBORCovarianceModelStochasticVolatility*>(&cov)) {   // StochasticVolatility.java:103-121
        std::vector<double> p = getParameter(*w->getBaseCovarianceModel());
        if (w->isCalibrateable()) { p.push_back(w->getNu()); p.push_back(w->getRho()); }
        return p;
    }
    if (auto* w = dynamic_cast<const DisplacedLocalVolatilityModel*>(&cov)) {   // DisplacedLocalVolatilityModel.java:121-137
        std::vector<double> p = getParameter(*w->getBaseCovarianceModel());
        if (w->isCalibrateable()) p.push_back(w->getDisplacement());
        return p;
    }
    if (auto* w = dynamic_cast<const BlendedLocalVolatilityModel*>(&cov)) {     // BlendedLocalVolatilityModel.java:211-228
        std::vector<double> p = getParameter(*w->getBaseCovarianceModel());
        if (w->isCalibrateable()) p.push_back(w->getDisplacement());
        return p;
    }
    std::vector<double> p = cov.getVolatilityModel()->getParameter();
    if (cov.getCorrelationModel()->isCalibrateable()) p.push_back(cov.getCorrelationModel()->getParameter());
    return p;
}

CovarianceModelPtr getCloneWithModifiedParameters(const LIBORCovarianceModelFromVolatilityAndCorrelation& cov,
                                                  const std::vector<double>& parameters) {
    // wrappers: the base model takes the leading parameters, a calibrateable displacement or
    // blend the last one (DisplacedLocalVolatilityModel.java:140-166, BlendedLocalVolatilityModel.java:240-257)
    auto split = [&](bool calibrateable, double current, std::vector<double>& base) {
        if (!calibrateable) { base = parameters; return current; }
        if (parameters.empty()) throw IllegalArgumentException("the displacement parameter is missing");
        base.assign(parameters.begin(), parameters.end() - 1);
        return parameters.back();
    };
    if (auto* w = dynamic_cast<const LIBORCovarianceModelStochasticHestonVolatility*>(&cov)) {   // Heston:107-164
        std::vector<double> base = parameters;
        double kappa = w->getKappa(), theta = w->getTheta(), xi = w->getXi();
        if (w->isCalibrateable()) {
            if (parameters.size() < 3) throw IllegalArgumentException("kappa, theta and xi are missing");
            base.assign(parameters.begin(), parameters.end() - 3);
            kappa = parameters[parameters.size() - 3];
            theta = parameters[parameters.size() - 2];
            xi = parameters[parameters.size() - 1];
        }
        return std::make_shared<LIBORCovarianceModelStochasticHestonVolatility>(
            getCloneWithModifiedParameters(*w->getBaseCovarianceModel(), base), w->getBrownianMotion(), kappa, theta, xi,
            w->isCalibrateable());
    }
    if (auto* w = dynamic_cast<const LIBORCovarianceModelStochasticVolatility*>(&cov)) {   // :123-160
        std::vector<double> base = parameters;
        double nu = w->getNu(), rho = w->getRho();
        if (w->isCalibrateable()) {
            if (parameters.size() < 2) throw IllegalArgumentException("nu and rho are missing");
            base.assign(parameters.begin(), parameters.end() - 2);
            nu = parameters[parameters.size() - 2];
            rho = parameters[parameters.size() - 1];
        }
        return std::make_shared<LIBORCovarianceModelStochasticVolatility>(
            getCloneWithModifiedParameters(*w->getBaseCovarianceModel(), base), w->getBrownianMotion(), nu, rho,
            w->isCalibrateable());
    }
    if (auto* w = dynamic_cast<const DisplacedLocalVolatilityModel*>(&cov)) {
        std::vector<double> base;
        const double d = split(w->isCalibrateable(), w->getDisplacement(), base);
        return std::make_shared<DisplacedLocalVolatilityModel>(getCloneWithModifiedParameters(*w->getBaseCovarianceModel(), base),
                                                               d, w->isCalibrateable());
    }
    if (auto* w = dynamic_cast<const BlendedLocalVolatilityModel*>(&cov)) {
        std::vector<double> base;
        const double a = split(w->isCalibrateable(), w->getDisplacement(), base);
        return std::make_shared<BlendedLocalVolatilityModel>(getCloneWithModifiedParameters(*w->getBaseCovarianceModel(), base),
                                                             w->getForwardCurve(), a, w->isCalibrateable());
    }
    size_t used = 0;
    VolatilityModelPtr vol = cov.getVolatilityModel();
    auto four = std::dynamic_pointer_cast<const LIBORVolatilityModelFourParameterExponentialForm>(vol);
    if (four && four->isCalibrateable) {
        if (parameters.size() < 4) throw IllegalArgumentException("four volatility parameters expected");
        vol = std::make_shared<LIBORVolatilityModelFourParameterExponentialForm>(
            cov.getTimeDiscretization(), cov.getLiborPeriodDiscretization(), parameters[0], parameters[1], parameters[2],
            parameters[3], true);
        used = 4;
    } else if (auto pw = std::dynamic_pointer_cast<const LIBORVolatilityModelPiecewiseConstant>(vol);
               pw && pw->isCalibrateable()) {
        const size_t np = pw->getParameter().size();
        if (parameters.size() < np) throw IllegalArgumentException("piecewise-constant volatilities expected");
        vol = std::make_shared<LIBORVolatilityModelPiecewiseConstant>(
            cov.getTimeDiscretization(), cov.getLiborPeriodDiscretization(), pw->getSimulationTimeDiscretization(),
            pw->getTimeToMaturityDiscretization(), std::vector<double>(parameters.begin(), parameters.begin() + np), true);
        used = np;
    } else if (!vol->getParameter().empty()) {
        throw UnsupportedOperationException("only the four-parameter and piecewise-constant volatilities are calibrated");
    }
    CorrelationModelPtr corr = cov.getCorrelationModel();
    if (corr->isCalibrateable()) {
        if (parameters.size() < used + 1) throw IllegalArgumentException("correlation parameter expected");
        corr = std::make_shared<LIBORCorrelationModelExponentialDecay>(cov.getTimeDiscretization(),
                                                                      cov.getLiborPeriodDiscretization(),
                                                                      corr->getNumberOfFactors(), parameters[used], true);
        used++;
    }
    if (used != parameters.size()) throw IllegalArgumentException("number of parameters does not match the model");
    return std::make_shared<LIBORCovarianceModelFromVolatilityAndCorrelation>(cov.getTimeDiscretization(),
                                                                             cov.getLiborPeriodDiscretization(), vol, corr);
}

CalibrationResult getCloneCalibrated(const LIBORCovarianceModelFromVolatilityAndCorrelation& cov,
                                     LIBORMarketModelPtr model, const std::vector<CalibrationProduct>& products,
                                     const std::map<std::string, double>& params) {
    auto opt = [&](const char* k, double def) { auto it = params.find(k); return it == params.end() ? def : it->second; };
    const int numberOfPaths = (int)opt("numberOfPaths", 2000);
    const int seed = (int)opt("seed", 31415);
    const int maxIterations = (int)opt("maxIterations", 400);
    const double accuracy = opt("accuracy", 1e-7);
    const double parameterStep = opt("parameterStep", 1e-4);
    const int precision = (int)opt("precision", FM_PRECISION_FP64);
    const int numberOfThreads = (int)opt("numberOfThreads", 4);
    if (!model) throw IllegalArgumentException("calibration model must be given");
    if (products.empty()) throw IllegalArgumentException("no calibration products");
    const std::vector<double> initial = getParameter(cov);
    if (initial.empty()) throw IllegalArgumentException("the covariance model has no calibrateable parameters");

    // one context, one Brownian motion on the covariance model's time discretization
    const auto td = cov.getTimeDiscretization();
    const auto& times = td->getAsDoubleArray();
    const auto& tenor = model->getLiborPeriodDiscretization()->getAsDoubleArray();
    fm_lmm_config cfg{};
    cfg.num_time_steps = (int32_t)times.size() - 1;
    cfg.times = times.data();
    cfg.num_libors = (int32_t)tenor.size() - 1;
    cfg.tenor = tenor.data();
    cfg.num_factors = cov.getNumberOfFactors();
    cfg.num_paths = numberOfPaths;
    cfg.seed = seed;
    cfg.measure = model->measure;
    cfg.state_space = model->stateSpace;
    cfg.scheme = FM_SCHEME_EULER_FUNCTIONAL;
    cfg.precision = precision;
    cfg.libor_cap = model->liborCap;
    cfg.regenerate_increments = 0;
    // one context per worker: the same seed gives every worker the same Brownian increments, and
    // each context has its own stream, so the Jacobian's passes run concurrently on the device
    if (numberOfThreads < 1) throw IllegalArgumentException("numberOfThreads must be positive");
    const int workers = std::min(numberOfThreads, (int)initial.size());
    std::vector<std::unique_ptr<fm_lmm, void (*)(fm_lmm*)>> contexts;
    const std::vector<double> L0 = model->getInitialForwards();
    std::vector<double> pd;
    if (model->getDiscountCurve()) {
        pd.resize(tenor.size());
        for (size_t k = 0; k < tenor.size(); k++) pd[k] = model->getDiscountCurve()->getDiscountFactor(tenor[k]);
    }
    for (int w = 0; w < workers; w++) {
        fm_lmm* ctx = nullptr;
        check(fm_lmm_create(&cfg, &ctx));
        contexts.emplace_back(ctx, fm_lmm_destroy);
        check(fm_lmm_set_initial_forwards(ctx, L0.data()));
        if (!pd.empty()) check(fm_lmm_set_discount_curve(ctx, pd.data()));
    }
    std::deque<std::vector<double>> arena;
    std::vector<fm_product> lowered(products.size());
    for (size_t q = 0; q < products.size(); q++) {
        if (!products[q].product) throw IllegalArgumentException("calibration product is null");
        products[q].product->lower(arena, lowered[q]);
    }
    std::atomic<int> valuations{0};
    auto setValues = [&](const std::vector<double>& p, std::vector<double>& out, int worker) {
        fm_lmm* ctx = contexts[(size_t)worker].get();
        std::vector<double> lam, var, lvOffsets, vals(products.size());
        out.assign(products.size(), NAN);
        try {
            const CovarianceModelPtr c = getCloneWithModifiedParameters(cov, p);
            c->tables(*td, lam, var);
            check(fm_lmm_set_covariance(ctx, lam.data(), var.data()));
            double lvAlpha = 0.0;
            const bool lv = c->localVolatility(*td, lvAlpha, lvOffsets);
            check(fm_lmm_set_local_volatility(ctx, lvAlpha, lv ? lvOffsets.data() : nullptr));
            applyStochasticVolatility(ctx, *c, tenor, numberOfPaths);
            check(fm_lmm_value_products(ctx, lowered.data(), (int32_t)lowered.size(), vals.data(), nullptr));
            valuations++;
            for (size_t q = 0; q < products.size(); q++)
                out[q] = (products[q].product->toValueUnit(vals[q], *model) - products[q].targetValue) * products[q].weight;
        } catch (const CalculationException&) {
            // a failed valuation (e.g. a non-finite price) is a rejected point
        } catch (const IllegalArgumentException&) {
            // parameters whose tables are not finite: rejected as well
        }
    };
    LevenbergMarquardt lm(initial, std::vector<double>(products.size(), 0.0), maxIterations);
    lm.setErrorTolerance(accuracy).setParameterSteps(std::vector<double>(initial.size(), parameterStep));
    lm.setNumberOfThreads(workers);
    lm.run(LevenbergMarquardt::WorkerObjectiveFunction(setValues));
    CalibrationResult r;
    r.parameters = lm.getBestFitParameters();
    r.covarianceModel = getCloneWithModifiedParameters(cov, r.parameters);
    r.iterations = lm.getIterations();
    r.rootMeanSquaredError = lm.getRootMeanSquaredError();
    r.valuations = valuations.load();
    return r;
}

} // namespace finmath

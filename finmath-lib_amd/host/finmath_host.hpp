// Host-side mirror of the finmath-lib API for the LMM hot path, implemented over
// the C ABI of include/finmath_lmm.h. Class and method names follow the reference
// (paths relative to src/main/java/net/finmath/), so code written against the
// reference reads the same against this library.
//
// What runs where: the model set-up (time grids, curves, volatility/correlation
// tables, PCA) runs on the host as in the reference; path generation, the Euler
// scheme, payoffs and their Monte Carlo reductions run on the GPU.
#pragma once

#include <cmath>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "finmath_lmm.h"

namespace finmath {

// Exceptions of the reference, raised from ABI status codes.
struct CalculationException : std::runtime_error { using std::runtime_error::runtime_error; };
struct IllegalArgumentException : std::invalid_argument { using std::invalid_argument::invalid_argument; };
struct IllegalStateException : std::logic_error { using std::logic_error::logic_error; };
struct UnsupportedOperationException : std::logic_error { using std::logic_error::logic_error; };
void check(int status);

// ---- time/TimeDiscretizationFromArray.java -----------------------------------
class TimeDiscretizationFromArray {
public:
    static constexpr double kTickSize = 1.0 / (365.0 * 24.0);
    TimeDiscretizationFromArray(double initial, int numberOfTimeSteps, double deltaT);   // :210-217
    explicit TimeDiscretizationFromArray(std::vector<double> times);
    int getNumberOfTimes() const { return (int)t_.size(); }
    int getNumberOfTimeSteps() const { return (int)t_.size() - 1; }
    double getTime(int i) const { return t_.at((size_t)i); }
    double getTimeStep(int i) const { return t_.at((size_t)i + 1) - t_.at((size_t)i); }
    int getTimeIndex(double time) const;                                                // :272-274
    int getTimeIndexNearestLessOrEqual(double time) const;
    const std::vector<double>& getAsDoubleArray() const { return t_; }
    static double roundToTimeTickSize(double t) { return std::rint(t / kTickSize) * kTickSize; }
private:
    std::vector<double> t_;
};
using TimeDiscretizationPtr = std::shared_ptr<const TimeDiscretizationFromArray>;

// ---- marketdata/model/curves --------------------------------------------------
// ForwardCurveInterpolation: LINEAR interpolation of the forward VALUE with CONSTANT
// extrapolation (marketdata/model/curves/ForwardCurveInterpolation.java:224-234,298-347).
class ForwardCurveInterpolation {
public:
    static std::shared_ptr<ForwardCurveInterpolation> createForwardCurveFromForwards(
        const std::string& name, const std::vector<double>& times, const std::vector<double>& forwards,
        double paymentOffset);
    double getForward(double fixingTime) const;
    double getForward(double fixingTime, double paymentOffset) const;
    double getPaymentOffset(double) const { return paymentOffset_; }
    const std::string& getName() const { return name_; }
private:
    std::string name_;
    std::vector<double> t_, f_;
    double paymentOffset_ = 0.0;
};
using ForwardCurvePtr = std::shared_ptr<const ForwardCurveInterpolation>;

// DiscountCurveFromForwardCurve (marketdata/model/curves/DiscountCurveFromForwardCurve.java:100-143).
class DiscountCurveFromForwardCurve {
public:
    explicit DiscountCurveFromForwardCurve(ForwardCurvePtr forwardCurve) : fc_(std::move(forwardCurve)) {}
    double getDiscountFactor(double maturity) const;
    ForwardCurvePtr getForwardCurve() const { return fc_; }
private:
    ForwardCurvePtr fc_;
};
using DiscountCurvePtr = std::shared_ptr<const DiscountCurveFromForwardCurve>;

// Single-curve par swap rate (marketdata/products/Swap.java:146-174).
double getForwardSwapRate(const TimeDiscretizationFromArray& fixTenor, const TimeDiscretizationFromArray& floatTenor,
                          const ForwardCurveInterpolation& forwardCurve);

// ---- montecarlo/interestrate/models/covariance -----------------------------------
class LIBORVolatilityModel {
public:
    LIBORVolatilityModel(TimeDiscretizationPtr td, TimeDiscretizationPtr tenor) : td_(std::move(td)), tenor_(std::move(tenor)) {}
    virtual ~LIBORVolatilityModel() = default;
    virtual double getVolatility(int timeIndex, int liborIndex) const = 0;
    virtual std::vector<double> getParameter() const { return {}; }
    TimeDiscretizationPtr getTimeDiscretization() const { return td_; }
    TimeDiscretizationPtr getLiborPeriodDiscretization() const { return tenor_; }
protected:
    TimeDiscretizationPtr td_, tenor_;
};
using VolatilityModelPtr = std::shared_ptr<const LIBORVolatilityModel>;

// sigma_i(t_j) = (a + b (T_i - t_j)) exp(-c (T_i - t_j)) + d
// (covariance/LIBORVolatilityModelFourParameterExponentialForm.java:168-195).
class LIBORVolatilityModelFourParameterExponentialForm : public LIBORVolatilityModel {
public:
    LIBORVolatilityModelFourParameterExponentialForm(TimeDiscretizationPtr td, TimeDiscretizationPtr tenor,
                                                     double a, double b, double c, double d, bool isCalibrateable);
    double getVolatility(int timeIndex, int liborIndex) const override;
    std::vector<double> getParameter() const override;
    double a, b, c, d;
    bool isCalibrateable;
};

// Volatility table given by a matrix (covariance/LIBORVolatilityModelFromGivenMatrix.java).
class LIBORVolatilityModelFromGivenMatrix : public LIBORVolatilityModel {
public:
    LIBORVolatilityModelFromGivenMatrix(TimeDiscretizationPtr td, TimeDiscretizationPtr tenor,
                                        std::vector<double> volatility /* [nT][n] */);
    double getVolatility(int timeIndex, int liborIndex) const override;
    std::vector<double> getParameter() const override { return vol_; }
private:
    std::vector<double> vol_;
};

// Volatility constant on buckets of (simulation time, time to maturity): the parameter of the
// buckets starting at the nearest grid points at or below t_i and T_j - t_i; 0 once fixed; one
// parameter per bucket pair whose start lies within the last tenor date
// (covariance/LIBORVolatilityModelPiecewiseConstant.java:36-76,113-195).
class LIBORVolatilityModelPiecewiseConstant : public LIBORVolatilityModel {
public:
    LIBORVolatilityModelPiecewiseConstant(TimeDiscretizationPtr td, TimeDiscretizationPtr tenor,
                                          TimeDiscretizationPtr simulationTimes, TimeDiscretizationPtr timesToMaturity,
                                          std::vector<double> volatility, bool isCalibrateable = true);
    double getVolatility(int timeIndex, int liborIndex) const override;
    std::vector<double> getParameter() const override { return calibrateable_ ? vol_ : std::vector<double>(); }
    TimeDiscretizationPtr getSimulationTimeDiscretization() const { return sim_; }
    TimeDiscretizationPtr getTimeToMaturityDiscretization() const { return ttm_; }
    bool isCalibrateable() const { return calibrateable_; }
private:
    TimeDiscretizationPtr sim_, ttm_;
    std::vector<std::vector<int>> index_;     // [simulation bucket][maturity bucket] -> parameter, -1 none
    std::vector<double> vol_;
    bool calibrateable_;
};

// rho_ij = exp(-a |T_i - T_j|), PCA-reduced to m factors
// (covariance/LIBORCorrelationModelExponentialDecay.java:102-134, functions/LinearAlgebra.java:380-482).
class LIBORCorrelationModelExponentialDecay {
public:
    LIBORCorrelationModelExponentialDecay(TimeDiscretizationPtr td, TimeDiscretizationPtr tenor, int numberOfFactors,
                                          double a, bool isCalibrateable = false);
    double getFactorLoading(int timeIndex, int factor, int component) const { return factor_[(size_t)component * m_ + factor]; }
    double getCorrelation(int, int c1, int c2) const { return corr_[(size_t)c1 * n_ + c2]; }
    int getNumberOfFactors() const { return m_; }
    double getParameter() const { return a_; }
    bool isCalibrateable() const { return calibrateable_; }
    const std::vector<double>& factorMatrix() const { return factor_; }
private:
    int n_, m_;
    double a_;
    bool calibrateable_ = false;
    std::vector<double> factor_, corr_;
};
using CorrelationModelPtr = std::shared_ptr<const LIBORCorrelationModelExponentialDecay>;

// PCA of a symmetric matrix with the reference's sign/scale conventions, returned
// as an n x m row-major factor matrix (functions/LinearAlgebra.java:392-482).
std::vector<double> factorReduction(const std::vector<double>& correlation, int n, int numberOfFactors);

struct BrownianMotionFromMersenneRandomNumbers;

// lambda_jk(t_i) = sigma_j(t_i) f_jk; covariance sigma_j sigma_l rho_jl
// (covariance/LIBORCovarianceModelFromVolatilityAndCorrelation.java:48-93).
class LIBORCovarianceModelFromVolatilityAndCorrelation {
public:
    LIBORCovarianceModelFromVolatilityAndCorrelation(TimeDiscretizationPtr td, TimeDiscretizationPtr tenor,
                                                     VolatilityModelPtr vol, CorrelationModelPtr corr)
        : td_(std::move(td)), tenor_(std::move(tenor)), vol_(std::move(vol)), corr_(std::move(corr)) {}
    virtual ~LIBORCovarianceModelFromVolatilityAndCorrelation() = default;
    std::vector<double> getFactorLoading(int timeIndex, int component) const;
    double getCovariance(int timeIndex, int component1, int component2) const;
    int getNumberOfFactors() const { return corr_->getNumberOfFactors(); }
    TimeDiscretizationPtr getTimeDiscretization() const { return td_; }
    TimeDiscretizationPtr getLiborPeriodDiscretization() const { return tenor_; }
    VolatilityModelPtr getVolatilityModel() const { return vol_; }
    CorrelationModelPtr getCorrelationModel() const { return corr_; }
    // tables for the engine: factor loading [nT][n][F], variance [nT][n] on the simulation grid
    void tables(const TimeDiscretizationFromArray& simulationTimes, std::vector<double>& factorLoading,
                std::vector<double>& variance) const;
    // Local-volatility factor of a wrapper for the engine, on the simulation grid: the
    // loadings of component j at simulation time index i are multiplied by
    // (L_j - L_j alpha) + offsets[i n + j] (L_j + offsets[i n + j] for alpha = 0). False
    // for a model without wrapper.
    virtual bool localVolatility(const TimeDiscretizationFromArray& simulationTimes, double& alpha,
                                 std::vector<double>& offsets) const;
    // Stochastic volatility scaling of a wrapper: its parameters and Brownian motion. False for
    // a model without scaling.
    virtual bool stochasticVolatility(double& nu, double& rho,
                                      std::shared_ptr<const BrownianMotionFromMersenneRandomNumbers>& brownianMotion) const {
        return false;
    }
    // Heston variance scaling of a wrapper: its parameters and Brownian motion. False without.
    virtual bool hestonVolatility(double& kappa, double& theta, double& xi,
                                  std::shared_ptr<const BrownianMotionFromMersenneRandomNumbers>& brownianMotion) const {
        return false;
    }
protected:
    // round-down index of a simulation time on the covariance model's grid
    // (AbstractLIBORCovarianceModel.getFactorLoading(double time, ...), :54-60)
    int covarianceTimeIndex(double time) const;
private:
    TimeDiscretizationPtr td_, tenor_;
    VolatilityModelPtr vol_;
    CorrelationModelPtr corr_;
};
using CovarianceModelPtr = std::shared_ptr<const LIBORCovarianceModelFromVolatilityAndCorrelation>;

// Factor loadings of the base model times L_j(t) + d
// (covariance/DisplacedLocalVolatilityModel.java:67-71,175-186); the parameters are the base
// model's, followed by the displacement when it is calibrateable (:120-166).
class DisplacedLocalVolatilityModel : public LIBORCovarianceModelFromVolatilityAndCorrelation {
public:
    DisplacedLocalVolatilityModel(CovarianceModelPtr covarianceModel, double displacement, bool isCalibrateable);
    CovarianceModelPtr getBaseCovarianceModel() const { return base_; }
    double getDisplacement() const { return displacement_; }
    bool isCalibrateable() const { return calibrateable_; }
    bool localVolatility(const TimeDiscretizationFromArray& simulationTimes, double& alpha,
                         std::vector<double>& offsets) const override;
private:
    CovarianceModelPtr base_;
    double displacement_;
    bool calibrateable_;
};

// Factor loadings of the base model times (L_j(t) - L_j(t) a) + a f0(max(T_j - t, 0)), with
// f0 the given forward curve (1 without curve) (covariance/BlendedLocalVolatilityModel.java:
// 79-86,265-283); the parameters are the base model's, followed by the blend a when it is
// calibrateable (:211-262; the clone leaves this model unchanged).
class BlendedLocalVolatilityModel : public LIBORCovarianceModelFromVolatilityAndCorrelation {
public:
    BlendedLocalVolatilityModel(CovarianceModelPtr covarianceModel, ForwardCurvePtr forwardCurve, double displacement,
                                bool isCalibrateable);
    CovarianceModelPtr getBaseCovarianceModel() const { return base_; }
    ForwardCurvePtr getForwardCurve() const { return fc_; }
    double getDisplacement() const { return displacement_; }
    bool isCalibrateable() const { return calibrateable_; }
    bool localVolatility(const TimeDiscretizationFromArray& simulationTimes, double& alpha,
                         std::vector<double>& offsets) const override;
private:
    CovarianceModelPtr base_;
    ForwardCurvePtr fc_;
    double displacement_;
    bool calibrateable_;
};

// ---- montecarlo/interestrate/models/LIBORMarketModelFromCovarianceModel.java -----
class LIBORMarketModelFromCovarianceModel {
public:
    // LIBORMarketModelFromCovarianceModel.of(...) (LMM:283-322); properties: measure,
    // stateSpace, liborCap (LMM:400-414). Calibration products are not supported.
    static std::shared_ptr<LIBORMarketModelFromCovarianceModel> of(
        TimeDiscretizationPtr liborPeriodDiscretization, ForwardCurvePtr forwardRateCurve,
        DiscountCurvePtr discountCurve, CovarianceModelPtr covarianceModel,
        const std::map<std::string, std::string>& properties = {});
    TimeDiscretizationPtr getLiborPeriodDiscretization() const { return tenor_; }
    ForwardCurvePtr getForwardRateCurve() const { return fc_; }
    DiscountCurvePtr getDiscountCurve() const { return dc_; }
    CovarianceModelPtr getCovarianceModel() const { return cov_; }
    int getNumberOfComponents() const { return tenor_->getNumberOfTimeSteps(); }
    int getNumberOfLibors() const { return tenor_->getNumberOfTimeSteps(); }
    double getLiborPeriod(int i) const { return tenor_->getTime(i); }
    int getLiborPeriodIndex(double time) const { return tenor_->getTimeIndex(time); }
    std::vector<double> getInitialForwards() const;             // L_j(0) (LMM:1081-1093)
    int measure = FM_MEASURE_SPOT;
    int stateSpace = FM_STATE_SPACE_LOGNORMAL;
    double liborCap = 1e5;
private:
    TimeDiscretizationPtr tenor_;
    ForwardCurvePtr fc_;
    DiscountCurvePtr dc_;
    CovarianceModelPtr cov_;
};
using LIBORMarketModelPtr = std::shared_ptr<const LIBORMarketModelFromCovarianceModel>;

// ---- montecarlo/BrownianMotionFromMersenneRandomNumbers.java ------------------------
struct BrownianMotionFromMersenneRandomNumbers {
    BrownianMotionFromMersenneRandomNumbers(TimeDiscretizationPtr td, int numberOfFactors, int numberOfPaths, int seed);
    TimeDiscretizationPtr getTimeDiscretization() const { return td; }
    int getNumberOfFactors() const { return numberOfFactors; }
    int getNumberOfPaths() const { return numberOfPaths; }
    int getSeed() const { return seed; }
    TimeDiscretizationPtr td;
    int numberOfFactors, numberOfPaths, seed;
};
using BrownianMotionPtr = std::shared_ptr<const BrownianMotionFromMersenneRandomNumbers>;

// Factor loadings of the base model (plain or with a local-volatility wrapper) times S(t), the
// exponential of a functional Euler process with drift -nu^2/2 and loadings (rho nu,
// sqrt(1 - rho^2) nu) on the first two factors of the given Brownian motion
// (covariance/LIBORCovarianceModelStochasticVolatility.java:72-99,174-256); the parameters are the
// base model's, followed by nu and rho when calibrateable (:103-160).
class LIBORCovarianceModelStochasticVolatility : public LIBORCovarianceModelFromVolatilityAndCorrelation {
public:
    LIBORCovarianceModelStochasticVolatility(CovarianceModelPtr covarianceModel, BrownianMotionPtr brownianMotion,
                                             double nu, double rho, bool isCalibrateable);
    CovarianceModelPtr getBaseCovarianceModel() const { return base_; }
    BrownianMotionPtr getBrownianMotion() const { return bm_; }
    double getNu() const { return nu_; }
    double getRho() const { return rho_; }
    bool isCalibrateable() const { return calibrateable_; }
    bool localVolatility(const TimeDiscretizationFromArray& simulationTimes, double& alpha,
                         std::vector<double>& offsets) const override {
        return base_->localVolatility(simulationTimes, alpha, offsets);
    }
    bool stochasticVolatility(double& nu, double& rho, BrownianMotionPtr& brownianMotion) const override {
        nu = nu_; rho = rho_; brownianMotion = bm_;
        return true;
    }
private:
    CovarianceModelPtr base_;
    BrownianMotionPtr bm_;
    double nu_, rho_;
    bool calibrateable_;
};

// Factor loadings of the base model (plain or with a local-volatility wrapper) times
// sqrt(max(V(t), 0)), V a square-root variance process from 1 with drift (V - theta)(-kappa)
// and loading sqrt(max(V, 0)) xi on the first factor of the given Brownian motion
// (covariance/LIBORCovarianceModelStochasticHestonVolatility.java:60-82,172-256); the parameters
// are the base model's, followed by kappa, theta, xi when calibrateable (:84-164).
class LIBORCovarianceModelStochasticHestonVolatility : public LIBORCovarianceModelFromVolatilityAndCorrelation {
public:
    LIBORCovarianceModelStochasticHestonVolatility(CovarianceModelPtr covarianceModel, BrownianMotionPtr brownianMotion,
                                                   double kappa, double theta, double xi, bool isCalibrateable);
    CovarianceModelPtr getBaseCovarianceModel() const { return base_; }
    BrownianMotionPtr getBrownianMotion() const { return bm_; }
    double getKappa() const { return kappa_; }
    double getTheta() const { return theta_; }
    double getXi() const { return xi_; }
    bool isCalibrateable() const { return calibrateable_; }
    bool localVolatility(const TimeDiscretizationFromArray& simulationTimes, double& alpha,
                         std::vector<double>& offsets) const override {
        return base_->localVolatility(simulationTimes, alpha, offsets);
    }
    bool hestonVolatility(double& kappa, double& theta, double& xi, BrownianMotionPtr& brownianMotion) const override {
        kappa = kappa_; theta = theta_; xi = xi_; brownianMotion = bm_;
        return true;
    }
private:
    CovarianceModelPtr base_;
    BrownianMotionPtr bm_;
    double kappa_, theta_, xi_;
    bool calibrateable_;
};

// ---- a RandomVariable holding realisations on the host ------------------------------
// getAverage is the Kahan mean (montecarlo/RandomVariableFromDoubleArray.java:287-307).
class RandomVariable {
public:
    RandomVariable() = default;
    RandomVariable(double time, double value) : time_(time), det_(true), value_(value) {}
    RandomVariable(double time, std::vector<double> realizations) : time_(time), det_(false), r_(std::move(realizations)) {}
    bool isDeterministic() const { return det_; }
    int size() const { return det_ ? 1 : (int)r_.size(); }
    double get(int i) const { return det_ ? value_ : r_.at((size_t)i); }
    double getFiltrationTime() const { return time_; }
    double getAverage() const;
    double getVariance() const;
    double getStandardError() const;
    double doubleValue() const { return det_ ? value_ : getAverage(); }
    std::vector<double> getRealizations(int numberOfPaths) const {
        return det_ ? std::vector<double>((size_t)numberOfPaths, value_) : r_;
    }
private:
    double time_ = 0.0;
    bool det_ = true;
    double value_ = 0.0;
    std::vector<double> r_;
};

// ---- montecarlo/process/EulerSchemeFromProcessModel.java ------------------------------
class Product;

// Result of a pathwise adjoint sweep (RandomVariableDifferentiable.getGradient for the
// model leaves FORWARD(T_j) and the volatility parameters).
struct Gradient {
    double value = 0.0;
    std::vector<double> delta;            // dV/dL_j(0), j < n
    std::vector<double> volatility;       // dV/dsigma_j(t_i), [nT][n]
    std::vector<double> forwardVega;      // sum_i dV/dsigma_j(t_i) (parallel bump of sigma_j)
    std::vector<double> volatilityModel;  // dV/d(volatility model entry [c][j]), [nTc + 1][n] on its own time grid
    std::vector<double> parameterVega;    // dV/d(a,b,c,d) for the four-parameter form, else empty
};

class EulerSchemeFromProcessModel {
public:
    enum Scheme { EULER = FM_SCHEME_EULER, PREDICTOR_CORRECTOR = FM_SCHEME_PREDICTOR_CORRECTOR,
                  EULER_FUNCTIONAL = FM_SCHEME_EULER_FUNCTIONAL,
                  PREDICTOR_CORRECTOR_FUNCTIONAL = FM_SCHEME_PREDICTOR_CORRECTOR_FUNCTIONAL };
    // The default scheme for a model with an inverse state-space transform is
    // EULER_FUNCTIONAL (Euler:108-123). precision selects the GPU arithmetic.
    EulerSchemeFromProcessModel(LIBORMarketModelPtr model, BrownianMotionPtr driver, Scheme scheme = EULER_FUNCTIONAL,
                                int precision = FM_PRECISION_FP64);
    ~EulerSchemeFromProcessModel();
    EulerSchemeFromProcessModel(const EulerSchemeFromProcessModel&) = delete;
    EulerSchemeFromProcessModel& operator=(const EulerSchemeFromProcessModel&) = delete;

    RandomVariable getProcessValue(int timeIndex, int componentIndex);        // Euler:131-146
    RandomVariable getMonteCarloWeights(int) const { return RandomVariable(0.0, 1.0 / getNumberOfPaths()); }
    RandomVariable getBrownianIncrement(int timeIndex, int factor);
    RandomVariable getNumeraire(double time);
    int getNumberOfPaths() const { return driver_->getNumberOfPaths(); }
    int getNumberOfFactors() const { return driver_->getNumberOfFactors(); }
    int getNumberOfComponents() const { return model_->getNumberOfComponents(); }
    TimeDiscretizationPtr getTimeDiscretization() const { return driver_->getTimeDiscretization(); }
    double getTime(int i) const { return getTimeDiscretization()->getTime(i); }
    int getTimeIndex(double t) const { return getTimeDiscretization()->getTimeIndex(t); }
    LIBORMarketModelPtr getModel() const { return model_; }
    BrownianMotionPtr getStochasticDriver() const { return driver_; }
    Scheme getScheme() const { return scheme_; }
    int getPrecision() const { return precision_; }
    // 1: regenerate the Brownian increments on every GPU pass instead of keeping them
    void setRegenerateIncrements(bool on) { regenerate_ = on; if (ctx_) { fm_lmm_destroy(ctx_); ctx_ = nullptr; simulated_ = false; } }
    // Engine access: value products in one fused GPU pass; adjoint sweep.
    void valueProducts(const std::vector<std::shared_ptr<const Product>>& products, std::vector<double>& values,
                       std::vector<double>& errors);
    std::vector<double> lastTimings();
    Gradient getGradient(const std::vector<std::shared_ptr<const Product>>& products, const std::vector<double>& weights,
                         int checkpointStride = 4);
    // The derivative of a Bermudan's exercise indicator in getGradient: the properties
    // "diracDeltaApproximationMethod" (DISCRETE_DELTA, ONE or ZERO) and
    // "diracDeltaApproximationWidthPerStdDev" of the reference's AAD factory
    // (RandomVariableDifferentiableAADFactory.java:60-67), defaults DISCRETE_DELTA and 0.05.
    void setDiracDeltaApproximation(const std::string& method, double widthPerStdDev = 0.05);
    fm_lmm* handle();
private:
    void ensureContext(bool storePaths);
    LIBORMarketModelPtr model_;
    BrownianMotionPtr driver_;
    Scheme scheme_;
    int precision_;
    fm_lmm* ctx_ = nullptr;
    bool ctxStores_ = false;
    bool simulated_ = false;
    bool regenerate_ = false;
    int ckptStride_ = 4;
    int diracMethod_ = FM_DIRAC_DISCRETE_DELTA;
    double diracWidth_ = 0.05;
};
using ProcessPtr = std::shared_ptr<EulerSchemeFromProcessModel>;

// ---- montecarlo/interestrate/LIBORMonteCarloSimulationFromLIBORModel.java -------------
class LIBORMonteCarloSimulationFromLIBORModel {
public:
    explicit LIBORMonteCarloSimulationFromLIBORModel(ProcessPtr process) : process_(std::move(process)) {}
    RandomVariable getLIBOR(int timeIndex, int liborIndex) { return process_->getProcessValue(timeIndex, liborIndex); }
    RandomVariable getNumeraire(double time) { return process_->getNumeraire(time); }
    RandomVariable getMonteCarloWeights(int timeIndex) const { return process_->getMonteCarloWeights(timeIndex); }
    RandomVariable getBrownianIncrement(int timeIndex, int factor) { return process_->getBrownianIncrement(timeIndex, factor); }
    int getNumberOfPaths() const { return process_->getNumberOfPaths(); }
    int getNumberOfFactors() const { return process_->getNumberOfFactors(); }
    TimeDiscretizationPtr getTimeDiscretization() const { return process_->getTimeDiscretization(); }
    TimeDiscretizationPtr getLiborPeriodDiscretization() const { return process_->getModel()->getLiborPeriodDiscretization(); }
    LIBORMarketModelPtr getModel() const { return process_->getModel(); }
    ProcessPtr getProcess() const { return process_; }
    // value several products in one pass: {value, error} per product
    std::vector<std::pair<double, double>> getValues(const std::vector<std::shared_ptr<const Product>>& products);
private:
    ProcessPtr process_;
};
using SimulationPtr = std::shared_ptr<LIBORMonteCarloSimulationFromLIBORModel>;

// ---- montecarlo/interestrate/products -------------------------------------------------
class LIBORMarketModelFromCovarianceModel;

class Product : public std::enable_shared_from_this<Product> {
public:
    virtual ~Product() = default;
    // The product's value unit applied to the average value of the pass (identity: VALUE).
    virtual double toValueUnit(double value, const LIBORMarketModelFromCovarianceModel&) const { return value; }
    virtual bool isValueUnitValue() const { return true; }
    // AbstractMonteCarloProduct.getValue(model) = getValue(0, model).getAverage() (AbstractMonteCarloProduct.java:82-84)
    double getValue(LIBORMonteCarloSimulationFromLIBORModel& model) const;
    // AbstractMonteCarloProduct.getValues: {"value", "error"} (AbstractMonteCarloProduct.java:87-104)
    std::map<std::string, double> getValues(LIBORMonteCarloSimulationFromLIBORModel& model) const;
    virtual void lower(std::deque<std::vector<double>>& arena, fm_product& out) const = 0;
};
using ProductPtr = std::shared_ptr<const Product>;

// Swaption (montecarlo/interestrate/products/Swaption.java:58-199)
class Swaption : public Product {
public:
    Swaption(double exerciseDate, std::vector<double> fixingDates, std::vector<double> paymentDates,
             std::vector<double> periodLengths, std::vector<double> swaprates, double notional = 1.0);
    Swaption(double exerciseDate, const TimeDiscretizationFromArray& swapTenor, double swaprate);
    void lower(std::deque<std::vector<double>>& arena, fm_product& out) const override;
    double exerciseDate, notional;
    std::vector<double> fixingDates, paymentDates, periodLengths, swaprates;
};

// SimpleSwap (montecarlo/interestrate/products/SimpleSwap.java:137-170)
class SimpleSwap : public Product {
public:
    SimpleSwap(std::vector<double> fixingDates, std::vector<double> paymentDates, std::vector<double> swaprates,
               bool isPayFix = true, std::vector<double> notional = {});
    void lower(std::deque<std::vector<double>>& arena, fm_product& out) const override;
    std::vector<double> fixingDates, paymentDates, swaprates, notional;
    bool isPayFix;
};

// SwaptionSimple (montecarlo/interestrate/products/SwaptionSimple.java:38-111): a payer swaption on
// the tenor {T_0, ..., T_n} at one swap rate, exercised at T_0, valued in a value unit: VALUE, or
// the implied volatility of the Monte Carlo price (VOLATILITYLOGNORMAL / VOLATILITY,
// VOLATILITYNORMAL) or its integrated variance (INTEGRATEDVARIANCELOGNORMAL, ...NORMAL), with the
// par swap rate and annuity of the model's forward curve (:68-104). A value in a volatility unit
// is a constant: its error is 0.
class SwaptionSimple : public Product {
public:
    SwaptionSimple(double swaprate, std::vector<double> swapTenor, const std::string& valueUnit = "VALUE");
    void lower(std::deque<std::vector<double>>& arena, fm_product& out) const override;
    double toValueUnit(double value, const LIBORMarketModelFromCovarianceModel& model) const override;
    bool isValueUnitValue() const override { return unit_ == 0; }
    double swaprate;
    std::vector<double> swapTenor;
private:
    int unit_;       // 0 VALUE, 1 lognormal volatility, 2 normal volatility, 3, 4 the integrated variances
};

// AnalyticFormulas (functions/AnalyticFormulas.java): Black and Bachelier option values and their
// implied volatilities, inverted by a Newton iteration safeguarded by bisection to full precision.
namespace AnalyticFormulas {
// (forward N(d+) - strike N(d-)) payoffUnit (:54-83); the intrinsic value for T <= 0 or vol <= 0
double blackScholesGeneralizedOptionValue(double forward, double volatility, double optionMaturity, double optionStrike,
                                          double payoffUnit);
// ((forward - strike) N(d) + vol sqrt(T) phi(d)) payoffUnit, d = (forward - strike)/(vol sqrt(T)) (:1046-1076)
double bachelierOptionValue(double forward, double volatility, double optionMaturity, double optionStrike,
                            double payoffUnit);
// blackScholesOptionImpliedVolatility (:667-713): 0 for a strike <= 0
double blackScholesOptionImpliedVolatility(double forward, double optionMaturity, double optionStrike, double payoffUnit,
                                           double optionValue);
// bachelierOptionImpliedVolatility (:1102-1110)
double bachelierOptionImpliedVolatility(double forward, double optionMaturity, double optionStrike, double payoffUnit,
                                        double optionValue);
} // namespace AnalyticFormulas

// SwapAnnuity.getSwapAnnuity(tenor, forwardCurve) (marketdata/products/SwapAnnuity.java:70-129): sum of
// period length times the forward curve's discount factor at the payment dates after 0.
double getSwapAnnuity(const std::vector<double>& tenor, const ForwardCurveInterpolation& forwardCurve);

// Caplet (montecarlo/interestrate/products/Caplet.java:115-140), value unit VALUE
class Caplet : public Product {
public:
    Caplet(double maturity, double periodLength, double strike, double daycountFraction, bool isFloorlet = false);
    void lower(std::deque<std::vector<double>>& arena, fm_product& out) const override;
    double maturity, periodLength, strike, daycountFraction;
    bool isFloorlet;
};

// RegressionBasisFunctionsProvider
// (montecarlo/conditionalexpectation/RegressionBasisFunctionsProvider.java): the engine
// evaluates the regression basis on the GPU inside the path pass, so a provider names
// one of the engine's bases.
class RegressionBasisFunctionsProvider {
public:
    virtual ~RegressionBasisFunctionsProvider() = default;
    virtual int basisId() const = 0;
};
using RegressionBasisFunctionsProviderPtr = std::shared_ptr<const RegressionBasisFunctionsProvider>;

// Basis 1, S, S^2, N(T_e): S the par swap rate of the remaining periods at the exercise
// date, N the numeraire.
class RegressionBasisFunctionsSwapRate : public RegressionBasisFunctionsProvider {
public:
    int basisId() const override { return FM_BASIS_SWAP_RATE; }
};

// BermudanSwaption (montecarlo/interestrate/products/BermudanSwaption.java:49-157): without a
// provider the basis is the product's own getBasisFunctions (:215-252).
class BermudanSwaption : public Product {
public:
    BermudanSwaption(std::vector<bool> isPeriodStartDateExerciseDate, std::vector<double> fixingDates,
                     std::vector<double> periodLengths, std::vector<double> paymentDates,
                     std::vector<double> periodNotionals, std::vector<double> swaprates, bool isCallable = true,
                     RegressionBasisFunctionsProviderPtr regressionBasisFunctionsProvider = nullptr);
    void lower(std::deque<std::vector<double>>& arena, fm_product& out) const override;
    std::vector<double> getExerciseTimes() const;                               // :259-268
    std::vector<double> fixingDates, periodLengths, paymentDates, periodNotionals, swaprates;
    bool isCallable;
    RegressionBasisFunctionsProviderPtr provider;
private:
    std::vector<int32_t> exerciseFlags_;
};

// ---- optimizer/LevenbergMarquardt.java -----------------------------------------------------
// Minimises sum_k w_k (f_k(x) - y_k)^2. Each step solves (J^T W J with its diagonal scaled by
// 1 + lambda) dx = J^T W (y - f) by the pseudo-inverse; lambda is divided by 3 after an
// improvement, multiplied by 2 after a rejected point and by 16 when the step cannot be
// solved; the Jacobian is a forward difference with the given parameter steps; the solver
// stops when the RMS error improves by no more than the tolerance, after maxIteration
// iterations, or when lambda becomes infinite (LevenbergMarquardt.java:623-815). A point
// with a NaN value is rejected.
class LevenbergMarquardt {
public:
    using ObjectiveFunction = std::function<void(const std::vector<double>& parameters, std::vector<double>& values)>;
    // objective evaluated by worker `worker` (0 .. numberOfThreads-1); workers run concurrently
    using WorkerObjectiveFunction =
        std::function<void(const std::vector<double>& parameters, std::vector<double>& values, int worker)>;
    LevenbergMarquardt(std::vector<double> initialParameters, std::vector<double> targetValues, int maxIteration);
    LevenbergMarquardt& setWeights(std::vector<double> weights);
    LevenbergMarquardt& setParameterSteps(std::vector<double> parameterSteps);
    LevenbergMarquardt& setErrorTolerance(double errorTolerance);
    // numberOfThreads: the Jacobian's columns are evaluated by that many workers at once
    // (LevenbergMarquardt.java's numberOfThreads, :222-247,550-620)
    LevenbergMarquardt& setNumberOfThreads(int numberOfThreads);
    void run(const ObjectiveFunction& setValues);                // one worker
    void run(const WorkerObjectiveFunction& setValues);
    std::vector<double> getBestFitParameters() const { return parameterCurrent_; }
    int getIterations() const { return iteration_; }
    double getRootMeanSquaredError() const { return std::sqrt(errorMeanSquaredCurrent_); }
private:
    double meanSquaredError(const std::vector<double>& values) const;
    std::vector<double> initial_, targets_, weights_, steps_;
    int maxIteration_;
    int threads_ = 1;
    double tolerance_ = 0.0;
    double lambda_ = 0.001, lambdaDivisor_ = 3.0, lambdaMultiplicator_ = 2.0;
    std::vector<double> parameterCurrent_;
    double errorMeanSquaredCurrent_ = INFINITY;
    int iteration_ = 0;
};

// montecarlo/interestrate/CalibrationProduct.java: a product, its target value and weight.
struct CalibrationProduct {
    ProductPtr product;
    double targetValue = 0.0;
    double weight = 1.0;
};

struct CalibrationResult {
    CovarianceModelPtr covarianceModel;   // the calibrated clone
    std::vector<double> parameters;       // its calibrated parameters
    int iterations = 0;
    double rootMeanSquaredError = 0.0;
    int valuations = 0;                   // pricing passes on the device
};

// The parameters of the covariance model, in the reference's order: the four-parameter
// volatility (a, b, c, d) when it is calibrateable, then the correlation decay when that is
// calibrateable (LIBORCovarianceModelFromVolatilityAndCorrelation.getParameter), then a
// local-volatility wrapper's displacement or blend when that is calibrateable.
std::vector<double> getParameter(const LIBORCovarianceModelFromVolatilityAndCorrelation& covarianceModel);
CovarianceModelPtr getCloneWithModifiedParameters(const LIBORCovarianceModelFromVolatilityAndCorrelation& covarianceModel,
                                                  const std::vector<double>& parameters);

// AbstractLIBORCovarianceModelParametric.getCloneCalibrated (covariance/
// AbstractLIBORCovarianceModelParametric.java:156-276): Levenberg-Marquardt on the calibration
// errors (value - target) * weight, every evaluation one fused pricing pass of all products on
// the device, with one Brownian motion (its increments kept in device memory) for all
// evaluations. calibrationParameters: numberOfPaths (2000), seed (31415), maxIterations (400),
// accuracy (1e-7), parameterStep (1e-4), precision (FM_PRECISION_FP64), numberOfThreads (4: the
// Jacobian's passes run on that many contexts at once; results do not depend on it).
CalibrationResult getCloneCalibrated(const LIBORCovarianceModelFromVolatilityAndCorrelation& covarianceModel,
                                     LIBORMarketModelPtr calibrationModel,
                                     const std::vector<CalibrationProduct>& calibrationProducts,
                                     const std::map<std::string, double>& calibrationParameters);

} // namespace finmath

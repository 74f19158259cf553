// Python bindings (pybind11) of the host-side mirror of the finmath-lib API.
// Method names are the reference's, so tests read like its JUnit tests.
#include <pybind11/pybind11.h>
#include <pybind11/numpy.h>
#include <pybind11/stl.h>

#include "finmath_host.hpp"

namespace py = pybind11;
using namespace finmath;

namespace {

py::array_t<double> to_array(const std::vector<double>& v) {
    py::array_t<double> a((py::ssize_t)v.size());
    std::copy(v.begin(), v.end(), a.mutable_data());
    return a;
}

int precision_of(const std::string& s) {
    if (s == "FP64") return FM_PRECISION_FP64;
    if (s == "FP32_MIXED") return FM_PRECISION_FP32_MIXED;
    if (s == "FP64_LOGSTATE") return FM_PRECISION_FP64_LOGSTATE;
    throw IllegalArgumentException("unknown precision " + s);
}

EulerSchemeFromProcessModel::Scheme scheme_of(const std::string& s) {
    if (s == "EULER") return EulerSchemeFromProcessModel::EULER;
    if (s == "PREDICTOR_CORRECTOR") return EulerSchemeFromProcessModel::PREDICTOR_CORRECTOR;
    if (s == "EULER_FUNCTIONAL") return EulerSchemeFromProcessModel::EULER_FUNCTIONAL;
    if (s == "PREDICTOR_CORRECTOR_FUNCTIONAL") return EulerSchemeFromProcessModel::PREDICTOR_CORRECTOR_FUNCTIONAL;
    throw IllegalArgumentException("unknown scheme " + s);
}

} // namespace

PYBIND11_MODULE(_finmath_host, m) {
    m.doc() = "Host-side mirror of finmath-lib's LMM Monte Carlo API over the MI355X engine";

    py::register_exception<CalculationException>(m, "CalculationException", PyExc_RuntimeError);
    py::register_exception<IllegalArgumentException>(m, "IllegalArgumentException", PyExc_ValueError);
    py::register_exception<IllegalStateException>(m, "IllegalStateException", PyExc_RuntimeError);
    py::register_exception<UnsupportedOperationException>(m, "UnsupportedOperationException", PyExc_NotImplementedError);

    py::class_<TimeDiscretizationFromArray, std::shared_ptr<TimeDiscretizationFromArray>>(m, "TimeDiscretizationFromArray")
        .def(py::init<double, int, double>(), py::arg("initial"), py::arg("numberOfTimeSteps"), py::arg("deltaT"))
        .def(py::init<std::vector<double>>(), py::arg("times"))
        .def("getTime", &TimeDiscretizationFromArray::getTime)
        .def("getTimeStep", &TimeDiscretizationFromArray::getTimeStep)
        .def("getTimeIndex", &TimeDiscretizationFromArray::getTimeIndex)
        .def("getTimeIndexNearestLessOrEqual", &TimeDiscretizationFromArray::getTimeIndexNearestLessOrEqual)
        .def("getNumberOfTimeSteps", &TimeDiscretizationFromArray::getNumberOfTimeSteps)
        .def("getNumberOfTimes", &TimeDiscretizationFromArray::getNumberOfTimes)
        .def("getAsDoubleArray", [](const TimeDiscretizationFromArray& t) { return to_array(t.getAsDoubleArray()); });

    py::class_<ForwardCurveInterpolation, std::shared_ptr<ForwardCurveInterpolation>>(m, "ForwardCurveInterpolation")
        .def_static("createForwardCurveFromForwards", &ForwardCurveInterpolation::createForwardCurveFromForwards,
                    py::arg("name"), py::arg("times"), py::arg("givenForwards"), py::arg("paymentOffset"))
        .def("getForward", py::overload_cast<double>(&ForwardCurveInterpolation::getForward, py::const_))
        .def("getForward", py::overload_cast<double, double>(&ForwardCurveInterpolation::getForward, py::const_))
        .def("getPaymentOffset", &ForwardCurveInterpolation::getPaymentOffset)
        .def("getName", &ForwardCurveInterpolation::getName);

    py::class_<DiscountCurveFromForwardCurve, std::shared_ptr<DiscountCurveFromForwardCurve>>(m, "DiscountCurveFromForwardCurve")
        .def(py::init([](std::shared_ptr<ForwardCurveInterpolation> fc) {
            return std::make_shared<DiscountCurveFromForwardCurve>(fc);
        }))
        .def("getDiscountFactor", &DiscountCurveFromForwardCurve::getDiscountFactor);

    m.def("getForwardSwapRate", &getForwardSwapRate, py::arg("fixTenor"), py::arg("floatTenor"), py::arg("forwardCurve"));
    m.def("factorReduction", [](py::array_t<double, py::array::c_style | py::array::forcecast> corr, int m_) {
        const int n = (int)corr.shape(0);
        std::vector<double> c(corr.data(), corr.data() + (size_t)n * n);
        std::vector<double> f = factorReduction(c, n, m_);
        py::array_t<double> out({n, m_});
        std::copy(f.begin(), f.end(), out.mutable_data());
        return out;
    });

    py::class_<LIBORVolatilityModel, std::shared_ptr<LIBORVolatilityModel>>(m, "LIBORVolatilityModel")
        .def("getVolatility", &LIBORVolatilityModel::getVolatility)
        .def("getParameter", [](const LIBORVolatilityModel& v) { return to_array(v.getParameter()); });
    py::class_<LIBORVolatilityModelFourParameterExponentialForm, LIBORVolatilityModel,
               std::shared_ptr<LIBORVolatilityModelFourParameterExponentialForm>>(m, "LIBORVolatilityModelFourParameterExponentialForm")
        .def(py::init<TimeDiscretizationPtr, TimeDiscretizationPtr, double, double, double, double, bool>(),
             py::arg("timeDiscretization"), py::arg("liborPeriodDiscretization"), py::arg("a"), py::arg("b"),
             py::arg("c"), py::arg("d"), py::arg("isCalibrateable"));
    py::class_<LIBORVolatilityModelFromGivenMatrix, LIBORVolatilityModel,
               std::shared_ptr<LIBORVolatilityModelFromGivenMatrix>>(m, "LIBORVolatilityModelFromGivenMatrix")
        .def(py::init([](TimeDiscretizationPtr td, TimeDiscretizationPtr tenor,
                         py::array_t<double, py::array::c_style | py::array::forcecast> vol) {
            std::vector<double> v(vol.data(), vol.data() + vol.size());
            return std::make_shared<LIBORVolatilityModelFromGivenMatrix>(td, tenor, v);
        }));

    py::class_<LIBORVolatilityModelPiecewiseConstant, LIBORVolatilityModel,
               std::shared_ptr<LIBORVolatilityModelPiecewiseConstant>>(m, "LIBORVolatilityModelPiecewiseConstant")
        .def(py::init([](TimeDiscretizationPtr td, TimeDiscretizationPtr tenor, TimeDiscretizationPtr sim,
                         TimeDiscretizationPtr ttm, py::object volatility, bool isCalibrateable) {
            std::vector<double> v;
            if (py::isinstance<py::float_>(volatility) || py::isinstance<py::int_>(volatility)) v = {volatility.cast<double>()};
            else v = volatility.cast<std::vector<double>>();
            return std::make_shared<LIBORVolatilityModelPiecewiseConstant>(td, tenor, sim, ttm, v, isCalibrateable);
        }), py::arg("timeDiscretization"), py::arg("liborPeriodDiscretization"), py::arg("simulationTimeDiscretization"),
           py::arg("timeToMaturityDiscretization"), py::arg("volatility"), py::arg("isCalibrateable") = true)
        .def("getVolatility", &LIBORVolatilityModelPiecewiseConstant::getVolatility)
        .def("getParameter", &LIBORVolatilityModelPiecewiseConstant::getParameter);

    py::class_<LIBORCorrelationModelExponentialDecay, std::shared_ptr<LIBORCorrelationModelExponentialDecay>>(
        m, "LIBORCorrelationModelExponentialDecay")
        .def(py::init<TimeDiscretizationPtr, TimeDiscretizationPtr, int, double, bool>(), py::arg("timeDiscretization"),
             py::arg("liborPeriodDiscretization"), py::arg("numberOfFactors"), py::arg("a"),
             py::arg("isCalibrateable") = false)
        .def("getFactorLoading", &LIBORCorrelationModelExponentialDecay::getFactorLoading)
        .def("getCorrelation", &LIBORCorrelationModelExponentialDecay::getCorrelation)
        .def("getNumberOfFactors", &LIBORCorrelationModelExponentialDecay::getNumberOfFactors)
        .def("factorMatrix", [](const LIBORCorrelationModelExponentialDecay& c) {
            const auto& f = c.factorMatrix();
            const int m_ = c.getNumberOfFactors();
            py::array_t<double> out({(int)(f.size() / m_), m_});
            std::copy(f.begin(), f.end(), out.mutable_data());
            return out;
        });

    py::class_<LIBORCovarianceModelFromVolatilityAndCorrelation,
               std::shared_ptr<LIBORCovarianceModelFromVolatilityAndCorrelation>>(m, "LIBORCovarianceModelFromVolatilityAndCorrelation")
        .def(py::init([](TimeDiscretizationPtr td, TimeDiscretizationPtr tenor, std::shared_ptr<LIBORVolatilityModel> v,
                         std::shared_ptr<LIBORCorrelationModelExponentialDecay> c) {
            return std::make_shared<LIBORCovarianceModelFromVolatilityAndCorrelation>(td, tenor, v, c);
        }))
        .def("getFactorLoading", [](const LIBORCovarianceModelFromVolatilityAndCorrelation& c, int i, int j) {
            return to_array(c.getFactorLoading(i, j));
        })
        .def("getCovariance", &LIBORCovarianceModelFromVolatilityAndCorrelation::getCovariance)
        .def("getNumberOfFactors", &LIBORCovarianceModelFromVolatilityAndCorrelation::getNumberOfFactors)
        .def("getParameter", [](const LIBORCovarianceModelFromVolatilityAndCorrelation& c) { return getParameter(c); })
        .def("getCloneWithModifiedParameters", [](const LIBORCovarianceModelFromVolatilityAndCorrelation& c,
                                                  std::vector<double> p) {
            return std::const_pointer_cast<LIBORCovarianceModelFromVolatilityAndCorrelation>(getCloneWithModifiedParameters(c, p));
        })
        .def("getCloneCalibrated", [](const LIBORCovarianceModelFromVolatilityAndCorrelation& c,
                                      std::shared_ptr<LIBORMarketModelFromCovarianceModel> model,
                                      const std::vector<CalibrationProduct>& products, std::map<std::string, double> params) {
            CalibrationResult r;
            {
                py::gil_scoped_release rel;
                r = getCloneCalibrated(c, model, products, params);
            }
            py::dict info;
            info["parameters"] = r.parameters;
            info["iterations"] = r.iterations;
            info["rootMeanSquaredError"] = r.rootMeanSquaredError;
            info["valuations"] = r.valuations;
            return py::make_tuple(std::const_pointer_cast<LIBORCovarianceModelFromVolatilityAndCorrelation>(r.covarianceModel), info);
        }, py::arg("calibrationModel"), py::arg("calibrationProducts"),
           py::arg("calibrationParameters") = std::map<std::string, double>{})
        .def("tables", [](const LIBORCovarianceModelFromVolatilityAndCorrelation& c, const TimeDiscretizationFromArray& sim) {
            std::vector<double> lam, var;
            c.tables(sim, lam, var);
            const int nT = sim.getNumberOfTimeSteps(), n = c.getLiborPeriodDiscretization()->getNumberOfTimeSteps(),
                      F = c.getNumberOfFactors();
            py::array_t<double> L({nT, n, F}), V({nT, n});
            std::copy(lam.begin(), lam.end(), L.mutable_data());
            std::copy(var.begin(), var.end(), V.mutable_data());
            return py::make_tuple(L, V);
        })
        .def("localVolatility", [](const LIBORCovarianceModelFromVolatilityAndCorrelation& c,
                                   const TimeDiscretizationFromArray& sim) -> py::object {
            double alpha = 0.0;
            std::vector<double> offsets;
            if (!c.localVolatility(sim, alpha, offsets)) return py::none();
            const int nT = sim.getNumberOfTimeSteps(), n = c.getLiborPeriodDiscretization()->getNumberOfTimeSteps();
            py::array_t<double> B({nT, n});
            std::copy(offsets.begin(), offsets.end(), B.mutable_data());
            return py::make_tuple(alpha, B);
        });

    py::class_<DisplacedLocalVolatilityModel, LIBORCovarianceModelFromVolatilityAndCorrelation,
               std::shared_ptr<DisplacedLocalVolatilityModel>>(m, "DisplacedLocalVolatilityModel")
        .def(py::init([](std::shared_ptr<LIBORCovarianceModelFromVolatilityAndCorrelation> cov, double displacement,
                         bool isCalibrateable) {
            return std::make_shared<DisplacedLocalVolatilityModel>(cov, displacement, isCalibrateable);
        }), py::arg("covarianceModel"), py::arg("displacement"), py::arg("isCalibrateable"))
        .def("getBaseCovarianceModel", [](const DisplacedLocalVolatilityModel& c) {
            return std::const_pointer_cast<LIBORCovarianceModelFromVolatilityAndCorrelation>(c.getBaseCovarianceModel());
        })
        .def("getDisplacement", &DisplacedLocalVolatilityModel::getDisplacement);

    py::class_<BlendedLocalVolatilityModel, LIBORCovarianceModelFromVolatilityAndCorrelation,
               std::shared_ptr<BlendedLocalVolatilityModel>>(m, "BlendedLocalVolatilityModel")
        .def(py::init([](std::shared_ptr<LIBORCovarianceModelFromVolatilityAndCorrelation> cov,
                         std::shared_ptr<ForwardCurveInterpolation> forwardCurve, double displacement, bool isCalibrateable) {
            return std::make_shared<BlendedLocalVolatilityModel>(cov, forwardCurve, displacement, isCalibrateable);
        }), py::arg("covarianceModel"), py::arg("forwardCurve"), py::arg("displacement"), py::arg("isCalibrateable"))
        .def(py::init([](std::shared_ptr<LIBORCovarianceModelFromVolatilityAndCorrelation> cov, double displacement,
                         bool isCalibrateable) {
            return std::make_shared<BlendedLocalVolatilityModel>(cov, nullptr, displacement, isCalibrateable);
        }), py::arg("covarianceModel"), py::arg("displacement"), py::arg("isCalibrateable"))
        .def("getBaseCovarianceModel", [](const BlendedLocalVolatilityModel& c) {
            return std::const_pointer_cast<LIBORCovarianceModelFromVolatilityAndCorrelation>(c.getBaseCovarianceModel());
        })
        .def("getDisplacement", &BlendedLocalVolatilityModel::getDisplacement);

    py::class_<LIBORCovarianceModelStochasticVolatility, LIBORCovarianceModelFromVolatilityAndCorrelation,
               std::shared_ptr<LIBORCovarianceModelStochasticVolatility>>(m, "LIBORCovarianceModelStochasticVolatility")
        .def(py::init([](std::shared_ptr<LIBORCovarianceModelFromVolatilityAndCorrelation> cov,
                         std::shared_ptr<BrownianMotionFromMersenneRandomNumbers> bm, double nu, double rho,
                         bool isCalibrateable) {
            return std::make_shared<LIBORCovarianceModelStochasticVolatility>(cov, bm, nu, rho, isCalibrateable);
        }), py::arg("covarianceModel"), py::arg("brownianMotion"), py::arg("nu"), py::arg("rho"),
           py::arg("isCalibrateable"))
        .def("getBaseCovarianceModel", [](const LIBORCovarianceModelStochasticVolatility& c) {
            return std::const_pointer_cast<LIBORCovarianceModelFromVolatilityAndCorrelation>(c.getBaseCovarianceModel());
        })
        .def("getNu", &LIBORCovarianceModelStochasticVolatility::getNu)
        .def("getRho", &LIBORCovarianceModelStochasticVolatility::getRho);

    py::class_<LIBORCovarianceModelStochasticHestonVolatility, LIBORCovarianceModelFromVolatilityAndCorrelation,
               std::shared_ptr<LIBORCovarianceModelStochasticHestonVolatility>>(
        m, "LIBORCovarianceModelStochasticHestonVolatility")
        .def(py::init([](std::shared_ptr<LIBORCovarianceModelFromVolatilityAndCorrelation> cov,
                         std::shared_ptr<BrownianMotionFromMersenneRandomNumbers> bm, double kappa, double theta,
                         double xi, bool isCalibrateable) {
            return std::make_shared<LIBORCovarianceModelStochasticHestonVolatility>(cov, bm, kappa, theta, xi,
                                                                                    isCalibrateable);
        }), py::arg("covarianceModel"), py::arg("brownianMotion"), py::arg("kappa"), py::arg("theta"), py::arg("xi"),
           py::arg("isCalibrateable"))
        .def("getBaseCovarianceModel", [](const LIBORCovarianceModelStochasticHestonVolatility& c) {
            return std::const_pointer_cast<LIBORCovarianceModelFromVolatilityAndCorrelation>(c.getBaseCovarianceModel());
        })
        .def("getKappa", &LIBORCovarianceModelStochasticHestonVolatility::getKappa)
        .def("getTheta", &LIBORCovarianceModelStochasticHestonVolatility::getTheta)
        .def("getXi", &LIBORCovarianceModelStochasticHestonVolatility::getXi);

    py::class_<LevenbergMarquardt>(m, "LevenbergMarquardt")
        .def(py::init<std::vector<double>, std::vector<double>, int>(), py::arg("initialParameters"),
             py::arg("targetValues"), py::arg("maxIteration"))
        .def("setWeights", &LevenbergMarquardt::setWeights, py::return_value_policy::reference_internal)
        .def("setParameterSteps", &LevenbergMarquardt::setParameterSteps, py::return_value_policy::reference_internal)
        .def("setErrorTolerance", &LevenbergMarquardt::setErrorTolerance, py::return_value_policy::reference_internal)
        .def("setNumberOfThreads", &LevenbergMarquardt::setNumberOfThreads, py::return_value_policy::reference_internal)
        .def("run", [](LevenbergMarquardt& lm, py::function f) {
            lm.run([&](const std::vector<double>& p, std::vector<double>& v) {
                py::list out = f(p).cast<py::list>();
                v.assign(out.size(), 0.0);
                for (size_t i = 0; i < out.size(); i++) v[i] = out[i].cast<double>();
            });
        }, py::arg("setValues"))
        .def("getBestFitParameters", &LevenbergMarquardt::getBestFitParameters)
        .def("getIterations", &LevenbergMarquardt::getIterations)
        .def("getRootMeanSquaredError", &LevenbergMarquardt::getRootMeanSquaredError);

    py::class_<CalibrationProduct>(m, "CalibrationProduct")
        .def(py::init([](std::shared_ptr<Product> p, double target, double weight) {
            return CalibrationProduct{p, target, weight};
        }), py::arg("product"), py::arg("targetValue"), py::arg("weight") = 1.0)
        .def_readonly("targetValue", &CalibrationProduct::targetValue)
        .def_readonly("weight", &CalibrationProduct::weight);

    py::class_<LIBORMarketModelFromCovarianceModel, std::shared_ptr<LIBORMarketModelFromCovarianceModel>>(
        m, "LIBORMarketModelFromCovarianceModel")
        .def_static("of", [](TimeDiscretizationPtr tenor, py::object analyticModel, std::shared_ptr<ForwardCurveInterpolation> fc,
                             std::shared_ptr<DiscountCurveFromForwardCurve> dc, py::object factory,
                             std::shared_ptr<LIBORCovarianceModelFromVolatilityAndCorrelation> cov, py::object calibrationItems,
                             std::map<std::string, std::string> properties) {
            if (!analyticModel.is_none()) throw UnsupportedOperationException("analytic models are not supported");
            if (!factory.is_none()) throw UnsupportedOperationException("random variable factories are chosen by precision");
            if (!calibrationItems.is_none() && py::len(calibrationItems) > 0)
                throw UnsupportedOperationException("calibration is not part of this engine");
            return LIBORMarketModelFromCovarianceModel::of(tenor, fc, dc, cov, properties);
        }, py::arg("liborPeriodDiscretization"), py::arg("analyticModel"), py::arg("forwardRateCurve"),
           py::arg("discountCurve"), py::arg("randomVariableFactory"), py::arg("covarianceModel"),
           py::arg("calibrationItems"), py::arg("properties"))
        .def("getInitialForwards", [](const LIBORMarketModelFromCovarianceModel& mm) { return to_array(mm.getInitialForwards()); })
        .def("getNumberOfComponents", &LIBORMarketModelFromCovarianceModel::getNumberOfComponents)
        .def("getNumberOfLibors", &LIBORMarketModelFromCovarianceModel::getNumberOfLibors)
        .def("getLiborPeriod", &LIBORMarketModelFromCovarianceModel::getLiborPeriod)
        .def("getLiborPeriodIndex", &LIBORMarketModelFromCovarianceModel::getLiborPeriodIndex)
        .def("getDiscountCurve", &LIBORMarketModelFromCovarianceModel::getDiscountCurve)
        .def("getForwardRateCurve", &LIBORMarketModelFromCovarianceModel::getForwardRateCurve)
        .def("getCovarianceModel", &LIBORMarketModelFromCovarianceModel::getCovarianceModel);

    py::class_<BrownianMotionFromMersenneRandomNumbers, std::shared_ptr<BrownianMotionFromMersenneRandomNumbers>>(
        m, "BrownianMotionFromMersenneRandomNumbers")
        .def(py::init<TimeDiscretizationPtr, int, int, int>(), py::arg("timeDiscretization"), py::arg("numberOfFactors"),
             py::arg("numberOfPaths"), py::arg("seed"))
        .def("getNumberOfFactors", &BrownianMotionFromMersenneRandomNumbers::getNumberOfFactors)
        .def("getNumberOfPaths", &BrownianMotionFromMersenneRandomNumbers::getNumberOfPaths)
        .def("getSeed", &BrownianMotionFromMersenneRandomNumbers::getSeed);

    py::class_<RandomVariable>(m, "RandomVariable")
        .def("getAverage", &RandomVariable::getAverage)
        .def("getVariance", &RandomVariable::getVariance)
        .def("getStandardError", &RandomVariable::getStandardError)
        .def("isDeterministic", &RandomVariable::isDeterministic)
        .def("size", &RandomVariable::size)
        .def("get", &RandomVariable::get)
        .def("doubleValue", &RandomVariable::doubleValue)
        .def("getFiltrationTime", &RandomVariable::getFiltrationTime)
        .def("getRealizations", [](const RandomVariable& r, int P) { return to_array(r.getRealizations(P)); },
             py::arg("numberOfPaths"));

    py::class_<Product, std::shared_ptr<Product>>(m, "AbstractMonteCarloProduct")
        .def("getValue", &Product::getValue)
        .def("getValues", &Product::getValues);
    py::class_<Swaption, Product, std::shared_ptr<Swaption>>(m, "Swaption")
        .def(py::init([](double ex, std::vector<double> fix, std::vector<double> pay, std::vector<double> k,
                         std::vector<double> pl, double notional) {
            return std::make_shared<Swaption>(ex, fix, pay, pl, k, notional);
        }), py::arg("exerciseDate"), py::arg("fixingDates"), py::arg("paymentDates"), py::arg("swaprates"),
            py::arg("periodLengths") = std::vector<double>{}, py::arg("notional") = 1.0)
        .def(py::init([](double ex, const TimeDiscretizationFromArray& tenor, double k) {
            return std::make_shared<Swaption>(ex, tenor, k);
        }), py::arg("exerciseDate"), py::arg("swapTenor"), py::arg("swaprate"));
    py::class_<SwaptionSimple, Product, std::shared_ptr<SwaptionSimple>>(m, "SwaptionSimple")
        .def(py::init<double, std::vector<double>, const std::string&>(), py::arg("swaprate"), py::arg("swapTenor"),
             py::arg("valueUnit") = "VALUE");
    m.def("blackScholesGeneralizedOptionValue", &AnalyticFormulas::blackScholesGeneralizedOptionValue);
    m.def("bachelierOptionValue", &AnalyticFormulas::bachelierOptionValue);
    m.def("blackScholesOptionImpliedVolatility", &AnalyticFormulas::blackScholesOptionImpliedVolatility);
    m.def("bachelierOptionImpliedVolatility", &AnalyticFormulas::bachelierOptionImpliedVolatility);
    m.def("getSwapAnnuity", [](std::vector<double> tenor, std::shared_ptr<ForwardCurveInterpolation> fc) {
        return getSwapAnnuity(tenor, *fc);
    });

    py::class_<SimpleSwap, Product, std::shared_ptr<SimpleSwap>>(m, "SimpleSwap")
        .def(py::init([](std::vector<double> fix, std::vector<double> pay, std::vector<double> k, bool payFix,
                         std::vector<double> notional) {
            return std::make_shared<SimpleSwap>(fix, pay, k, payFix, notional);
        }), py::arg("fixingDates"), py::arg("paymentDates"), py::arg("swaprates"), py::arg("isPayFix") = true,
            py::arg("notional") = std::vector<double>{});
    py::class_<Caplet, Product, std::shared_ptr<Caplet>>(m, "Caplet")
        .def(py::init([](double mat, double pl, double k, py::object dcf, bool floorlet) {
            return std::make_shared<Caplet>(mat, pl, k, dcf.is_none() ? pl : dcf.cast<double>(), floorlet);
        }), py::arg("maturity"), py::arg("periodLength"), py::arg("strike"), py::arg("daycountFraction") = py::none(),
            py::arg("isFloorlet") = false);
    py::class_<RegressionBasisFunctionsProvider, std::shared_ptr<RegressionBasisFunctionsProvider>>(
        m, "RegressionBasisFunctionsProvider");
    py::class_<RegressionBasisFunctionsSwapRate, RegressionBasisFunctionsProvider,
               std::shared_ptr<RegressionBasisFunctionsSwapRate>>(m, "RegressionBasisFunctionsSwapRate")
        .def(py::init<>());
    py::class_<BermudanSwaption, Product, std::shared_ptr<BermudanSwaption>>(m, "BermudanSwaption")
        .def(py::init([](std::vector<bool> ex, std::vector<double> fix, std::vector<double> pl, std::vector<double> pay,
                         std::vector<double> notionals, std::vector<double> k, bool callable,
                         std::shared_ptr<RegressionBasisFunctionsProvider> prov) {
            return std::make_shared<BermudanSwaption>(ex, fix, pl, pay, notionals, k, callable, prov);
        }), py::arg("isPeriodStartDateExerciseDate"), py::arg("fixingDates"), py::arg("periodLength"),
            py::arg("paymentDates"), py::arg("periodNotionals"), py::arg("swaprates"), py::arg("isCallable") = true,
            py::arg("regressionBasisFunctionsProvider") = nullptr)
        .def("getExerciseTimes", &BermudanSwaption::getExerciseTimes);

    py::class_<EulerSchemeFromProcessModel, std::shared_ptr<EulerSchemeFromProcessModel>>(m, "EulerSchemeFromProcessModel")
        .def(py::init([](std::shared_ptr<LIBORMarketModelFromCovarianceModel> model,
                         std::shared_ptr<BrownianMotionFromMersenneRandomNumbers> bm, const std::string& scheme,
                         const std::string& precision) {
            return std::make_shared<EulerSchemeFromProcessModel>(model, bm, scheme_of(scheme), precision_of(precision));
        }), py::arg("model"), py::arg("stochasticDriver"), py::arg("scheme") = "EULER_FUNCTIONAL",
            py::arg("precision") = "FP64")
        .def("getProcessValue", &EulerSchemeFromProcessModel::getProcessValue, py::call_guard<py::gil_scoped_release>())
        .def("getMonteCarloWeights", &EulerSchemeFromProcessModel::getMonteCarloWeights)
        .def("getBrownianIncrement", &EulerSchemeFromProcessModel::getBrownianIncrement, py::call_guard<py::gil_scoped_release>())
        .def("getNumeraire", &EulerSchemeFromProcessModel::getNumeraire, py::call_guard<py::gil_scoped_release>())
        .def("getNumberOfPaths", &EulerSchemeFromProcessModel::getNumberOfPaths)
        .def("getNumberOfFactors", &EulerSchemeFromProcessModel::getNumberOfFactors)
        .def("getNumberOfComponents", &EulerSchemeFromProcessModel::getNumberOfComponents)
        .def("getTime", &EulerSchemeFromProcessModel::getTime)
        .def("getTimeIndex", &EulerSchemeFromProcessModel::getTimeIndex)
        .def("lastTimings", [](EulerSchemeFromProcessModel& p) { return to_array(p.lastTimings()); })
        .def("setRegenerateIncrements", &EulerSchemeFromProcessModel::setRegenerateIncrements)
        .def("setDiracDeltaApproximation", &EulerSchemeFromProcessModel::setDiracDeltaApproximation,
             py::arg("method"), py::arg("widthPerStdDev") = 0.05)
        .def("getGradient", [](EulerSchemeFromProcessModel& p, const std::vector<std::shared_ptr<Product>>& ps,
                               std::vector<double> weights, int stride) {
            std::vector<ProductPtr> v(ps.begin(), ps.end());
            Gradient g;
            {
                py::gil_scoped_release rel;
                g = p.getGradient(v, weights, stride);
            }
            const int n = (int)g.delta.size();
            const int nT = n ? (int)(g.volatility.size() / n) : 0;
            py::dict d;
            d["value"] = g.value;
            d["delta"] = to_array(g.delta);
            py::array_t<double> vol({nT, n});
            std::copy(g.volatility.begin(), g.volatility.end(), vol.mutable_data());
            d["volatility"] = vol;
            d["forwardVega"] = to_array(g.forwardVega);
            d["parameterVega"] = to_array(g.parameterVega);
            const int nTc = n ? (int)(g.volatilityModel.size() / n) : 0;
            py::array_t<double> volm({nTc, n});
            std::copy(g.volatilityModel.begin(), g.volatilityModel.end(), volm.mutable_data());
            d["volatilityModel"] = volm;
            return d;
        }, py::arg("products"), py::arg("weights") = std::vector<double>{}, py::arg("checkpointStride") = 4);

    py::class_<LIBORMonteCarloSimulationFromLIBORModel, std::shared_ptr<LIBORMonteCarloSimulationFromLIBORModel>>(
        m, "LIBORMonteCarloSimulationFromLIBORModel")
        .def(py::init<ProcessPtr>(), py::arg("process"))
        .def("getLIBOR", &LIBORMonteCarloSimulationFromLIBORModel::getLIBOR, py::call_guard<py::gil_scoped_release>())
        .def("getNumeraire", &LIBORMonteCarloSimulationFromLIBORModel::getNumeraire, py::call_guard<py::gil_scoped_release>())
        .def("getMonteCarloWeights", &LIBORMonteCarloSimulationFromLIBORModel::getMonteCarloWeights)
        .def("getBrownianIncrement", &LIBORMonteCarloSimulationFromLIBORModel::getBrownianIncrement,
             py::call_guard<py::gil_scoped_release>())
        .def("getNumberOfPaths", &LIBORMonteCarloSimulationFromLIBORModel::getNumberOfPaths)
        .def("getProcess", &LIBORMonteCarloSimulationFromLIBORModel::getProcess)
        .def("getModel", &LIBORMonteCarloSimulationFromLIBORModel::getModel)
        .def("getValues", [](LIBORMonteCarloSimulationFromLIBORModel& s, const std::vector<std::shared_ptr<Product>>& ps) {
            std::vector<ProductPtr> v(ps.begin(), ps.end());
            std::vector<std::pair<double, double>> r;
            {
                py::gil_scoped_release rel;
                r = s.getValues(v);
            }
            py::list out;
            for (auto& x : r) { py::dict d; d["value"] = x.first; d["error"] = x.second; out.append(d); }
            return out;
        });
}

"""ctypes binding of the C ABI declared in include/finmath_lmm.h and include/finmath_rv.h.

The binding is the Python face of the drop-in boundary: plain pointers and sizes,
status codes mapped onto the exception types the reference raises. It loads the
in-tree shared library and fails loudly when it has not been built.
"""
from __future__ import annotations

import ctypes
import math
import os
import struct
from ctypes import POINTER, c_char_p, c_double, c_int32, c_int64, c_uint32, c_uint64, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libfinmath_lmm.so")

FM_OK = 0
FM_ERR_INVALID_ARGUMENT = 1
FM_ERR_OUT_OF_MEMORY = 2
FM_ERR_DEVICE = 3
FM_ERR_NUMERICAL = 4
FM_ERR_STATE = 5
FM_ERR_INTERNAL = 6
FM_ERR_UNSUPPORTED = 7

MEASURE = {"SPOT": 0, "TERMINAL": 1}
STATE_SPACE = {"NORMAL": 0, "LOGNORMAL": 1}
SCHEME = {"EULER": 0, "PREDICTOR_CORRECTOR": 1, "EULER_FUNCTIONAL": 2, "PREDICTOR_CORRECTOR_FUNCTIONAL": 3}
PRECISION = {"FP64": 0, "FP32_MIXED": 1, "FP64_LOGSTATE": 2}
# RandomVariableDifferentiableAADFactory.DiracDeltaApproximationMethod ordinals
DIRAC = {"DISCRETE_DELTA": 0, "REGRESSION_ON_DENSITY": 1, "REGRESSION_ON_DISTRIBUITON": 2, "ONE": 3, "ZERO": 4}
PRODUCT_SWAPTION, PRODUCT_SWAP, PRODUCT_CAPLET, PRODUCT_BERMUDAN_SWAPTION = 1, 2, 3, 4
BASIS_REFERENCE, BASIS_SWAP_RATE = 0, 1

# Every symbol include/finmath_lmm.h declares.
EXPORTED_SYMBOLS = (
    "fm_lmm_create", "fm_lmm_destroy", "fm_lmm_set_initial_forwards", "fm_lmm_set_covariance",
    "fm_lmm_set_discount_curve", "fm_lmm_set_local_volatility", "fm_lmm_set_stochastic_volatility",
    "fm_lmm_set_heston_volatility",
    "fm_lmm_simulate", "fm_lmm_get_process_value",
    "fm_lmm_process_value_device_ptr", "fm_lmm_get_numeraire", "fm_lmm_get_brownian_increment",
    "fm_lmm_value_products", "fm_lmm_value_products_noalloc", "fm_lmm_set_volatility_factorization", "fm_lmm_adjoint",
    "fm_lmm_set_dirac_delta",
    "fm_lmm_bermudan_exercise_times", "fm_lmm_serialize", "fm_lmm_deserialize",
    "fm_mt19937_words", "fm_mt19937_doubles",
    "fm_last_error", "fm_lmm_last_timings", "fm_device_info",
    # include/finmath_rv.h
    "fm_rv_from_host", "fm_rv_constant", "fm_rv_release", "fm_rv_info", "fm_rv_to_host", "fm_rv_device_ptr",
    "fm_rv_evaluate", "fm_rv_evaluate_average", "fm_rv_average", "fm_rv_variance",
    "fm_lmm_process_value_rv", "fm_lmm_numeraire_rv", "fm_lmm_weights_rv",
)


class CalculationException(Exception):
    """Counterpart of net.finmath.exception.CalculationException."""


class IllegalArgumentException(ValueError):
    """Counterpart of java.lang.IllegalArgumentException."""


class IllegalStateException(RuntimeError):
    """Counterpart of java.lang.IllegalStateException."""


class UnsupportedOperationException(NotImplementedError):
    """Counterpart of java.lang.UnsupportedOperationException."""


_EXC = {
    FM_ERR_INVALID_ARGUMENT: IllegalArgumentException,
    FM_ERR_OUT_OF_MEMORY: CalculationException,
    FM_ERR_DEVICE: CalculationException,
    FM_ERR_NUMERICAL: CalculationException,
    FM_ERR_STATE: IllegalStateException,
    FM_ERR_INTERNAL: CalculationException,
    FM_ERR_UNSUPPORTED: UnsupportedOperationException,
}


class FmLmmConfig(ctypes.Structure):
    _fields_ = [
        ("num_time_steps", c_int32), ("times", POINTER(c_double)),
        ("num_libors", c_int32), ("tenor", POINTER(c_double)),
        ("num_factors", c_int32), ("num_paths", c_int32), ("seed", c_int32),
        ("measure", c_int32), ("state_space", c_int32), ("scheme", c_int32), ("precision", c_int32),
        ("libor_cap", c_double), ("store_paths", c_int32), ("checkpoint_stride", c_int32),
        ("regenerate_increments", c_int32), ("path_offset", c_int64),
    ]


class FmRvInstr(ctypes.Structure):
    _fields_ = [("op", c_int32), ("dst", ctypes.c_int16), ("a", ctypes.c_int16), ("b", ctypes.c_int16),
                ("c", ctypes.c_int16), ("imm", c_double)]


class FmProduct(ctypes.Structure):
    _fields_ = [
        ("type", c_int32), ("num_periods", c_int32), ("exercise_date", c_double),
        ("fixing_dates", POINTER(c_double)), ("payment_dates", POINTER(c_double)),
        ("period_lengths", POINTER(c_double)), ("swaprates", POINTER(c_double)),
        ("notionals", POINTER(c_double)), ("discounting_adjustment", POINTER(c_double)),
        ("notional", c_double), ("is_pay_fix", c_int32), ("is_floorlet", c_int32),
        ("exercise_flags", POINTER(c_int32)), ("is_cancelable", c_int32), ("basis", c_int32),
    ]


_lib = None


def library_path() -> str:
    return LIB_PATH


def load() -> ctypes.CDLL:
    """Load the HIP library; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} not built: run `python finmath-lib_amd/build.py` (or __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    dp = POINTER(c_double)
    sig = {
        "fm_lmm_create": ([POINTER(FmLmmConfig), POINTER(c_void_p)], c_int32),
        "fm_lmm_destroy": ([c_void_p], None),
        "fm_lmm_set_initial_forwards": ([c_void_p, dp], c_int32),
        "fm_lmm_set_covariance": ([c_void_p, dp, dp], c_int32),
        "fm_lmm_set_discount_curve": ([c_void_p, dp], c_int32),
        "fm_lmm_set_local_volatility": ([c_void_p, c_double, dp], c_int32),
        "fm_lmm_set_stochastic_volatility": ([c_void_p, c_double, c_double, c_void_p, c_int32, c_int32], c_int32),
        "fm_lmm_set_heston_volatility": ([c_void_p, c_double, c_double, c_double, c_void_p, c_int32], c_int32),
        "fm_lmm_simulate": ([c_void_p], c_int32),
        "fm_lmm_get_process_value": ([c_void_p, c_int32, c_int32, dp], c_int32),
        "fm_lmm_process_value_device_ptr": ([c_void_p, c_int32, c_int32, POINTER(c_void_p)], c_int32),
        "fm_lmm_get_numeraire": ([c_void_p, c_double, dp], c_int32),
        "fm_lmm_get_brownian_increment": ([c_void_p, c_int32, c_int32, dp], c_int32),
        "fm_lmm_value_products": ([c_void_p, POINTER(FmProduct), c_int32, dp, dp], c_int32),
        "fm_lmm_value_products_noalloc": ([c_void_p, POINTER(FmProduct), c_int32, dp, dp], c_int32),
        "fm_lmm_set_volatility_factorization": ([c_void_p, dp, dp, dp], c_int32),
        "fm_lmm_adjoint": ([c_void_p, POINTER(FmProduct), c_int32, dp, dp, dp, dp], c_int32),
        "fm_lmm_set_dirac_delta": ([c_void_p, c_int32, c_double], c_int32),
        "fm_lmm_bermudan_exercise_times": ([c_void_p, c_int32, dp], c_int32),
        "fm_lmm_serialize": ([c_void_p, c_void_p, POINTER(c_int64)], c_int32),
        "fm_lmm_deserialize": ([c_void_p, c_int64, POINTER(c_void_p)], c_int32),
        "fm_mt19937_words": ([c_int64, c_uint64, c_uint64, POINTER(c_uint32)], c_int32),
        "fm_mt19937_doubles": ([c_int64, c_uint64, c_uint64, dp], c_int32),
        "fm_last_error": ([], c_char_p),
        "fm_lmm_last_timings": ([c_void_p, dp, c_int32], c_int32),
        "fm_device_info": ([ctypes.c_char_p, c_int32], c_int32),
        "fm_rv_from_host": ([c_double, c_int64, dp, POINTER(c_void_p)], c_int32),
        "fm_rv_constant": ([c_double, c_double, POINTER(c_void_p)], c_int32),
        "fm_rv_release": ([c_void_p], None),
        "fm_rv_info": ([c_void_p, dp, POINTER(c_int64), POINTER(c_int32), dp], c_int32),
        "fm_rv_to_host": ([c_void_p, c_int64, dp], c_int32),
        "fm_rv_device_ptr": ([c_void_p, POINTER(c_void_p)], c_int32),
        "fm_rv_evaluate": ([c_int32, POINTER(c_void_p), c_int32, POINTER(FmRvInstr), c_int32, c_double,
                            POINTER(c_void_p)], c_int32),
        "fm_rv_evaluate_average": ([c_int32, POINTER(c_void_p), c_int32, POINTER(FmRvInstr), c_int32, c_void_p,
                                    dp, dp], c_int32),
        "fm_rv_average": ([c_void_p, dp], c_int32),
        "fm_rv_variance": ([c_void_p, dp], c_int32),
        "fm_lmm_process_value_rv": ([c_void_p, c_int32, c_int32, POINTER(c_void_p)], c_int32),
        "fm_lmm_numeraire_rv": ([c_void_p, c_double, POINTER(c_void_p)], c_int32),
        "fm_lmm_weights_rv": ([c_void_p, POINTER(c_void_p)], c_int32),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != FM_OK:
        msg = (load().fm_last_error() or b"").decode(errors="replace")
        raise _EXC.get(status, CalculationException)(f"[status {status}] {msg}")


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(POINTER(c_double))


def as_f64(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


class Product:
    """Owns the arrays behind an FmProduct."""

    def __init__(self, type_: int, exercise_date: float = 0.0, fixing_dates=None, payment_dates=None,
                 period_lengths=None, swaprates=None, notionals=None, discounting_adjustment=None,
                 notional: float = 1.0, is_pay_fix: bool = True, is_floorlet: bool = False):
        self._keep = []
        self.s = FmProduct()
        self.s.type = type_
        self.s.exercise_date = float(exercise_date)
        self.s.notional = float(notional)
        self.s.is_pay_fix = 1 if is_pay_fix else 0
        self.s.is_floorlet = 1 if is_floorlet else 0
        n = 0
        for fname, arr in (("fixing_dates", fixing_dates), ("payment_dates", payment_dates),
                           ("period_lengths", period_lengths), ("swaprates", swaprates),
                           ("notionals", notionals), ("discounting_adjustment", discounting_adjustment)):
            if arr is None:
                setattr(self.s, fname, POINTER(c_double)())
                continue
            a = as_f64(arr)
            self._keep.append(a)
            setattr(self.s, fname, _dptr(a))
            n = max(n, a.size)
        self.s.num_periods = n
        self.s.exercise_flags = POINTER(c_int32)()

    @staticmethod
    def bermudan(fixing_dates, payment_dates, period_lengths, swaprates, exercise_flags, notionals=None,
                 is_callable=True, basis=BASIS_REFERENCE):
        p = Product(PRODUCT_BERMUDAN_SWAPTION, 0.0, fixing_dates, payment_dates, period_lengths, swaprates,
                    notionals)
        flags = np.ascontiguousarray(np.asarray(exercise_flags, dtype=np.int32))
        p._keep.append(flags)
        p.s.exercise_flags = flags.ctypes.data_as(POINTER(c_int32))
        p.s.is_cancelable = 0 if is_callable else 1
        p.s.basis = int(basis)
        return p

    @staticmethod
    def swaption(exercise_date, fixing_dates, payment_dates, swaprates, period_lengths=None, notional=1.0,
                 discounting_adjustment=None):
        return Product(PRODUCT_SWAPTION, exercise_date, fixing_dates, payment_dates, period_lengths, swaprates,
                       None, discounting_adjustment, notional)

    @staticmethod
    def swap(fixing_dates, payment_dates, swaprates, notionals=None, is_pay_fix=True):
        return Product(PRODUCT_SWAP, 0.0, fixing_dates, payment_dates, None, swaprates, notionals,
                       is_pay_fix=is_pay_fix)

    @staticmethod
    def caplet(maturity, period_length, strike, daycount_fraction=None, is_floorlet=False):
        dcf = period_length if daycount_fraction is None else daycount_fraction
        return Product(PRODUCT_CAPLET, maturity, None, None, [period_length], [strike], None, None, dcf,
                       is_floorlet=is_floorlet)


def product_array(products):
    arr = (FmProduct * len(products))()
    for i, p in enumerate(products):
        arr[i] = p.s
    return arr


class LmmContext:
    """One simulation context of the engine (a model + Brownian driver + Euler process)."""

    def __init__(self, times, tenor, num_factors: int, num_paths: int, seed: int, *, measure="SPOT",
                 state_space="LOGNORMAL", scheme="EULER_FUNCTIONAL", precision="FP64", libor_cap=1e5,
                 store_paths=False, checkpoint_stride=4, regenerate_increments=False, path_offset=0):
        self.lib = load()
        self.times = as_f64(times)
        self.tenor = as_f64(tenor)
        cfg = FmLmmConfig()
        cfg.num_time_steps = self.times.size - 1
        cfg.times = _dptr(self.times)
        cfg.num_libors = self.tenor.size - 1
        cfg.tenor = _dptr(self.tenor)
        cfg.num_factors = int(num_factors)
        cfg.num_paths = int(num_paths)
        cfg.seed = int(seed)
        cfg.measure = MEASURE[measure]
        cfg.state_space = STATE_SPACE[state_space]
        cfg.scheme = SCHEME[scheme]
        cfg.precision = PRECISION[precision]
        cfg.libor_cap = float(libor_cap) if libor_cap is not None else math.inf
        cfg.store_paths = 1 if store_paths else 0
        cfg.checkpoint_stride = int(checkpoint_stride)
        cfg.regenerate_increments = 1 if regenerate_increments else 0
        cfg.path_offset = int(path_offset)
        self.cfg = cfg
        h = c_void_p()
        check(self.lib.fm_lmm_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self.discount_factors = None
        self.factor_loading = None       # [nT][n][F], as set
        self.local_volatility = None     # (alpha, offsets[nT][n]), as set
        self._sv_driver = None           # driver of a stochastic volatility scaling, as set
        self.n = cfg.num_libors
        self.nT = cfg.num_time_steps
        self.F = cfg.num_factors
        self.P = cfg.num_paths

    def close(self):
        if getattr(self, "h", None):
            self.lib.fm_lmm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_initial_forwards(self, L0):
        a = as_f64(L0)
        check(self.lib.fm_lmm_set_initial_forwards(self.h, _dptr(a)))

    def set_covariance(self, factor_loading, variance):
        lam = as_f64(factor_loading).reshape(-1)
        var = as_f64(variance).reshape(-1)
        if lam.size != self.nT * self.n * self.F or var.size != self.nT * self.n:
            raise IllegalArgumentException("covariance table shapes do not match the configuration")
        check(self.lib.fm_lmm_set_covariance(self.h, _dptr(lam), _dptr(var)))
        self.factor_loading = lam.reshape(self.nT, self.n, self.F).copy()

    def set_local_volatility(self, alpha, offsets):
        """Local-volatility factor f = (L - L alpha) + b_ij (L + b_ij for alpha = 0) on every
        loading of component j at time index i; offsets is b as [nT][n] (a scalar is
        broadcast), None removes the wrapper."""
        if offsets is None:
            check(self.lib.fm_lmm_set_local_volatility(self.h, 0.0, POINTER(c_double)()))
            self.local_volatility = None
            return
        b = np.broadcast_to(as_f64(offsets), (self.nT, self.n)).copy()
        check(self.lib.fm_lmm_set_local_volatility(self.h, float(alpha), _dptr(b)))
        self.local_volatility = (float(alpha), b)

    def set_stochastic_volatility(self, nu, rho, driver, factor_a=0, factor_b=1):
        """Scale every loading by S(t_i) = exp(X), X an Euler process with drift -nu^2/2 and loadings
        (rho nu, sqrt(1 - rho^2) nu) on factors (factor_a, factor_b) of the context `driver`'s
        Brownian motion (None removes the scaling)."""
        h = None if driver is None else driver.h
        check(self.lib.fm_lmm_set_stochastic_volatility(self.h, float(nu), float(rho), h, int(factor_a),
                                                         int(factor_b)))
        self._sv_driver = driver

    def set_heston_volatility(self, kappa, theta, xi, driver, factor=0):
        """Scale every loading by sqrt(max(V(t_i), 0)), V a square-root process from 1 with mean
        reversion kappa to theta and volatility xi on factor `factor` of `driver`'s Brownian motion."""
        h = None if driver is None else driver.h
        check(self.lib.fm_lmm_set_heston_volatility(self.h, float(kappa), float(theta), float(xi), h, int(factor)))
        self._sv_driver = driver

    def set_discount_curve(self, df):
        if df is None:
            check(self.lib.fm_lmm_set_discount_curve(self.h, POINTER(c_double)()))
            self.discount_factors = None
        else:
            a = as_f64(df)
            check(self.lib.fm_lmm_set_discount_curve(self.h, _dptr(a)))
            self.discount_factors = a.copy()

    def simulate(self):
        check(self.lib.fm_lmm_simulate(self.h))

    def process_value(self, t: int, j: int) -> np.ndarray:
        out = np.empty(self.P)
        check(self.lib.fm_lmm_get_process_value(self.h, t, j, _dptr(out)))
        return out

    def numeraire(self, time: float) -> np.ndarray:
        out = np.empty(self.P)
        check(self.lib.fm_lmm_get_numeraire(self.h, float(time), _dptr(out)))
        return out

    def brownian_increment(self, t: int, k: int) -> np.ndarray:
        out = np.empty(self.P)
        check(self.lib.fm_lmm_get_brownian_increment(self.h, t, k, _dptr(out)))
        return out

    def value_products(self, products):
        arr = product_array(products)
        vals = np.empty(len(products))
        errs = np.empty(len(products))
        check(self.lib.fm_lmm_value_products(self.h, arr, len(products), _dptr(vals), _dptr(errs)))
        return vals, errs

    def value_products_noalloc(self, products):
        """value_products without device allocation (single-date products; the context must
        have been used by value_products once)."""
        arr = product_array(products)
        vals = np.empty(len(products))
        errs = np.empty(len(products))
        check(self.lib.fm_lmm_value_products_noalloc(self.h, arr, len(products), _dptr(vals), _dptr(errs)))
        return vals, errs

    def set_volatility_factorization(self, volatility, factors, variance_factor=None):
        v = as_f64(volatility).reshape(-1)
        f = as_f64(factors).reshape(-1)
        c = None if variance_factor is None else as_f64(variance_factor)
        check(self.lib.fm_lmm_set_volatility_factorization(self.h, _dptr(v), _dptr(f),
                                                           POINTER(c_double)() if c is None else _dptr(c)))

    def set_dirac_delta(self, method="DISCRETE_DELTA", width=0.05):
        """Derivative of a Bermudan's exercise indicator in its trigger within adjoint()
        (RandomVariableDifferentiableAADFactory's diracDeltaApproximationMethod and
        diracDeltaApproximationWidthPerStdDev): DISCRETE_DELTA (default), ONE or ZERO."""
        check(self.lib.fm_lmm_set_dirac_delta(self.h, DIRAC[method], float(width)))

    def serialize(self) -> bytes:
        """Configuration and tables as bytes (fm_lmm_serialize); the paths are regenerated."""
        n = c_int64(0)
        check(self.lib.fm_lmm_serialize(self.h, None, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        check(self.lib.fm_lmm_serialize(self.h, buf, ctypes.byref(n)))
        return buf.raw[:n.value]

    @classmethod
    def deserialize(cls, blob: bytes) -> "LmmContext":
        """A context that regenerates the serialised one's increments and paths (fm_lmm_deserialize)."""
        lib = load()
        h = c_void_p()
        buf = ctypes.create_string_buffer(bytes(blob), len(blob))
        check(lib.fm_lmm_deserialize(buf, len(blob), ctypes.byref(h)))
        self = cls.__new__(cls)
        self.lib = lib
        self.h = h
        # shape from the blob header (version 1: magic, version, then nT, n, F, P as int32)
        nT, n, F, P = struct.unpack_from("<4i", bytes(blob), 12)
        self.nT, self.n, self.F, self.P = nT, n, F, P
        self.times = self.tenor = None
        self.cfg = None
        self.discount_factors = None
        self.factor_loading = None
        self.local_volatility = None
        return self

    def bermudan_exercise_times(self, bermudan: int = 0) -> np.ndarray:
        """exerciseTime of BermudanSwaption.getValues per path (+inf: never exercised)."""
        out = np.empty(self.P)
        check(self.lib.fm_lmm_bermudan_exercise_times(self.h, bermudan, _dptr(out)))
        return out

    def adjoint(self, products, weights=None):
        """(value, dV/dL(0) [n], dV/dsigma [nT][n])"""
        arr = product_array(products)
        w = None if weights is None else as_f64(weights)
        value = np.zeros(1)
        dL0 = np.zeros(self.n)
        dsig = np.zeros(self.nT * self.n)
        check(self.lib.fm_lmm_adjoint(self.h, arr, len(products), POINTER(c_double)() if w is None else _dptr(w),
                                      _dptr(value), _dptr(dL0), _dptr(dsig)))
        return value[0], dL0, dsig.reshape(self.nT, self.n)

    def timings(self) -> np.ndarray:
        ms = np.zeros(5)
        check(self.lib.fm_lmm_last_timings(self.h, _dptr(ms), 5))
        return ms


def mt19937_words(seed: int, offset: int, count: int) -> np.ndarray:
    out = np.empty(count, dtype=np.uint32)
    check(load().fm_mt19937_words(seed, offset, count, out.ctypes.data_as(POINTER(c_uint32))))
    return out


def mt19937_doubles(seed: int, offset: int, count: int) -> np.ndarray:
    out = np.empty(count)
    check(load().fm_mt19937_doubles(seed, offset, count, _dptr(out)))
    return out


def device_info() -> str:
    buf = ctypes.create_string_buffer(256)
    check(load().fm_device_info(buf, 256))
    return buf.value.decode()

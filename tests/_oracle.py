"""ctypes access to the CPU oracle (oracle/lmm_oracle.cpp). Test infrastructure only."""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
from ctypes import POINTER, c_double, c_int, c_int32, c_int64, c_uint32, c_uint64, c_void_p

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "build", "liblmm_oracle.so")

_lib = None


def build_oracle() -> str:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    return ORACLE_LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            build_oracle()
        L = ctypes.CDLL(ORACLE_LIB)
        dp = POINTER(c_double)
        L.orc_mt_words.argtypes = [c_int64, c_uint64, c_uint64, POINTER(c_uint32)]
        L.orc_mt_words_from_state.argtypes = [POINTER(c_uint32), c_int, c_uint64, POINTER(c_uint32)]
        L.orc_mt_words_by_array.argtypes = [POINTER(c_uint32), c_int, c_uint64, POINTER(c_uint32)]
        L.orc_mt_seeded_state.argtypes = [c_int64, POINTER(c_uint32)]
        L.orc_mt_doubles.argtypes = [c_int64, c_uint64, c_uint64, dp]
        L.orc_icdf.argtypes = [c_double]
        L.orc_icdf.restype = c_double
        L.orc_icdf_array.argtypes = [dp, c_int64, dp]
        L.orc_round_time.argtypes = [c_double]
        L.orc_round_time.restype = c_double
        L.orc_brownian.argtypes = [dp, c_int, c_int, c_int, c_int32, dp]
        L.orc_eigen_sym.argtypes = [dp, c_int, dp, dp]
        L.orc_factor_reduction.argtypes = [dp, c_int, c_int, dp]
        L.orc_corr_expdecay.argtypes = [dp, c_int, c_double, c_int, dp, dp]
        L.orc_vol_fourparam.argtypes = [dp, c_int, dp, c_int, c_double, c_double, c_double, c_double, dp]
        L.orc_forward_curve.argtypes = [dp, dp, c_int, c_double]
        L.orc_forward_curve.restype = c_double
        L.orc_df_from_forward_curve.argtypes = [dp, dp, c_int, c_double, c_double]
        L.orc_df_from_forward_curve.restype = c_double
        L.orc_lmm_create.argtypes = [c_void_p]
        L.orc_lmm_create.restype = c_void_p
        L.orc_lmm_destroy.argtypes = [c_void_p]
        L.orc_lmm_process_value.argtypes = [c_void_p, c_int, c_int, dp]
        L.orc_lmm_brownian_increment.argtypes = [c_void_p, c_int, c_int, dp]
        L.orc_lmm_numeraire.argtypes = [c_void_p, c_double, dp]
        L.orc_lmm_numeraire_unadjusted.argtypes = [c_void_p, c_int, dp]
        L.orc_lmm_forward_rate.argtypes = [c_void_p, c_double, c_double, c_double, dp]
        L.orc_lmm_numeraire_adjustment.argtypes = [c_void_p, c_int]
        L.orc_lmm_numeraire_adjustment.restype = c_double
        L.orc_lmm_swaption.argtypes = [c_void_p, c_double, c_int, dp, dp, dp, dp, c_double, dp, dp, dp, dp]
        L.orc_lmm_swap.argtypes = [c_void_p, c_int, dp, dp, dp, dp, c_int, dp, dp, dp]
        L.orc_lmm_caplet.argtypes = [c_void_p, c_double, c_double, c_double, c_double, c_int, dp, dp]
        L.orc_lmm_bermudan.argtypes = [c_void_p, c_int, POINTER(c_int), dp, dp, dp, dp, dp, c_int, c_int,
                                       dp, dp, dp, dp, dp]
        L.orc_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(POINTER(c_double))


def f64(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def mt_words(seed, offset, count):
    out = np.empty(count, dtype=np.uint32)
    lib().orc_mt_words(seed, offset, count, out.ctypes.data_as(POINTER(c_uint32)))
    return out


def mt_doubles(seed, offset, count):
    out = np.empty(count)
    lib().orc_mt_doubles(seed, offset, count, _dp(out))
    return out


def icdf(p):
    p = f64(p).ravel()
    out = np.empty_like(p)
    lib().orc_icdf_array(_dp(p), p.size, _dp(out))
    return out


def brownian(times, F, P, seed):
    times = f64(times)
    nT = times.size - 1
    out = np.empty((nT, F, P))
    lib().orc_brownian(_dp(times), nT, F, P, seed, _dp(out))
    return out


def eigen_sym(a):
    a = f64(a)
    n = a.shape[0]
    ev = np.empty(n)
    V = np.empty((n, n))
    lib().orc_eigen_sym(_dp(a), n, _dp(ev), _dp(V))
    return ev, V


def factor_reduction(corr, m):
    corr = f64(corr)
    n = corr.shape[0]
    out = np.empty((n, m))
    lib().orc_factor_reduction(_dp(corr), n, m, _dp(out))
    return out


def corr_expdecay(tenor_fix, a, m):
    t = f64(tenor_fix)
    n = t.size
    fac = np.empty((n, m))
    corr = np.empty((n, n))
    lib().orc_corr_expdecay(_dp(t), n, a, m, _dp(fac), _dp(corr))
    return fac, corr


def vol_fourparam(times, tenor_fix, a, b, c, d):
    times = f64(times)
    tf = f64(tenor_fix)
    nT = times.size - 1
    out = np.empty((nT, tf.size))
    lib().orc_vol_fourparam(_dp(times), nT, _dp(tf), tf.size, a, b, c, d, _dp(out))
    return out


def df_from_forward_curve(node_t, node_f, offset, maturity):
    nt, nf = f64(node_t), f64(node_f)
    return lib().orc_df_from_forward_curve(_dp(nt), _dp(nf), nt.size, offset, maturity)


def forward_curve(node_t, node_f, t):
    nt, nf = f64(node_t), f64(node_f)
    return lib().orc_forward_curve(_dp(nt), _dp(nf), nt.size, t)


class Spec(ctypes.Structure):
    _fields_ = [
        ("nT", c_int), ("times", POINTER(c_double)), ("n", c_int), ("tenor", POINTER(c_double)),
        ("F", c_int), ("P", c_int), ("seed", c_int32), ("L0", POINTER(c_double)), ("vol", POINTER(c_double)),
        ("factors", POINTER(c_double)), ("cap", c_double), ("scheme", c_int), ("measure", c_int),
        ("state_space", c_int), ("pd", POINTER(c_double)), ("threads", c_int),
        ("lv_alpha", c_double), ("lv_offset", POINTER(c_double)),
        ("fc_n", c_int), ("fc_t", POINTER(c_double)), ("fc_f", POINTER(c_double)), ("fc_offset", c_double),
        ("interpolation", c_int),
        ("sv_on", c_int), ("sv_nu", c_double), ("sv_rho", c_double), ("sv_seed", c_int32),
        ("sv_F", c_int), ("sv_a", c_int), ("sv_b", c_int),
        ("sv_kind", c_int), ("sv_kappa", c_double), ("sv_theta", c_double), ("sv_xi", c_double),
    ]


class ModelTables:
    """Model inputs computed by the oracle's restatement of the reference classes."""

    def __init__(self, wl, *, vol_params=None, L0=None):
        self.wl = wl
        self.times = f64(wl.times())
        self.tenor = f64(wl.tenor())
        n = wl.num_libors
        a, b, c, d = vol_params if vol_params is not None else wl.vol_params
        self.vol = vol_fourparam(self.times, self.tenor[:n], a, b, c, d)        # [nT][n]
        self.factors, self.corr = corr_expdecay(self.tenor[:n], wl.corr_decay, wl.num_factors)
        self.lam = self.vol[:, :, None] * self.factors[None, :, :]             # sigma * f (Scalar.mult)
        self.var = (self.vol * self.vol) * 1.0
        node_t = self.tenor[:n]
        node_f = f64(wl.forwards)
        self.node_t, self.node_f = node_t, node_f
        if L0 is None:
            self.L0 = np.array([forward_curve(node_t, node_f, self.tenor[j]) for j in range(n)])
        else:
            self.L0 = f64(L0)
        self.pd = np.array([df_from_forward_curve(node_t, node_f, wl.period, self.tenor[k]) for k in range(n + 1)])


class OracleModel:
    def __init__(self, tables: ModelTables, num_paths=None, *, cap=1e5, scheme=2, measure=0, state_space=1,
                 with_discount=True, threads=8, seed=None, local_vol=None, interpolation=1, stochastic_vol=None,
                 heston_vol=None):
        """local_vol: (alpha, offsets) of a local-volatility wrapper, offsets [nT][n] or a scalar;
        stochastic_vol: (nu, rho, seed, F, a, b) of a stochastic volatility scaling driven by
        factors a, b of an F-factor Brownian motion with that seed; heston_vol: (kappa, theta, xi,
        seed, F, a) of a Heston variance scaling driven by factor a."""
        wl = tables.wl
        self.t = tables
        self.P = int(num_paths or wl.num_paths)
        self.n = wl.num_libors
        self._keep = [tables.times, tables.tenor, tables.L0, f64(tables.vol), f64(tables.factors), tables.pd]
        s = Spec()
        s.nT = tables.times.size - 1
        s.times = _dp(self._keep[0])
        s.n = self.n
        s.tenor = _dp(self._keep[1])
        s.F = wl.num_factors
        s.P = self.P
        s.seed = wl.seed if seed is None else seed
        s.L0 = _dp(self._keep[2])
        s.vol = _dp(self._keep[3])
        s.factors = _dp(self._keep[4])
        s.cap = math.inf if cap is None else cap
        s.scheme, s.measure, s.state_space = scheme, measure, state_space
        s.pd = _dp(self._keep[5]) if with_discount else POINTER(c_double)()
        s.threads = threads
        if local_vol is not None:
            nT = tables.times.size - 1
            self._keep.append(np.ascontiguousarray(np.broadcast_to(f64(local_vol[1]), (nT, self.n))))
            s.lv_alpha = float(local_vol[0])
            s.lv_offset = _dp(self._keep[-1])
        else:
            s.lv_alpha = 0.0
            s.lv_offset = POINTER(c_double)()
        # the model's forward curve (analytic adjustment of interpolated forwards)
        self._keep += [f64(tables.node_t), f64(tables.node_f)]
        s.fc_n = self._keep[-1].size
        s.fc_t, s.fc_f = _dp(self._keep[-2]), _dp(self._keep[-1])
        s.fc_offset = float(wl.period)
        s.interpolation = interpolation          # LINEAR 0, LOG_LINEAR_UNCORRECTED 1, LOG_LINEAR_CORRECTED 2
        s.sv_kind = 0
        if stochastic_vol is not None:
            s.sv_on = 1
            s.sv_nu, s.sv_rho, s.sv_seed, s.sv_F, s.sv_a, s.sv_b = stochastic_vol
        elif heston_vol is not None:
            s.sv_on, s.sv_kind, s.sv_b = 1, 1, 0
            s.sv_kappa, s.sv_theta, s.sv_xi, s.sv_seed, s.sv_F, s.sv_a = heston_vol
        else:
            s.sv_on, s.sv_F, s.sv_a, s.sv_b = 0, 2, 0, 1
        self.spec = s
        self.h = lib().orc_lmm_create(ctypes.byref(s))
        if not self.h:
            raise RuntimeError(lib().orc_last_error().decode())

    def close(self):
        if self.h:
            lib().orc_lmm_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def process_value(self, t, j):
        out = np.empty(self.P)
        lib().orc_lmm_process_value(self.h, t, j, _dp(out))
        return out

    def brownian_increment(self, t, k):
        out = np.empty(self.P)
        lib().orc_lmm_brownian_increment(self.h, t, k, _dp(out))
        return out

    def numeraire(self, T):
        out = np.empty(self.P)
        if lib().orc_lmm_numeraire(self.h, T, _dp(out)) != 0:
            raise RuntimeError(lib().orc_last_error().decode())
        return out

    def forward_rate(self, time, start, end):
        """getForwardRate(time, start, end), off-grid ends interpolated (LOG_LINEAR_UNCORRECTED)."""
        out = np.empty(self.P)
        if lib().orc_lmm_forward_rate(self.h, time, start, end, _dp(out)) != 0:
            raise RuntimeError(lib().orc_last_error().decode())
        return out

    def swaption(self, spec, per_path=False):
        fix, pay, K = f64(spec["fixing"]), f64(spec["payment"]), f64(spec["strike"])
        v, e = np.zeros(1), np.zeros(1)
        pp = np.empty(self.P) if per_path else None
        rc = lib().orc_lmm_swaption(self.h, spec["exercise"], fix.size, _dp(fix), _dp(pay), POINTER(c_double)(),
                                    _dp(K), spec.get("notional", 1.0), POINTER(c_double)(), _dp(v), _dp(e),
                                    _dp(pp) if per_path else POINTER(c_double)())
        if rc != 0:
            raise RuntimeError(lib().orc_last_error().decode())
        return (v[0], e[0], pp) if per_path else (v[0], e[0])

    def swap(self, spec):
        fix, pay, K = f64(spec["fixing"]), f64(spec["payment"]), f64(spec["strike"])
        v, e = np.zeros(1), np.zeros(1)
        rc = lib().orc_lmm_swap(self.h, fix.size, _dp(fix), _dp(pay), _dp(K), POINTER(c_double)(),
                                1 if spec.get("pay_fix", True) else 0, _dp(v), _dp(e), POINTER(c_double)())
        if rc != 0:
            raise RuntimeError(lib().orc_last_error().decode())
        return v[0], e[0]

    def caplet(self, maturity, period, strike, dcf=None, floorlet=False):
        v, e = np.zeros(1), np.zeros(1)
        rc = lib().orc_lmm_caplet(self.h, maturity, period, strike, period if dcf is None else dcf,
                                  1 if floorlet else 0, _dp(v), _dp(e))
        if rc != 0:
            raise RuntimeError(lib().orc_last_error().decode())
        return v[0], e[0]

    def bermudan(self, spec, per_path=False, coefficients=False):
        """BermudanSwaption.getValues; returns (value, error[, per-path values, exercise times][, coefficients])."""
        fix, pay, K = f64(spec["fixing"]), f64(spec["payment"]), f64(spec["strike"])
        pl = f64(spec.get("period_lengths", pay - fix))
        flags = np.ascontiguousarray(np.asarray(spec["exercise_flags"], dtype=np.int32))
        notionals = f64(spec["notionals"]) if spec.get("notionals") is not None else None
        v, e = np.zeros(1), np.zeros(1)
        pp, ex = np.empty(self.P), np.empty(self.P)
        nex = int(flags.sum())
        kb = 6 if spec.get("basis", 0) == 0 else 4
        coef = np.zeros((max(nex, 1), kb))
        rc = lib().orc_lmm_bermudan(self.h, fix.size, flags.ctypes.data_as(POINTER(c_int)), _dp(fix), _dp(pl),
                                    _dp(pay), _dp(notionals) if notionals is not None else POINTER(c_double)(),
                                    _dp(K), 1 if spec.get("callable", True) else 0, int(spec.get("basis", 0)),
                                    _dp(v), _dp(e), _dp(pp), _dp(ex), _dp(coef))
        if rc != 0:
            raise RuntimeError(lib().orc_last_error().decode())
        out = (v[0], e[0])
        if per_path:
            out = out + (pp, ex)
        if coefficients:
            out = out + (coef[:nex],)
        return out

    def value(self, spec):
        if spec["type"] == "bermudan":
            return self.bermudan(spec)
        return self.swaption(spec) if spec["type"] == "swaption" else self.swap(spec)


# ---- tape AAD of the oracle (oracle/lmm_aad.cpp) -------------------------------------------

class AadSpec(ctypes.Structure):
    _fields_ = [
        ("nT", c_int), ("times", POINTER(c_double)), ("n", c_int), ("tenor", POINTER(c_double)),
        ("F", c_int), ("P", c_int), ("seed", c_int32), ("L0", POINTER(c_double)), ("factors", POINTER(c_double)),
        ("cap", c_double), ("scheme", c_int), ("state_space", c_int), ("pd", POINTER(c_double)),
        ("vol_mode", c_int), ("vol", POINTER(c_double)), ("abcd", POINTER(c_double)), ("vol_time", POINTER(c_double)),
        ("dirac", c_int), ("dirac_width", c_double),
    ]


AAD_OPS = {"add": 1, "sub": 2, "mult": 3, "div": 4, "exp": 5, "log": 6, "invert": 7, "cap": 8, "floor": 9,
           "addProduct": 10, "discount": 11, "accrue": 12, "choose": 13, "condexp": 14}
DIRAC = {"ONE": 0, "ZERO": 1, "DISCRETE_DELTA": 2}


def _aad_lib():
    L = lib()
    if not hasattr(L, "_aad_ready"):
        dp, ip = POINTER(c_double), POINTER(c_int)
        L.orc_aad_swaptions.argtypes = [c_void_p, c_int, dp, ip, dp, dp, dp, dp, dp, dp, dp, dp]
        L.orc_aad_bermudan.argtypes = [c_void_p, c_int, ip, dp, dp, dp, dp, c_int, c_int, dp, dp, dp, dp]
        L.orc_aad_last_error.restype = ctypes.c_char_p
        L.orc_aad_tape_new.argtypes = [c_int, c_double]
        L.orc_aad_tape_new.restype = c_void_p
        L.orc_aad_tape_free.argtypes = [c_void_p]
        L.orc_aad_input.argtypes = [c_void_p, c_int, c_int, dp]
        L.orc_aad_estimator.argtypes = [c_void_p, c_int, c_int, dp]
        L.orc_aad_op.argtypes = [c_void_p, c_int, c_int, c_int, c_int, c_double]
        L.orc_aad_value.argtypes = [c_void_p, c_int, c_int, dp]
        L.orc_aad_derivative.argtypes = [c_void_p, c_int, c_int, c_int, dp]
        L._aad_ready = True
    return L


class AadTape:
    """A bare tape of the oracle's AAD: inputs, operations and per-path derivatives."""

    def __init__(self, size, dirac="DISCRETE_DELTA", width=0.05):
        self.size = int(size)
        self.h = _aad_lib().orc_aad_tape_new(DIRAC[dirac], width)

    def close(self):
        if self.h:
            _aad_lib().orc_aad_tape_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def input(self, values, differentiable=True):
        v = f64(np.atleast_1d(values))
        return _aad_lib().orc_aad_input(self.h, 1 if differentiable else 0, v.size, _dp(v))

    def estimator(self, basis):
        b = f64(basis)
        return _aad_lib().orc_aad_estimator(self.h, b.shape[0], b.shape[1], _dp(b))

    def op(self, name, a, b=-1, c=-1, imm=0.0):
        r = _aad_lib().orc_aad_op(self.h, AAD_OPS[name], a, b, c, float(imm))
        if r < 0:
            raise RuntimeError(_aad_lib().orc_aad_last_error().decode())
        return r

    def value(self, h):
        out = np.empty(self.size)
        _aad_lib().orc_aad_value(self.h, h, self.size, _dp(out))
        return out

    def derivative(self, result, wrt):
        out = np.empty(self.size)
        _aad_lib().orc_aad_derivative(self.h, result, wrt, self.size, _dp(out))
        return out


def _aad_spec(tables, P, *, with_discount, vol_mode, abcd, vol_time, vol, state_space, scheme, cap, dirac, width,
              seed, L0, keep):
    wl = tables.wl
    s = AadSpec()
    times, tenor = f64(tables.times), f64(tables.tenor)
    L0 = f64(tables.L0 if L0 is None else L0)
    fac = f64(tables.factors)
    volt = f64(tables.vol if vol is None else vol)
    keep += [times, tenor, L0, fac, volt, f64(tables.pd)]
    s.nT, s.times, s.n, s.tenor = times.size - 1, _dp(times), wl.num_libors, _dp(tenor)
    s.F, s.P, s.seed = wl.num_factors, int(P), wl.seed if seed is None else seed
    s.L0, s.factors = _dp(L0), _dp(fac)
    s.cap = math.inf if cap is None else cap
    s.scheme, s.state_space = scheme, state_space
    s.pd = _dp(keep[5]) if with_discount else POINTER(c_double)()
    s.vol_mode = vol_mode
    s.vol = _dp(volt)
    if vol_mode == 1:
        ab = f64(abcd if abcd is not None else wl.vol_params)
        vt = f64(times[:-1] if vol_time is None else vol_time)
        keep += [ab, vt]
        s.abcd, s.vol_time = _dp(ab), _dp(vt)
    s.dirac, s.dirac_width = DIRAC[dirac], width
    return s


def aad_swaptions(tables, P, specs, weights=None, *, with_discount=True, vol_mode=0, abcd=None, vol_time=None,
                  vol=None, state_space=1, scheme=2, cap=1e5, seed=None, L0=None):
    """Sum_q w_q Swaption_q.getValue and its tape gradient: (value, dY0[n], dvol[nT][n] | None,
    dabcd[4] | None). vol_mode 0: the volatility entries are the leaves; 1: a, b, c, d."""
    keep = []
    s = _aad_spec(tables, P, with_discount=with_discount, vol_mode=vol_mode, abcd=abcd, vol_time=vol_time, vol=vol,
                  state_space=state_space, scheme=scheme, cap=cap, dirac="ZERO", width=0.05, seed=seed, L0=L0,
                  keep=keep)
    ex = f64([sp["exercise"] for sp in specs])
    nper = np.ascontiguousarray([len(sp["fixing"]) for sp in specs], dtype=np.int32)
    fix = f64(np.concatenate([sp["fixing"] for sp in specs]))
    pay = f64(np.concatenate([sp["payment"] for sp in specs]))
    K = f64(np.concatenate([sp["strike"] for sp in specs]))
    w = f64(np.ones(len(specs)) if weights is None else weights)
    n, nT = tables.wl.num_libors, tables.times.size - 1
    v, dY, dv, dp_ = np.zeros(1), np.zeros(n), np.zeros((nT, n)), np.zeros(4)
    rc = _aad_lib().orc_aad_swaptions(ctypes.byref(s), len(specs), _dp(ex), nper.ctypes.data_as(POINTER(c_int)),
                                      _dp(fix), _dp(pay), _dp(K), _dp(w), _dp(v), _dp(dY), _dp(dv), _dp(dp_))
    if rc != 0:
        raise RuntimeError(_aad_lib().orc_aad_last_error().decode())
    return v[0], dY, (dv if vol_mode == 0 else None), (dp_ if vol_mode == 1 else None)


def aad_bermudan(tables, P, spec, *, with_discount=True, vol_mode=0, abcd=None, vol_time=None, vol=None,
                 state_space=1, scheme=2, cap=1e5, dirac="DISCRETE_DELTA", width=0.05, seed=None, L0=None):
    """BermudanSwaption.getValues and its tape gradient (outputs as aad_swaptions)."""
    keep = []
    s = _aad_spec(tables, P, with_discount=with_discount, vol_mode=vol_mode, abcd=abcd, vol_time=vol_time, vol=vol,
                  state_space=state_space, scheme=scheme, cap=cap, dirac=dirac, width=width, seed=seed, L0=L0,
                  keep=keep)
    fix, pay, K = f64(spec["fixing"]), f64(spec["payment"]), f64(spec["strike"])
    pl = f64(spec.get("period_lengths", pay - fix))
    flags = np.ascontiguousarray(np.asarray(spec["exercise_flags"], dtype=np.int32))
    n, nT = tables.wl.num_libors, tables.times.size - 1
    v, dY, dv, dp_ = np.zeros(1), np.zeros(n), np.zeros((nT, n)), np.zeros(4)
    rc = _aad_lib().orc_aad_bermudan(ctypes.byref(s), fix.size, flags.ctypes.data_as(POINTER(c_int)), _dp(fix),
                                     _dp(pl), _dp(pay), _dp(K), 1 if spec.get("callable", True) else 0,
                                     int(spec.get("basis", 0)), _dp(v), _dp(dY), _dp(dv), _dp(dp_))
    if rc != 0:
        raise RuntimeError(_aad_lib().orc_aad_last_error().decode())
    return v[0], dY, (dv if vol_mode == 0 else None), (dp_ if vol_mode == 1 else None)

"""The GPU pathwise adjoint against the oracle's tape AAD (oracle/lmm_aad.cpp), a restatement
of the reference's RandomVariableDifferentiableAAD (RandomVariableDifferentiableAAD.java:
165-219, 230-452), on the same Brownian paths.

Both hold every path average constant (getAverage returns a double: the numeraire adjustment
avg(N(0)/N(T)) of getNumeraire, LIBORMarketModelFromCovarianceModel.java:860-876), so the
discount-curve case is compared as it is, without freezing anything by hand. The Bermudan is
compared under the ZERO Dirac method (exercise decisions held fixed) and under the reference's
default DISCRETE_DELTA (and ONE), where the exercise indicator's trigger is differentiated.

Bars: the FP64 kernels within 1e-10 of the tape's largest entry (value within 1e-12); the
FP32-mixed kernels within 2e-3 of it (their paths round to float at every step).

The tape's leaves are the initial states Y_j(0) = log L_j(0) in the log-normal state space, so
dV/dL_j(0) = dV/dY_j(0) / L_j(0). The volatility leaves are the entries sigma_j(t_i) of the
simulation grid (or the four parameters a, b, c, d).
"""
import numpy as np
import pytest

import _oracle as O
from finmath_hip import _abi
from finmath_hip import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def gpu_context(wl, t, P, *, precision="FP64", with_discount=True, stride=4, state_space="LOGNORMAL", cap=1e5,
                dirac=None, width=0.05):
    ctx = _abi.LmmContext(t.times, t.tenor, wl.num_factors, P, wl.seed, precision=precision,
                          checkpoint_stride=stride, state_space=state_space, libor_cap=cap)
    ctx.set_initial_forwards(t.L0)
    ctx.set_covariance(t.vol[:, :, None] * t.factors[None, :, :], t.vol * t.vol * 1.0)
    ctx.set_discount_curve(t.pd if with_discount else None)
    ctx.set_volatility_factorization(t.vol, t.factors)
    if dirac is not None:
        ctx.set_dirac_delta(dirac, width)
    return ctx


def swaptions(specs):
    return [_abi.Product.swaption(s["exercise"], s["fixing"], s["payment"], s["strike"]) for s in specs]


def bermudan(spec):
    return _abi.Product.bermudan(spec["fixing"], spec["payment"], spec["period_lengths"], spec["strike"],
                                 spec["exercise_flags"], is_callable=spec["callable"], basis=spec["basis"])


def close(gpu, tape, tol, what):
    scale = np.max(np.abs(tape))
    err = np.max(np.abs(gpu - tape)) / scale
    say(f"  {what}: max |gpu - tape| / max |tape| = {err:.3e} (scale {scale:.3e})")
    assert err <= tol, what


@pytest.mark.parametrize("with_discount", [True, False])
@pytest.mark.parametrize("stride", [4, 1])
def test_strip_fp64_against_tape(with_discount, stride):
    wl = W.c1()
    t = O.ModelTables(wl)
    P = 4000
    say(f"case aad tape c1 strip P={P} discount={with_discount} stride={stride}")
    rng = np.random.default_rng(11)
    w = rng.uniform(-1.0, 2.0, len(wl.products))
    v, dL, dsig = gpu_context(wl, t, P, with_discount=with_discount, stride=stride).adjoint(swaptions(wl.products), w)
    vt, dY, dvol, _ = O.aad_swaptions(t, P, wl.products, w, with_discount=with_discount)
    assert v == pytest.approx(vt, rel=1e-12)
    close(dL * t.L0, dY, 1e-10, "deltas (log-forward leaves)")
    close(dsig, dvol, 1e-10, "volatility entries")


def test_strip_fp32_mixed_against_tape():
    wl = W.c3()
    t = O.ModelTables(wl)
    P = 4000
    say(f"case aad tape c3 strip FP32_MIXED P={P}")
    v, dL, dsig = gpu_context(wl, t, P, precision="FP32_MIXED").adjoint(swaptions(wl.products))
    vt, dY, dvol, _ = O.aad_swaptions(t, P, wl.products, with_discount=True)
    assert v == pytest.approx(vt, rel=2e-5)
    close(dL * t.L0, dY, 2e-3, "deltas")
    close(dsig, dvol, 2e-3, "volatility entries")


def test_normal_state_space_against_tape():
    # the set-up of the reference's own AAD test: NORMAL state space, liborCap infinite,
    # no discount curve (LIBORMarketModelNormalAADSensitivitiesTest.java:288-402); the
    # leaves are the forwards themselves
    wl = W.c1()
    t = O.ModelTables(wl)
    t.vol = t.vol * 0.02
    P = 4000
    say(f"case aad tape c1 normal state space P={P}")
    ctx = gpu_context(wl, t, P, with_discount=False, state_space="NORMAL", cap=float("inf"))
    v, dL, dsig = ctx.adjoint(swaptions(wl.products))
    vt, dY, dvol, _ = O.aad_swaptions(t, P, wl.products, with_discount=False, state_space=0, cap=None)
    assert v == pytest.approx(vt, rel=1e-12)
    close(dL, dY, 1e-10, "deltas")
    close(dsig, dvol, 1e-10, "volatility entries")


def test_parameter_vegas_against_tape():
    # the host mirror's getGradient (chain rule of the engine's dV/dsigma to a, b, c, d)
    # against the tape with a, b, c, d as its leaves
    from test_gpu_greeks import host_process, host_strip
    wl = W.c1().with_paths(4000)
    t = O.ModelTables(wl)
    say("case aad tape parameter vegas c1 P=4000")
    H, process = host_process(wl, tuple(wl.vol_params), discount=True, precision="FP64", paths=4000)
    g = process.getGradient(host_strip(H, wl), [], 4)
    vt, dY, _, dabcd = O.aad_swaptions(t, 4000, wl.products, with_discount=True, vol_mode=1)
    assert g["value"] == pytest.approx(vt, rel=1e-11)
    close(np.asarray(g["parameterVega"]), dabcd, 1e-9, "parameter vegas (a, b, c, d)")
    close(np.asarray(g["delta"]) * t.L0, dY, 1e-9, "deltas through the host mirror")


@pytest.mark.parametrize("basis", [0, 1])
def test_bermudan_fp64_against_tape_zero_method(basis):
    wl = W.c0()
    wl.forwards = 0.02 + 0.001 * np.arange(wl.num_libors)
    t = O.ModelTables(wl)
    P = 4000
    spec = W.bermudan_spec(2.0, 10.0, 0.5, 0.03, 1.0, True, basis=basis)
    say(f"case aad tape bermudan c0 basis={basis} P={P}")
    v, dL, dsig = gpu_context(wl, t, P, dirac="ZERO").adjoint([bermudan(spec)])
    vt, dY, dvol, _ = O.aad_bermudan(t, P, spec, with_discount=True, dirac="ZERO")
    assert v == pytest.approx(vt, rel=1e-11)
    close(dL * t.L0, dY, 1e-10, "deltas")
    close(dsig, dvol, 1e-10, "volatility entries")


@pytest.mark.parametrize("precision,tol", [("FP64", 1e-10), ("FP32_MIXED", 2e-3)])
def test_bermudan_80_forwards_10_factors_against_tape(precision, tol):
    # the c4 model and product: lmm_adjoint_kernel at 80 forwards and 10 factors
    wl = W.c4()
    t = O.ModelTables(wl)
    P = 2048
    spec = wl.products[0]
    say(f"case aad tape bermudan c4 {precision} P={P}")
    v, dL, dsig = gpu_context(wl, t, P, precision=precision, dirac="ZERO").adjoint([bermudan(spec)])
    vt, dY, dvol, _ = O.aad_bermudan(t, P, spec, with_discount=True, dirac="ZERO")
    assert v == pytest.approx(vt, rel=1e-11 if precision == "FP64" else 1e-4)
    close(dL * t.L0, dY, tol, "deltas")
    close(dsig, dvol, tol, "volatility entries")


@pytest.mark.parametrize("dirac,basis,callable_", [("DISCRETE_DELTA", 0, True), ("DISCRETE_DELTA", 1, True),
                                                   ("DISCRETE_DELTA", 0, False), ("ONE", 0, True)])
def test_bermudan_fp64_against_tape_trigger_methods(dirac, basis, callable_):
    # the reference's default DISCRETE_DELTA (and ONE): the exercise indicator's derivative in its
    # trigger, whose adjoint is regressed on the exercise date's basis
    wl = W.c0()
    wl.forwards = 0.02 + 0.001 * np.arange(wl.num_libors)
    t = O.ModelTables(wl)
    P = 4000
    spec = W.bermudan_spec(2.0, 10.0, 0.5, 0.03, 1.0, callable_, basis=basis)
    say(f"case aad tape bermudan c0 {dirac} basis={basis} callable={callable_} P={P}")
    v, dL, dsig = gpu_context(wl, t, P, dirac=dirac).adjoint([bermudan(spec)])
    vt, dY, dvol, _ = O.aad_bermudan(t, P, spec, with_discount=True, dirac=dirac)
    _, dY0, dvol0, _ = O.aad_bermudan(t, P, spec, with_discount=True, dirac="ZERO")
    # the trigger's term must matter for the comparison to pin it
    say(f"  trigger term: max |tape - tape(ZERO)| / max |tape| = {np.max(np.abs(dY - dY0)) / np.max(np.abs(dY)):.3e}")
    assert np.max(np.abs(dY - dY0)) > 1e-6 * np.max(np.abs(dY))
    assert v == pytest.approx(vt, rel=1e-11)
    # the trigger's adjoint is regressed on a nearly collinear basis (1, ds, ds^2, dl, dl^2, 1/N),
    # so two pseudo-inverse solves of its Gram matrix (Jacobi here, the tape's eigen solve) agree to
    # about 1e-8 of the gradient; the bar stays far below the trigger term (3e-2 of it here)
    close(dL * t.L0, dY, 1e-7, "deltas")
    close(dsig, dvol, 1e-7, "volatility entries")


@pytest.mark.parametrize("precision,tol", [("FP64", 1e-7), ("FP32_MIXED", 5e-3)])
def test_bermudan_80_forwards_10_factors_discrete_delta_against_tape(precision, tol):
    # the c4 model and product under the reference's default DISCRETE_DELTA
    wl = W.c4()
    t = O.ModelTables(wl)
    P = 2048
    spec = wl.products[0]
    say(f"case aad tape bermudan c4 DISCRETE_DELTA {precision} P={P}")
    v, dL, dsig = gpu_context(wl, t, P, precision=precision).adjoint([bermudan(spec)])
    vt, dY, dvol, _ = O.aad_bermudan(t, P, spec, with_discount=True, dirac="DISCRETE_DELTA")
    assert v == pytest.approx(vt, rel=1e-11 if precision == "FP64" else 1e-4)
    close(dL * t.L0, dY, tol, "deltas")
    close(dsig, dvol, tol, "volatility entries")

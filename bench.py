"""Benchmark: LIBOR market model Monte Carlo path-steps/s on one MI355X.

A step is one pricing pass of the hot path over the whole workload: generate the
MT19937 stream and the Brownian increments, run the Euler scheme for every path,
value the product strip (30 payer swaptions: expiries 1..10y into 10y swaps at
ATM and +-100bp) and reduce value and Monte Carlo error. Model and tables are
set up before the timed region; nothing is cached between steps (the Brownian
increments are regenerated every step).

The headline line is the plain FP64 form at 40 semiannual forwards x 5 factors x
1,000,000 paths. Beside it:
  fp32_mixed     the FP32-mixed form at 4,000,000 paths and its per-price difference to
                 the FP64 form on the same paths, in basis points
  fp64_logstate  FP64 carrying the log-forwards between steps
  variants       the general path kernels on the same workload: terminal measure with the
                 predictor-corrector, the displaced local-volatility and the stochastic
                 volatility wrappers (normal state space)
  bermudan       the Bermudan swaption by regression at 80 forwards x 10 factors x
                 1,000,000 paths (FP32-mixed), its FP64 price, and its adjoint Greeks
  greeks         adjoint Greeks of the strip (1,000,000 paths, FP32 paths) over one
                 pricing pass, checkpoints every 4 steps and every step
  roofline       the FP64 path kernel against the FP64 peak, with its measured HBM traffic
  cpu_baseline   the CPU restatement of the reference timed on the host cores

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--no-cpu] [--no-fp32]
                    [--no-logstate] [--no-variants] [--no-bermudan] [--no-greeks] [--only-fp64] [--copy-probe]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "finmath-lib_amd"))

import numpy as np  # noqa: E402

FP64_PEAK_TFLOPS = 78.6      # MI355X published FP64 peak (vector and matrix)
HBM_PEAK_GBPS = 8000.0       # MI355X published HBM3E bandwidth
# HBM bytes per launch of lmm_spot_kernel<5,5,double,true,0> on the 40 x 5 x 1,000,000-path strip:
# FETCH_SIZE and WRITE_SIZE counter passes scaled by a plain 1 GiB copy
# (profiles/lmm_spot_fp64_c1_hbm_traffic.md)
PATH_KERNEL_TRAFFIC_BYTES = 2.933e9
# VALU-busy fraction of the same kernel from its SQ counter pass: 4 SQ_ACTIVE_INST_VALU over
# 1,024 SIMDs x GRBM_GUI_ACTIVE / 8 (profiles/lmm_spot_c1_sq_counters.md)
PATH_KERNEL_VALU_BUSY = 0.673
FP32_PEAK_TFLOPS = 157.3     # MI355X published FP32 peak


def build_simulation(wl, precision="FP64", regenerate=True, measure="SPOT", state_space="LOGNORMAL",
                     scheme="EULER_FUNCTIONAL", vol_params=None, wrap=None):
    """The workload's model through the host mirror. wrap(H, cov, td) may return a covariance
    wrapper (local or stochastic volatility) around the plain model."""
    from finmath_hip import _finmath_host as H
    td = H.TimeDiscretizationFromArray(0.0, wl.num_time_steps, wl.period)
    tenor = H.TimeDiscretizationFromArray(0.0, wl.num_libors, wl.period)
    fc = H.ForwardCurveInterpolation.createForwardCurveFromForwards(
        "forwardCurve", list(tenor.getAsDoubleArray()[:-1]), list(wl.forwards), wl.period)
    a, b, c, d = wl.vol_params if vol_params is None else vol_params
    vol = H.LIBORVolatilityModelFourParameterExponentialForm(td, tenor, a, b, c, d, False)
    corr = H.LIBORCorrelationModelExponentialDecay(td, tenor, wl.num_factors, wl.corr_decay)
    cov = H.LIBORCovarianceModelFromVolatilityAndCorrelation(td, tenor, vol, corr)
    if wrap is not None:
        cov = wrap(H, cov, td)
    model = H.LIBORMarketModelFromCovarianceModel.of(tenor, None, fc, H.DiscountCurveFromForwardCurve(fc), None, cov,
                                                     [], {"measure": measure, "stateSpace": state_space})
    bm = H.BrownianMotionFromMersenneRandomNumbers(td, wl.num_factors, wl.num_paths, wl.seed)
    process = H.EulerSchemeFromProcessModel(model, bm, scheme, precision)
    process.setRegenerateIncrements(regenerate)
    sim = H.LIBORMonteCarloSimulationFromLIBORModel(process)
    products = []
    for s in wl.products:
        if s["type"] == "swaption":
            products.append(H.Swaption(s["exercise"], list(s["fixing"]), list(s["payment"]), list(s["strike"])))
        elif s["type"] == "bermudan":
            provider = H.RegressionBasisFunctionsSwapRate() if s.get("basis", 0) == 1 else None
            products.append(H.BermudanSwaption([bool(f) for f in s["exercise_flags"]], list(s["fixing"]),
                                               list(s["period_lengths"]), list(s["payment"]), [], list(s["strike"]),
                                               s.get("callable", True), provider))
        else:
            products.append(H.SimpleSwap(list(s["fixing"]), list(s["payment"]), list(s["strike"]), s.get("pay_fix", True)))
    return sim, process, products


def step_kernel_flops_per_path(wl) -> int:
    """Algorithmic flops of the path kernel per path (FMA = 2; add/mul/div/exp/log/min = 1):
    6F+11 per alive forward-step, 3 per numeraire date, 7 per swaption period."""
    F, n = wl.num_factors, wl.num_libors
    alive = n * (n - 1) // 2
    periods = sum(len(s["fixing"]) for s in wl.products if s["type"] == "swaption")
    return alive * (6 * F + 11) + 3 * n + 7 * periods


def run_gpu(wl, precision, steps, warmup, **model):
    import torch
    sim, process, products = build_simulation(wl, precision, regenerate=True, **model)
    results = None
    for _ in range(warmup):
        results = sim.getValues(products)
    torch.cuda.synchronize()
    stage = np.zeros(5)
    t0 = time.perf_counter()
    for _ in range(steps):
        results = sim.getValues(products)
        stage += np.asarray(process.lastTimings())
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / steps
    values = np.array([r["value"] for r in results])
    errors = np.array([r["error"] for r in results])
    return ms_per_step, stage / steps, values, errors


def run_greeks(wl, steps, stride, dirac="DISCRETE_DELTA"):
    """Time of the adjoint call (forward pass with checkpoints + reverse sweep) over the pricing pass;
    dirac is the Bermudan exercise convention (the reference's AAD factory default DISCRETE_DELTA)."""
    import torch
    sim, process, products = build_simulation(wl, wl.precision, regenerate=True)
    process.setDiracDeltaApproximation(dirac)
    sim.getValues(products)
    process.getGradient(products, [], stride)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.getValues(products)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    stage = np.zeros(5)
    for _ in range(steps):
        g = process.getGradient(products, [], stride)
        stage += np.asarray(process.lastTimings())
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    price_ms = (t1 - t0) * 1e3 / steps
    greeks_ms = (t2 - t1) * 1e3 / steps
    return price_ms, greeks_ms, g, stage / steps


def copy_probe(nbytes=1 << 30, reps=3):
    """A plain device-to-device copy of known size: its FETCH_SIZE / WRITE_SIZE counter
    readings give the scale factor from counter units to HBM bytes."""
    import torch
    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()


def host_threads() -> int:
    """Host cores this process may use: its CPU affinity, capped by OMP_NUM_THREADS when set."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(omp))) if omp.isdigit() and int(omp) > 0 else n


def cpu_baseline(wl, v_gpu, all_paths, one_paths, reps=5):
    """The oracle (C++ restatement of the reference CPU path) on the host cores.

    - Full-size parity: the oracle simulates every path of the workload once (all host cores)
      and its prices are compared with the FP64 GPU prices of the same paths.
    - Throughput: the median of `reps` timed runs (simulate + value the strip) on a bounded
      sample of the same workload, on all host cores and on one core.
    """
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import _oracle as O
    threads = host_threads()

    def run(paths, th):
        w = wl.with_paths(paths)
        t = O.ModelTables(w)
        t0 = time.perf_counter()
        om = O.OracleModel(t, paths, threads=th)
        vals = np.array([om.value(s)[0] for s in w.products])
        dt = time.perf_counter() - t0
        om.close()
        return dt, vals

    dt_full, v_oracle = run(wl.num_paths, threads)
    rel = np.abs(v_gpu - v_oracle) / np.abs(v_oracle)
    all_times = sorted(run(all_paths, threads)[0] for _ in range(reps))
    one_times = sorted(run(one_paths, 1)[0] for _ in range(reps))
    t_all, t_one = all_times[reps // 2], one_times[reps // 2]
    nT = wl.num_time_steps
    return {
        "value": all_paths * nT / t_all, "unit": "path-steps/s", "cores": threads, "kind": "port",
        "sample": f"{wl.name} model, {all_paths} paths, simulate + {len(wl.products)} swaptions, median of {reps} "
                  f"runs ({t_all:.2f} s) on {threads} threads",
        "host_nproc": os.cpu_count(),
        "one_core": {"value": one_paths * nT / t_one, "unit": "path-steps/s", "cores": 1,
                     "sample": f"{one_paths} paths, median of {reps} runs ({t_one:.2f} s)"},
        "full_size_parity": {"paths": wl.num_paths, "oracle_seconds": round(dt_full, 3), "threads": threads,
                             "max_rel_diff_vs_oracle": float(rel.max()),
                             "within_1e-10": bool(rel.max() <= 1e-10)},
    }


def run_rv_algebra(steps):
    """Device random-variable programs (include/finmath_rv.h) on 2^26 realisations: one
    step of the swaption recursion, (x * y + z) / (1.0 + y * 0.5) capped at 5, applied once
    and four times, materialised (3 inputs read, 1 written) and averaged without
    materialising (3 inputs read). Algorithmic bytes over the kernel time on the library's
    stream, timed around `steps` calls with one synchronisation at the end."""
    from finmath_hip import randomvariable as R
    n = 1 << 26
    rng = np.random.default_rng(1)
    X = R.RandomVariable.from_array(rng.standard_normal(n))
    Y = R.RandomVariable.from_array(rng.uniform(0.5, 2.0, n))
    Z = R.RandomVariable.from_array(rng.uniform(-1.0, 1.0, n))

    def prog(k):
        e = X
        for _ in range(k):
            e = e.mult(Y).add(Z).discount(Y, 0.5).cap(5.0)
        return e

    out = {"workload": "2^26 double realisations, program (x*y + z)/(1.0 + y*0.5) capped, 3 inputs",
           "peak_GBps": HBM_PEAK_GBPS}
    for k in (1, 4):
        prog(k).cache().getAverage()
        t0 = time.perf_counter()
        for _ in range(steps):
            r = prog(k).cache()
        r.getAverage()                       # synchronises the stream
        t1 = time.perf_counter()
        for _ in range(steps):
            prog(k).getAverage()
        t2 = time.perf_counter()
        ms_store = (t1 - t0) * 1e3 / steps
        ms_avg = (t2 - t1) * 1e3 / steps
        out[f"ops{4 * k}"] = {
            "materialise_ms": ms_store, "materialise_GBps": 32 * n / (ms_store * 1e-3) / 1e9,
            "average_ms": ms_avg, "average_GBps": 24 * n / (ms_avg * 1e-3) / 1e9}
    return out


def run_calibration(paths):
    """getCloneCalibrated of the four-parameter volatility to the c1 strip (30 swaptions),
    targets priced under (0.2, 0.05, 0.3, 0.1) on the calibration's Brownian motion, start
    (0.15, 0, 0.25, 0.15); every Levenberg-Marquardt evaluation is one fused pricing pass."""
    from finmath_hip import _finmath_host as H
    from finmath_hip import workloads as W
    wl = W.c1()
    td = H.TimeDiscretizationFromArray(0.0, wl.num_time_steps, wl.period)
    tenor = H.TimeDiscretizationFromArray(0.0, wl.num_libors, wl.period)
    fc = H.ForwardCurveInterpolation.createForwardCurveFromForwards(
        "forwardCurve", list(tenor.getAsDoubleArray()[:-1]), list(wl.forwards), wl.period)

    def model(params):
        vol = H.LIBORVolatilityModelFourParameterExponentialForm(td, tenor, *params, True)
        corr = H.LIBORCorrelationModelExponentialDecay(td, tenor, wl.num_factors, wl.corr_decay)
        cov = H.LIBORCovarianceModelFromVolatilityAndCorrelation(td, tenor, vol, corr)
        return cov, H.LIBORMarketModelFromCovarianceModel.of(tenor, None, fc, H.DiscountCurveFromForwardCurve(fc), None,
                                                             cov, [], {"measure": "SPOT", "stateSpace": "LOGNORMAL"})
    products = [H.Swaption(s["exercise"], list(s["fixing"]), list(s["payment"]), list(s["strike"])) for s in wl.products]
    _, mtrue = model([0.2, 0.05, 0.3, 0.1])
    bm = H.BrownianMotionFromMersenneRandomNumbers(td, wl.num_factors, paths, 31415)
    sim = H.LIBORMonteCarloSimulationFromLIBORModel(H.EulerSchemeFromProcessModel(mtrue, bm))
    targets = [r["value"] for r in sim.getValues(products)]
    cov0, m0 = model([0.15, 0.0, 0.25, 0.15])
    calib = [H.CalibrationProduct(p, v, 1.0) for p, v in zip(products, targets)]
    t0 = time.perf_counter()
    _, info = cov0.getCloneCalibrated(m0, calib, {"numberOfPaths": paths, "accuracy": 1e-12, "maxIterations": 200})
    dt = time.perf_counter() - t0
    return {"paths": paths, "seconds": dt, "iterations": info["iterations"], "valuations": info["valuations"],
            "ms_per_valuation": 1e3 * dt / max(1, info["valuations"]), "rms_error": info["rootMeanSquaredError"],
            "parameters": info["parameters"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-paths", type=int, default=100_000, help="oracle timing sample on all host cores")
    ap.add_argument("--cpu-paths-one", type=int, default=20_000, help="oracle timing sample on one core")
    ap.add_argument("--no-greeks", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-logstate", action="store_true")
    ap.add_argument("--no-variants", action="store_true", help="skip the general-kernel variants leg")
    ap.add_argument("--no-bermudan", action="store_true")
    ap.add_argument("--no-rv", action="store_true", help="skip the device random-variable algebra leg")
    ap.add_argument("--no-calibration", action="store_true", help="skip the covariance calibration leg")
    ap.add_argument("--only-fp64", action="store_true", help="the headline leg only (for counter passes)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a plain 1 GiB device copy (calibrates the HBM byte counters)")
    args = ap.parse_args()
    if args.only_fp64:
        args.no_fp32 = args.no_logstate = args.no_bermudan = args.no_greeks = args.no_cpu = args.no_rv = True
        args.no_variants = True
        args.no_calibration = True
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); paths would shard across GPUs as "
              "contiguous path ranges (path_offset), but this measures one MI355X.", file=sys.stderr)
        sys.exit(2)

    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    from finmath_hip import workloads as W

    wl = W.c1()
    ms, stage, v64, e64 = run_gpu(wl, "FP64", args.steps, args.warmup)
    path_steps = wl.path_steps
    value = path_steps / (ms * 1e-3)
    flops = step_kernel_flops_per_path(wl) * wl.num_paths
    step_ms = stage[1]
    achieved = flops / (step_ms * 1e-3) / 1e12
    line = {
        "metric": "LMM path-steps/s (plain FP64, 40 fwd x 5 factors x 1M paths, 30 swaptions)",
        "value": value, "unit": "path-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "c1: LMM spot measure, log-Euler, 40 semiannual forwards, 5 factors, 1,000,000 paths, "
                               "seed 3141, 30 payer swaptions (1..10y into 10y, ATM and +-100bp)",
                   "num_paths": wl.num_paths, "num_forwards": wl.num_libors, "num_factors": wl.num_factors,
                   "time_steps": wl.num_time_steps},
        "stage_ms": {"generator": stage[0], "path_kernel": stage[1], "reductions": stage[2], "call": stage[4]},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": PATH_KERNEL_TRAFFIC_BYTES,
                     "alu": "FP64 vector (MI355X FP64 vector and matrix peaks are both 78.6 TFLOP/s)",
                     "hbm_gbps_at_traffic": PATH_KERNEL_TRAFFIC_BYTES / (step_ms * 1e-3) / 1e9,
                     "kernel": "lmm_spot_kernel<5,5,double,true,0>", "flops_per_launch": flops,
                     "valu_busy": PATH_KERNEL_VALU_BUSY,
                     "valu_busy_source": "SQ counter pass, profiles/lmm_spot_c1_sq_counters.md"},
    }
    if not args.no_fp32:
        wl2 = W.c2()
        ms32, stage32, v32, e32 = run_gpu(wl2, "FP32_MIXED", max(2, args.steps // 2), 1)
        ms64b, _, v64b, e64b = run_gpu(wl2, "FP64", 1, 0)
        err_bp = np.abs(v32 - v64b) * 1e4
        line["fp32_mixed"] = {
            "workload": "c2: c1 model, 4,000,000 paths, FP32 path arithmetic, FP64 numeraire and sums",
            "value": wl2.path_steps / (ms32 * 1e-3), "unit": "path-steps/s", "ms_per_step": ms32,
            "fp64_same_paths_ms_per_step": ms64b,
            "fp64_same_paths_value": wl2.path_steps / (ms64b * 1e-3),
            "max_abs_err_bp_vs_fp64": float(err_bp.max()), "max_mc_stderr_bp": float(e64b.max() * 1e4),
            "within_0.1bp": bool(err_bp.max() < 0.1),
            # per price, in the strip's order (expiry 1..10y, strikes ATM-100bp, ATM, ATM+100bp)
            "err_bp_vs_fp64": [round(float(x), 6) for x in err_bp],
            "mc_stderr_bp": [round(float(x), 4) for x in e64b * 1e4],
        }
    if not args.no_logstate:
        ms_ls, stage_ls, v_ls, _ = run_gpu(wl, "FP64_LOGSTATE", args.steps, args.warmup)
        line["fp64_logstate"] = {
            "workload": "c1, FP64 arithmetic carrying the log-forwards between steps, fused multiply-add",
            "value": path_steps / (ms_ls * 1e-3), "unit": "path-steps/s", "ms_per_step": ms_ls,
            "path_kernel_ms": stage_ls[1], "max_rel_diff_vs_fp64": float(np.max(np.abs(v_ls - v64) / np.abs(v64)))}
    if not args.no_variants:
        # the general path kernels on the headline workload: terminal measure with the
        # predictor-corrector, and the local / stochastic volatility wrappers (normal state space)
        variants = {
            "terminal_predictor_corrector": dict(measure="TERMINAL", scheme="PREDICTOR_CORRECTOR"),
            "displaced_normal": dict(state_space="NORMAL", vol_params=(0.005, 0.0, 0.25, 0.002),
                                     wrap=lambda H, cov, td: H.DisplacedLocalVolatilityModel(cov, 4.0, False)),
            "stochastic_volatility_normal": dict(
                state_space="NORMAL",
                wrap=lambda H, cov, td: H.LIBORCovarianceModelStochasticVolatility(
                    H.BlendedLocalVolatilityModel(cov, 0.0, False),
                    H.BrownianMotionFromMersenneRandomNumbers(td, 2, wl.num_paths, 31415), 0.3, -0.3, False)),
        }
        line["variants"] = {"workload": "c1 strip, 1,000,000 paths, FP64, general path kernels"}
        for name, model in variants.items():
            msv, stv, _, _ = run_gpu(wl, "FP64", 3, 1, **model)
            line["variants"][name] = {"ms_per_step": msv, "value": path_steps / (msv * 1e-3),
                                      "unit": "path-steps/s", "path_kernel_ms": stv[1]}
    if not args.no_bermudan:
        wl4 = W.c4()
        ms4, st4, v4, e4 = run_gpu(wl4, "FP32_MIXED", max(2, args.steps // 2), 1)
        ms4d, st4d, v4d, e4d = run_gpu(wl4, "FP64", 1, 1)
        line["bermudan"] = {
            "workload": "c4: 80 semiannual forwards, 10 factors, flat 3.5 %, 1,000,000 paths; Bermudan payer swaption "
                        "10y no-call then callable annually to 30y at 3.5 %, regression basis 1, S, S^2, N",
            "value": wl4.path_steps / (ms4 * 1e-3), "unit": "path-steps/s", "ms_per_step": ms4,
            "stage_ms": {"generator": st4[0], "path_kernel": st4[1], "induction_and_reductions": st4[2]},
            "price": float(v4[0]), "mc_stderr": float(e4[0]),
            "fp64_ms_per_step": ms4d, "fp64_price": float(v4d[0]),
            "err_bp_vs_fp64": float(abs(v4[0] - v4d[0]) * 1e4), "mc_stderr_bp": float(e4d[0] * 1e4)}
        if not args.no_greeks:
            pm4, gm4, g4, gst4 = run_greeks(wl4, 5, 4)
            pm4z, gm4z, g4z, _ = run_greeks(wl4, 3, 4, dirac="ZERO")
            line["bermudan"]["greeks"] = {
                "outputs": "80 deltas, 80 forward vegas, 4 vol-parameter vegas, dV/dsigma table; exercise under the "
                           "reference's default DISCRETE_DELTA (trigger term regressed on each date's basis, width "
                           "0.05 sd)",
                "pricing_ms": pm4, "adjoint_ms_stride4": gm4, "ratio_stride4": gm4 / pm4,
                "adjoint_stage_ms": {"generator": gst4[0], "forward_path_kernel": gst4[1],
                                     "induction_and_reductions": gst4[2], "reverse_sweep": gst4[3]},
                "sum_delta": float(np.sum(g4["delta"])), "parameter_vega": [float(x) for x in g4["parameterVega"]],
                "zero_method": {"adjoint_ms_stride4": gm4z, "ratio_stride4": gm4z / pm4z,
                                "sum_delta": float(np.sum(g4z["delta"]))}}
    if not args.no_greeks:
        wl3 = W.c3()
        pm, gm, g, gst = run_greeks(wl3, max(2, args.steps // 3), 4)
        pm1, gm1, _, _ = run_greeks(wl3, max(2, args.steps // 3), 1)
        line["greeks"] = {
            "workload": "c3: c1 strip, 1,000,000 paths, FP32 paths, FP64 gradient sums; 40 deltas, 40 forward vegas, "
                        "4 vol-parameter vegas, dV/dsigma table",
            "pricing_ms": pm, "adjoint_ms_stride4": gm, "ratio_stride4": gm / pm,
            "adjoint_ms_stride1": gm1, "ratio_stride1": gm1 / pm1,
            "adjoint_stage_ms_stride4": {"generator": gst[0], "forward_path_kernel": gst[1], "reverse_sweep": gst[3],
                                         "call": gst[4]},
            "sum_delta": float(np.sum(g["delta"])), "parameter_vega": [float(x) for x in g["parameterVega"]]}
    if not args.no_rv:
        line["rv_algebra"] = run_rv_algebra(args.steps)
    if not args.no_calibration:
        line["calibration"] = {
            "workload": "c1 model, 4 volatility parameters to the 30-swaption strip, Levenberg-Marquardt with "
                        "forward-difference Jacobian, one fused pricing pass per evaluation, cached increments",
            "runs": [run_calibration(2000), run_calibration(100_000)]}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(wl, v64, args.cpu_paths, args.cpu_paths_one)
    if args.copy_probe:
        copy_probe()
    print(json.dumps(line))


if __name__ == "__main__":
    main()

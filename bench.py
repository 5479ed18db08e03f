"""Benchmark of the device Monte-Carlo engine on one MI355X.

    python bench.py [--steps K] [--warmup W] [--paths P] [--leg NAME ...]

A step is one device sweep of the 5-asset LSV worst-of basket (2^20 Sobol
paths, 365 two-day steps over 2y, 100 x 50 leverage surface, FP32 stepping
with FP64 accumulation), the book prepared and resident in HBM before the
timed region (grid, LSV leverage calibration and packing are setup).  Prints
ONE JSON line: paths/s of that sweep with its roofline, beside the other
workloads (legs): the basket with Cholesky-correlated spots and in FP64, the
basket's bump and pathwise Greeks, the Heston down-and-out call with bump and
pathwise Greeks, the vanilla grids in
one persistent graph-replayed launch, the learned control variate as
specified, the time to a 1bp confidence interval measured on the host clock,
and the CPU oracle timed on this host.  `--leg` runs only the named legs
(for profiling); the JSON line then holds those legs alone.
"""
import argparse
import json
import math
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from thoth_amd import abi, configs, mcl  # noqa: E402
from thoth_amd.book import BuiltRequest  # noqa: E402

FP32_PEAK_TFLOPS = 157.3   # MI355X vector FP32 (256 CU x 128 lanes x 2 x 2.4 GHz)
FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0  # dense matrix cores
HBM_PEAK_GBS = 8000.0

# Algorithmic FLOPs (FMA = 2; add/mul/min/max/transcendental = 1; integer Sobol
# work not counted), charged only for work the kernel does:
#   LSV asset-step: 2 normals x 35 + QE 22 + Andersen 21 + leverage lookup 9 = 122,
#     + 6 (the asset's share of a 5 x 5 lower Cholesky, 15 FMA / 5) when the
#     spots are correlated; worst-of payoff 16 per path;
#   Heston + discrete barrier step: 2 x 35 + 22 + 11 + 2 = 105;
#   vanilla-grid payoffs: 82 cells x (sub + max + FMA) = 328 per maturity;
#   one-step Black-Scholes path: 35 (FP64).
LSV_STEP = 122
CHOLESKY_SHARE = 6
WORST_OF = 16
HESTON_BARRIER_STEP = 105
GRID_MATURITY = 328
MLP_MACS = 9216            # 11 -> 64 -> 64 -> 64 -> 5 per evaluation
T975_15 = 2.131449545559323  # Student t, 15 degrees of freedom, two-sided 95 %


def basket(paths, precision=abi.FP32, correlated=False, greeks=0):
    return BuiltRequest(configs.lsv_worst_of_basket(paths=paths, precision=precision,
                                                    correlate_stochastic=correlated, greeks=greeks))


def shape_of(sh, tune_s):
    return {"lev_group": sh.lev_group, "workgroups_per_cu": sh.workgroups_per_cu, "tuned_ms": sh.device_ms,
            "default_shape_ms": sh.default_ms, "candidates": sh.candidates, "tune_s": tune_s}


# --shape G:W forces the plans' launch shape instead of tuning it (profiling
# runs, whose kernel statistics would otherwise include the tuner's candidates)
FORCED_SHAPE = None


def apply_shape(plan):
    """Tune the plan's launch shape (or force FORCED_SHAPE); returns its description."""
    t0 = time.perf_counter()
    if FORCED_SHAPE:
        plan.set_shape(*FORCED_SHAPE)
        sh = plan.shape()
    else:
        sh = plan.tune()
    return shape_of(sh, time.perf_counter() - t0)


def timed_plan(built, steps, warmup):
    """Sweeps of a prepared plan, its launch shape picked by the tuner first:
    (setup s, host-clock s per sweep list, device ms list, last result, shape)."""
    t0 = time.perf_counter()
    plan = mcl.Plan(built)
    setup = time.perf_counter() - t0
    shape = apply_shape(plan)
    for _ in range(warmup):
        plan.run()
    wall, dev = [], []
    res = None
    for _ in range(steps):
        t = time.perf_counter()
        res = plan.run()
        wall.append(time.perf_counter() - t)
        dev.append(res.sweep_ms)
    return setup, wall, dev, res, shape


def roofline(flops, kernel_ms, peak, kernel):
    achieved = flops / (kernel_ms * 1e-3) / 1e12
    return {"bound": "valu", "kernel": kernel, "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
            "frac": achieved / peak, "flops_per_launch": flops}


def basket_leg(paths, steps, warmup, precision=abi.FP32, correlated=False):
    """The basket sweep (uniform LSV kernel; correlated: its Cholesky form; FP64: the oracle-matched kernel)."""
    b = basket(paths, precision, correlated)
    setup, wall, dev, res, shape = timed_plan(b, steps, warmup)
    _, t, _ = mcl.grid(b)
    m = len(t) - 1
    na = len(b.spec.equities)
    per_step = LSV_STEP + (CHOLESKY_SHARE if correlated else 0)
    flops = paths * (m * na * per_step + WORST_OF)
    kernel_ms = sum(dev) / len(dev)
    fp32 = precision == abi.FP32
    kernel = ("sweep_kernel<float, 5, 8, 0, %d>" % (3 if correlated else 1)) if fp32 else "sweep_kernel<double, 5, 8>"
    return {"paths": paths, "steps": m, "assets": na, "correlated_spots": correlated,
            "precision": "fp32 stepping, fp64 accumulation" if fp32 else "fp64",
            "paths_per_s": paths * len(wall) / sum(wall), "ms_per_sweep": 1e3 * sum(wall) / len(wall),
            "kernel_ms": kernel_ms, "setup_s": setup, "premium": res.premium, "premium_trust": res.premium_trust,
            "launch_shape": shape,
            "roofline": roofline(flops, kernel_ms, FP32_PEAK_TFLOPS if fp32 else FP64_PEAK_TFLOPS, kernel)}


def cpu_baseline(sample_paths):
    """The CPU FP64 oracle (a restatement of the reference engine, which is
    single-threaded) timed on this host on a bounded sample of the basket."""
    try:
        from oracle import pyoracle
        pyoracle.lib()
    except Exception as e:  # the oracle is built on the CPU side; report if absent
        return {"value": None, "unit": "paths/s", "cores": 1, "kind": "port",
                "sample": "oracle unavailable: %s" % e}
    sec, _ = pyoracle.sweep_seconds(basket(sample_paths, abi.FP64))
    return {"value": sample_paths / sec, "unit": "paths/s", "cores": 1, "kind": "port",
            "sample": "%d paths of the same book (5-asset LSV worst-of, 365 steps), path sweep only (grid, LSV "
                      "calibration excluded), %.1f s on %s" % (sample_paths, sec,
                                                                platform.processor() or platform.machine())}


def _oracle_block(args):
    """One CPU worker of the all-cores baseline: the oracle's sweep of a
    disjoint Sobol block (run_cluster.cpp:308-339 splits paths the same way)."""
    paths, skip = args
    from oracle import pyoracle
    spec = configs.lsv_worst_of_basket(paths=paths, precision=abi.FP64)
    spec.mcl.sobol_skip = skip
    sec, _ = pyoracle.sweep_seconds(BuiltRequest(spec))
    return sec


def cpu_baseline_all_cores(paths_per_core):
    """The same oracle on every host core this process may use (at most 16, the
    host share of one GPU), one process per core on disjoint Sobol blocks;
    throughput = all paths / the slowest worker's sweep time."""
    import multiprocessing as mp
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(16, cores))
    try:
        with mp.get_context("spawn").Pool(cores) as pool:
            secs = pool.map(_oracle_block, [(paths_per_core, k * paths_per_core) for k in range(cores)])
    except Exception as e:
        return {"value": None, "unit": "paths/s", "cores": cores, "kind": "port", "sample": "failed: %s" % e}
    return {"value": cores * paths_per_core / max(secs), "unit": "paths/s", "cores": cores, "kind": "port",
            "sample": "%d processes x %d paths on disjoint Sobol blocks, path sweep only, slowest %.1f s"
                      % (cores, paths_per_core, max(secs))}


def _counter_kib(csv_path, kernel):
    """Mean per-dispatch value of the one counter in a rocprofv3 csv for kernels whose name contains `kernel`."""
    import csv
    vals = []
    with open(csv_path) as f:
        for row in csv.DictReader(f):
            if kernel in row["Kernel_Name"]:
                vals.append(float(row["Counter_Value"]))
    return sum(vals) / len(vals) if vals else None


def _profile_counter(counter, cmd, kernel, workdir):
    """Runs cmd under `rocprofv3 --pmc counter` (a child process, its own time
    limit) and returns the counter's KiB per dispatch of `kernel`."""
    import glob
    import signal
    import subprocess
    out = os.path.join(workdir, counter + "_" + kernel.strip("_"))
    full = ["rocprofv3", "--pmc", counter, "-d", out, "-o", "run", "-f", "csv", "--"] + cmd
    p = subprocess.Popen(full, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, start_new_session=True,
                         cwd=workdir)
    try:
        p.wait(timeout=150)
    except subprocess.TimeoutExpired:
        os.killpg(p.pid, signal.SIGKILL)
        p.wait()
        return None
    files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
    return _counter_kib(files[0], kernel) if files else None


def hbm_traffic(paths, shape=None, kernel="sweep_kernel"):
    """HBM bytes per launch of the basket sweep: rocprofv3 FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch) of a one-step run at the measured launch
    shape (forced, so that no tuning candidates enter the average), each
    scaled by bytes / counter of a device-to-device copy of known size."""
    import shutil
    import tempfile
    if not shutil.which("rocprofv3"):
        return None, "rocprofv3 not found"
    probe_mib = 256
    me = [sys.executable, str(Path(__file__).resolve())]
    run = me + ["--leg", "c3", "--paths", str(paths), "--steps", "1", "--warmup", "0"]
    if shape:
        run += ["--shape", "%d:%d" % shape]
    probe = me + ["--copy-probe", str(probe_mib)]
    detail = {}
    total = 0.0
    with tempfile.TemporaryDirectory() as wd:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            kib = _profile_counter(counter, run, kernel, wd)
            ref = _profile_counter(counter, probe, "copyBuffer", wd)
            if kib is None or not ref:
                return None, "%s pass failed" % counter
            factor = probe_mib * 1024 / ref
            detail[counter] = {"kib_per_dispatch": kib, "copy_probe_factor": factor}
            total += kib * 1024 * factor
    return total, detail


# ---- learned control variate and the time to a 1bp confidence interval ---------
def learned_cv(cfg, eval_replicates=256, eval_paths=1 << 12, seed=2026):
    """Train the hedge network on 2^18 engine paths (points of a disjoint Sobol
    block) and evaluate it on 2^20 fresh paths (random digital shifts of Sobol
    points 0 .. eval_paths - 1), against plain MC on the same paths."""
    b = basket(eval_paths, abi.FP32)
    w, rep = mcl.cv_train(b, cfg)
    r = mcl.cv_price(b, cfg, w, eval_replicates, seed)
    # the in-loop sweep: 365 x 5 LSV asset-steps per path plus the network at J dates
    J = cfg.hedge_dates
    n = eval_replicates * eval_paths
    out = {"activation": "silu" if cfg.activation == abi.CV_SILU else "smoothstep-gated linear",
           "network": "11 -> 64 -> 64 -> 64 -> 5, BF16 weights, FP32 accumulation",
           "adam": {"learning_rate": round(cfg.learning_rate, 6), "steps": rep.steps, "batch_paths": cfg.batch_paths},
           "train_paths": cfg.train_paths, "hedge_dates": J, "hedge_spacing": cfg.hedge_spacing,
           "train_ms": rep.train_ms, "train_data_ms": rep.data_ms, "loss_first": rep.loss_first,
           "loss_last": rep.loss_last,
           "eval_paths": n, "plain_mean": r.plain_mean, "plain_halfwidth": r.plain_halfwidth,
           "cv_mean": r.cv_mean, "cv_halfwidth": r.cv_halfwidth, "variance_ratio": r.plain_var / r.cv_var,
           "paired_cv_term": {"mean": r.cv_term_mean, "halfwidth": r.cv_term_halfwidth},
           "sweep_ms": r.sweep_ms, "paths_per_s": n / (r.sweep_ms * 1e-3),
           "inference_tflops_bf16": n * J * 2 * MLP_MACS / (r.sweep_ms * 1e-3) / 1e12}
    return w, rep, out


def _hw16(means):
    """Half-widths of consecutive groups of 16 replicate means."""
    g = np.asarray(means).reshape(-1, 16)
    return T975_15 * g.std(axis=1, ddof=1) / 4.0


def time_to_1bp(cfg, w, train_ms, target=0.01, batch=64, seed=101):
    """Host-clock time to a 95 % half-width of 1bp of notional (0.01 on the
    rebased 100 notional) on the basket, for plain MC in the FP64
    oracle-matched mode, plain MC in FP32 and the learned control variate
    (FP32), each a randomised-QMC estimate over 16 random digital shifts.

    For each estimator the paths per replicate N are the smallest of the grid
    256 ceil(2^(k/4)) whose expected 16-replicate half-width (the RMS over 4
    independent groups of 16 shifts) is at most the target, searched upward
    from 2^(3/2) below the 1/sqrt(N) prediction of a 4096-point pilot.  Then:
    * single: one pricing of 16 replicates at that N, alone on the GPU: the
      median host-clock time of 5 runs with the book resident (launch, sweep,
      inference, reduction; the book setup is reported apart);
    * batched: `batch` pricings (16 replicates each) in one launch, the
      host-clock time per pricing and the half-widths each pricing reached.
    The training time of the network is reported beside, with the number of
    pricings after which it has paid for itself."""
    b64 = lambda n: basket(n, abi.FP64)
    b32 = lambda n: basket(n, abi.FP32)
    cfg2 = mcl.cv_config(hedge_dates=cfg.hedge_dates, train_paths=cfg.train_paths,
                         train_sobol_skip=cfg.train_sobol_skip, adam_steps=cfg.adam_steps,
                         batch_paths=cfg.batch_paths, learning_rate=cfg.learning_rate, init_seed=cfg.init_seed,
                         eval_mode=abi.CV_TWO_PHASE, hedge_spacing=cfg.hedge_spacing, activation=cfg.activation)
    runs = {
        "plain_fp64": lambda n, R, s: mcl.price_rqmc(b64(n), R, s, per_replicate=True),
        "plain_fp32": lambda n, R, s: mcl.price_rqmc(b32(n), R, s, per_replicate=True),
        "learned_cv_fp32": lambda n, R, s: mcl.cv_price(b32(n), cfg, w, R, s, per_replicate=True),
        # the same estimator with the network evaluated after the sweep on its
        # samples (the sweep of a pricing this small runs in the equity-wave form)
        "learned_cv_fp32_two_phase": lambda n, R, s: mcl.cv_price(b32(n), cfg2, w, R, s, per_replicate=True),
    }

    def result(name, out):
        if name.startswith("learned_cv_fp32"):
            r, _, cm = out
            wall = r.run_wall_ms if r.run_wall_ms > 0 else r.sweep_ms + r.eval_ms
            return r, cm, wall, r.sweep_ms + r.eval_ms, r.setup_ms
        r, m = out
        return r, m, r.run_wall_ms, r.sweep_ms, r.setup_ms

    grid_n = lambda k: 256 * math.ceil(2 ** (k / 4))
    res = {"target_halfwidth": target, "replicates": 16}
    for name, run in runs.items():
        progress("time to 1bp: " + name)

        def expected_hw(n):
            _, means, _, _, _ = result(name, run(n, 64, seed))
            return float(np.sqrt(np.mean(_hw16(means) ** 2)))
        if name == "learned_cv_fp32_two_phase":
            # the same estimator as the in-loop one: the same paths per replicate
            n, hw_expected = res["learned_cv_fp32"]["single"]["paths_per_replicate"], \
                res["learned_cv_fp32"]["single"]["expected_halfwidth"]
        else:
            pilot = 4096
            k = max(0, math.ceil(4 * math.log2(pilot / 256 * (expected_hw(pilot) / target) ** 2)) - 6)
            while True:
                n = grid_n(k)
                hw_expected = expected_hw(n)
                if hw_expected <= target or n >= 1 << 22:
                    break
                k += 1
        walls = []
        for k in range(5):
            r, means, wall, dev, setup = result(name, run(n, 16, seed + 1 + k))
            walls.append(wall)
        single = {"paths_per_replicate": n, "paths": 16 * n, "expected_halfwidth": hw_expected,
                  "halfwidth": float(_hw16(means)[0]), "wall_ms": float(np.median(walls)), "device_ms": dev,
                  "book_setup_ms": setup}
        if name == "learned_cv_fp32_two_phase":
            res[name] = {"single": single}
            continue
        r, means, wall, dev, _ = result(name, run(n, 16 * batch, seed + 100))
        hws = _hw16(means)
        batched = {"pricings": batch, "wall_ms_per_pricing": wall / batch, "device_ms_per_pricing": dev / batch,
                   "halfwidth_mean": float(hws.mean()), "halfwidth_max": float(hws.max()),
                   "fraction_at_target": float(np.mean(hws <= target))}
        res[name] = {"single": single, "batched": batched}
    cv, p64, p32 = res["learned_cv_fp32"], res["plain_fp64"], res["plain_fp32"]
    # a single pricing takes the faster evaluation of the network (in the loop,
    # or after the sweep on its samples); a batch the in-loop one
    cv2 = res["learned_cv_fp32_two_phase"]
    res["single_cv_evaluation"] = "two_phase" if cv2["single"]["wall_ms"] < cv["single"]["wall_ms"] else "in_loop"
    for regime, key in (("single", "wall_ms"), ("batched", "wall_ms_per_pricing")):
        t64, t32, tcv = p64[regime][key], p32[regime][key], cv[regime][key]
        if regime == "single":
            tcv = min(tcv, cv2["single"]["wall_ms"])
        res["speedup_" + regime] = {
            "learned_cv_fp32_vs_plain_fp64": t64 / tcv,
            "learned_cv_fp32_vs_plain_fp32": t32 / tcv,
            "plain_fp32_vs_plain_fp64": t64 / t32,
            "training_break_even_pricings_vs_plain_fp64": train_ms / (t64 - tcv) if t64 > tcv else None,
        }
    res["training_ms"] = train_ms
    return res


# ---- other workloads ---------------------------------------------------------------
def heston_barrier(paths=1 << 20, steps=252, repeats=3):
    """The single-asset Heston down-and-out call (v0 = theta = 0.04, kappa
    1.5, xi 0.5, rho -0.7, K 100, H 80 monitored daily) on 252 uniform steps:
    FP32 device time of the plain sweep with the bump Greeks (base + delta,
    gamma and vega scenario books in one launch) and of the pathwise kernel
    (delta, vega, gamma from pathwise deltas), and the FP32 - FP64 difference
    of every figure on the same Sobol points."""
    greeks = abi.GREEK_DELTA | abi.GREEK_GAMMA | abi.GREEK_VEGA
    out = {"workload": "heston_down_out_call", "paths": paths, "steps": steps}
    res = {}
    for prec, name in ((abi.FP32, "fp32"), (abi.FP64, "fp64")):
        spec = configs.heston_down_out_call(paths=paths, precision=prec, uniform_steps=steps, greeks=greeks)
        b = BuiltRequest(spec)
        plan = mcl.Plan(b)
        shape = apply_shape(plan)
        plan.run()
        ms = [plan.run().sweep_ms for _ in range(repeats)]
        r = plan.run()
        pw = [mcl.price_pathwise(b) for _ in range(repeats)]
        res[name] = {"bump": {"premium": r.premium, "premium_trust": r.premium_trust, "delta": r.delta,
                              "gamma": r.gamma, "vega": r.vega, "ms": min(ms), "books": r.n_scenarios,
                              "launch_shape": shape},
                     "pathwise": {"premium": pw[-1].premium, "delta": pw[-1].delta, "delta_trust": pw[-1].delta_trust,
                                  "gamma": pw[-1].gamma, "vega": pw[-1].vega, "vega_trust": pw[-1].vega_trust,
                                  "ms": min(p.sweep_ms for p in pw)}}
    out.update(res["fp32"])
    books = res["fp32"]["bump"]["books"]
    out["paths_per_s_bump"] = paths * books / (res["fp32"]["bump"]["ms"] * 1e-3)
    out["paths_per_s_pathwise"] = paths / (res["fp32"]["pathwise"]["ms"] * 1e-3)
    out["bump"]["roofline"] = roofline(books * paths * steps * HESTON_BARRIER_STEP, res["fp32"]["bump"]["ms"],
                                       FP32_PEAK_TFLOPS, "sweep_kernel<float, 1, 8, 0, 2>")
    out["fp32_minus_fp64"] = {
        kind: {k: res["fp32"][kind][k] - res["fp64"][kind][k] for k in ("premium", "delta", "gamma", "vega")}
        for kind in ("bump", "pathwise")}
    out["fp64_ms"] = {kind: res["fp64"][kind]["ms"] for kind in ("bump", "pathwise")}
    return out


def basket_greeks(paths=1 << 20, repeats=2, fp64_paths=1 << 14):
    """Greeks of the 5-asset LSV worst-of basket (365 steps, 100 x 50 leverage
    surface, FP32 stepping): the reference's bump Greeks (+1 % spot per
    equity, +-10 % for gamma, +1 vol point, the leverage recalibrated in every
    scenario: 17 books in one launch) beside the pathwise basket Greeks (the
    five deltas and the vega from one book, gamma from the pathwise deltas of
    the +-10 % books: 11 books in one launch); device time of each launch and
    host setup (calibrations, packing, upload), and the FP32 - FP64 difference
    of the pathwise figures on the same points (fp64_paths)."""
    greeks = abi.GREEK_DELTA | abi.GREEK_GAMMA | abi.GREEK_VEGA
    b = basket(paths, greeks=greeks)
    t0 = time.perf_counter()
    plan = mcl.Plan(b)
    setup = time.perf_counter() - t0
    plan.run()
    ms = [plan.run().sweep_ms for _ in range(repeats)]
    r = plan.run()
    pw = [mcl.price_pathwise(b) for _ in range(repeats)]
    p = pw[-1]
    n = p.n_assets
    out = {"workload": "lsv_worst_of_basket_greeks", "paths": paths, "steps": 365, "assets": n,
           "bump": {"premium": r.premium, "delta": r.delta, "gamma": r.gamma, "vega": r.vega,
                    "ms": min(ms), "books": r.n_scenarios, "setup_s": setup},
           "pathwise": {"premium": p.premium, "delta": p.delta, "delta_trust": p.delta_trust,
                        "asset_delta": [p.asset_delta[a] for a in range(n)],
                        "asset_delta_trust": [p.asset_delta_trust[a] for a in range(n)],
                        "gamma": p.gamma, "asset_gamma": [p.asset_gamma[a] for a in range(n)],
                        "vega": p.vega, "vega_trust": p.vega_trust,
                        "ms": min(x.sweep_ms for x in pw), "books": p.n_books,
                        "setup_s": min(x.setup_ms for x in pw) * 1e-3}}
    # delta and vega alone: one pathwise book against the 7 bump books
    dv = abi.GREEK_DELTA | abi.GREEK_VEGA
    b2 = basket(paths, greeks=dv)
    plan2 = mcl.Plan(b2)
    plan2.run()
    ms2 = [plan2.run().sweep_ms for _ in range(repeats)]
    pw2 = [mcl.price_pathwise(b2) for _ in range(repeats)]
    out["delta_vega_only"] = {"bump_ms": min(ms2), "bump_books": plan2.run().n_scenarios,
                              "pathwise_ms": min(x.sweep_ms for x in pw2), "pathwise_books": pw2[-1].n_books}
    small = {prec: mcl.price_pathwise(basket(fp64_paths, precision=prec, greeks=greeks))
             for prec in (abi.FP32, abi.FP64)}
    a, c = small[abi.FP32], small[abi.FP64]
    out["pathwise_fp32_minus_fp64"] = {
        "paths": fp64_paths, "premium": a.premium - c.premium, "delta": a.delta - c.delta,
        "asset_delta": [a.asset_delta[k] - c.asset_delta[k] for k in range(n)],
        "gamma": a.gamma - c.gamma, "vega": a.vega - c.vega}
    return out


def american_leg(repeats=3):
    """The American put of the reference's tests (S = K = 100, r 8 %, vol 30 %,
    1y in 30-day steps, pseudo-random draws, 300,000 paths) with delta, gamma,
    vega and rho: the sweep records the equity path of each of the 6 scenario
    books, then the Longstaff-Schwartz fit and the frozen-policy apply run on
    the device.  sweep_ms is the device time of both."""
    from thoth_amd.book import BookSpec, Equity, McConfig, Vanilla
    greeks = abi.GREEK_DELTA | abi.GREEK_GAMMA | abi.GREEK_VEGA | abi.GREEK_RHO
    out = {"workload": "american_put_lsm", "paths": 300000, "exercise_dates": None}
    for prec, name in ((abi.FP32, "fp32"), (abi.FP64, "fp64")):
        spec = BookSpec(mcl=McConfig(paths=300000, max_day_step=30, min_day_step=-1, vol_year_step=0.01,
                                     use_sobol=False),
                        equities=[Equity(spot=100.0, vol_kind=abi.VOL_BS, volatility=30.0)],
                        contracts=[Vanilla(0, 365, 100.0, is_call=False, american=True)],
                        rate_days=[0], rate_values=[8.0], greeks=greeks, precision=prec)
        plan = mcl.Plan(BuiltRequest(spec))
        plan.run()
        walls, rs = [], []
        for _ in range(repeats):
            t = time.perf_counter()
            rs.append(plan.run())
            walls.append(time.perf_counter() - t)
        r = rs[-1]
        out["exercise_dates"] = r.n_dates
        out[name] = {"premium": r.premium, "trust": r.contract_trust[0], "delta": r.delta, "gamma": r.gamma,
                     "vega": r.vega, "rho": r.rho, "device_ms": min(x.sweep_ms for x in rs),
                     "wall_ms": 1e3 * min(walls), "books": r.n_scenarios}
    return out


def prng_leg(paths=1 << 20):
    """Pseudo-random draws: the 8-equity LSV worst-of on 365 daily steps (16
    factors, 5,840 dimensions: 1,743 per path past the Sobol table, drawn
    from the device xoshiro stream), and the 5-asset basket with use_sobol
    false (every normal from the stream, no bridge), FP32, device ms per sweep."""
    out = {}
    spec = configs.lsv_worst_of_basket(paths=paths, n_assets=8, maturity_days=365, max_day_step=1)
    _, _, dev, res, shape = timed_plan(BuiltRequest(spec), 3, 1)
    out["lsv8_365_sobol_tail"] = {"paths": paths, "dimensions": 5840, "tail_dimensions": 1743,
                                  "device_ms": min(dev), "premium": res.premium, "trust": res.premium_trust}
    spec = configs.lsv_worst_of_basket(paths=paths)
    spec.mcl.use_sobol = False
    _, _, dev, res, shape = timed_plan(BuiltRequest(spec), 3, 1)
    out["basket_use_sobol_false"] = {"paths": paths, "device_ms": min(dev), "premium": res.premium,
                                     "trust": res.premium_trust}
    return out


def vanilla_grids(paths, replays=5):
    """The 5 LSV underlyings' vanilla grids (41 strikes x 12 monthly maturities,
    calls and puts, base + spot-bump + vol-bump books = 15 books of `paths`
    paths) in one persistent launch replayed from a hipGraph; setup includes
    the 15 leverage calibrations and the autotuner."""
    books = configs.lsv_vanilla_grids(paths=paths, precision=abi.FP32)
    t0 = time.perf_counter()
    plan = mcl.GridPlan([BuiltRequest(s) for s, _ in books], [mcl.vanilla_grid(**g) for _, g in books])
    setup = time.perf_counter() - t0
    plan.run()
    ms = [plan.run()[1] for _ in range(replays)]
    t = plan.tuning()
    dev = sum(ms) / len(ms)
    n_books = 3 * len(books)
    # per book-path: 365 LSV steps (one equity: no Cholesky) + 12 maturities x 82 cells
    flops = n_books * paths * (365 * LSV_STEP + 12 * GRID_MATURITY)
    # K books to a lane: the K equities of the grouped kernel (kind 5)
    k = t.books_per_lane
    kernel = "sweep_persistent<float, %d, 8, 5>" % k if k > 1 else "sweep_persistent<float, 1, 8, 1>"
    return {"workload": "lsv_vanilla_grids", "underlyings": len(books), "books": n_books,
            "cells_per_underlying": 2 * len(configs.GRID_STRIKES) * len(configs.GRID_MATURITIES),
            "paths_per_book": paths, "replays": replays, "device_ms": dev,
            "book_paths_per_s": n_books * paths / (dev * 1e-3),
            "roofline": roofline(flops, dev, FP32_PEAK_TFLOPS, kernel),
            "autotuned": {"books_per_lane": k, "workgroups_per_cu": t.blocks_per_cu,
                          "leverage_layers_per_barrier": t.lev_group,
                          "candidates": t.candidates, "best_ms": t.device_ms},
            "setup_s": setup}


LEGS = ("c3", "c3corr", "c3fp64", "c3greeks", "c2", "c4", "c5", "t1bp", "american", "prng", "cpu", "traffic")


def progress(msg):
    """One line per leg on stderr, so that a long run shows it is alive."""
    print("bench: %s" % msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--paths", type=int, default=1 << 20)
    ap.add_argument("--leg", action="append", choices=LEGS,
                    help="run only these legs (repeatable); default: all")
    ap.add_argument("--cpu-sample", type=int, default=1 << 16)
    ap.add_argument("--cpu-all-cores-paths", type=int, default=1 << 14,
                    help="paths per core of the all-cores CPU baseline (0 skips it)")
    ap.add_argument("--grid-paths", type=int, default=1 << 18)
    ap.add_argument("--no-gated", action="store_true", help="skip the smoothstep-gated network variant")
    ap.add_argument("--adam-steps", type=int, default=2000,
                    help="Adam steps of the hedge-network training (2000 as specified; fewer only for profiling)")
    ap.add_argument("--shape", default=None, metavar="G:W",
                    help="force the plans' launch shape (steps per LDS barrier G, persistent workgroups per CU W, "
                         "0 = one workgroup per Sobol block) instead of tuning it")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device (calibrates rocprof's FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.shape:
        global FORCED_SHAPE
        FORCED_SHAPE = tuple(int(x) for x in args.shape.split(":"))
    if args.copy_probe:
        import torch
        n = args.copy_probe * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": 4 * n}))
        return
    if args.gpus != 1:
        print("bench.py: this build is single-GPU (one MI355X); --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    legs = set(args.leg or LEGS)
    line = {}
    if "c3" in legs:
        progress("c3 basket")
        c3 = basket_leg(args.paths, args.steps, args.warmup)
        line.update({
            "metric": "paths/s & time-to-1bp CI: LSV basket, learned-CV FP32 vs plain MC, 1 MI355X",
            "value": c3["paths_per_s"],
            "unit": "paths/s",
            "n_gpus": 1,
            "steps": args.steps,
            "warmup": args.warmup,
            "ms_per_step": c3["ms_per_sweep"],
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f32",
            "data": "synthetic (seeded books, Joe-Kuo Sobol paths)",
            "config": {"workload": "lsv_worst_of_basket", "paths": args.paths, "steps": c3["steps"],
                       "assets": c3["assets"], "leverage_surface": "100 spot x 50 time",
                       "precision": c3["precision"], "correlated_spots": False},
            "kernel_ms": c3["kernel_ms"],
            "setup_s": c3["setup_s"],
            "premium": c3["premium"],
            "premium_trust": c3["premium_trust"],
            "roofline": dict(c3["roofline"], traffic=None),
            "launch_shape": c3["launch_shape"],
            "device": mcl.device_info(),
        })
    if "c3corr" in legs:
        progress("c3 correlated")
        line["basket_correlated"] = basket_leg(args.paths, max(3, args.steps // 2), 1, correlated=True)
    if "c3fp64" in legs:
        progress("c3 fp64")
        line["basket_fp64"] = basket_leg(args.paths, 3, 1, precision=abi.FP64)
    if "c3greeks" in legs:
        progress("c3 greeks: bump and pathwise")
        line["basket_greeks"] = basket_greeks(args.paths)
    if "c4" in legs:
        progress("c4 vanilla grids")
        line["vanilla_grids"] = vanilla_grids(args.grid_paths)
    if "c2" in legs:
        progress("c2 heston barrier")
        line["heston_barrier"] = heston_barrier()
    if "c5" in legs or "t1bp" in legs:
        progress("c5 learned control variate")
        cfg = mcl.cv_config(adam_steps=args.adam_steps)
        w, rep, c5 = learned_cv(cfg)
        line["learned_cv"] = c5
        if not args.no_gated:
            line["learned_cv_gated"] = learned_cv(mcl.cv_config_gated(adam_steps=args.adam_steps))[2]
        if "t1bp" in legs:
            line["time_to_1bp"] = time_to_1bp(cfg, w, rep.train_ms + rep.data_ms)
    if "american" in legs:
        progress("american put")
        line["american"] = american_leg()
    if "prng" in legs:
        progress("pseudo-random draws")
        line["pseudo_random"] = prng_leg(args.paths)
    if "traffic" in legs and "roofline" in line:
        progress("hbm traffic")
        sh = line.get("launch_shape") or {}
        traffic, detail = hbm_traffic(args.paths, (sh["lev_group"], sh["workgroups_per_cu"]) if sh else None)
        line["roofline"]["traffic"] = traffic
        line["roofline"]["traffic_detail"] = detail
        if traffic:
            line["roofline"]["hbm_gbs"] = traffic / (line["kernel_ms"] * 1e-3) / 1e9
            line["roofline"]["hbm_frac"] = line["roofline"]["hbm_gbs"] / HBM_PEAK_GBS
    if "cpu" in legs:
        progress("cpu baselines")
        line["cpu_baseline"] = cpu_baseline(args.cpu_sample)
        if args.cpu_all_cores_paths:
            line["cpu_baseline_all_cores"] = cpu_baseline_all_cores(args.cpu_all_cores_paths)
    print(json.dumps(line))


if __name__ == "__main__":
    main()

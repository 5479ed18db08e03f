"""Benchmark of the device Monte-Carlo sweep on one MI355X.

    python bench.py [--steps K] [--warmup W] [--paths P] [--workload NAME]

A step is one device sweep of the workload's book over `paths` Sobol paths
(the books are prepared and resident in HBM before the timed region: the LSV
leverage calibration, grid and packing are setup).  Prints ONE JSON line.
"""
import argparse
import json
import math
import os
import platform
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from thoth_amd import abi, configs, mcl  # noqa: E402
from thoth_amd.book import BuiltRequest  # noqa: E402

FP32_PEAK_TFLOPS = 157.3   # MI355X vector FP32 (256 CU x 128 lanes x 2 x 2.4 GHz)
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0

# Algorithmic FLOPs per path (FMA = 2; add/mul/min/max/transcendental = 1;
# integer Sobol work not counted): per LSV asset-step 2 normals x 35 + QE 22 +
# Andersen 21 + leverage lookup 9 + Cholesky share 6 = 128; worst-of payoff 16;
# Heston + discrete barrier step 105; one-step BS path 35 (FP64).
FLOPS = {
    "lsv_worst_of_basket": lambda steps, assets: steps * assets * 128 + 16,
    "heston_down_out_call": lambda steps, assets: steps * 105,
    "black_scholes_call": lambda steps, assets: 35,
}


def workload(name, paths):
    if name == "lsv_worst_of_basket":
        return configs.lsv_worst_of_basket(paths=paths, precision=abi.FP32)
    if name == "heston_down_out_call":
        return configs.heston_down_out_call(paths=paths, precision=abi.FP32)
    if name == "black_scholes_call":
        return configs.black_scholes_call(paths=paths)
    raise SystemExit("unknown workload %s" % name)


def cpu_baseline(name, sample_paths):
    """The CPU FP64 oracle (a restatement of the reference engine, which is
    single-threaded) timed on this host on a bounded sample of the workload."""
    try:
        sys.path.insert(0, str(ROOT))
        from oracle import pyoracle
        pyoracle.lib()
    except Exception as e:  # the oracle is built on the CPU side; report if absent
        return {"value": None, "unit": "paths/s", "cores": 1, "kind": "port",
                "sample": "oracle unavailable: %s" % e}
    spec = workload(name, sample_paths)
    spec.precision = abi.FP64
    sec, _ = pyoracle.sweep_seconds(BuiltRequest(spec))
    return {"value": sample_paths / sec, "unit": "paths/s", "cores": 1, "kind": "port",
            "sample": "%d paths of the same book, path sweep only (grid, LSV calibration excluded), "
                      "%.1f s on %s" % (sample_paths, sec, platform.processor() or platform.machine())}


def _oracle_block(args):
    """One CPU worker of the all-cores baseline: the oracle's sweep of a
    disjoint Sobol block (run_cluster.cpp:308-339 splits paths the same way)."""
    name, paths, skip = args
    from oracle import pyoracle
    spec = workload(name, paths)
    spec.precision = abi.FP64
    spec.mcl.sobol_skip = skip
    sec, _ = pyoracle.sweep_seconds(BuiltRequest(spec))
    return sec


def cpu_baseline_all_cores(name, paths_per_core):
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
            secs = pool.map(_oracle_block, [(name, paths_per_core, k * paths_per_core) for k in range(cores)])
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


def hbm_traffic(workload_name, paths, kernel="sweep_kernel"):
    """HBM bytes per launch of the sweep kernel: rocprofv3 FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch) of a one-step run of this workload, each
    scaled by bytes / counter of a device-to-device copy of known size."""
    import shutil
    import tempfile
    if not shutil.which("rocprofv3"):
        return None, "rocprofv3 not found"
    probe_mib = 256
    me = [sys.executable, str(Path(__file__).resolve())]
    run = me + ["--workload", workload_name, "--paths", str(paths), "--steps", "1", "--warmup", "0",
                "--no-cpu", "--no-t1bp", "--no-grid", "--no-traffic", "--no-heston"]
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


def time_to_1bp(replicates, paths, seed=101, target=0.01, other_regimes=((512, 1 << 14),)):
    """Wall time to a 95 % half-width of 1bp of notional (0.01 on the rebased
    100 notional) on the 5-asset LSV worst-of basket, plain Monte-Carlo in the
    FP64 oracle-matched mode against the FP32 learned control variate.  Each
    estimator is run at R randomised-QMC replicates of `paths` Sobol points
    (the same digital shifts for all) and its device time is projected to the
    target by the 1/sqrt(N) law of the half-width.

    The regime matters for plain MC, whose RQMC gain grows with the points per
    replicate: at the target it needs about 26 replicates of 4096 points (a CI
    needs at least 16 replicates, so larger replicates overshoot the target),
    and 4096 points per replicate is the default.  The learned control variate
    has no such gain (its residual is the high-frequency hedging error).  The
    same comparison at other (replicates, points) is reported beside it."""
    out = {"target_halfwidth": target, "replicates": replicates, "paths_per_replicate": paths,
           "projection": "device ms x (halfwidth / target)^2"}

    def entry(mean, hw, ms, **kw):
        e = {"premium": mean, "halfwidth": hw, "ms": ms, "projected_ms": ms * (hw / target) ** 2}
        e.update(kw)
        return e

    cfg = mcl.cv_config()
    w, rep = mcl.cv_train(BuiltRequest(configs.lsv_worst_of_basket(paths=paths, precision=abi.FP32)), cfg)
    train_ms = rep.data_ms + rep.train_ms

    def regime(R, N):
        res = {}
        b64 = BuiltRequest(configs.lsv_worst_of_basket(paths=N, precision=abi.FP64))
        r64 = mcl.price_rqmc(b64, R, seed)
        res["plain_fp64"] = entry(r64.mean, r64.halfwidth, r64.sweep_ms, per_path_var=r64.var)
        b32 = BuiltRequest(configs.lsv_worst_of_basket(paths=N, precision=abi.FP32))
        r32 = mcl.price_rqmc(b32, R, seed)
        res["plain_fp32"] = entry(r32.mean, r32.halfwidth, r32.sweep_ms, per_path_var=r32.var)
        rcv = mcl.cv_price(b32, cfg, w, R, seed)
        res["learned_cv_fp32"] = entry(rcv.cv_mean, rcv.cv_halfwidth, rcv.sweep_ms + rcv.eval_ms,
                                       per_path_var=rcv.cv_var, variance_ratio=rcv.plain_var / rcv.cv_var)
        res["speedup_vs_plain_fp64"] = res["plain_fp64"]["projected_ms"] / res["learned_cv_fp32"]["projected_ms"]
        return res

    out.update(regime(replicates, paths))
    out["learned_cv_fp32"].update(
        train_ms=train_ms,
        training={"paths": cfg.train_paths, "adam_steps": cfg.adam_steps, "batch_paths": cfg.batch_paths,
                  "hedge_dates": cfg.hedge_dates, "hedge_spacing": cfg.hedge_spacing,
                  "learning_rate": cfg.learning_rate, "loss_first": rep.loss_first, "loss_last": rep.loss_last})
    out["speedup_vs_plain_fp64_incl_training"] = out["plain_fp64"]["projected_ms"] / (
        out["learned_cv_fp32"]["projected_ms"] + train_ms)
    # one trained network serves every re-pricing of the book: pricings at the
    # target after which the training has paid for itself
    out["training_break_even_pricings"] = train_ms / (
        out["plain_fp64"]["projected_ms"] - out["learned_cv_fp32"]["projected_ms"])
    out["other_regimes"] = []
    for R, N in other_regimes:
        r = regime(R, N)
        out["other_regimes"].append({"replicates": R, "paths_per_replicate": N,
                                     "speedup_vs_plain_fp64": r["speedup_vs_plain_fp64"],
                                     "plain_fp64_halfwidth": r["plain_fp64"]["halfwidth"],
                                     "learned_cv_halfwidth": r["learned_cv_fp32"]["halfwidth"],
                                     "plain_fp64_ms": r["plain_fp64"]["ms"],
                                     "learned_cv_ms": r["learned_cv_fp32"]["ms"]})
    return out


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
        plan.run()
        ms = [plan.run().sweep_ms for _ in range(repeats)]
        r = plan.run()
        pw = [mcl.price_pathwise(b) for _ in range(repeats)]
        res[name] = {"bump": {"premium": r.premium, "premium_trust": r.premium_trust, "delta": r.delta,
                              "gamma": r.gamma, "vega": r.vega, "ms": min(ms), "books": r.n_scenarios},
                     "pathwise": {"premium": pw[-1].premium, "delta": pw[-1].delta, "delta_trust": pw[-1].delta_trust,
                                  "gamma": pw[-1].gamma, "vega": pw[-1].vega, "vega_trust": pw[-1].vega_trust,
                                  "ms": min(p.sweep_ms for p in pw)}}
    out.update(res["fp32"])
    out["paths_per_s_bump"] = paths * res["fp32"]["bump"]["books"] / (res["fp32"]["bump"]["ms"] * 1e-3)
    out["paths_per_s_pathwise"] = paths / (res["fp32"]["pathwise"]["ms"] * 1e-3)
    out["fp32_minus_fp64"] = {
        kind: {k: res["fp32"][kind][k] - res["fp64"][kind][k] for k in ("premium", "delta", "gamma", "vega")}
        for kind in ("bump", "pathwise")}
    out["fp64_ms"] = {kind: res["fp64"][kind]["ms"] for kind in ("bump", "pathwise")}
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
    # per book-path: 365 LSV steps x 128 FLOP + 12 maturities x 328 (82 cells)
    flops = n_books * paths * (365 * 128 + 12 * 328)
    achieved = flops / (dev * 1e-3) / 1e12
    return {"workload": "lsv_vanilla_grids", "underlyings": len(books), "books": n_books,
            "cells_per_underlying": 2 * len(configs.GRID_STRIKES) * len(configs.GRID_MATURITIES),
            "paths_per_book": paths, "replays": replays, "device_ms": dev,
            "book_paths_per_s": n_books * paths / (dev * 1e-3),
            "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": achieved / FP32_PEAK_TFLOPS},
            "autotuned": {"workgroups_per_cu": t.blocks_per_cu, "leverage_layers_per_barrier": t.lev_group,
                          "candidates": t.candidates, "best_ms": t.device_ms},
            "setup_s": setup}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--paths", type=int, default=1 << 20)
    ap.add_argument("--workload", default="lsv_worst_of_basket")
    ap.add_argument("--cpu-sample", type=int, default=1 << 16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-t1bp", action="store_true", help="skip the time-to-1bp comparison")
    ap.add_argument("--t1bp-replicates", type=int, default=2048)
    ap.add_argument("--t1bp-paths", type=int, default=1 << 12)
    ap.add_argument("--no-grid", action="store_true", help="skip the vanilla-grid batch")
    ap.add_argument("--grid-paths", type=int, default=1 << 18)
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 HBM-traffic passes")
    ap.add_argument("--no-heston", action="store_true", help="skip the Heston barrier (bump and pathwise Greeks)")
    ap.add_argument("--cpu-all-cores-paths", type=int, default=1 << 14,
                    help="paths per core of the all-cores CPU baseline (0 skips it)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device (calibrates rocprof's FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
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

    spec = workload(args.workload, args.paths)
    built = BuiltRequest(spec)
    t0 = time.perf_counter()
    plan = mcl.Plan(built)
    setup_s = time.perf_counter() - t0
    for _ in range(args.warmup):
        plan.run()
    wall, dev = [], []
    res = None
    for _ in range(args.steps):
        t = time.perf_counter()
        res = plan.run()
        wall.append(time.perf_counter() - t)
        dev.append(res.sweep_ms)
    total = sum(wall)
    ms_step = 1e3 * total / args.steps
    value = args.paths * args.steps / total
    kernel_ms = sum(dev) / len(dev)
    _, t, _ = mcl.grid(built)
    steps = len(t) - 1
    flops = args.paths * FLOPS[args.workload](steps, len(spec.equities))
    fp32 = spec.precision == abi.FP32
    peak = FP32_PEAK_TFLOPS if fp32 else FP64_PEAK_TFLOPS
    achieved = flops / (kernel_ms * 1e-3) / 1e12
    line = {
        "metric": "paths/s & time-to-1bp CI: LSV basket, learned-CV FP32 vs plain MC, 1 MI355X",
        "value": value,
        "unit": "paths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if fp32 else "f64",
        "data": "synthetic (seeded books, Joe-Kuo Sobol paths)",
        "config": {"workload": args.workload, "paths": args.paths, "steps": steps,
                   "assets": len(spec.equities), "precision": "fp32 stepping, fp64 accumulation" if fp32 else "fp64"},
        "kernel_ms": kernel_ms,
        "setup_s": setup_s,
        "premium": res.premium,
        "premium_trust": res.premium_trust,
        "roofline": {"bound": "valu", "kernel": "sweep_kernel<%s, %d, ...>" % ("float" if fp32 else "double",
                                                                            len(spec.equities)),
                     "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None},
        "device": mcl.device_info(),
    }
    if not args.no_grid:
        line["vanilla_grids"] = vanilla_grids(args.grid_paths)
    if not args.no_heston:
        line["heston_barrier"] = heston_barrier()
    if not args.no_t1bp and args.workload == "lsv_worst_of_basket":
        line["time_to_1bp"] = time_to_1bp(args.t1bp_replicates, args.t1bp_paths)
    if not args.no_traffic:
        traffic, detail = hbm_traffic(args.workload, args.paths)
        line["roofline"]["traffic"] = traffic
        line["roofline"]["traffic_detail"] = detail
        if traffic:
            line["roofline"]["hbm_gbs"] = traffic / (kernel_ms * 1e-3) / 1e9
            line["roofline"]["hbm_frac"] = line["roofline"]["hbm_gbs"] / HBM_PEAK_GBS
    if not args.no_cpu:
        cb = cpu_baseline(args.workload, args.cpu_sample)
        line["cpu_baseline"] = cb
        if args.cpu_all_cores_paths:
            line["cpu_baseline_all_cores"] = cpu_baseline_all_cores(args.workload, args.cpu_all_cores_paths)
    print(json.dumps(line))


if __name__ == "__main__":
    main()

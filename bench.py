#!/usr/bin/env python3
"""Benchmarks on one MI355X.  Prints ONE JSON line on stdout (progress on stderr).

Default workload (the headline metric): FP64 RK4 steps/s of the 2048x2048
Rossler LLE map -- Benettin LLE of x'=-y-z, y'=x+a*y, z'=b+z*(x-c), b=0.2,
a over linspace(0.05,0.35,2048), c over linspace(2,10,2048), RK4 h=0.01 from
(0.1,0.1,0.1), 500 time units of transient, then 4000 windows of 0.5 time
units on the base and the perturbed trajectory (perturbation 1e-8 along a
seeded unit vector): 450,000 RK4 steps per point, 1.887e12 per map.  One
benchmark "step" is one full map written to a device buffer, so nothing
crosses PCIe in the timed region.

--workload selects the other configurations (reported the same way):
  bif1d   1-D Rossler bifurcation, c in [2,6] x 1000, 1000 + 1000 t.u.
  bif2d   Lorenz peak-count map 1024^2 (rho x beta), 500 + 500 t.u., 1-D DBSCAN
  basins  Lorenz basins 2048^2 ICs, rho=24.5, 300 + 300 t.u., 3-D DBSCAN
  stream  the served mixed-job stream, engine vs one launch per job
  copy    streaming copy of --bytes (reference for HBM traffic counters)
  ls      legacy LS2D Lyapunov spectrum, Lorenz 1024^2 (rho x beta), 50 + 100 t.u.
  regression  src/main.cu-style sweeps (resolution 100..600, modelling time
          1000..6000) of bifurcation2D / LLE2D / basinsOfAttraction_2 through
          chaoslibrary_amd/chaos_regression, rows Test,Parameter,Value,Library,
          ExecutionTime_ms with Library cpu | hip-per-job | hip-batched
--semantics 0 runs bif2d / basins with the reference's own analyses (the
per-point dbscan on peak intervals; mean-peak / mean-interval features and
single linkage) instead of the configured ones.

The roofline block prices the dominant kernel's algorithmic FP64 work
(63 FLOP per Rossler RK4 step, 71 per Lorenz step, FMA = 2) against MI355X's
vector-FP64 peak (78.6 TFLOP/s) using the kernel's own HIP-event time;
cpu_baseline times the host oracle (the same per-point code, OpenMP over grid
points) on a bounded sample of the same workload.
"""
import argparse
import ctypes as C
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64 RHS steps/sec on 2048² LLE map (% FP64 peak); p50/p99 latency, mixed job stream"
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
FLOP_PER_STEP = {"rossler_rk4": 63, "lorenz_rk4": 71}
SYS_ID = {"rossler_rk4": 5, "lorenz_rk4": 6}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_threads():
    """OpenMP threads of the CPU baseline: OMP_NUM_THREADS when the host sets
    it (a shared GPU host reserves a CPU share per GPU that way), else every
    core this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:
        return max(1, os.cpu_count() or 1)


def cpu_info():
    """Model name, logical CPUs and threads per core of the host."""
    model, siblings, cores = None, None, None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "siblings" and siblings is None:
                    siblings = int(v)
                elif k == "cpu cores" and cores is None:
                    cores = int(v)
    except OSError:
        pass
    smt = siblings // cores if siblings and cores else None
    return {"cpu_model": model, "host_logical_cpus": os.cpu_count(), "threads_per_core": smt}


def oracle():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle
    return _oracle


def profile_figure(summary, key):
    """A figure recorded in a rocprofv3 summary under profiles/ (a line
    `key: value`), or None when the summary or the line is absent."""
    path = os.path.join(ROOT, "profiles", summary + ".md")
    try:
        with open(path) as f:
            for line in f:
                m = re.match(r"\s*" + re.escape(key) + r"\s*:\s*([-+0-9.eE]+)\s*$", line)
                if m:
                    return float(m.group(1))
    except OSError:
        pass
    return None


def roofline(system, steps, kernel_s, name, kernel_ms, summary=None):
    """FP64 VALU roofline of the dominant kernel.  `traffic` is the kernel's
    HBM bytes per launch from the FETCH_SIZE / WRITE_SIZE passes recorded in
    profiles/<summary>.md (scaled by the copy-kernel factors noted there)."""
    achieved = steps * FLOP_PER_STEP[system] / kernel_s / 1e12
    traffic = profile_figure(summary, "traffic_bytes_per_launch") if summary else None
    return {"bound": "valu_fp64", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": traffic, "kernel": name,
            "kernel_ms": round(kernel_ms, 3)}


def base_result(args, value, wall, steps, dtype="f64"):
    return {"metric": METRIC, "value": value, "unit": "RK4 steps/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": dtype}


def timed(args, fn):
    import chaoslibrary_amd as ch
    for i in range(args.warmup):
        st = fn()
        log(f"warmup {i}: kernel {st[0]:.1f} ms analysis {st[1]:.1f} ms")
    ch._check(ch.lib.chaos_device_synchronize(0))
    kms, ams, steps = [], [], 0.0
    t0 = time.perf_counter()
    for i in range(args.steps):
        k, a, s = fn()
        kms.append(k)
        ams.append(a)
        steps += s
        log(f"step {i}: kernel {k:.1f} ms analysis {a:.1f} ms")
    ch._check(ch.lib.chaos_device_synchronize(0))
    wall = time.perf_counter() - t0
    return wall, sum(kms) / len(kms), sum(ams) / len(ams), steps


# ---------------------------------------------------------------------------
# LLE map (headline)
# ---------------------------------------------------------------------------

def run_lle(args):
    import numpy as np
    import chaoslibrary_amd as ch
    n = args.grid or 2048
    w = dict(system="rossler_rk4", tMax=2000.0, NT=0.5, h=0.01, eps=1e-8, ic=[0.1, 0.1, 0.1],
             ranges=[0.05, 0.35, 2.0, 10.0], idx=[0, 2], transient=500.0, values=[0.2, 0.2, 5.7], seed=12345)
    stats = ch.Stats()
    dptr = C.c_void_p()
    ch._check(ch.lib.chaos_device_malloc(C.byref(dptr), n * n * 8, 0))

    def one():
        ch.lle2d(w["tMax"], w["NT"], n, w["h"], w["eps"], w["ic"], w["ranges"], w["idx"], w["transient"],
                 w["values"], system=w["system"], semantics=1, seed=w["seed"], block=args.block, stats=stats,
                 out_device_ptr=dptr.value)
        return stats.kernel_ms, 0.0, stats.rk4_steps

    wall, kms, _, steps = timed(args, one)
    host = np.empty(n * n)
    ch._check(ch.lib.chaos_memcpy_d2h(host.ctypes.data, dptr, n * n * 8))
    ch._check(ch.lib.chaos_device_free(dptr))
    if not np.isfinite(host).all():
        log("warning: non-finite exponents in the map")
    per_map = steps / args.steps
    res = base_result(args, steps / wall, wall, steps)
    res["data"] = "synthetic: seeded 2048x2048 Rossler (a,c) parameter grid, no input data"
    res["config"] = {"workload": f"lle2d_rossler_rk4_{n}x{n}", "grid": n, "h": w["h"],
                     "transient_time": w["transient"], "tMax": w["tMax"], "NT": w["NT"], "eps": w["eps"],
                     "seed": w["seed"], "rk4_steps_per_map": per_map, "flop_per_step": 63}
    res["roofline"] = roofline(w["system"], per_map, kms / 1e3, "lle_map_kernel<5>", kms,
                               f"lle_map_kernel_rossler_{n}" if n == 2048 else None)
    # the same rate as RHS evaluations (4 per RK4 step) and algorithmic FLOP
    res["rhs_evals_per_s"] = 4.0 * res["value"]
    res["tflop_per_s"] = res["value"] * 63 / 1e12
    res["pct_fp64_peak"] = 100.0 * res["tflop_per_s"] / 78.6
    if not args.no_cpu:
        log("cpu baseline ...")
        # threads spread over the cores (OpenMP reads this when the oracle
        # library loads its runtime); a caller's own setting wins
        os.environ.setdefault("OMP_PROC_BIND", "spread")
        O = oracle()
        ic = np.asarray(w["ic"])
        rg = np.asarray(w["ranges"])
        ix = np.asarray(w["idx"], dtype=np.int32)
        vals = np.asarray(w["values"])
        steps_pt = round(w["transient"] / w["h"]) + 2 * round(w["tMax"] / w["h"])
        threads = cpu_threads()

        def cpu(rows, stride):
            out = np.zeros(rows * n)
            t0 = time.perf_counter()
            used = O.lib.orc_lle2d_omp(SYS_ID[w["system"]], 1, w["tMax"], w["NT"], n, w["h"], w["eps"],
                                       O.dptr(ic), 3, O.dptr(rg), O.iptr(ix), w["transient"], O.dptr(vals), 3,
                                       w["seed"], O.dptr(out), 0, rows * stride, stride, threads)
            return time.perf_counter() - t0, used

        dt1, _ = cpu(1, n)  # probe one row to size the sample
        # three runs of a third of the CPU budget each; the median is reported
        rows = max(1, min(n, int(args.cpu_seconds / 3 / dt1)))
        stride = max(1, n // rows)
        rows = min(rows, n // stride)
        runs = [cpu(rows, stride) for _ in range(3)]
        dts = sorted(r[0] for r in runs)
        dt, used = dts[1], runs[0][1]
        res["cpu_baseline"] = {
            "value": rows * n * steps_pt / dt, "unit": "RK4 steps/s", "cores": used, "kind": "port",
            "sample": f"{rows} rows (every {stride}th) x {n} points of the same map = {rows * n} points, "
                      f"{steps_pt} RK4 steps each; median of 3 runs "
                      f"({', '.join(f'{d:.2f}' for d in dts)} s) on {used} OpenMP threads, "
                      f"OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}",
            **cpu_info()}
    if not args.no_stream:
        res["stream"] = stream_supplement(args)
    return res


ENGINE_TUNED = ("block", "waves_per_simd", "steps_per_launch", "points_per_tile", "use_graph")


def engine_params(args):
    """Engine parameters: the library defaults, overridden by the autotuner's
    choice recorded in profiles/engine_autotune_served_stream.md (lines
    `tuned_<field>: value`; --untuned skips them), then --spl, then --engine
    "key=value,..." (block, waves_per_simd, steps_per_launch,
    points_per_tile, use_graph, max_jobs, max_tiles)."""
    import chaoslibrary_amd as ch
    prm = ch.engine_params()
    if not getattr(args, "untuned", False):
        for k in ENGINE_TUNED:
            v = profile_figure("engine_autotune_served_stream", "tuned_" + k)
            if v is not None:
                setattr(prm, k, int(v))
    if args.spl:
        prm.steps_per_launch = args.spl
    for kv in filter(None, args.engine.split(",")):
        k, v = kv.split("=")
        setattr(prm, k.strip(), int(v))
    return prm


def stream_supplement(args):
    """The other half of the metric: the served mixed-job stream (Poisson
    arrivals at 200 jobs/s, the configurations' kinds and size ranges) served
    by the engine and by one launch per job, in this process on this GPU.
    Horizons x0.1 and a 10 s arrival window keep it to ~25 s; the stream at
    full horizons offers ~5.5x the GPU's FP64 peak and only measures the
    queue."""
    import chaoslibrary_amd as ch
    prm = engine_params(args)
    out = {}
    for mode, name in ((0, "batched"), (1, "per_job")):
        cfg = ch.stream_config(lambda_=200.0, duration_s=args.stream_duration, horizon_scale=args.stream_scale,
                               seed=2024, mode=mode)
        log(f"stream {name}: lambda 200/s for {args.stream_duration} s, horizons x{args.stream_scale}")
        rep, _ = ch.stream_run(cfg, prm)
        out[name] = {"jobs": rep.jobs, "makespan_s": rep.makespan_s, "rk4_steps_per_s": rep.steps_per_s,
                     "jobs_per_s": rep.jobs_per_s, "p50_ms": rep.p50_ms, "p99_ms": rep.p99_ms}
        log(f"  makespan {rep.makespan_s:.2f} s, {rep.steps_per_s:.3e} steps/s, p50 {rep.p50_ms:.0f} ms, "
            f"p99 {rep.p99_ms:.0f} ms")
    b, p = out["batched"], out["per_job"]
    return {"workload": f"served_stream_lambda200_{args.stream_duration:g}s_horizons_x{args.stream_scale:g}",
            "engine": {"block": prm.block, "waves_per_simd": prm.waves_per_simd,
                       "steps_per_launch": prm.steps_per_launch, "points_per_tile": prm.points_per_tile,
                       "use_graph": prm.use_graph},
            "batched": b, "per_job": p,
            "throughput_ratio": b["rk4_steps_per_s"] / p["rk4_steps_per_s"],
            "p50_ratio": b["p50_ms"] / p["p50_ms"], "p99_ratio": b["p99_ms"] / p["p99_ms"]}


# ---------------------------------------------------------------------------
# Bifurcation / basins configurations
# ---------------------------------------------------------------------------

def run_bif1d(args):
    import chaoslibrary_amd as ch
    stats = ch.Stats()
    n = 1000

    def one():
        ch.bif1d(1000.0, n, 0.01, [0.1, 0.1, 0.1], [2.0, 6.0], [2], 0, 0.0, 1000.0, [0.2, 0.2, 5.7], 1,
                 system="rossler_rk4", semantics=1, stats=stats)
        return stats.kernel_ms, stats.analysis_ms, stats.rk4_steps

    wall, kms, ams, steps = timed(args, one)
    res = base_result(args, steps / wall, wall, steps)
    res["metric"] = "FP64 RK4 steps/s, 1-D Rossler bifurcation (1000 points)"
    res["data"] = "synthetic: c = linspace(2,6,1000), a = b = 0.2"
    res["config"] = {"workload": "bif1d_rossler_rk4_1000", "points": n, "transient_time": 1000.0, "tMax": 1000.0,
                     "h": 0.01, "peaks": stats.peaks}
    res["roofline"] = roofline("rossler_rk4", steps / args.steps, kms / 1e3, "sweep_kernel<5,0,0>", kms)
    res["roofline"]["note"] = "1000 lanes = 16 waves on a 1024-SIMD GPU: latency-bound by design"
    res["analysis_ms"] = ams
    if not args.no_cpu:
        O = oracle()
        t0 = time.perf_counter()
        O.bif1d(O.SYS["rossler_rk4"], 1, 1000.0, 100, 0.01, [0.1, 0.1, 0.1], [2.0, 6.0], [2], 0, 0.0, 1000.0,
                [0.2, 0.2, 5.7], 1)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": 100 * 200000 / dt, "unit": "RK4 steps/s", "cores": 1, "kind": "port",
                               "sample": f"100 of the points (c over [2,6] x 100), {dt:.1f} s single-threaded"}
    return res


def run_bif2d(args):
    import chaoslibrary_amd as ch
    stats = ch.Stats()
    n = args.grid or 1024
    sem = 1 if args.semantics is None else args.semantics
    dptr = C.c_void_p()
    ch._check(ch.lib.chaos_device_malloc(C.byref(dptr), n * n * 4, 0))

    def one():
        ch.bif2d(500.0, n, 0.01, [1.0, 1.0, 1.0], [20.0, 200.0, 0.5, 4.0], [1, 2], 0, 0.0, 500.0,
                 [10.0, 28.0, 8 / 3], 1, 1e-2, system="lorenz_rk4", semantics=sem, stats=stats,
                 out_device_ptr=dptr.value)
        return stats.kernel_ms, stats.analysis_ms, stats.rk4_steps

    wall, kms, ams, steps = timed(args, one)
    ch._check(ch.lib.chaos_device_free(dptr))
    res = base_result(args, steps / wall, wall, steps)
    res["metric"] = f"FP64 RK4 steps/s, Lorenz peak-count map {n}^2 (whole pipeline)" + (
        ", reference per-point dbscan" if sem == 0 else "")
    res["data"] = "synthetic: rho = linspace(20,200), beta = linspace(0.5,4), sigma = 10"
    res["config"] = {"workload": f"bif2d_lorenz_rk4_{n}x{n}", "grid": n, "transient_time": 500.0, "tMax": 500.0,
                     "h": 0.01, "dbscan_eps": 1e-2, "peaks": stats.peaks, "semantics": sem}
    summary = f"sweep_kernel_bif2d_lorenz_{n}" if n == 1024 and sem == 1 else None
    res["roofline"] = roofline("lorenz_rk4", steps / args.steps, kms / 1e3,
                               "sweep_kernel<6,0,3>" if sem == 1 else "sweep_kernel<6,0,0>", kms, summary)
    res["analysis_ms"] = ams
    # peak output and 1-D DBSCAN stages: measured HBM bytes over their live time
    ab = profile_figure(summary, "analysis_bytes_per_run") if summary else None
    res["analysis_hbm_gbs"] = ab / (ams / 1e3) / 1e9 if ab and ams > 0 else None
    return res


def run_basins(args):
    import chaoslibrary_amd as ch
    stats = ch.Stats()
    n = args.grid or 2048
    sem = 1 if args.semantics is None else args.semantics
    dptr = C.c_void_p()
    ch._check(ch.lib.chaos_device_malloc(C.byref(dptr), n * n * 4, 0))

    def one():
        o = ch.make_options("lorenz_rk4", sem, stats=stats, device_buffers=True)
        ic, rg, ix = ch._f64([0.0, 0.0, 25.0]), ch._f64([-20.0, 20.0, -20.0, 20.0]), ch._i32([0, 1])
        vals = ch._f64([10.0, 24.5, 8 / 3])
        ch._check(ch.lib.chaos_basins2d(C.byref(o), 300.0, n, 0.01, 3, ch._dp(ic), ch._dp(rg), ch._ip(ix), 0, 0.0,
                                        300.0, ch._dp(vals), 3, 1, 0.5, dptr, None, None))
        return stats.kernel_ms, stats.analysis_ms, stats.rk4_steps

    wall, kms, ams, steps = timed(args, one)
    ch._check(ch.lib.chaos_device_free(dptr))
    res = base_result(args, steps / wall, wall, steps)
    res["metric"] = f"FP64 RK4 steps/s, Lorenz basins {n}^2 " + (
        "(features + z-score + 3-D DBSCAN)" if sem == 1 else "(reference features + single linkage)")
    res["data"] = "synthetic: IC grid x0, y0 in [-20,20], z0 = 25; rho = 24.5"
    res["config"] = {"workload": f"basins_lorenz_rk4_{n}x{n}", "grid": n, "transient_time": 300.0, "tMax": 300.0,
                     "h": 0.01, "eps": 0.5, "min_pts": 20, "semantics": sem}
    summary = f"sweep_kernel_basins_lorenz_{n}" if n == 2048 and sem == 1 else None
    res["roofline"] = roofline("lorenz_rk4", steps / args.steps, kms / 1e3,
                               "sweep_kernel<6,0,2>" if sem == 1 else "sweep_kernel<6,0,1>", kms, summary)
    res["analysis_ms"] = ams
    # z-score + 3-D DBSCAN stages: measured HBM bytes over their live time
    ab = profile_figure(summary, "analysis_bytes_per_run") if summary else None
    res["dbscan_hbm_gbs"] = ab / (ams / 1e3) / 1e9 if ab and ams > 0 else None
    return res


def run_ls(args):
    """Legacy LS2D: Lyapunov spectrum of Lorenz over (rho, beta), 4 trajectories per point."""
    import numpy as np
    import chaoslibrary_amd as ch
    n = args.grid or 1024
    stats = ch.Stats()
    dptr = C.c_void_p()
    ch._check(ch.lib.chaos_device_malloc(C.byref(dptr), n * n * 3 * 8, 0))
    ic, rg, ix, vals = (ch._f64([1.0, 1.0, 1.0]), ch._f64([20.0, 200.0, 0.5, 4.0]), ch._i32([1, 2]),
                        ch._f64([10.0, 28.0, 8 / 3]))

    def one():
        o = ch.make_options("lorenz_rk4", 0, stats=stats, device_buffers=True)
        ch._check(ch.lib.chaos_legacy_ls(C.byref(o), 100.0, 0.5, n, 0.01, 1e-8, ch._dp(ic), 3, ch._dp(rg),
                                         ch._ip(ix), 0, 0.0, 50.0, ch._dp(vals), 3, 2, dptr))
        return stats.kernel_ms, 0.0, stats.rk4_steps

    wall, kms, _, steps = timed(args, one)
    host = np.empty(n * n * 3)
    ch._check(ch.lib.chaos_memcpy_d2h(host.ctypes.data, dptr, host.nbytes))
    ch._check(ch.lib.chaos_device_free(dptr))
    res = base_result(args, steps / wall, wall, steps)
    res["metric"] = f"FP64 RK4 steps/s, legacy LS2D Lyapunov spectrum, Lorenz {n}^2 (4 trajectories per point)"
    res["data"] = "synthetic: rho = linspace(20,200), beta = linspace(0.5,4), sigma = 10"
    res["config"] = {"workload": f"ls2d_lorenz_rk4_{n}x{n}", "grid": n, "transient_time": 50.0, "tMax": 100.0,
                     "NT": 0.5, "h": 0.01, "eps": 1e-8}
    res["roofline"] = roofline("lorenz_rk4", steps / args.steps, kms / 1e3, "legacy_ls_kernel<6>", kms)
    res["spectrum_sum_mean"] = float(host.reshape(-1, 3).sum(axis=1).mean())
    return res


def run_stream(args):
    import numpy as np
    import chaoslibrary_amd as ch
    prm = engine_params(args)
    if args.tune:
        prm, rate = ch.autotune()
        log(f"autotune: block {prm.block} waves {prm.waves_per_simd} steps {prm.steps_per_launch} ({rate:.3e} steps/s)")
    out = {}
    modes = [(m, n) for m, n in ((0, "batched"), (1, "per_job")) if args.mode in ("both", n)]
    for mode, name in modes:
        cfg = ch.stream_config(lambda_=args.lam, duration_s=args.duration, horizon_scale=args.scale, seed=args.seed,
                               mode=mode, progress_s=30.0)
        log(f"stream {name}: lambda {args.lam}/s for {args.duration} s, horizons x{args.scale}")
        rep, lat = ch.stream_run(cfg, prm)
        out[name] = {k: v for k, v in rep.as_dict().items()}
        log(f"  jobs {rep.jobs} makespan {rep.makespan_s:.2f} s steps/s {rep.steps_per_s:.3e} "
            f"p50 {rep.p50_ms:.1f} ms p99 {rep.p99_ms:.1f} ms")
    both = "batched" in out and "per_job" in out
    main = out.get("batched") or out["per_job"]
    res = {"metric": "served mixed-job stream: RK4 steps/s (engine) with p50/p99 job latency",
           "value": main["steps_per_s"], "unit": "RK4 steps/s", "n_gpus": 1, "steps": 1, "warmup": 0,
           "ms_per_step": main["makespan_s"] * 1e3, "higher_is_better": True, "scaling": "strong",
           "vs_baseline": None, "dtype": "f64",
           "data": "synthetic seeded Poisson stream (bifurcation / LLE / basins jobs, log-uniform sizes)",
           "config": {"workload": "served_stream", "lambda": args.lam, "duration_s": args.duration,
                      "horizon_scale": args.scale, "seed": args.seed, "engine": {"block": prm.block,
                                                              "waves_per_simd": prm.waves_per_simd,
                                                              "steps_per_launch": prm.steps_per_launch,
                                                              "points_per_tile": prm.points_per_tile,
                                                              "use_graph": prm.use_graph}},
           "batched": out.get("batched"), "per_job": out.get("per_job"),
           "throughput_ratio": out["batched"]["steps_per_s"] / out["per_job"]["steps_per_s"] if both else None,
           "p99_ratio": out["batched"]["p99_ms"] / out["per_job"]["p99_ms"] if both else None}
    return res


def run_tune(args):
    """The engine autotuner on this GPU: each candidate serves the same short
    seeded stream (200 jobs/s over 1.5 s, full horizons) and is scored by its
    power, RK4 steps/s over the mean job latency; every candidate's rate and
    latency, and the chosen parameters (also written to --tune-out)."""
    import chaoslibrary_amd as ch
    path = args.tune_out
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    t0 = time.perf_counter()
    prm, rate = ch.autotune(cache_path=path)
    wall = time.perf_counter() - t0
    cands, mean_ms = [], None
    with open(path) as f:
        for line in f:
            if line.startswith("# candidate"):
                kv = line.split()[2:]
                cands.append({kv[i]: float(kv[i + 1]) for i in range(0, len(kv), 2)})
            elif line.startswith("mean_latency_ms"):
                mean_ms = float(line.split()[1])
    log(f"autotune: {len(cands)} candidates in {wall:.1f} s; chosen block {prm.block} waves {prm.waves_per_simd} "
        f"steps {prm.steps_per_launch} points/tile {prm.points_per_tile} graph {prm.use_graph} "
        f"(mean latency {mean_ms} ms, {rate:.3e} steps/s)")
    power = rate / mean_ms if mean_ms else None
    return {"metric": "engine autotuner: power (RK4 steps/s per ms of mean job latency) of a short served stream",
            "value": power, "unit": "RK4 steps/s per ms", "n_gpus": 1, "steps": len(cands), "warmup": 1,
            "ms_per_step": wall * 1e3 / max(1, len(cands)), "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic seeded stream (seed 7): 200 jobs/s over 1.5 s, bifurcation / LLE / basins kinds and "
                    "sizes of the served stream, full horizons",
            "config": {"workload": "engine_autotune_served_stream"},
            "chosen": {"block": prm.block, "waves_per_simd": prm.waves_per_simd,
                       "steps_per_launch": prm.steps_per_launch, "points_per_tile": prm.points_per_tile,
                       "use_graph": prm.use_graph, "steps_per_s": rate},
            "candidates": cands}


def _parse_suite(line):
    """A `SUITE <name> key=value ...` line of chaos_regression."""
    parts = line.split()
    d = {"name": parts[1]}
    for kv in parts[2:]:
        k, v = kv.split("=", 1)
        if k in ("init", "values", "ranges"):
            d[k] = [float(x) for x in v.split(",")]
        elif k == "idx":
            d[k] = [int(x) for x in v.split(",")]
        elif k in ("system", "semantics", "preScaller", "writableVar", "seed"):
            d[k] = int(v)
        else:
            d[k] = float(v)
    return d


def _cpu_run_ms(O, s, n, TT, threads):
    """Oracle time of one harness run (same inputs), on whole grid rows
    around the middle of the map (at least 4 points per thread), scaled to
    the n x n grid.  Basins: the per-point features only (no clustering)."""
    import numpy as np
    rows = max(1, min(n, -(-4 * threads // n)))
    r0 = (n - rows) // 2
    ic, vals, rg = (np.asarray(s[k], dtype=np.float64) for k in ("init", "values", "ranges"))
    ix = np.asarray(s["idx"], dtype=np.int32)
    t0 = time.perf_counter()
    if s["name"] == "LLE":
        out = np.zeros(rows * n)
        used = O.lib.orc_lle2d_omp(s["system"], s["semantics"], s["CT"], s["NT"], n, s["h"], s["eps"], O.dptr(ic),
                                   ic.size, O.dptr(rg), O.iptr(ix), TT, O.dptr(vals), vals.size, s["seed"],
                                   O.dptr(out), r0, r0 + rows, 1, threads)
        assert used > 0, used
    elif s["name"] == "Bifurcation":
        out = np.zeros(rows * n, dtype=np.int32)
        rc = O.lib.orc_bif2d(s["system"], s["semantics"], s["CT"], n, s["h"], O.dptr(ic), ic.size, O.dptr(rg),
                             O.iptr(ix), s["writableVar"], s["maxValue"], TT, O.dptr(vals), vals.size,
                             s["preScaller"], s["eps"], O.iptr(out), r0, r0 + rows)
        assert rc == 0, rc
    else:
        O.basins_features(s["system"], s["semantics"], s["CT"], n, s["h"], ic, rg, ix, s["writableVar"],
                          s["maxValue"], TT, vals, s["preScaller"], r0 * n, (r0 + rows) * n)
    dt = time.perf_counter() - t0
    return dt * 1e3 * n / rows, rows


def run_regression(args):
    """src/main.cu-style harness: chaoslibrary_amd/chaos_regression times the
    C++ API (hip-per-job) and the batched engine (hip-batched) over the
    resolution and modelling-time sweeps and writes
    <results>/<module>_performance_results.csv; the cpu rows appended here
    time the oracle on the same runs."""
    import shutil
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "chaoslibrary_amd", "chaos_regression")
    if not os.path.exists(exe):
        raise SystemExit("chaoslibrary_amd/chaos_regression is not built; run __graft_entry__.build()")
    results = os.path.abspath(args.results_dir)
    maps = tempfile.mkdtemp(prefix="chaos_regression_maps_")
    cmd = [exe, "--params", args.params, "--results", results, "--out", maps, "--semantics",
           "config" if args.semantics == 1 else "reference"]
    if args.resolutions:
        cmd += ["--resolutions", args.resolutions]
    if args.times:
        cmd += ["--times", args.times]
    if args.fixed_resolution:
        cmd += ["--fixed-resolution", str(args.fixed_resolution)]
    log("+ " + " ".join(cmd))
    suites, runs, summary = {}, [], ""
    t0 = time.perf_counter()
    proc = subprocess.Popen(cmd, stdout=subprocess.PIPE, text=True)
    for line in proc.stdout:
        line = line.rstrip("\n")
        log(line)
        if line.startswith("SUITE "):
            d = _parse_suite(line)
            suites[d["name"]] = d
        elif line.startswith("RUN "):
            _, test, par, val, lib, ms, steps, ok = line.split()
            runs.append({"test": test, "parameter": par, "value": int(val), "library": lib, "ms": float(ms),
                         "rk4_steps": float(steps), "ok": ok == "1"})
        elif line.startswith("rows "):
            summary = line
    rc = proc.wait()
    harness_s = time.perf_counter() - t0
    shutil.rmtree(maps, ignore_errors=True)
    if rc != 0:
        raise SystemExit(f"chaos_regression exited with {rc}: {summary}")

    cpu = {}
    if not args.no_cpu:
        O = oracle()
        threads = cpu_threads()
        for r in runs:
            if r["library"] != "hip-per-job" or r["parameter"] not in ("Resolution", "ModelingTime"):
                continue
            s = suites[r["test"]]
            n = r["value"] if r["parameter"] == "Resolution" else (args.fixed_resolution or 300)
            TT = s["TT"] if r["parameter"] == "Resolution" else float(r["value"])
            ms, rows = _cpu_run_ms(O, s, n, TT, threads)
            cpu[(r["test"], r["parameter"], r["value"])] = ms
            log(f"cpu {r['test']} {r['parameter']}={r['value']}: {ms:.0f} ms (from {rows} of {n} rows)")
        files = {"Bifurcation": "bifurcation_performance_results.csv", "LLE": "lle_performance_results.csv",
                 "Basins": "basins_performance_results.csv"}
        for (test, par, val), ms in cpu.items():
            with open(os.path.join(results, files[test]), "a") as f:
                f.write(f"{test},{par},{val},cpu,{ms:.3f}" + (",2D" if test == "Bifurcation" else "") + "\n")

    per_job = [r for r in runs if r["library"] == "hip-per-job" and r["ok"]
               and r["parameter"] in ("Resolution", "ModelingTime")]
    steps = sum(r["rk4_steps"] for r in per_job)
    ms = sum(r["ms"] for r in per_job)
    by_suite = {}
    for name in suites:
        d = {}
        for lib in ("hip-per-job", "hip-batched"):
            rr = [r for r in runs if r["test"] == name and r["library"] == lib]
            single = [r for r in rr if r["parameter"] in ("Resolution", "ModelingTime")]
            sweep = [r for r in rr if r["parameter"].endswith("Sweep")]
            if single:
                d[lib + "_ms"] = round(sum(r["ms"] for r in single), 3)
            if sweep:
                d[lib + "_sweeps_ms"] = round(sum(r["ms"] for r in sweep), 3)
        c_ms = [v for (t, _, _), v in cpu.items() if t == name]
        if c_ms:
            d["cpu_ms"] = round(sum(c_ms), 1)
            d["speedup_vs_cpu"] = round(sum(c_ms) / d["hip-per-job_ms"], 1) if d.get("hip-per-job_ms") else None
        if d.get("hip-batched_sweeps_ms") and d.get("hip-per-job_sweeps_ms"):
            d["batched_sweep_speedup"] = round(d["hip-per-job_sweeps_ms"] / d["hip-batched_sweeps_ms"], 3)
        by_suite[name] = d
    return {"metric": "src/main.cu-style regression harness: RK4 steps/s over the hip-per-job runs "
                      "(C++ API, map files included)",
            "value": steps / (ms / 1e3) if ms else 0.0, "unit": "RK4 steps/s", "n_gpus": 1, "steps": len(per_job),
            "warmup": 0, "ms_per_step": ms / max(1, len(per_job)), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic parameter / initial-condition grids",
            "config": {"workload": f"regression_{args.params}", "resolutions": args.resolutions or "100..600",
                       "modelling_times": args.times or "1000..6000", "semantics": args.semantics or 0},
            "suites": by_suite, "harness": summary, "harness_s": round(harness_s, 1), "results_dir": results,
            "cpu_rows": (f"oracle on whole middle grid rows (>= 4 points per thread, {cpu_threads()} OpenMP "
                         f"threads) scaled to the full grid; basins: per-point features only")
            if cpu else None}


def run_copy(args):
    import chaoslibrary_amd as ch
    nb = args.bytes
    a, b = C.c_void_p(), C.c_void_p()
    ch._check(ch.lib.chaos_device_malloc(C.byref(a), nb, 0))
    ch._check(ch.lib.chaos_device_malloc(C.byref(b), nb, 0))
    for _ in range(args.warmup):
        ch._check(ch.lib.chaos_device_copy(b, a, nb, 0))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ch._check(ch.lib.chaos_device_copy(b, a, nb, 0))
    wall = time.perf_counter() - t0
    ch._check(ch.lib.chaos_device_free(a))
    ch._check(ch.lib.chaos_device_free(b))
    gbs = 2 * nb * args.steps / wall / 1e9
    return {"metric": "streaming copy GB/s (read + write)", "value": gbs, "unit": "GB/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
            "data": "device buffers", "config": {"workload": "copy", "bytes": nb}}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed runs")
    ap.add_argument("--warmup", type=int, default=1, help="untimed runs")
    ap.add_argument("--workload", default="lle", choices=["lle", "bif1d", "bif2d", "basins", "stream", "copy", "ls",
                                                           "regression", "tune"])
    ap.add_argument("--tune-out", default=os.path.join("workspace", "engine_tune.txt"),
                    help="tune: where the autotuner writes its cache (chosen parameters + candidates)")
    ap.add_argument("--grid", type=int, default=0,
                    help="grid side (default: 2048 for lle and basins, 1024 for bif2d and ls)")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline sample time")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-stream", action="store_true", help="lle: skip the served-stream half of the metric")
    ap.add_argument("--stream-duration", type=float, default=10.0, help="lle: stream arrival window (s)")
    ap.add_argument("--stream-scale", type=float, default=0.1, help="lle: stream horizon scale")
    ap.add_argument("--lam", type=float, default=200.0, help="stream arrival rate (jobs/s)")
    ap.add_argument("--seed", type=int, default=2024, help="stream generator seed")
    ap.add_argument("--duration", type=float, default=60.0, help="stream arrival window (s)")
    ap.add_argument("--scale", type=float, default=1.0, help="stream horizon scale")
    ap.add_argument("--spl", type=int, default=0, help="engine steps per launch (0 = default)")
    ap.add_argument("--engine", default="", help='engine parameters "key=value,..." over the defaults')
    ap.add_argument("--untuned", action="store_true", help="engine: library defaults, not the recorded autotune")
    ap.add_argument("--tune", action="store_true", help="autotune the engine before the stream")
    ap.add_argument("--mode", default="both", choices=["both", "batched", "per_job"], help="stream modes to run")
    ap.add_argument("--bytes", type=int, default=1 << 30)
    ap.add_argument("--params", default="rk4", choices=["rk4", "reference"], help="regression parameter sets")
    ap.add_argument("--semantics", type=int, default=None, choices=[0, 1],
                    help="0 reference, 1 config (default: config for bif2d / basins, reference for regression)")
    ap.add_argument("--resolutions", default="", help="regression: comma list (default 100..600)")
    ap.add_argument("--times", default="", help="regression: comma list of modelling times (default 1000..6000)")
    ap.add_argument("--fixed-resolution", type=int, default=0, help="regression: resolution of the time sweep")
    ap.add_argument("--results-dir", default=os.path.join("workspace", "regression"),
                    help="regression: where the *_performance_results.csv files go")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only: the maps shard by grid rows with no "
                          "exchange, but multi-GPU runs are not part of it; use --gpus 1"}))
        sys.exit(2)
    res = {"lle": run_lle, "bif1d": run_bif1d, "bif2d": run_bif2d, "basins": run_basins, "stream": run_stream,
           "copy": run_copy, "ls": run_ls, "regression": run_regression, "tune": run_tune}[args.workload](args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Benchmark: FP64 RK4 steps/s of the 2048x2048 Rossler LLE map on one MI355X.

Workload (the headline metric): Benettin LLE of the Rossler system
x'=-y-z, y'=x+a*y, z'=b+z*(x-c) with b=0.2, a over linspace(0.05,0.35,2048)
and c over linspace(2,10,2048), FP64 RK4 h=0.01 from (0.1,0.1,0.1), 500 time
units of transient, then 4000 windows of 0.5 time units on the base and the
perturbed trajectory (perturbation 1e-8 along a seeded unit vector), i.e.
450,000 RK4 steps per grid point and 1.887e12 per map.  One "step" of this
benchmark is one full map; its output (33.5 MB of FP64 exponents) is written
to a device buffer, so nothing crosses PCIe inside the timed region.

Prints ONE JSON line on stdout (progress goes to stderr).  The roofline block
prices the map kernel's algorithmic FP64 work (63 FLOP per Rossler RK4 step,
FMA = 2) against MI355X's vector-FP64 peak (78.6 TFLOP/s), using the kernel's
own HIP-event time; cpu_baseline times the host oracle (the same per-point
code, OpenMP over grid points) on a strided sample of the grid rows.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FP64 RHS steps/sec on 2048² LLE map (% FP64 peak); p50/p99 latency, mixed job stream"
FP64_PEAK_TFLOPS = 78.6
FLOP_PER_STEP = {"rossler_rk4": 63, "lorenz_rk4": 71}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def lle_workload(args):
    return dict(system="rossler_rk4", nPts=args.grid, tMax=2000.0, NT=0.5, h=0.01, eps=1e-8,
                ic=[0.1, 0.1, 0.1], ranges=[0.05, 0.35, 2.0, 10.0], idx=[0, 2], transient=500.0,
                values=[0.2, 0.2, 5.7], seed=12345, semantics=1)


def cpu_threads():
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(w, target_s):
    """Oracle LLE over a strided row sample, OpenMP over points."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import _oracle as O
    n = w["nPts"]
    threads = cpu_threads()
    sid = O.SYS[w["system"]]
    ic = np.asarray(w["ic"], dtype=np.float64)
    rg = np.asarray(w["ranges"], dtype=np.float64)
    ix = np.asarray(w["idx"], dtype=np.int32)
    vals = np.asarray(w["values"], dtype=np.float64)
    steps_pt = round(w["transient"] / w["h"]) + 2 * round(w["tMax"] / w["h"])

    def run(rows, stride, cols_only=None):
        out = np.zeros(rows * n)
        t0 = time.perf_counter()
        used = O.lib.orc_lle2d_omp(sid, w["semantics"], w["tMax"], w["NT"], n, w["h"], w["eps"], O.dptr(ic),
                                   3, O.dptr(rg), O.iptr(ix), w["transient"], O.dptr(vals), 3, w["seed"],
                                   O.dptr(out), 0, rows * stride, stride, threads)
        return time.perf_counter() - t0, used

    # probe: one row, to size the sample to about target_s seconds
    dt1, used = run(1, n)
    rate = n * steps_pt / dt1
    rows = max(1, min(n, int(target_s * rate / (n * steps_pt))))
    stride = max(1, n // rows)
    rows = min(rows, n // stride)
    dt, used = run(rows, stride)
    pts = rows * n
    return {"value": pts * steps_pt / dt, "unit": "RK4 steps/s", "cores": used, "kind": "port",
            "sample": f"{rows} rows (every {stride}th) x {n} points = {pts} points of the same map, "
                      f"{steps_pt} RK4 steps each, {dt:.1f} s on {used} OpenMP threads"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed maps")
    ap.add_argument("--warmup", type=int, default=1, help="untimed maps")
    ap.add_argument("--grid", type=int, default=2048)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU-baseline sample time")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only: the LLE map shards by grid rows "
                          "with no exchange, but multi-GPU runs are not part of it; use --gpus 1"}))
        sys.exit(2)

    import numpy as np
    import chaoslibrary_amd as ch

    w = lle_workload(args)
    n = w["nPts"]
    stats = ch.Stats()
    dptr = C.c_void_p()
    ch._check(ch.lib.chaos_device_malloc(C.byref(dptr), n * n * 8, 0))

    def one_map():
        ch.lle2d(w["tMax"], w["NT"], n, w["h"], w["eps"], w["ic"], w["ranges"], w["idx"], w["transient"],
                 w["values"], system=w["system"], semantics=w["semantics"], seed=w["seed"],
                 block=args.block, stats=stats, out_device_ptr=dptr.value)
        return stats.kernel_ms, stats.rk4_steps

    for i in range(args.warmup):
        kms, _ = one_map()
        log(f"warmup {i}: kernel {kms:.1f} ms")
    ch._check(ch.lib.chaos_device_synchronize(0))
    kernel_ms = []
    steps_total = 0.0
    t0 = time.perf_counter()
    for i in range(args.steps):
        kms, st = one_map()
        kernel_ms.append(kms)
        steps_total += st
        log(f"step {i}: kernel {kms:.1f} ms")
    ch._check(ch.lib.chaos_device_synchronize(0))
    wall = time.perf_counter() - t0

    host = np.empty(n * n)
    ch._check(ch.lib.chaos_memcpy_d2h(host.ctypes.data, dptr, n * n * 8))
    ch._check(ch.lib.chaos_device_free(dptr))
    if not np.isfinite(host).all():
        log("warning: non-finite exponents in the map")

    value = steps_total / wall
    avg_kernel_s = sum(kernel_ms) / len(kernel_ms) / 1e3
    steps_per_map = steps_total / args.steps
    achieved = steps_per_map * FLOP_PER_STEP[w["system"]] / avg_kernel_s / 1e12
    roof = {"bound": "valu_fp64", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": None,
            "kernel": f"lle_map_kernel<{5 if w['system'] == 'rossler_rk4' else 6}>",
            "kernel_ms": round(avg_kernel_s * 1e3, 3)}
    res = {
        "metric": METRIC, "value": value, "unit": "RK4 steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: seeded 2048x2048 Rossler (a,c) parameter grid, no input data",
        "config": {"workload": f"lle2d_rossler_rk4_{n}x{n}", "grid": n, "h": w["h"],
                   "transient_time": w["transient"], "tMax": w["tMax"], "NT": w["NT"], "eps": w["eps"],
                   "seed": w["seed"], "rk4_steps_per_map": steps_per_map,
                   "flop_per_step": FLOP_PER_STEP[w["system"]]},
        "roofline": roof,
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        res["cpu_baseline"] = cpu_baseline(w, args.cpu_seconds)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()

"""The served mixed-job stream at the configured workload: Poisson arrivals
at 200 jobs/s, kinds uniform over 1-D Rossler bifurcation (256-8192 points),
Rossler LLE maps (64^2-1024^2) and Lorenz basins (128^2-1024^2), sizes
log-uniform, full horizons (1000 + 1000 t.u.; 500 + 2000 t.u. with 0.5-t.u.
windows; 300 + 300 t.u.), over a one-second arrival window.

* Continuous batching gives every job the result it has alone: all jobs of
  the window are served together by the engine, and sampled jobs of each
  kind are compared bit-for-bit with the one-shot entry points.
* Served through the engine the stream's p99 latency is below the plain
  one-launch-per-job form's on the same GPU, and its throughput above it.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WINDOW_S = 1.0


def _cfg(chaos, mode=0):
    return chaos.stream_config(lambda_=200.0, duration_s=WINDOW_S, horizon_scale=1.0, seed=2024, mode=mode)


def test_stream_jobs_match_the_configuration(chaos):
    jobs, arr = chaos.stream_jobs(_cfg(chaos))
    assert 150 <= len(jobs) <= 260 and np.all(np.diff(arr) >= 0) and arr[-1] <= WINDOW_S * 1e3
    for j in jobs:
        if j.kind == chaos.JOB_BIF1D:
            assert 256 <= j.nPts <= 8192 and (j.transientTime, j.tMax) == (1000.0, 1000.0)
        elif j.kind == chaos.JOB_LLE2D:
            assert 64 <= j.nPts <= 1024 and (j.transientTime, j.tMax, j.NT) == (500.0, 2000.0, 0.5)
        else:
            assert j.kind == chaos.JOB_BASINS2D and 128 <= j.nPts <= 1024
            assert (j.transientTime, j.tMax) == (300.0, 300.0)


def _one_shot(chaos, j):
    sysname = {5: "rossler_rk4", 6: "lorenz_rk4"}[j.system]
    ic = list(j.initialConditions)[:j.amountOfInitialConditions]
    vals = list(j.values)[:j.amountOfValues]
    rg, idx = list(j.ranges), list(j.indicesOfMutVars)
    if j.kind == chaos.JOB_LLE2D:
        return chaos.lle2d(j.tMax, j.NT, j.nPts, j.h, j.eps, ic, rg, idx, j.transientTime, vals,
                           system=sysname, semantics=j.semantics, seed=j.seed)
    if j.kind == chaos.JOB_BIF1D:
        return chaos.bif1d(j.tMax, j.nPts, j.h, ic, rg[:2], idx[:1], j.writableVar, j.maxValue, j.transientTime,
                           vals, j.preScaller, system=sysname, semantics=j.semantics)
    return chaos.basins2d(j.tMax, j.nPts, j.h, ic, rg, idx, j.writableVar, j.maxValue, j.transientTime, vals,
                          j.preScaller, j.eps, system=sysname, semantics=j.semantics, min_pts=j.min_pts)


def _same(a, b):
    if isinstance(a, tuple):
        return all(_same(x, y) for x, y in zip(a, b))
    return np.array_equal(np.asarray(a), np.asarray(b))


def test_configured_stream_alone_equals_mix(chaos):
    jobs, _ = chaos.stream_jobs(_cfg(chaos))
    print(f"configured stream window: {len(jobs)} jobs through the engine", flush=True)
    srv = chaos.Server()
    ids = [srv.submit(j) for j in jobs]
    srv.drain()
    for jid in ids:
        assert srv.info(jid).status == 2
    # per kind: the smallest job, the largest, and one in between
    picks = []
    for kind in (chaos.JOB_BIF1D, chaos.JOB_LLE2D, chaos.JOB_BASINS2D):
        ks = sorted((k for k, j in enumerate(jobs) if j.kind == kind), key=lambda k: jobs[k].nPts)
        if ks:
            picks += sorted({ks[0], ks[len(ks) // 2], ks[-1]})
    for k in picks:
        j = jobs[k]
        print(f"  alone vs mix: job {k} kind {j.kind} nPts {j.nPts}", flush=True)
        assert _same(srv.result(ids[k], j), _one_shot(chaos, j)), (k, j.kind, j.nPts)
    srv.close()


def test_configured_stream_batched_beats_per_job(chaos):
    reps = {}
    for mode, name in ((0, "engine"), (1, "per-job")):
        print(f"configured stream window, {name}", flush=True)
        rep, lat = chaos.stream_run(_cfg(chaos, mode))
        assert rep.jobs == len(lat) and np.all(lat > 0)
        reps[name] = rep
        print(f"  {name}: {rep.jobs} jobs, makespan {rep.makespan_s:.2f} s, {rep.steps_per_s:.3e} RK4 steps/s, "
              f"p50 {rep.p50_ms:.0f} ms, p99 {rep.p99_ms:.0f} ms", flush=True)
    e, p = reps["engine"], reps["per-job"]
    print(f"throughput x{e.steps_per_s / p.steps_per_s:.2f}, p99 x{e.p99_ms / p.p99_ms:.2f}", flush=True)
    assert e.rk4_steps == p.rk4_steps
    assert e.p99_ms < p.p99_ms
    assert e.steps_per_s > p.steps_per_s

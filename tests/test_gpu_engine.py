"""Batched engine / job server on the GPU.

Every job of a mixed batch must give exactly the result of the same job run
alone through the one-shot entry points (bit-exact), whatever the steps per
launch (tiles stop mid-transient, mid-window and mid-record), with and
without hipGraph replay.  The stream driver runs in both modes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROS = dict(ic=[0.1, 0.1, 0.1], values=[0.2, 0.2, 5.7])
LOR = dict(ic=[1.0, 1.0, 1.0], values=[10.0, 28.0, 8 / 3])


def mixed_jobs(chaos):
    J = chaos.make_job
    jobs = [
        J(chaos.JOB_LLE2D, "rossler_rk4", 1, 12, ROS["ic"], ROS["values"], [0.05, 0.35, 2.0, 10.0], [0, 2],
          tMax=20.0, NT=0.5, eps=1e-8, transientTime=5.3, seed=7),
        J(chaos.JOB_LLE2D, "rossler_rk4", 0, 7, ROS["ic"], ROS["values"], [0.05, 0.35, 2.0, 10.0], [0, 2],
          tMax=10.0, NT=0.5, eps=1e-8, transientTime=2.0),
        J(chaos.JOB_LLE2D, "lorenz_rk4", 1, 9, LOR["ic"], LOR["values"], [20.0, 200.0, 0.5, 4.0], [1, 2],
          tMax=12.0, NT=0.5, eps=1e-8, transientTime=3.0),
        J(chaos.JOB_BIF1D, "rossler_rk4", 1, 100, ROS["ic"], ROS["values"], [2.0, 6.0], [2],
          tMax=60.0, transientTime=30.0),
        J(chaos.JOB_BIF1D, "lorenz_rk4", 0, 70, LOR["ic"], LOR["values"], [0.5, 30.0], [1],
          tMax=40.0, transientTime=0.5, maxValue=15.0, preScaller=2),
        J(chaos.JOB_BIF2D, "lorenz_rk4", 1, 10, LOR["ic"], LOR["values"], [20.0, 200.0, 0.5, 4.0], [1, 2],
          tMax=20.0, transientTime=10.0, eps=1e-2),
        J(chaos.JOB_BIF2D, "rossler_rk4", 0, 9, ROS["ic"], ROS["values"], [0.05, 0.35, 2.0, 10.0], [0, 2],
          tMax=60.0, transientTime=30.0, eps=0.05),
        J(chaos.JOB_BASINS2D, "lorenz_rk4", 0, 16, [0.0, 0.0, 9.0], [10.0, 10.0, 8 / 3],
          [-10.0, 10.0, -10.0, 10.0], [0, 1], tMax=15.0, transientTime=40.0, eps=0.05, preScaller=3),
        J(chaos.JOB_BASINS2D, "lorenz_rk4", 1, 24, [0.0, 0.0, 25.0], [10.0, 24.5, 8 / 3],
          [-20.0, 20.0, -20.0, 20.0], [0, 1], tMax=20.0, transientTime=20.0, eps=0.5, min_pts=5),
    ]
    return jobs


def one_shot(chaos, j):
    sysname = {5: "rossler_rk4", 6: "lorenz_rk4"}[j.system]
    ic = list(j.initialConditions)[:j.amountOfInitialConditions]
    vals = list(j.values)[:j.amountOfValues]
    rg = list(j.ranges)
    idx = list(j.indicesOfMutVars)
    if j.kind == chaos.JOB_LLE2D:
        return chaos.lle2d(j.tMax, j.NT, j.nPts, j.h, j.eps, ic, rg, idx, j.transientTime, vals,
                           system=sysname, semantics=j.semantics, seed=j.seed)
    if j.kind == chaos.JOB_BIF1D:
        return chaos.bif1d(j.tMax, j.nPts, j.h, ic, rg[:2], idx[:1], j.writableVar, j.maxValue, j.transientTime,
                           vals, j.preScaller, system=sysname, semantics=j.semantics)
    if j.kind == chaos.JOB_BIF2D:
        return chaos.bif2d(j.tMax, j.nPts, j.h, ic, rg, idx, j.writableVar, j.maxValue, j.transientTime, vals,
                           j.preScaller, j.eps, system=sysname, semantics=j.semantics)
    return chaos.basins2d(j.tMax, j.nPts, j.h, ic, rg, idx, j.writableVar, j.maxValue, j.transientTime, vals,
                          j.preScaller, j.eps, system=sysname, semantics=j.semantics, min_pts=j.min_pts)


def same(a, b):
    if isinstance(a, tuple):
        return all(same(x, y) for x, y in zip(a, b))
    return np.array_equal(np.asarray(a), np.asarray(b))


def tiny_jobs(chaos):
    J = chaos.make_job
    return [
        # records of one to five samples (the detector warm-up meets the
        # first and last samples), h = 0.1
        *[J(chaos.JOB_BIF1D, "rossler_rk4", sem, 40, ROS["ic"], ROS["values"], [2.0, 6.0], [2], h=0.1,
            tMax=round(0.1 * b + 0.001, 4), transientTime=5.0) for b in (1, 2, 5) for sem in (0, 1)],
        # LLE horizons shorter than one window
        J(chaos.JOB_LLE2D, "rossler_rk4", 0, 3, ROS["ic"], ROS["values"], [0.05, 0.35, 2.0, 10.0], [0, 2],
          tMax=0.3, NT=0.5, eps=1e-8, transientTime=1.0),
        J(chaos.JOB_LLE2D, "rossler_rk4", 1, 3, ROS["ic"], ROS["values"], [0.05, 0.35, 2.0, 10.0], [0, 2],
          tMax=0.51, NT=0.5, eps=1e-8, transientTime=1.0),
        # 2x2 and 3x3 grids
        J(chaos.JOB_BIF2D, "lorenz_rk4", 1, 2, LOR["ic"], LOR["values"], [20.0, 200.0, 0.5, 4.0], [1, 2], h=0.05,
          tMax=5.05, transientTime=5.0, eps=1e-2),
        J(chaos.JOB_BASINS2D, "lorenz_rk4", 0, 3, [0.0, 0.0, 25.0], [10.0, 24.5, 8 / 3], [-20.0, 20.0, -20.0, 20.0],
          [0, 1], tMax=5.0, transientTime=5.0, eps=0.5),
        J(chaos.JOB_BASINS2D, "lorenz_rk4", 1, 3, [0.0, 0.0, 25.0], [10.0, 24.5, 8 / 3], [-20.0, 20.0, -20.0, 20.0],
          [0, 1], tMax=0.01, transientTime=5.0, eps=0.5, min_pts=2),
    ]


@pytest.mark.parametrize("steps", [3, 700])
def test_tiny_jobs_equal_one_shot(chaos, steps):
    """Jobs of a few points and a few samples, tiles of 3 ticks (every tile
    boundary falls inside the warm-up, the first or the last sample)."""
    jobs = tiny_jobs(chaos)
    srv = chaos.Server(params=chaos.engine_params(steps_per_launch=steps))
    ids = [srv.submit(j) for j in jobs]
    print(f"engine tiny jobs steps={steps}", flush=True)
    srv.drain()
    for jid, j in zip(ids, jobs):
        assert srv.info(jid).status == 2
        got = srv.result(jid, j)
        want = one_shot(chaos, j)
        assert same(got, want), (j.kind, j.semantics, j.nPts, j.tMax)
    srv.close()


@pytest.mark.parametrize("steps,graph", [(700, 1), (20000, 1), (700, 0)])
def test_mixed_batch_equals_one_shot(chaos, steps, graph):
    jobs = mixed_jobs(chaos)
    srv = chaos.Server(params=chaos.engine_params(steps_per_launch=steps, use_graph=graph))
    ids = [srv.submit(j) for j in jobs]
    print(f"engine mixed batch steps={steps} graph={graph}", flush=True)
    srv.drain()
    st = srv.stats()
    assert st.active_jobs == 0 and st.pending_jobs == 0 and st.launches >= 1
    if graph:
        assert st.graph_replays == st.launches
    for jid, j in zip(ids, jobs):
        info = srv.info(jid)
        assert info.status == 2 and info.reran == 0
        got = srv.result(jid, j)
        want = one_shot(chaos, j)
        assert same(got, want), (j.kind, j.system, j.semantics)
    srv.close()


def test_alone_vs_mix(chaos):
    jobs = mixed_jobs(chaos)
    params = chaos.engine_params(steps_per_launch=900)
    mix = chaos.Server(params=params)
    ids = [mix.submit(j) for j in jobs]
    mix.drain()
    for jid, j in zip(ids, jobs):
        alone = chaos.Server(params=params)
        a = alone.submit(j)
        alone.drain()
        assert same(alone.result(a, j), mix.result(jid, j))
        alone.close()
    mix.close()


def test_admission_by_hbm_budget(chaos):
    # a budget that fits one small job at a time still completes every job
    jobs = mixed_jobs(chaos)[:4]
    srv = chaos.Server(params=chaos.engine_params(steps_per_launch=5000, max_jobs=2))
    ids = [srv.submit(j) for j in jobs]
    n_active_max = 0
    while True:
        srv.step()
        st = srv.stats()
        n_active_max = max(n_active_max, st.active_jobs)
        if st.active_jobs == 0 and st.pending_jobs == 0 and st.finishing_jobs == 0:
            break
    assert n_active_max <= 2
    for jid, j in zip(ids, jobs):
        assert srv.info(jid).status == 2
        assert same(srv.result(jid, j), one_shot(chaos, j))
    srv.close()


def test_engine_rejects_non_rk4_and_bad_jobs(chaos):
    srv = chaos.Server()
    j = chaos.make_job(chaos.JOB_LLE2D, "chameleon", 0, 8, [3.0, 3.0, 0.0], [0.5, 3, 0, 2, 1, 18, 1],
                       [0, 5, 0, 20], [3, 5], tMax=10.0, NT=0.5, eps=1e-6, transientTime=1.0)
    with pytest.raises(chaos.ChaosError):
        srv.submit(j)
    bad = chaos.make_job(chaos.JOB_LLE2D, "rossler_rk4", 1, 0, ROS["ic"], ROS["values"], [0, 1, 0, 1], [0, 2],
                         tMax=10.0, NT=0.5, eps=1e-8)
    with pytest.raises(chaos.ChaosError):
        srv.submit(bad)
    srv.close()


@pytest.mark.parametrize("mode", [0, 1])
def test_stream_run_small(chaos, mode):
    cfg = chaos.stream_config(lambda_=40.0, duration_s=0.5, horizon_scale=0.004, mode=mode, seed=3,
                              bif_max=2048, lle_max=128 * 128, basins_max=192 * 192)
    print(f"stream mode={mode}", flush=True)
    rep, lat = chaos.stream_run(cfg, chaos.engine_params(steps_per_launch=2000))
    assert rep.jobs == len(lat) and rep.jobs > 3
    assert sum(rep.jobs_kind) == rep.jobs
    assert np.all(lat > 0) and rep.p99_ms >= rep.p50_ms > 0
    assert rep.steps_per_s > 0

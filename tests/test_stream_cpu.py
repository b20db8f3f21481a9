"""CPU checks of the served stream's generator (chaos_stream_jobs, the jobs
chaos_stream_run serves; host code, no device needed): the seeded Poisson
arrivals, the uniform choice of kind, the log-uniform sizes and each kind's
physics, against the configured stream (200 jobs/s for 60 s; 1-D Rossler
bifurcations of 256-8192 points, Rossler LLE maps of 64^2-1024^2 points,
Lorenz basins of 128^2-1024^2 points)."""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def configured(chaos):
    cfg = chaos.stream_config(lambda_=200.0, duration_s=60.0, horizon_scale=1.0, seed=2024, mode=0)
    return chaos.stream_jobs(cfg)


def test_arrivals_are_a_poisson_process(configured):
    jobs, arr = configured
    n = len(jobs)
    # 12000 expected, standard deviation 110; this seed gives 12098
    assert n == 12098 and abs(n - 12000) < 4 * math.sqrt(12000)
    assert np.all(np.diff(arr) > 0) and 0 < arr[0] and arr[-1] <= 60000.0
    gaps = np.diff(np.concatenate([[0.0], arr]))
    # exponential inter-arrival times of mean 5 ms: mean, and P(gap > mean) = 1/e
    assert abs(gaps.mean() - 5.0) < 0.15
    assert abs((gaps > 5.0).mean() - math.exp(-1)) < 0.02
    # arrivals spread evenly over the window
    assert abs(np.mean(arr) / 60000.0 - 0.5) < 0.02


def test_kinds_uniform_and_sizes_log_uniform(chaos, configured):
    jobs, _ = configured
    kinds = np.array([j.kind for j in jobs])
    ranges = {chaos.JOB_BIF1D: (256, 8192, 1), chaos.JOB_LLE2D: (64 * 64, 1024 * 1024, 2),
              chaos.JOB_BASINS2D: (128 * 128, 1024 * 1024, 2)}
    for kind, (lo, hi, dims) in ranges.items():
        sel = [j for j in jobs if j.kind == kind]
        assert abs(len(sel) / len(jobs) - 1 / 3) < 0.02
        pts = np.array([j.nPts ** dims for j in sel], dtype=np.float64)
        side_lo, side_hi = round(lo ** (1 / dims)), round(hi ** (1 / dims))
        assert all(side_lo <= j.nPts <= side_hi for j in sel)
        # log-uniform: the size's position on the log scale is uniform on [0, 1]
        x = (np.log(pts) - math.log(lo)) / (math.log(hi) - math.log(lo))
        assert abs(x.mean() - 0.5) < 0.025 and abs(x.var() - 1 / 12) < 0.01
        assert x.min() < 0.01 and x.max() > 0.99
    assert set(kinds.tolist()) == set(ranges)


def test_each_kind_carries_the_configured_physics(chaos, configured):
    jobs, _ = configured
    for j in jobs[:600]:
        assert j.h == 0.01 and j.semantics == 1
        if j.kind == chaos.JOB_BIF1D:
            assert j.system == 5 and (j.transientTime, j.tMax) == (1000.0, 1000.0)
            assert list(j.ranges)[:2] == [2.0, 6.0] and list(j.initialConditions)[:3] == [0.1, 0.1, 0.1]
        elif j.kind == chaos.JOB_LLE2D:
            assert j.system == 5 and (j.transientTime, j.tMax, j.NT, j.eps) == (500.0, 2000.0, 0.5, 1e-8)
            assert list(j.ranges) == [0.05, 0.35, 2.0, 10.0]
        else:
            assert j.system == 6 and (j.transientTime, j.tMax) == (300.0, 300.0)
            assert list(j.ranges) == [-20.0, 20.0, -20.0, 20.0] and j.initialConditions[2] == 25.0
            assert j.values[1] == 24.5


def test_seeded_and_scaled(chaos):
    def gen(**kw):
        base = dict(lambda_=200.0, duration_s=2.0, horizon_scale=1.0, seed=2024, mode=0)
        base.update(kw)
        return chaos.stream_jobs(chaos.stream_config(**base))
    (a, ta), (b, tb), (c, tc) = gen(), gen(), gen(seed=7)
    assert np.array_equal(ta, tb) and [j.nPts for j in a] == [j.nPts for j in b]
    assert not (len(ta) == len(tc) and np.array_equal(ta, tc))
    s, ts = gen(horizon_scale=0.3)
    assert np.array_equal(ts, ta)  # the same arrivals and jobs, shorter horizons
    for x, y in zip(a, s):
        assert (x.kind, x.nPts) == (y.kind, y.nPts)
        assert y.tMax == pytest.approx(0.3 * x.tMax) and y.transientTime == pytest.approx(0.3 * x.transientTime)
    m, tm = gen(max_jobs=10)
    assert len(m) == 10 and np.array_equal(tm, ta[:10])

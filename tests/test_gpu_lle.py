"""GPU parity of the LLE map (chaos_lle2d / chaos_LLE2D) against the oracle.

The trajectories are bit-exact; the log of each window's growth is the
device FP64 log vs the host libm log (not fed back), so the exponents agree
to ~1e-15 -- the tests use 1e-12 absolute, far inside the 1e-3 budget."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROS_RANGES = [0.05, 0.35, 2.0, 10.0]


@pytest.mark.parametrize("sem", [0, 1])
def test_lle_small_grid_matches_oracle(chaos, oracle, sem):
    args = dict(tMax=20.0, NT=0.5, nPts=16, h=0.01, eps=1e-8)
    print(f"lle2d rossler 16x16 sem={sem}", flush=True)
    got = chaos.lle2d(args["tMax"], args["NT"], args["nPts"], args["h"], args["eps"], [0.1, 0.1, 0.1],
                      ROS_RANGES, [0, 2], 5.0, [0.2, 0.2, 5.7], system="rossler_rk4", semantics=sem)
    want = oracle.lle2d(oracle.SYS["rossler_rk4"], sem, 20.0, 0.5, 16, 0.01, 1e-8, [0.1, 0.1, 0.1],
                        ROS_RANGES, [0, 2], 5.0, [0.2, 0.2, 5.7], 12345)
    assert np.all(np.abs(got - want) <= 1e-12), np.abs(got - want).max()


def test_lle_full_horizon_lorenz(chaos, oracle):
    # configs-like horizon on a small grid (Lorenz over rho x beta)
    print("lle2d lorenz 6x6 full horizon", flush=True)
    r = [20.0, 200.0, 0.5, 4.0]
    got = chaos.lle2d(2000.0, 0.5, 6, 0.01, 1e-8, [1, 1, 1], r, [1, 2], 500.0, [10.0, 28.0, 8 / 3],
                      system="lorenz_rk4", semantics=1, seed=7)
    want = oracle.lle2d(oracle.SYS["lorenz_rk4"], 1, 2000.0, 0.5, 6, 0.01, 1e-8, [1, 1, 1], r, [1, 2],
                        500.0, [10.0, 28.0, 8 / 3], 7)
    assert np.all(np.abs(got - want) <= 1e-12), np.abs(got - want).max()


def test_lle_2048_sampled_points_match_oracle(chaos, oracle):
    """Full configs size (2048^2, full horizon) on the GPU; 24 sampled points
    recomputed by the oracle must agree within 1e-12."""
    n = 2048
    print("lle2d rossler 2048x2048 full horizon", flush=True)
    got = chaos.lle2d(2000.0, 0.5, n, 0.01, 1e-8, [0.1, 0.1, 0.1], ROS_RANGES, [0, 2], 500.0,
                      [0.2, 0.2, 5.7], system="rossler_rk4", semantics=1, seed=12345)
    assert got.shape == (n, n) and np.isfinite(got).all()
    rng = np.random.default_rng(0)
    for _ in range(24):
        i, j = (int(v) for v in rng.integers(0, n, 2))
        a = oracle.lib.orc_grid_value(i, n, 0.05, 0.35)
        c = oracle.lib.orc_grid_value(j, n, 2.0, 10.0)
        want = oracle.lle_point(oracle.SYS["rossler_rk4"], 1, 2000.0, 0.5, 0.01, 1e-8, [0.1, 0.1, 0.1],
                                [a, 0.2, c], 500.0, 12345, j * n + i)
        assert abs(got[j, i] - want) <= 1e-12, (i, j, got[j, i], want)
    # property: periodic-window region (small c) has |LLE| small, chaotic band positive
    assert np.median(got[:64, :]) < 0.02
    assert got.max() > 0.05


def test_LLE2D_writes_reference_csv(chaos, oracle, tmp_path):
    path = tmp_path / "lle.csv"
    ranges = [0.05, 0.35, 2.0, 10.0]
    opt = chaos.make_options("rossler_rk4", 1)
    print("LLE2D csv 8x8", flush=True)
    chaos.LLE2D(10.0, 0.5, 8, 0.01, 1e-8, [0.1, 0.1, 0.1], 3, ranges, [0, 2], 0, 0.0, 2.0,
                [0.2, 0.2, 5.7], 3, str(path), options=opt)
    vals = chaos.lle2d(10.0, 0.5, 8, 0.01, 1e-8, [0.1, 0.1, 0.1], ranges, [0, 2], 2.0, [0.2, 0.2, 5.7],
                       system="rossler_rk4", semantics=1)
    # expected bytes per LLE/src/LLECUDA.cu:252-253,383-398 (%g == ostream default)
    exp = "%g %g\n%g %g\n" % tuple(ranges)
    flat = vals.reshape(-1)
    body = []
    cnt = 0
    for v in flat:
        s = ""
        if cnt != 0:
            s += ", "
        if cnt == 8:
            s += "\n"
            cnt = 0
        s += "0" if np.isnan(v) else "%g" % v
        body.append(s)
        cnt += 1
    exp += "".join(body)
    assert path.read_text() == exp

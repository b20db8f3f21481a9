"""Edge cases against the oracle: records of one to five samples (the
detector's warm-up, the first and last samples and the two-samples-per-trip
record loop meet), LLE horizons shorter than one window, 1x1 to 3x3 grids,
and the inputs the entry points reject before any kernel runs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROS = [0.2, 0.2, 5.7]


@pytest.mark.parametrize("sem", [0, 1])
@pytest.mark.parametrize("samples", [1, 2, 3, 4, 5])
def test_bif1d_short_records_match_oracle(chaos, oracle, sem, samples):
    # h = 0.1: 256 Rossler points whose first peak can fall inside the window
    # (it is stored, then dropped by the reference's rule)
    tmax = round(0.1 * samples + 0.001, 4)
    args = (tmax, 256, 0.1, [0.1, 0.1, 0.1], [2.0, 6.0], [2], 0, 0.0, 50.0, ROS, 1)
    print(f"bif1d short record samples={samples} sem={sem}", flush=True)
    c, f, pk, iv = chaos.bif1d(*args, system="rossler_rk4", semantics=sem)
    oc, of, opk, oiv = oracle.bif1d(oracle.SYS["rossler_rk4"], sem, *args)
    assert np.array_equal(c, oc)
    assert np.array_equal(f[c == -1], of[c == -1])
    assert np.array_equal(pk, opk) and np.array_equal(iv, oiv)


@pytest.mark.parametrize("tmax", [0.5, 5.0, 5.05])  # 10, 100 and 101 recorded samples at h = 0.05
@pytest.mark.parametrize("sem", [0, 1])
def test_bif2d_two_by_two(chaos, oracle, sem, tmax):
    args = (tmax, 2, 0.05, [1.0, 1.0, 1.0], [20.0, 200.0, 0.5, 4.0], [1, 2], 0, 0.0, 5.0, [10.0, 28.0, 8 / 3], 1,
            1e-2)
    print(f"bif2d 2x2 sem={sem} tMax={tmax}", flush=True)
    got = chaos.bif2d(*args, system="lorenz_rk4", semantics=sem)
    want = oracle.bif2d(oracle.SYS["lorenz_rk4"], sem, *args)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("sem", [0, 1])
def test_bif2d_one_point_grid_matches_oracle(chaos, oracle, sem):
    # n = 1: the grid step (end - start) / (n - 1) is infinite and 0 * inf is
    # NaN, as in the reference's getValueByIdx; the point diverges at once
    args = (1.0, 1, 0.01, [1.0, 1.0, 1.0], [20.0, 200.0, 0.5, 4.0], [1, 2], 0, 0.0, 1.0, [10.0, 28.0, 8 / 3], 1,
            1e-2)
    print(f"bif2d 1x1 sem={sem}", flush=True)
    got = chaos.bif2d(*args, system="lorenz_rk4", semantics=sem)
    want = oracle.bif2d(oracle.SYS["lorenz_rk4"], sem, *args)
    assert np.array_equal(got, want)


ROS_RANGES = [0.05, 0.35, 2.0, 10.0]


@pytest.mark.parametrize("sem", [0, 1])
@pytest.mark.parametrize("tmax,nt", [(0.3, 0.5), (0.5, 0.5), (0.51, 0.5), (0.05, 0.01)])
def test_lle_short_horizons_match_oracle(chaos, oracle, sem, tmax, nt):
    # tMax < NT: no full window (CONFIG: 0 windows and a zero sum; REFERENCE:
    # the two extra windows of LLECUDA.cu:227)
    print(f"lle2d 3x3 tMax={tmax} NT={nt} sem={sem}", flush=True)
    got = chaos.lle2d(tmax, nt, 3, 0.01, 1e-8, [0.1, 0.1, 0.1], ROS_RANGES, [0, 2], 1.0, ROS, system="rossler_rk4",
                      semantics=sem)
    want = oracle.lle2d(oracle.SYS["rossler_rk4"], sem, tmax, nt, 3, 0.01, 1e-8, [0.1, 0.1, 0.1], ROS_RANGES, [0, 2],
                        1.0, ROS, 12345)
    assert np.all(np.abs(got - want) <= 1e-12), np.abs(got - want).max()


@pytest.mark.parametrize("sem", [0, 1])
def test_lle_one_point_grid_matches_oracle(chaos, oracle, sem):
    print(f"lle2d 1x1 sem={sem}", flush=True)
    got = chaos.lle2d(5.0, 0.5, 1, 0.01, 1e-8, [0.1, 0.1, 0.1], ROS_RANGES, [0, 2], 1.0, ROS, system="rossler_rk4",
                      semantics=sem)
    want = oracle.lle2d(oracle.SYS["rossler_rk4"], sem, 5.0, 0.5, 1, 0.01, 1e-8, [0.1, 0.1, 0.1], ROS_RANGES, [0, 2],
                        1.0, ROS, 12345)
    assert np.all(np.abs(got - want) <= 1e-12), np.abs(got - want).max()


@pytest.mark.parametrize("tmax", [0.01, 5.0])
@pytest.mark.parametrize("n", [2, 3])
@pytest.mark.parametrize("sem", [0, 1])
def test_basins_tiny_grids_match_oracle(chaos, oracle, sem, n, tmax):
    # the 3x3 grid's centre starts on the invariant z axis (a fixed point)
    lor = [10.0, 24.5, 8 / 3]
    args = (tmax, n, 0.01, [0.0, 0.0, 25.0], [-20.0, 20.0, -20.0, 20.0], [0, 1], 0, 0.0, 5.0, lor, 1)
    print(f"basins {n}x{n} tMax={tmax} sem={sem}", flush=True)
    lab, feat, flags = chaos.basins2d(*args, eps=0.5, system="lorenz_rk4", semantics=sem, min_pts=2)
    f0, f1, f2, ofl = oracle.basins_features(oracle.SYS["lorenz_rk4"], sem, *args)
    assert np.array_equal(flags.reshape(-1), ofl)
    for k, f in enumerate((f0, f1) if sem == 0 else (f0, f1, f2)):
        assert np.array_equal(feat[k].reshape(-1), f), k
    if sem == 0:
        want = oracle.basins_cluster_ref(f0, f1, ofl, 0.5)
    else:
        z0, z1, z2, _ = oracle.zscore(f0, f1, f2, ofl)
        want = oracle.dbscan_bruteforce(np.column_stack([z0, z1, z2]), 0.5, 2, include=(ofl != 0))
    assert np.array_equal(lab.reshape(-1), want)


def test_rejected_inputs(chaos):
    lor = [10.0, 28.0, 8 / 3]
    bad = [
        # no recorded sample
        lambda: chaos.bif1d(0.0, 8, 0.01, [1.0, 1.0, 1.0], [20.0, 30.0], [1], 0, 0.0, 1.0, lor, 1,
                            system="lorenz_rk4"),
        # an empty grid
        lambda: chaos.bif2d(1.0, 0, 0.01, [1.0, 1.0, 1.0], [20.0, 200.0, 0.5, 4.0], [1, 2], 0, 0.0, 1.0, lor,
                            system="lorenz_rk4"),
        # no preScaller steps per sample
        lambda: chaos.bif1d(1.0, 8, 0.01, [1.0, 1.0, 1.0], [20.0, 30.0], [1], 0, 0.0, 1.0, lor, 0,
                            system="lorenz_rk4"),
        # non-positive step
        lambda: chaos.bif1d(1.0, 8, 0.0, [1.0, 1.0, 1.0], [20.0, 30.0], [1], 0, 0.0, 1.0, lor, 1,
                            system="lorenz_rk4"),
        # a basins grid needs two initial conditions per side
        lambda: chaos.basins2d(5.0, 1, 0.01, [0.0, 0.0, 25.0], [-20.0, 20.0, -20.0, 20.0], [0, 1], 0, 0.0, 5.0,
                               [10.0, 24.5, 8 / 3], 1, system="lorenz_rk4"),
        # LLE without a positive horizon
        lambda: chaos.lle2d(0.0, 0.5, 4, 0.01, 1e-8, [0.1, 0.1, 0.1], [0.05, 0.35, 2.0, 10.0], [0, 2], 1.0, ROS,
                            system="rossler_rk4"),
        # swept index outside the parameter vector
        lambda: chaos.bif1d(1.0, 8, 0.01, [1.0, 1.0, 1.0], [20.0, 30.0], [3], 0, 0.0, 1.0, lor, 1,
                            system="lorenz_rk4"),
    ]
    for i, f in enumerate(bad):
        with pytest.raises(chaos.ChaosError):
            f()
        print(f"rejected case {i}", flush=True)

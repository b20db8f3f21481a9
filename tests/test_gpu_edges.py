"""Edge cases of the sweep paths against the oracle: records of one to five
samples (the detector's warm-up, the first and last samples and the
two-samples-per-trip record loop meet), and the inputs the entry points
reject before any kernel runs."""
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
        # swept index outside the parameter vector
        lambda: chaos.bif1d(1.0, 8, 0.01, [1.0, 1.0, 1.0], [20.0, 30.0], [3], 0, 0.0, 1.0, lor, 1,
                            system="lorenz_rk4"),
    ]
    for i, f in enumerate(bad):
        with pytest.raises(chaos.ChaosError):
            f()
        print(f"rejected case {i}", flush=True)

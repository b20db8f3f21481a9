"""GPU parity of the legacy entry points (namespace old_library) against the
oracle: LLE in both legacy forms (parameter and initial-condition sweeps,
1-D and 2-D), the early stop on divergence, the Lyapunov spectrum, the
time-domain recorder, and the legacy file formats."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROS = dict(ic=[0.1, 0.1, 0.1], values=[0.2, 0.2, 5.7])
LOR = dict(ic=[1.0, 1.0, 1.0], values=[10.0, 28.0, 8 / 3])


def test_legacy_lle1d_params_matches_oracle(chaos, oracle):
    args = (30.0, 0.5, 40, 0.01, 1e-8, ROS["ic"], [2.0, 6.0], [2], 10.0, ROS["values"])
    print("legacy LLE 1-D rossler", flush=True)
    got = chaos.legacy_lle(*args, dims=1, system="rossler_rk4", seed=5)
    want = oracle.legacy_lle(oracle.SYS["rossler_rk4"], 0, 0, 1, 30.0, 0.5, 40, 0.01, 1e-8, ROS["ic"], [2.0, 6.0],
                             [2], 0.0, 10.0, ROS["values"], 5)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("dims", [1, 2])
def test_legacy_lle_ic_sweep_matches_oracle(chaos, oracle, dims):
    n = 24 if dims == 1 else 10
    rg = [-15.0, 15.0, -20.0, 20.0]
    print(f"legacy LLE IC sweep dims={dims}", flush=True)
    got = chaos.legacy_lle(20.0, 0.5, n, 0.01, 1e-8, LOR["ic"], rg, [0, 1], 5.0, LOR["values"], dims=dims,
                           sweep_ic=True, system="lorenz_rk4", seed=9)
    want = oracle.legacy_lle(oracle.SYS["lorenz_rk4"], 0, 1, dims, 20.0, 0.5, n, 0.01, 1e-8, LOR["ic"], rg, [0, 1],
                             0.0, 5.0, LOR["values"], 9)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-14)


def test_legacy_lle_early_stop_matches_oracle(chaos, oracle):
    # maxValue = 15 cuts the windows of every trajectory whose |x| exceeds it
    print("legacy LLE early stop", flush=True)
    got = chaos.legacy_lle(10.0, 0.5, 16, 0.01, 1e-8, LOR["ic"], [20.0, 40.0], [1], 2.0, LOR["values"], dims=1,
                           maxValue=15.0, system="lorenz_rk4", seed=2)
    want = oracle.legacy_lle(oracle.SYS["lorenz_rk4"], 0, 0, 1, 10.0, 0.5, 16, 0.01, 1e-8, LOR["ic"], [20.0, 40.0],
                             [1], 15.0, 2.0, LOR["values"], 2)
    free = chaos.legacy_lle(10.0, 0.5, 16, 0.01, 1e-8, LOR["ic"], [20.0, 40.0], [1], 2.0, LOR["values"], dims=1,
                            system="lorenz_rk4", seed=2)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-14, equal_nan=True)
    assert not np.allclose(got, free)


@pytest.mark.parametrize("dims", [1, 2])
def test_legacy_ls_matches_oracle(chaos, oracle, dims):
    n = 20 if dims == 1 else 8
    rg = [20.0, 40.0, 8.0, 12.0]
    print(f"legacy LS dims={dims}", flush=True)
    got = chaos.legacy_ls(20.0, 0.5, n, 0.01, 1e-8, LOR["ic"], rg, [1, 0], 5.0, LOR["values"], dims=dims,
                          system="lorenz_rk4", seed=4)
    want = oracle.legacy_ls(oracle.SYS["lorenz_rk4"], 0, dims, 20.0, 0.5, n, 0.01, 1e-8, LOR["ic"], rg, [1, 0], 0.0,
                            5.0, LOR["values"], 4)
    assert got.shape == want.shape
    assert np.allclose(got, want, rtol=1e-10, atol=1e-12)
    # dissipative: every point's exponents sum to about -(sigma + 1 + beta)
    assert np.all(got.sum(axis=1) < 0)


@pytest.mark.parametrize("pre,maxv", [(1, 0.0), (2, 0.0), (1, 12.0)])
def test_time_domain_matches_oracle(chaos, oracle, pre, maxv):
    print(f"time domain pre={pre} maxValue={maxv}", flush=True)
    got, gfl = chaos.time_domain(3.0, 12, 0.01, LOR["ic"], [10.0, 40.0], [1], 0, maxv, 1.0, LOR["values"], pre,
                                 system="lorenz_rk4")
    want, wfl = oracle.time_domain(oracle.SYS["lorenz_rk4"], 0, 3.0, 12, 0.01, LOR["ic"], [10.0, 40.0], [1], 0, maxv,
                                   1.0, LOR["values"], pre)
    assert np.array_equal(gfl, wfl)
    assert np.array_equal(got, want)
    if maxv:
        assert (wfl == 0).any()


def test_legacy_files(chaos, tmp_path):
    opt = chaos.make_options("rossler_rk4", 0, seed=5)
    ic, vals = ROS["ic"], ROS["values"]
    n = 6
    # LLE1D: "<c>, <lle>\n" at precision 12
    f = tmp_path / "lle1d.csv"
    chaos.LLE1D(10.0, 0.5, n, 0.01, 1e-8, ic, 3, [2.0, 6.0], [2], 0, 0.0, 5.0, vals, 3, str(f), options=opt)
    v = chaos.legacy_lle(10.0, 0.5, n, 0.01, 1e-8, ic, [2.0, 6.0], [2], 5.0, vals, dims=1, system="rossler_rk4",
                         seed=5)
    grid = [2.0 + k * ((6.0 - 2.0) / (n - 1)) for k in range(n)]
    assert f.read_text() == "".join("%.12g, %.12g\n" % (g, x) for g, x in zip(grid, v))
    # LS2D: one file per exponent, two range lines, the 2-D layout
    base = tmp_path / "ls"
    chaos.LS2D(10.0, 0.5, 3, 0.01, 1e-8, ic, 3, [0.1, 0.3, 4.0, 6.0], [0, 2], 0, 0.0, 5.0, vals, 3, str(base),
               options=opt)
    ls = chaos.legacy_ls(10.0, 0.5, 3, 0.01, 1e-8, ic, [0.1, 0.3, 4.0, 6.0], [0, 2], 5.0, vals, dims=2,
                         system="rossler_rk4", seed=5)
    for k in range(3):
        text = "0.1 0.3\n4 6\n"
        for i, x in enumerate(ls[:, k]):
            text += (", " if i else "") + ("\n" if i and i % 3 == 0 else "") + "%g" % x
        assert (tmp_path / f"ls{k + 1}.csv").read_text() == text
    # TimeDomainCalculation: precision 16, ", " after every sample, " \n" per row
    f = tmp_path / "td.csv"
    chaos.TimeDomainCalculation(0.2, 3, 0.01, 3, ic, [4.0, 6.0], [2], 0, 0.0, 0.5, vals, 3, 1, str(f), options=opt)
    td, _ = chaos.time_domain(0.2, 3, 0.01, ic, [4.0, 6.0], [2], 0, 0.0, 0.5, vals, 1, system="rossler_rk4")
    assert f.read_text() == "".join("".join("%.16g, " % x for x in row) + " \n" for row in td)


@pytest.mark.parametrize("pre,maxv", [(1, 0.0), (2, 0.0), (1, 18.0)])
def test_legacy_bif1d_ic_matches_oracle(chaos, oracle, pre, maxv):
    print(f"legacy bif1DIC pre={pre} maxValue={maxv}", flush=True)
    args = (30.0, 40, 0.01, LOR["ic"], [-15.0, 15.0], [0], 0, maxv, 5.0, LOR["values"], pre)
    counts, peaks = chaos.legacy_bif1d_ic(*args, system="lorenz_rk4")
    oc, op = oracle.legacy_bif1d_ic(oracle.SYS["lorenz_rk4"], 0, *args)
    assert np.array_equal(counts, oc)
    assert np.array_equal(peaks, op)
    assert peaks.size > 0


def test_legacy_bif1d_ic_fixed_points(chaos, oracle):
    # rho = 0.5: starts at the origin are recorded and end as fixed points
    vals = [10.0, 0.5, 8 / 3]
    counts, peaks = chaos.legacy_bif1d_ic(5.0, 4, 0.01, [0.0, 0.0, 0.0], [0.0, 0.0], [0], 0, 0.0, 1.0, vals,
                                          system="lorenz_rk4")
    assert list(counts) == [-1] * 4 and peaks.size == 0


@pytest.mark.parametrize("sem", [0, 1])
def test_legacy_bif2d_ic_matches_oracle(chaos, oracle, sem):
    print(f"legacy bif2DIC sem={sem}", flush=True)
    args = (20.0, 12, 0.01, [0.0, 0.0, 25.0], [-20.0, 20.0, -20.0, 20.0], [0, 1], 0, 0.0, 10.0,
            [10.0, 24.5, 8 / 3], 1)
    got = chaos.legacy_bif2d_ic(*args, eps=0.05, system="lorenz_rk4", semantics=sem)
    want = oracle.legacy_bif2d_ic(oracle.SYS["lorenz_rk4"], sem, *args, 0.05)
    assert np.array_equal(got, want)


def test_legacy_bifurcation_ic_files(chaos, tmp_path):
    opt = chaos.make_options("lorenz_rk4", 0)
    ic, vals = LOR["ic"], LOR["values"]
    f = tmp_path / "b1dic.csv"
    chaos.bifurcation1DIC(10.0, 5, 0.01, 3, ic, [-5.0, 5.0], [0], 0, 0.0, 2.0, vals, 3, 1, str(f), options=opt)
    counts, peaks = chaos.legacy_bif1d_ic(10.0, 5, 0.01, ic, [-5.0, 5.0], [0], 0, 0.0, 2.0, vals,
                                          system="lorenz_rk4")
    lines, k = [], 0
    for p in range(5):
        x = -5.0 + p * ((5.0 - -5.0) / 4)
        for _ in range(max(0, counts[p])):
            lines.append("%.12g, %.12g\n" % (x, peaks[k]))
            k += 1
    assert f.read_text() == "".join(lines) and k > 0
    f = tmp_path / "b2dic.csv"
    rg = [-10.0, 10.0, -10.0, 10.0]
    chaos.bifurcation2DIC(10.0, 4, 0.01, 3, ic, rg, [0, 1], 0, 0.0, 2.0, vals, 3, 1, 0.05, str(f), options=opt)
    m = chaos.legacy_bif2d_ic(10.0, 4, 0.01, ic, rg, [0, 1], 0, 0.0, 2.0, vals, 1, eps=0.05, system="lorenz_rk4")
    text = "-10 10\n-10 10\n" + "".join((", " if i else "") + ("\n" if i and i % 4 == 0 else "") + "%d" % v
                                        for i, v in enumerate(m))
    assert f.read_text() == text

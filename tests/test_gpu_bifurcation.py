"""GPU parity of Bifurcation 1-D / 2-D (chaos_bif1d, chaos_bif2d and the file
writers) against the oracle: bit-exact peaks, intervals, counts and maps."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROS = dict(system="rossler_rk4", ic=[0.1, 0.1, 0.1], values=[0.2, 0.2, 5.7])


def _oracle_bif1d(oracle, sysname, sem, tMax, nPts, h, ic, ranges, idx, wv, maxv, tT, values, pre):
    return oracle.bif1d(oracle.SYS[sysname], sem, tMax, nPts, h, ic, ranges, idx, wv, maxv, tT, values, pre)


# tMax 300 / 300.011: 30000 / 30001 recorded samples, i.e. an odd and an
# even number of detector-body samples (the record loop takes two per trip)
@pytest.mark.parametrize("tmax", [300.0, 300.011])
@pytest.mark.parametrize("sem", [0, 1])
def test_bif1d_rossler_matches_oracle(chaos, oracle, sem, tmax):
    args = (tmax, 96, 0.01, ROS["ic"], [2.0, 6.0], [2], 0, 0.0, 200.0, ROS["values"], 1)
    print(f"bif1d rossler 96 sem={sem} tMax={tmax}", flush=True)
    c, f, pk, iv = chaos.bif1d(*args, system="rossler_rk4", semantics=sem)
    oc, of, opk, oiv = _oracle_bif1d(oracle, "rossler_rk4", sem, *args)
    assert np.array_equal(c, oc)
    assert np.array_equal(f[c == -1], of[c == -1])
    assert np.array_equal(pk, opk) and np.array_equal(iv, oiv)
    assert c.sum() > 96  # periodic windows and chaos both produce peaks


def test_bif1d_flags_fixed_points_and_divergence(chaos, oracle):
    # Lorenz rho in [0.2, 30]: rho < 1 decays to the origin (flag -1, one
    # line with the first sample); maxValue 15 cuts the large-rho orbits (0)
    args = (60.0, 64, 0.01, [1.0, 1.0, 1.0], [0.2, 30.0], [1], 0, 15.0, 0.5, [10.0, 28.0, 8 / 3], 2)
    for sem in (0, 1):
        print(f"bif1d lorenz flags sem={sem}", flush=True)
        c, f, pk, iv = chaos.bif1d(*args, system="lorenz_rk4", semantics=sem)
        oc, of, opk, oiv = _oracle_bif1d(oracle, "lorenz_rk4", sem, *args)
        assert np.array_equal(c, oc)
        assert np.array_equal(f[c == -1], of[c == -1])
        assert np.array_equal(pk, opk) and np.array_equal(iv, oiv)
        assert (c == -1).any()


def test_bif1d_full_config(chaos, oracle):
    """configs: Rossler c in [2,6] x 1000, 1000 t.u. transient, 1000 recorded."""
    args = (1000.0, 1000, 0.01, ROS["ic"], [2.0, 6.0], [2], 0, 0.0, 1000.0, ROS["values"], 1)
    print("bif1d rossler 1000 full horizon", flush=True)
    c, f, pk, iv = chaos.bif1d(*args, system="rossler_rk4", semantics=1)
    oc, of, opk, oiv = _oracle_bif1d(oracle, "rossler_rk4", 1, *args)
    assert np.array_equal(c, oc)
    assert np.array_equal(pk, opk) and np.array_equal(iv, oiv)


def test_bifurcation1D_file(chaos, tmp_path):
    p = tmp_path / "bif1d.csv"
    ranges = [0.5, 30.0]
    args = (60.0, 24, 0.01, 3, [1.0, 1.0, 1.0], ranges, [1], 0, 15.0, 0.5, [10.0, 28.0, 8 / 3], 3, 2)
    opt = chaos.make_options("lorenz_rk4", 0)
    chaos.bifurcation1D(*args, str(p), options=opt)
    c, f, pk, iv = chaos.bif1d(60.0, 24, 0.01, [1.0, 1.0, 1.0], ranges, [1], 0, 15.0, 0.5,
                               [10.0, 28.0, 8 / 3], 2, system="lorenz_rk4", semantics=0)
    import _oracle as O
    lines = ["%g %g\n" % tuple(ranges)]
    k = 0
    for pnt in range(24):
        x = O.lib.orc_grid_value(pnt, 24, ranges[0], ranges[1])
        if c[pnt] == -1:
            lines.append("%.12g, %.12g, 0\n" % (x, f[pnt]))
        for _ in range(max(c[pnt], 0)):
            lines.append("%.12g, %.12g, %.12g\n" % (x, pk[k], iv[k]))
            k += 1
    assert p.read_text() == "".join(lines)


def _map_text(ranges, vals, n, fmt):
    s = "%g %g\n%g %g\n" % tuple(ranges)
    out = []
    cnt = 0
    for v in vals.reshape(-1):
        t = ""
        if cnt != 0:
            t += ", "
        if cnt == n:
            t += "\n"
            cnt = 0
        out.append(t + fmt(v))
        cnt += 1
    return s + "".join(out)


@pytest.mark.parametrize("tmax", [30.0, 30.011])  # 3000 / 3001 recorded samples
@pytest.mark.parametrize("sem", [0, 1])
def test_bif2d_lorenz_matches_oracle(chaos, oracle, sem, tmax):
    n = 12
    ranges = [20.0, 200.0, 0.5, 4.0]
    args = (tmax, n, 0.01, [1.0, 1.0, 1.0], ranges, [1, 2], 0, 0.0, 20.0, [10.0, 28.0, 8 / 3], 1, 1e-2)
    print(f"bif2d lorenz {n}x{n} sem={sem} tMax={tmax}", flush=True)
    got = chaos.bif2d(*args, system="lorenz_rk4", semantics=sem)
    want = oracle.bif2d(oracle.SYS["lorenz_rk4"], sem, tmax, n, 0.01, [1.0, 1.0, 1.0], ranges, [1, 2], 0, 0.0,
                        20.0, [10.0, 28.0, 8 / 3], 1, 1e-2)
    assert np.array_equal(got, want)


def test_bif2d_rossler_ref_quirky_dbscan(chaos, oracle):
    n = 16
    ranges = [0.05, 0.35, 2.0, 10.0]
    args = (200.0, n, 0.01, ROS["ic"], ranges, [0, 2], 0, 0.0, 100.0, ROS["values"], 1, 0.05)
    print("bif2d rossler ref 16x16", flush=True)
    got = chaos.bif2d(*args, system="rossler_rk4", semantics=0)
    want = oracle.bif2d(oracle.SYS["rossler_rk4"], 0, 200.0, n, 0.01, ROS["ic"], ranges, [0, 2], 0, 0.0, 100.0,
                        ROS["values"], 1, 0.05)
    assert np.array_equal(got, want)
    assert len(np.unique(got)) > 2


@pytest.mark.parametrize("sem", [0, 1])
@pytest.mark.parametrize("ct", [100.0, 500.0, 1200.0, 3000.0, 4500.0])
def test_bif2d_segment_classes_match_oracle(chaos, oracle, ct, sem):
    """Per-point 2-D analyses over segments of ~130 to ~6000 kept peaks:
    every size class (REFERENCE dbscan in LDS <= 256, 1024, 2048 and in
    global memory; CONFIG count sorted in registers up to 2048, in LDS up to
    4096 and in global memory)."""
    n = 5
    ranges = [24.0, 60.0, 1.5, 3.0]
    print(f"bif2d lorenz sem={sem} {n}x{n} CT={ct}", flush=True)
    st = chaos.Stats()
    for eps in (0.001, 0.05):
        got = chaos.bif2d(ct, n, 0.01, [1.0, 1.0, 1.0], ranges, [1, 2], 0, 0.0, 10.0, [10.0, 28.0, 8 / 3], 1, eps,
                          system="lorenz_rk4", semantics=sem, stats=st)
        want = oracle.bif2d(oracle.SYS["lorenz_rk4"], sem, ct, n, 0.01, [1.0, 1.0, 1.0], ranges, [1, 2], 0, 0.0,
                            10.0, [10.0, 28.0, 8 / 3], 1, eps)
        assert np.array_equal(got, want), eps
    per_point = st.peaks / (n * n)
    print(f"peaks per point {per_point:.0f}", flush=True)
    assert per_point > ct / 2


def test_bif2d_ref_keys32_size_classes(chaos, oracle):
    """REFERENCE map with a record of 65,000 samples (below 65,536: 32-bit
    sort keys) over rho x beta in [20,200] x [0.5,4]: points with 257 to
    ~1,700 kept intervals, i.e. the 512, 768, 1024, 1536 and 2048 classes
    (the 256 class is covered at CT 100 above)."""
    n = 8
    ranges = [20.0, 200.0, 0.5, 4.0]
    print(f"bif2d lorenz ref {n}x{n} CT=650", flush=True)
    st = chaos.Stats()
    got = chaos.bif2d(650.0, n, 0.01, [1.0, 1.0, 1.0], ranges, [1, 2], 0, 0.0, 20.0, [10.0, 28.0, 8 / 3], 1, 0.01,
                      system="lorenz_rk4", semantics=0, stats=st)
    want = oracle.bif2d(oracle.SYS["lorenz_rk4"], 0, 650.0, n, 0.01, [1.0, 1.0, 1.0], ranges, [1, 2], 0, 0.0, 20.0,
                        [10.0, 28.0, 8 / 3], 1, 0.01)
    assert np.array_equal(got, want)
    assert st.peaks / (n * n) > 700 and len(np.unique(got)) > 5


def test_bif2d_ref_harness_scale_sampled(chaos, oracle):
    """REFERENCE map at the regression harness's physics (Lorenz, CT 400,
    TT 2000, eps 0.001) on a 96^2 grid; 4 sampled rows recomputed by the
    oracle."""
    n = 96
    ranges = [20.0, 200.0, 0.5, 4.0]
    print(f"bif2d lorenz ref {n}x{n} CT=400 TT=2000", flush=True)
    got = chaos.bif2d(400.0, n, 0.01, [1.0, 1.0, 1.0], ranges, [1, 2], 0, 10000.0, 2000.0, [10.0, 28.0, 8 / 3], 1,
                      0.001, system="lorenz_rk4", semantics=0)
    for j in (0, 31, 64, 95):
        want = oracle.bif2d(oracle.SYS["lorenz_rk4"], 0, 400.0, n, 0.01, [1.0, 1.0, 1.0], ranges, [1, 2], 0,
                            10000.0, 2000.0, [10.0, 28.0, 8 / 3], 1, 0.001, row_begin=j, row_end=j + 1)
        assert np.array_equal(got[j], want[0]), j
    assert len(np.unique(got)) > 10


def test_bifurcation2D_file(chaos, tmp_path):
    p = tmp_path / "bif2d.csv"
    ranges = [20.0, 200.0, 0.5, 4.0]
    opt = chaos.make_options("lorenz_rk4", 1)
    chaos.bifurcation2D(20.0, 6, 0.01, 3, [1.0, 1.0, 1.0], ranges, [1, 2], 0, 0.0, 10.0,
                        [10.0, 28.0, 8 / 3], 3, 1, 1e-2, str(p), options=opt)
    vals = chaos.bif2d(20.0, 6, 0.01, [1.0, 1.0, 1.0], ranges, [1, 2], 0, 0.0, 10.0, [10.0, 28.0, 8 / 3], 1,
                       1e-2, system="lorenz_rk4", semantics=1)
    assert p.read_text() == _map_text(ranges, vals, 6, lambda v: "%d" % v)


def test_bif2d_full_config_sampled(chaos, oracle):
    """configs: Lorenz peak-count map 1024^2, 500 + 500 t.u.; 16 whole rows
    (16,384 points) recomputed by the oracle and compared point for point:
    the agreement fraction is reported against the 99.5 % tolerance for
    peak-count maps, and the map must in fact be equal on every point
    (same arithmetic on both sides)."""
    n = 1024
    ranges = [20.0, 200.0, 0.5, 4.0]
    print("bif2d lorenz 1024x1024 full horizon", flush=True)
    st = chaos.Stats()
    got = chaos.bif2d(500.0, n, 0.01, [1.0, 1.0, 1.0], ranges, [1, 2], 0, 0.0, 500.0, [10.0, 28.0, 8 / 3], 1,
                      1e-2, system="lorenz_rk4", semantics=1, stats=st)
    print(f"kernel {st.kernel_ms:.1f} ms analysis {st.analysis_ms:.1f} ms peaks {st.peaks}", flush=True)
    rng = np.random.default_rng(9)
    rows = sorted(set(int(v) for v in rng.choice(n, 16, replace=False)) | {0, n - 1})
    same = total = 0
    for j in rows:
        want = oracle.bif2d(oracle.SYS["lorenz_rk4"], 1, 500.0, n, 0.01, [1.0, 1.0, 1.0], ranges, [1, 2], 0, 0.0,
                            500.0, [10.0, 28.0, 8 / 3], 1, 1e-2, row_begin=j, row_end=j + 1)
        same += int((got[j] == want[0]).sum())
        total += n
    frac = same / total
    print(f"peak-count map: {same} of {total} points equal the oracle ({100 * frac:.3f} %)", flush=True)
    assert frac >= 0.995
    assert same == total
    assert (got > 0).mean() > 0.9

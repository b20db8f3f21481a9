"""GPU parity of Basins (chaos_basins2d / chaos_basinsOfAttraction_2) against
the oracle: bit-exact features and flags; labels equal to the oracle's
deterministic clustering (single linkage for REFERENCE, brute-force DBSCAN on
z-scored features for CONFIG)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LOR = [10.0, 24.5, 8.0 / 3.0]


def test_basins_ref_matches_oracle(chaos, oracle):
    n = 40
    ranges = [-20.0, 20.0, -20.0, 20.0]
    args = (20.0, n, 0.01, [0.0, 0.0, 25.0], ranges, [0, 1], 0, 0.0, 20.0, LOR, 1)
    print(f"basins ref lorenz {n}x{n}", flush=True)
    lab, feat, flags = chaos.basins2d(*args, eps=0.05, system="lorenz_rk4", semantics=0)
    f0, f1, f2, ofl = oracle.basins_features(oracle.SYS["lorenz_rk4"], 0, *args)
    assert np.array_equal(flags.reshape(-1), ofl)
    assert np.array_equal(feat[0].reshape(-1), f0) and np.array_equal(feat[1].reshape(-1), f1)
    want = oracle.basins_cluster_ref(f0, f1, ofl, 0.05)
    assert np.array_equal(lab.reshape(-1), want)


def test_basins_ref_prescaller_and_fixed_points(chaos, oracle):
    # rho = 10: the two stable equilibria attract everything -> fixed-point
    # flags and negative labels; preScaller 3 sub-steps per sample
    n = 24
    ranges = [-10.0, 10.0, -10.0, 10.0]
    args = (45.0, n, 0.01, [0.0, 0.0, 9.0], ranges, [0, 1], 0, 0.0, 150.0, [10.0, 10.0, 8 / 3], 3)
    print("basins ref lorenz rho=10 pre=3", flush=True)
    lab, feat, flags = chaos.basins2d(*args, eps=0.05, system="lorenz_rk4", semantics=0)
    f0, f1, f2, ofl = oracle.basins_features(oracle.SYS["lorenz_rk4"], 0, *args)
    assert np.array_equal(flags.reshape(-1), ofl)
    assert np.array_equal(feat[0].reshape(-1), f0) and np.array_equal(feat[1].reshape(-1), f1)
    want = oracle.basins_cluster_ref(f0, f1, ofl, 0.05)
    assert np.array_equal(lab.reshape(-1), want)


# tMax 30 / 30.011: 3000 / 3001 recorded samples (odd and even detector-body
# counts; the record loop takes two samples per trip)
@pytest.mark.parametrize("minpts,tmax", [(20, 30.0), (5, 30.0), (20, 30.011)])
def test_basins_config_matches_oracle(chaos, oracle, minpts, tmax):
    n = 64
    ranges = [-20.0, 20.0, -20.0, 20.0]
    args = (tmax, n, 0.01, [0.0, 0.0, 25.0], ranges, [0, 1], 0, 0.0, 30.0, LOR, 1)
    print(f"basins config lorenz {n}x{n} minpts={minpts} tMax={tmax}", flush=True)
    lab, feat, flags = chaos.basins2d(*args, eps=0.5, system="lorenz_rk4", semantics=1, min_pts=minpts)
    f0, f1, f2, ofl = oracle.basins_features(oracle.SYS["lorenz_rk4"], 1, *args)
    assert np.array_equal(flags.reshape(-1), ofl)
    for k, f in enumerate((f0, f1, f2)):
        assert np.array_equal(feat[k].reshape(-1), f), k
    z0, z1, z2, _ = oracle.zscore(f0, f1, f2, ofl)
    want = oracle.dbscan_bruteforce(np.column_stack([z0, z1, z2]), 0.5, minpts, include=(ofl != 0))
    assert np.array_equal(lab.reshape(-1), want)
    assert len(np.unique(want)) >= 2


def test_basinsOfAttraction_2_files(chaos, tmp_path):
    n = 8
    ranges = [-20.0, 20.0, -20.0, 20.0]
    base = tmp_path / "basins.csv"
    opt = chaos.make_options("lorenz_rk4", 0)
    chaos.basinsOfAttraction_2(10.0, n, 0.01, 3, [0.0, 0.0, 25.0], ranges, [0, 1], 0, 0.0, 10.0, LOR, 3, 1, 0.05,
                               str(base), options=opt)
    lab, feat, flags = chaos.basins2d(10.0, n, 0.01, [0.0, 0.0, 25.0], ranges, [0, 1], 0, 0.0, 10.0, LOR, 1,
                                      eps=0.05, system="lorenz_rk4", semantics=0)

    def text(vals, fmt):
        s = "%.15g %.15g\n%.15g %.15g\n" % tuple(ranges)
        cnt, out = 0, []
        for v in vals.reshape(-1):
            t = ", " if cnt != 0 else ""
            if cnt == n:
                t += "\n"
                cnt = 0
            out.append(t + fmt(v))
            cnt += 1
        return s + "".join(out)

    assert base.read_text() == text(lab, lambda v: "%d" % v)
    assert (tmp_path / "basins.csv_1.csv").read_text() == text(feat[0], lambda v: "%.15g" % v)
    assert (tmp_path / "basins.csv_2.csv").read_text() == text(feat[1], lambda v: "%.15g" % v)
    assert (tmp_path / "basins.csv_3.csv").read_text() == text(flags, lambda v: "%d" % v)


def _perm_agreement(a, b):
    """Fraction of points whose label agrees after mapping each label of a to
    the label of b it shares most points with (labels equal up to
    permutation); 0 (noise / divergent) maps to 0 only."""
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    pairs, counts = np.unique(np.stack([a, b]), axis=1, return_counts=True)
    best = {}
    for (la, lb), c in zip(pairs.T, counts):
        if la not in best or c > best[la][1]:
            best[la] = (lb, c)
    agree = sum(c for la, (lb, c) in best.items() if (la == 0) == (lb == 0))
    return agree / a.size


def test_basins_full_config(chaos, oracle):
    """configs: Lorenz rho=24.5 basins on a 2048^2 IC grid, 300 + 300 t.u.,
    DBSCAN(0.5, 20) on z-scored features.
    * 256 sampled points: raw features and flags bit-exact vs the oracle's
      per-point restatement;
    * all 4,194,304 labels: the GPU features z-scored by the oracle and
      clustered by its grid restatement of the DBSCAN (pinned against the
      brute force), compared up to permutation: the agreement fraction is
      reported against the 99 % tolerance for basin labels, and the labels
      must in fact be identical (same numbering rule on identical features).
    """
    n = 2048
    ranges = [-20.0, 20.0, -20.0, 20.0]
    st = chaos.Stats()
    print("basins config lorenz 2048x2048 full horizon", flush=True)
    lab, feat, flags = chaos.basins2d(300.0, n, 0.01, [0.0, 0.0, 25.0], ranges, [0, 1], 0, 0.0, 300.0, LOR, 1,
                                      eps=0.5, system="lorenz_rk4", semantics=1, stats=st)
    print(f"kernel {st.kernel_ms:.1f} ms analysis {st.analysis_ms:.1f} ms", flush=True)
    rng = np.random.default_rng(5)
    pts = rng.choice(n * n, 256, replace=False)
    for p in pts:
        p = int(p)
        f0, f1, f2, fl = oracle.basins_features(oracle.SYS["lorenz_rk4"], 1, 300.0, n, 0.01, [0.0, 0.0, 25.0],
                                                ranges, [0, 1], 0, 0.0, 300.0, LOR, 1, point_begin=p,
                                                point_end=p + 1)
        j, i = divmod(p, n)
        assert flags[j, i] == fl[0]
        assert feat[0, j, i] == f0[0] and feat[1, j, i] == f1[0] and feat[2, j, i] == f2[0]
    print("  256 sampled feature points bit-exact", flush=True)
    fl = flags.ravel().astype(np.int32)
    z0, z1, z2, _ = oracle.zscore(feat[0].ravel(), feat[1].ravel(), feat[2].ravel(), fl)
    want = oracle.dbscan_grid(np.stack([z0, z1, z2], 1), 0.5, 20, fl)
    frac = _perm_agreement(lab.ravel(), want)
    print(f"  labels: {100 * frac:.4f} % agree with the oracle up to permutation; "
          f"{int((lab.ravel() == want).sum())} of {n * n} identical", flush=True)
    assert frac >= 0.99
    assert np.array_equal(lab.ravel(), want)
    vals, counts = np.unique(lab, return_counts=True)
    top = np.sort(counts)[::-1]
    assert top[:3].sum() > 0.9 * n * n, dict(zip(vals.tolist(), counts.tolist()))

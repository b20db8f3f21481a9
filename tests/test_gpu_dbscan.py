"""GPU parity of the grid DBSCAN (chaos_dbscan) against the oracle's
brute-force DBSCAN and the reference-semantics single linkage, on point sets
built to reach every branch of the cell logic: crowded cells (a basin that
converges to one fixed point), neighbour cells wholly inside / wholly outside
eps of a point, distances of exactly eps, border points between clusters,
excluded points and fixed-point seeds."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def crowd_scene(rng, n_scatter, dim=3):
    eps = 0.5
    pts = [np.zeros((3000, dim))]                                  # one crowded cell at the origin
    pts.append(np.tile([1.2] + [0.0] * (dim - 1), (500, 1)))       # a second crowd > eps away
    ring = np.zeros((40, dim))
    ring[:, 0] = 0.5                                               # exactly eps from the origin crowd
    ring[20:, 0] = np.nextafter(0.5, 1.0)                          # one ulp beyond eps
    pts.append(ring)
    pts.append(rng.normal(0.0, 0.6, size=(n_scatter, dim)))       # partial overlaps everywhere
    pts.append(rng.normal(3.0, 0.05, size=(200, dim)))             # a tight blob, several cells
    pts.append(rng.uniform(-6, 6, size=(300, dim)))                # noise
    p = np.concatenate(pts)
    perm = rng.permutation(len(p))
    return p[perm], eps


@pytest.mark.parametrize("min_pts", [20, 5, 1])
def test_dbscan_3d_crowded_matches_bruteforce(chaos, oracle, min_pts):
    rng = np.random.default_rng(11 + min_pts)
    pts, eps = crowd_scene(rng, 4000)
    print(f"dbscan 3-D crowded n={len(pts)} min_pts={min_pts}", flush=True)
    got, k = chaos.dbscan(pts, eps, min_pts)
    want = oracle.dbscan_bruteforce(pts, eps, min_pts)
    assert np.array_equal(got, want)
    assert k == want.max()


def test_dbscan_2d_matches_bruteforce(chaos, oracle):
    rng = np.random.default_rng(3)
    pts, eps = crowd_scene(rng, 3000, dim=2)
    print(f"dbscan 2-D n={len(pts)}", flush=True)
    got, k = chaos.dbscan(pts, eps, 8)
    want = oracle.dbscan_bruteforce(pts, eps, 8)
    assert np.array_equal(got, want)


def test_dbscan_include_mask(chaos, oracle):
    rng = np.random.default_rng(4)
    pts, eps = crowd_scene(rng, 2000)
    inc = (rng.random(len(pts)) < 0.8).astype(np.int32)
    print(f"dbscan include mask n={len(pts)}", flush=True)
    got, _ = chaos.dbscan(pts, eps, 10, include=inc)
    want = oracle.dbscan_bruteforce(pts, eps, 10, include=inc)
    assert np.array_equal(got, want)
    assert np.all(got[inc == 0] == 0)


def test_dbscan_border_tie_takes_smallest_core(chaos, oracle):
    # two clusters of 6 points each; point 0 sits between them, within eps
    # of one core of each: it must join the cluster of the smaller core index
    a = np.array([[-0.9, 0.0]] + [[-0.9 - 0.01 * i, 0.0] for i in range(1, 6)])
    b = np.array([[0.9 + 0.01 * i, 0.0] for i in range(6)])
    mid = np.array([[0.0, 0.0]])
    for order in (np.r_[mid, b, a], np.r_[mid, a, b]):
        got, _ = chaos.dbscan(order, 0.95, 6)
        want = oracle.dbscan_bruteforce(order, 0.95, 6)
        assert np.array_equal(got, want)
        assert got[0] == got[1]   # index 1 is the smaller core neighbour of point 0


def test_dbscan_edge_cases(chaos, oracle):
    pts = np.zeros((1, 3))
    got, k = chaos.dbscan(pts, 0.5, 1)
    assert got.tolist() == [1] and k == 1
    got, k = chaos.dbscan(pts, 0.5, 2)
    assert got.tolist() == [0] and k == 0
    rng = np.random.default_rng(0)
    pts = rng.normal(size=(100, 3))
    got, k = chaos.dbscan(pts, 0.5, 3, include=np.zeros(100, np.int32))
    assert not got.any() and k == 0
    got, k = chaos.dbscan(np.zeros((0, 3)), 0.5, 3)
    assert got.size == 0 and k == 0
    with pytest.raises(chaos.ChaosError):
        chaos.dbscan(pts, -1.0, 3)
    with pytest.raises(chaos.ChaosError):
        chaos.dbscan(pts, 0.5, 0)


def test_dbscan_sign_rule_matches_reference_linkage(chaos, oracle):
    # Basins REFERENCE: single linkage, fixed points (-1) give negative ids
    rng = np.random.default_rng(9)
    n = 6000
    f0 = np.round(rng.normal(0, 1, n), 2)
    f1 = np.round(rng.normal(0, 1, n), 2)
    f0[:2000] = 999.0                           # the flag-0 sentinel crowd
    f1[:2000] = 999.0
    flags = np.ones(n, np.int32)
    flags[:2000] = 0
    flags[2000:2600] = -1
    perm = rng.permutation(n)
    f0, f1, flags = f0[perm], f1[perm], flags[perm]
    print("dbscan sign rule n=6000", flush=True)
    got, _ = chaos.dbscan(np.column_stack([f0, f1]), 0.05, 1, sign_rule=True, include=flags)
    want = oracle.basins_cluster_ref(f0, f1, flags, 0.05)
    assert np.array_equal(got, want)
    assert got.min() < 0 < got.max()


def test_dbscan_large_crowds_matches_bruteforce(chaos, oracle):
    # 20k points, 90% of them in three crowds (the shape of converged basins)
    rng = np.random.default_rng(21)
    crowds = [np.tile(c, (6000, 1)) for c in ([0.0, 0.0, 0.0], [0.7, 0.0, 0.0], [5.0, 5.0, 5.0])]
    halo = rng.normal(0.3, 0.4, size=(2000, 3))
    pts = np.concatenate(crowds + [halo])
    pts = pts[rng.permutation(len(pts))]
    print(f"dbscan large crowds n={len(pts)}", flush=True)
    got, _ = chaos.dbscan(pts, 0.5, 20)
    want = oracle.dbscan_bruteforce(pts, 0.5, 20)
    assert np.array_equal(got, want)


def test_dbscan_split_core_tests_match_bruteforce(chaos, oracle, capfd, monkeypatch):
    """Core tests too long for one wavefront: 300 candidates in a shell 0.42-0.5
    from a cloud of 20 000 points in one cell-sized box, minPts 300, so each
    candidate's count is decided only by scanning much of the cloud (a lone
    point at (-0.88, -0.88, -0.88) sets the grid origin so that the whole
    cloud falls in one cell of side 0.5 / sqrt(3)).  The
    candidates scan each cell's head themselves, then split the cloud into
    4096-member pieces; the piece list (n/16 + 64 entries) fills, so the last
    candidates fall back to scanning in their own wavefront.  Labels must
    equal the brute force, and the candidates must be a mix of core and
    non-core points."""
    import re
    rng = np.random.default_rng(5)
    cloud = rng.uniform(0, 0.25, (20000, 3))
    cands = []
    while len(cands) < 300:
        p = rng.uniform(-0.6, 0.85, 3)
        g = np.maximum(0, np.maximum(-p, p - 0.25))
        if 0.42 <= np.sqrt((g ** 2).sum()) <= 0.5:
            cands.append(p)
    cands = np.array(cands)
    pts = np.concatenate([cloud, cands, np.full((1, 3), -0.88)])
    pts = pts[rng.permutation(len(pts))]
    monkeypatch.setenv("CHAOS_DBSCAN_TRACE", "1")
    lo, side = pts.min(0), 0.5 / np.sqrt(3.0) * (1 - 1e-7)
    assert len(np.unique(np.floor((cloud - lo) / side), axis=0)) == 1
    print(f"dbscan split core tests n={len(pts)} min_pts=300", flush=True)
    got, k = chaos.dbscan(pts, 0.5, 300)
    line = [l for l in capfd.readouterr().err.splitlines() if l.startswith("dbscan:")][-1]
    want = oracle.dbscan_bruteforce(pts, 0.5, 300)
    # neighbour counts of the candidates, summed in coordinate order as the kernels do
    d = pts[None, :, :] - cands[:, None, :]
    cnt = (np.sqrt((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]) <= 0.5).sum(1)
    print(line, f"; core candidates {(cnt >= 300).sum()} of 300", flush=True)
    assert np.array_equal(got, want)
    pieces, cap = map(int, re.search(r"(\d+) split pieces of (\d+)", line).groups())
    assert pieces > 0 and pieces + 5 > cap  # the list was used and filled
    assert 0 < (cnt >= 300).sum() < 300

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

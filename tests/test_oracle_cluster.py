"""Pins for the oracle's clustering functions.

* 1-D DBSCAN count vs sklearn DBSCAN(min_samples=1) and the sorted-gap
  closed form.
* The reference per-point dbscan (Bifurcation/src/bifurcationCUDA.cu:516-572)
  on hand-traced inputs (its quirks make results differ from any textbook
  clustering).
* Basins single linkage vs connected components from scipy.
* Brute-force DBSCAN vs sklearn's core-sample partition.
* Deterministic sum vs math.fsum.
"""
import math

import numpy as np
import pytest
from scipy.sparse import csr_matrix
from scipy.sparse.csgraph import connected_components
from sklearn.cluster import DBSCAN

import _oracle as O


def test_dbscan1d_count_matches_sklearn():
    rng = np.random.default_rng(0)
    for trial in range(60):
        n = int(rng.integers(1, 300))
        k = int(rng.integers(1, 8))
        centers = rng.uniform(-20, 20, k)
        v = centers[rng.integers(0, k, n)] + rng.normal(0, rng.choice([1e-4, 3e-3, 2e-2]), n)
        want = len(set(DBSCAN(eps=1e-2, min_samples=1).fit(v.reshape(-1, 1)).labels_))
        assert O.dbscan1d_count(v, 1e-2) == want


def test_dbscan1d_count_edges():
    assert O.dbscan1d_count(np.array([]), 0.01) == 0
    assert O.dbscan1d_count(np.array([3.0]), 0.01) == 1
    assert O.dbscan1d_count(np.array([0.0, 0.01]), 0.01) == 1   # |d| == eps joins
    assert O.dbscan1d_count(np.array([0.0, 0.0100001]), 0.01) == 2
    assert O.dbscan1d_count(np.array([5.0, 1.0, 5.005, 1.002]), 0.01) == 2


@pytest.mark.parametrize("intervals,eps,want", [
    ([], 0.1, 0),
    ([1.0], 0.1, 1),
    # P=2: k only reaches 0 (k < P-1).  i=0: cluster 1, no k != 0 -> i=1:
    # helpful[1]==0 -> cluster 2 -> returns 1
    ([1.0, 1.0], 0.1, 1),
    # P=3, all equal: i=0 -> cluster 1, k=1 joins (push slot P+1=4, NumNeibor=1);
    # i=1 iteration pops slot P+0=3 (== 0) -> i=0, pass from 0 (nothing new),
    # i++ -> 1; helpful[1]=1 -> skip; i=2: helpful[2]==0 -> cluster 2 -> returns 1
    ([2.0, 2.0, 2.0], 0.1, 1),
    # P=3 all far apart: three clusters -> returns 2
    ([1.0, 2.0, 3.0], 0.1, 2),
    # P=4: 0 and 1 close, 2 and 3 close; k < 3 never reaches 3
    ([1.0, 1.05, 5.0, 5.05], 0.1, 2),
])
def test_dbscan_ref_hand_traced(intervals, eps, want):
    assert O.dbscan_ref(np.array(intervals, dtype=float), eps) == want


def test_dbscan_ref_ignores_peak_values():
    iv = np.array([1.0, 1.01, 3.0, 3.02, 7.0])
    a = O.dbscan_ref(iv, 0.05, peaks=np.array([0.0, 100.0, -5.0, 2.0, 9.0]))
    b = O.dbscan_ref(iv, 0.05)
    assert a == b


def test_dbscan_ref_distance_is_float32():
    # the distance is hypotf(0, (float)dy): a gap that is < eps only after
    # rounding to float still joins
    eps = 0.1
    dy = np.nextafter(np.float64(0.1), 1.0)  # double just above 0.1
    assert np.float32(dy) <= np.float32(0.1)
    # float(0.1) > 0.1 in double, so |float(dy)| < eps is False either way;
    # pick eps = float32(0.1) promoted: then float(dy) == eps -> not < eps
    res = O.dbscan_ref(np.array([0.0, dy, 10.0]), float(np.float32(0.1)) + 1e-12)
    assert res == 1  # {0,1} joined, {2} separate -> 2 clusters -> returns 1


def _components(f0, f1, flags, eps):
    pts = np.column_stack([f0, f1])
    inc = flags != 0
    idx = np.flatnonzero(inc)
    d = np.sqrt(((pts[idx, None, :] - pts[None, idx, :]) ** 2).sum(-1))
    ncomp, lab = connected_components(csr_matrix(d <= eps), directed=False)
    full = np.full(len(f0), -1)
    full[idx] = lab
    return full


def test_basins_cluster_ref_matches_components():
    rng = np.random.default_rng(1)
    n = 600
    centers = np.array([[0.0, 0.0], [5.0, 5.0], [999.0, 999.0], [-3.0, 4.0]])
    c = rng.integers(0, 4, n)
    f0 = centers[c, 0] + rng.normal(0, 0.05, n)
    f1 = centers[c, 1] + rng.normal(0, 0.05, n)
    flags = np.where(c == 2, 0, np.where(c == 3, -1, 1)).astype(np.int32)
    flags[rng.integers(0, n, 5)] = -1
    lab = O.basins_cluster_ref(f0, f1, flags, 0.3)
    comp = _components(f0, f1, flags, 0.3)
    assert np.all(lab[flags == 0] == 0)
    inc = flags != 0
    # same partition
    for a in np.unique(comp[inc]):
        members = np.flatnonzero(comp == a)
        assert len(set(lab[members])) == 1
        L = lab[members[0]]
        assert (L < 0) == bool(np.any(flags[members] == -1))
    assert len(set(lab[inc])) == len(set(comp[inc]))
    # numbering: positives 1..K by smallest member, negatives by smallest fixed index
    pos = [lab[i] for i in range(n) if lab[i] > 0]
    firsts = []
    for v in pos:
        if v not in firsts:
            firsts.append(v)
    assert firsts == list(range(1, len(firsts) + 1))


def test_bruteforce_dbscan_core_partition_matches_sklearn():
    rng = np.random.default_rng(2)
    pts = np.concatenate([rng.normal(0, 0.3, (400, 3)), rng.normal(3, 0.2, (300, 3)),
                          rng.uniform(-4, 6, (150, 3))])
    eps, mp = 0.5, 20
    lab = O.dbscan_bruteforce(pts, eps, mp)
    sk = DBSCAN(eps=eps, min_samples=mp).fit(pts)
    core = np.zeros(len(pts), bool)
    core[sk.core_sample_indices_] = True
    # same core partition (up to label permutation)
    pairs = set(zip(lab[core], sk.labels_[core]))
    assert len(pairs) == len(set(lab[core])) == len(set(sk.labels_[core]))
    # noise agrees
    assert np.array_equal(lab == 0, sk.labels_ == -1)
    # numbering by smallest core index
    order = []
    for i in np.flatnonzero(core):
        if lab[i] not in order:
            order.append(lab[i])
    assert order == list(range(1, len(order) + 1))


def test_bruteforce_dbscan_border_rule():
    # border point 3 is within eps of core points of two clusters: it joins
    # the cluster of the smaller-index core neighbour
    a = [[x, 0, 0] for x in (0.0, 0.05, 0.1, 0.15, 0.2)]
    b = [[x, 0, 0] for x in (1.2, 1.25, 1.3, 1.35, 1.4)]
    pts = np.array(b + a + [[0.7, 0, 0]])
    lab = O.dbscan_bruteforce(pts, 0.5, 4)
    assert list(lab[:10]) == [1] * 5 + [2] * 5
    # point 10 at 0.7 has 3 neighbours (not core) and is within eps of
    # 1.2 (index 0) and 0.2 (index 9): it takes index 0's cluster
    assert lab[10] == 1


def test_det_sum_accuracy_and_mask():
    rng = np.random.default_rng(3)
    for n in (1, 5, 1023, 1024, 1025, 5000):
        x = rng.normal(0, 10, n)
        m = (rng.random(n) < 0.8).astype(np.int32)
        assert abs(O.det_sum(x) - math.fsum(x)) <= 1e-12 * max(1.0, np.abs(x).sum())
        assert abs(O.det_sum(x, m) - math.fsum(x[m != 0])) <= 1e-12 * max(1.0, np.abs(x).sum())


def _pairwise(v):
    """The complete adjacent-pair tree, evaluated recursively: the sum of a
    run is the sum of its two halves."""
    if len(v) == 1:
        return float(v[0])
    h = len(v) // 2
    return _pairwise(v[:h]) + _pairwise(v[h:])


def test_det_sum_is_the_padded_pairwise_tree():
    """orc_det_sum pads to max(1024, next power of two) with +0 and sums by
    the complete pairwise tree; values of mixed magnitude make the order
    visible (sequential and blocked orders give other bits)."""
    rng = np.random.default_rng(11)
    differs = 0
    for n in (1, 7, 1000, 1024, 1025, 3000, 5000):
        x = rng.normal(0, 1, n) * 10.0 ** rng.integers(-8, 9, n)
        m = (rng.random(n) < 0.7).astype(np.int32)
        P = 1024
        while P < n:
            P *= 2
        for mask in (None, m):
            v = np.zeros(P)
            v[:n] = x if mask is None else np.where(mask != 0, x, 0.0)
            want = _pairwise(list(v))
            got = O.det_sum(x) if mask is None else O.det_sum(x, mask)
            assert got == want
            seq = 0.0
            for t in v:
                seq = seq + t
            differs += seq != want
    assert differs > 0


def test_zscore():
    rng = np.random.default_rng(4)
    f = [rng.normal(3, 2, 3000) for _ in range(3)]
    flags = np.where(rng.random(3000) < 0.1, 0, 1).astype(np.int32)
    z0, z1, z2, mom = O.zscore(*f, flags)
    inc = flags != 0
    for z, orig, k in ((z0, f[0], 0), (z1, f[1], 1), (z2, f[2], 2)):
        assert abs(z[inc].mean()) < 1e-12 and abs(z[inc].std() - 1) < 1e-12
        assert np.array_equal(z[~inc], orig[~inc])
        assert abs(mom[2 * k] - orig[inc].mean()) < 1e-12


def _grid_cases():
    rng = np.random.default_rng(17)
    cases = []
    for dim in (2, 3):
        # blobs + background noise, several densities around minPts
        for n, eps, mp in ((3000, 0.3, 5), (8000, 0.15, 12), (20000, 0.1, 20)):
            centres = rng.uniform(-3, 3, (5, dim))
            k = rng.integers(0, 6, n)
            pts = np.where((k < 5)[:, None], centres[np.minimum(k, 4)] + rng.normal(0, 0.3, (n, dim)),
                           rng.uniform(-5, 5, (n, dim)))
            inc = (rng.random(n) > 0.1).astype(np.int32)
            cases.append((pts, eps, mp, inc))
        # a converged basin: thousands of identical points beside a sparse cloud
        pts = np.concatenate([np.tile(rng.uniform(-1, 1, dim), (4000, 1)), rng.uniform(-2, 2, (3000, dim))])
        cases.append((pts, 0.25, 20, None))
        # a lattice at spacing exactly eps (ties sit on the eps boundary)
        g = np.arange(0, 30) * 0.5
        mesh = np.stack(np.meshgrid(*([g] * dim), indexing="ij"), -1).reshape(-1, dim)[:20000]
        cases.append((mesh, 0.5, 2 * dim + 1, None))
    return cases


@pytest.mark.parametrize("case", range(10))
def test_grid_dbscan_equals_bruteforce(case):
    pts, eps, mp, inc = _grid_cases()[case]
    want = O.dbscan_bruteforce(pts, eps, mp, inc)
    got = O.dbscan_grid(pts, eps, mp, inc)
    assert np.array_equal(got, want)
    assert want.max() >= 1

"""Pins for the oracle's peakFinder and integration-loop flags.

Expected outputs below are traced by hand through the reference text
(Bifurcation/src/bifurcationCUDA.cu:469-514 and :421-467), including its
plateau rule, the first-peak drop and the skipped index after a rejected
plateau.  The same cases are replayed on the GPU detector in
tests/test_gpu_peaks.py.
"""
import math

import numpy as np
import pytest

import _oracle as O

H = 0.01

# (samples, expected kept peaks, expected intervals in units of h)
HAND_CASES = [
    # plateau 2,2,2 then descent: peak value 2 at trunc((2+4)/2)=3; second
    # candidate at i=9 (3,3 then 1): peak 3 at trunc((9+10)/2)=9
    ([0, 1, 2, 2, 2, 1, 0, 1, 2, 3, 3, 1], [3.0], [6]),
    # peak 3 at 1; plateau 1,1 rejected at j=4 (1 < 2) -> i=5, ++i -> index 5
    # (the genuine maximum 2) is never a candidate; peak 5 at 7
    ([0, 3, 0, 1, 1, 2, 0, 5, 0, 0], [5.0], [6]),
    # single peak -> nothing kept
    ([0, 1, 1, 2, 0, 5, 0, 0], [], []),
    # plateau that runs to the end is not a peak
    ([0, 1, 2, 2, 2], [], []),
    # rises below the 1e-13 threshold are not candidates
    ([0, 1e-14, 0, 1, 0, 1, 0], [1.0], [2]),
    # a fall below the threshold extends the plateau
    ([0, 1, 1 - 1e-14, 0.5, 0, 2, 1], [2.0], [4]),
    # constant and monotone signals
    ([1.0] * 20, [], []),
    (list(range(20)), [], []),
    ([], [], []),
    ([1.0], [], []),
    ([0, 1], [], []),
]


@pytest.mark.parametrize("data,peaks,dts", HAND_CASES)
def test_peak_finder_hand_cases(data, peaks, dts):
    pk, iv = O.peak_finder(np.array(data, dtype=float), H)
    assert list(pk) == peaks
    assert np.array_equal(iv, np.array(dts, dtype=float) * H)


def test_peak_finder_sampled_sine():
    # maxima of sin(2 pi (i+0.3)/50) are at i = 12 + 50k (strict), 10 found,
    # the first dropped: 9 kept, all intervals 50 h
    i = np.arange(500)
    d = np.sin(2 * math.pi * (i + 0.3) / 50)
    pk, iv = O.peak_finder(d, H)
    assert len(pk) == 9
    assert np.allclose(pk, d[12])
    assert np.array_equal(iv, np.full(9, 50 * H))


def test_peak_finder_interval_uses_h():
    d = np.array([0, 1, 0, 0, 1, 0, 0, 0, 1, 0], dtype=float)
    _, iv = O.peak_finder(d, 0.5)
    assert list(iv) == [1.5, 2.0]


def test_loop_flags():
    ros = O.SYS["rossler_rk4"]
    lor = O.SYS["lorenz_rk4"]
    # stable origin of Lorenz (rho < 1) reached exactly -> -1
    flag, x, _ = O.loop(lor, [0.0, 0.0, 0.0], [10, 0.5, 8 / 3], 0.01, 100, 1, 0, 0.0, O.CHK_CONFIG)
    assert flag == -1
    # chaotic orbit -> 1
    flag, _, data = O.loop(lor, [1.0, 1.0, 1.0], [10, 28, 8 / 3], 0.01, 1000, 1, 0, 0.0, O.CHK_CONFIG)
    assert flag == 1 and data[0] == 1.0
    # exceeding maxValue -> 0 (|x| of Lorenz reaches ~18)
    flag, _, _ = O.loop(lor, [1.0, 1.0, 1.0], [10, 28, 8 / 3], 0.01, 1000, 1, 0, 10.0, O.CHK_CONFIG)
    assert flag == 0
    # blow-up (Rossler with c far outside) -> non-finite -> 0 in every mode
    for chk in (O.CHK_BIF_REF, O.CHK_BASINS_REF, O.CHK_CONFIG):
        flag, _, _ = O.loop(ros, [1e3, 1e3, 1e3], [0.2, 0.2, -50.0], 0.01, 5000, 1, 2, 0.0, chk)
        assert flag == 0, chk
    # amount 0 -> 1, data untouched
    flag, _, _ = O.loop(ros, [0.1, 0.1, 0.1], [0.2, 0.2, 5.7], 0.01, 0, 1, 0, 0.0, O.CHK_CONFIG)
    assert flag == 1


def test_loop_prescaller_steps():
    ros = O.SYS["rossler_rk4"]
    _, x1, d1 = O.loop(ros, [0.1, 0.1, 0.1], [0.2, 0.2, 5.7], 0.01, 50, 3, 0, 0.0, O.CHK_BASINS_REF)
    x2 = O.steps(ros, [0.1, 0.1, 0.1], [0.2, 0.2, 5.7], 0.01, 150)
    assert np.array_equal(x1, x2)
    assert d1[1] == O.steps(ros, [0.1, 0.1, 0.1], [0.2, 0.2, 5.7], 0.01, 3)[0]


def test_bif_ref_float_bound_quirk():
    # fabsf((float)v) > maxValue (bifurcationCUDA.cu:451): with maxValue set to
    # float(|v|) < |v|, the FP64 config test diverges while the reference's
    # float test does not
    ros = O.SYS["rossler_rk4"]
    x0 = [0.1, 0.1, 0.1]
    for n in range(1, 200):
        v = abs(O.steps(ros, x0, (0.2, 0.2, 5.7), 0.01, n)[0])
        f = float(np.float32(v))
        if f < v:
            break
    else:
        pytest.skip("no downward float rounding found")
    x_prev = O.steps(ros, x0, (0.2, 0.2, 5.7), 0.01, n - 1)
    f_ref, _, _ = O.loop(ros, x_prev, (0.2, 0.2, 5.7), 0.01, 1, 1, 0, f, O.CHK_BIF_REF)
    f_cfg, _, _ = O.loop(ros, x_prev, (0.2, 0.2, 5.7), 0.01, 1, 1, 0, f, O.CHK_CONFIG)
    assert f_cfg == 0 and f_ref == 1


def test_reference_dbscan_bucketed_walk_equals_oracle(oracle):
    """The per-point reference dbscan (Bifurcation/src/bifurcationCUDA.cu:
    516-572) decided in bulk, as the GPU kernel does: indices grouped by
    equal sample gap, a pass assigns whole buckets within eps of its centre
    (a contiguous run, walked through "nearest live bucket" links), a pass
    centred on an already expanded gap is skipped, and runs of such stack
    pops are taken up to `batch` at a time.  This model of that procedure
    must give the oracle's result on random interval sets (duplicates, near
    ties, eps below / at / above the sample spacing, eps = 0)."""
    import numpy as np

    def bucketed(dt, hp, eps, batch):
        P = len(dt)
        if P <= 1:
            return P
        y = [float(d) * hp for d in dt]
        order = sorted(range(P), key=lambda k: (dt[k], k))
        bstart, bucket_of, bdt = [], [0] * P, []
        for s, k in enumerate(order):
            if s == 0 or dt[k] != dt[order[s - 1]]:
                bstart.append(s)
                bdt.append(dt[k])
            bucket_of[k] = len(bstart) - 1
        nb = len(bstart)
        bstart.append(P)
        live = [any(order[s] < P - 1 for s in range(bstart[b], bstart[b + 1])) for b in range(nb)]
        lp = [b if live[b] else b - 1 for b in range(nb)]
        rp = [b if live[b] else b + 1 for b in range(nb)]
        expanded = [False] * nb
        h = [0] * (2 * P)
        st = {"cluster": 0, "num": 0}

        def find(links, b, end):
            while b != end and links[b] != b:
                b = links[b]
            return b

        def qualifies(b, yi):
            yb = float(bdt[b]) * hp
            return (0.0 if yb == yi else float(abs(np.float32(yb - yi)))) < eps

        def assign(b, i):
            found = 0
            for s in range(bstart[b], bstart[b + 1]):
                k = order[s]
                if k < P - 1 and k != i and h[k] == 0:
                    h[k] = st["cluster"]
                    h[P + k] = k
                    found += 1
            lp[b], rp[b] = b - 1, b + 1
            return found

        def run_pass(i):
            b0 = bucket_of[i]
            if expanded[b0]:
                return
            expanded[b0] = True
            found = 0
            if lp[b0] == b0 and qualifies(b0, y[i]):
                found += assign(b0, i)
            b = find(lp, b0 - 1, -1)
            while b >= 0 and qualifies(b, y[i]):
                found += assign(b, i)
                b = find(lp, b - 1, -1)
            b = find(rp, b0 + 1, nb)
            while b < nb and qualifies(b, y[i]):
                found += assign(b, i)
                b = find(rp, b + 1, nb)
            st["num"] += found

        i = 0
        while i < P:
            if st["num"] >= 1:
                m = st["num"]
                vals = [h[P + m - 1 - q] for q in range(min(batch, m))]
                eff = [not expanded[bucket_of[v]] for v in vals]
                q0 = eff.index(True) if True in eff else len(vals)
                for r in range(q0):
                    h[P + m - 1 - r] = 0
                if q0 > 0:
                    st["num"] -= q0
                    i = vals[q0 - 1] + 1
                    continue
                i = vals[0]
                h[P + m - 1] = 0
                st["num"] -= 1
                run_pass(i)
            else:
                while i < P and h[i] != 0:
                    i += 1
                if i >= P:
                    break
                st["cluster"] += 1
                h[i] = st["cluster"]
                run_pass(i)
            i += 1
        return st["cluster"] - 1

    rng = np.random.default_rng(11)
    choices = [lambda P: rng.integers(1, 5, P + 1), lambda P: rng.integers(60, 80, P + 1),
               lambda P: rng.choice([10, 11, 30, 31, 32], P + 1), lambda P: rng.integers(1, 200, P + 1),
               lambda P: rng.choice([5, 6, 7, 50, 51], P + 1)]
    for trial in range(1500):
        P = int(rng.integers(1, 90))
        t = np.cumsum(choices[trial % 5](P))
        dt = [int(t[k + 1]) - int(t[k]) for k in range(P)]
        hp = (0.01, 0.02, 0.005)[trial % 3]
        eps = (0.001, 0.01, 0.05, 0.2, 0.011, 0.0)[(trial // 3) % 6]
        want = oracle.dbscan_ref(np.array([float(d) * hp for d in dt]), eps)
        for batch in (1, 64):
            assert bucketed(dt, hp, eps, batch) == want, (trial, P, eps, hp, batch)

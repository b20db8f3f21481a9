"""Pins for the oracle's basin features (orc_basins_features).

REFERENCE semantics restate avgPeakFinderCUDA (Basins/src/basinsCUDA.cu:818-864)
after the Basins record loop (:318-380): flag 0 -> (999, 999); flag -1 ->
(last recorded sample, -1); no kept peak -> (1000, 1000); else the mean kept
peak and the mean interval, intervals in units of h * preScaller (:684).
CONFIG semantics: (mean X[0], mean X[NX-1] over the B recorded samples,
population std of the kept X[writableVar] peaks, 0 below two kept peaks).

Nothing here re-types the oracle's loops.  The expected values come from
  * closed forms: the Lorenz equilibria C+- = (+-sqrt(b(r-1)), ..., r-1);
  * the Lorenz mirror map (x, y, z) -> (-x, -y, z), which RK4 commutes with
    bit-exactly (tests/test_oracle_systems.py);
  * an independent evaluation in Python: the trajectory advanced one sample
    at a time with the pinned RK4 step, peaks found by a strict-maximum rule
    (valid because the test asserts the signals hold no near-plateau), sums
    by math.fsum, the std by numpy's two-pass np.std, the mean interval both
    as an fsum and telescoped from the first and last kept peak times.
Tolerances are forward-error bounds of recursive summation, so a mean over
B-1 instead of B samples, a std over all peaks instead of the kept ones, the
first instead of the last sample in the fixed-point sentinel, or h instead of
h * preScaller in the intervals each fail at least one test.
"""
import math

import numpy as np
import pytest

import _oracle as O

LOR = O.SYS["lorenz_rk4"]
ROS = O.SYS["rossler_rk4"]
U = 2.0 ** -53
H = 0.01
SIG, BETA = 10.0, 8.0 / 3.0
FAR = 1e300  # maxValue that never triggers


def features(sem, x0, y0, z0, rho, tT, tMax, pre=1, wv=0, maxv=FAR, sysid=LOR, values=None):
    """Oracle features of the single IC (x0, y0, z0): point 0 of a 2x2 grid."""
    values = [SIG, rho, BETA] if values is None else values
    f0, f1, f2, flags = O.basins_features(sysid, sem, tMax, 2, H, [0.0, 0.0, z0],
                                          [x0, x0 + 1.0, y0, y0 + 1.0], [0, 1], wv, maxv, tT,
                                          values, pre, 0, 1)
    return float(f0[0]), float(f1[0]), float(f2[0]), int(flags[0])


def record(x0, y0, z0, rho, S, B, pre, sysid=LOR, values=None):
    """Trajectory samples (the state before each sample's pre steps) and the
    state after the last sample, advanced with the pinned RK4 step."""
    a = [SIG, rho, BETA] if values is None else values
    x = O.steps(sysid, [x0, y0, z0], a, H, S)
    rec = np.empty((B, 3))
    for s in range(B):
        rec[s] = x
        x = O.steps(sysid, x, a, H, pre)
    return rec, x


def strict_peaks(d):
    """Indices of strict local maxima (rise and fall both above 1e-13): the
    reference's plateau-aware rule reduces to this when no neighbour pair is
    within 1e-13, which the caller asserts."""
    dd = np.diff(d)
    assert np.all(np.abs(dd) > 1e-13), "signal holds a near-plateau"
    return np.nonzero((dd[:-1] > 1e-13) & (-dd[1:] > 1e-13))[0] + 1


def mean_tol(x):
    """Bound on |recursive sum / n - exact mean| (n-1 roundings + 1 divide)."""
    x = np.asarray(x, dtype=float)
    n = x.size
    return (n + 1) * U * math.fsum(np.abs(x)) / n + 1e-300


def std_tol(x):
    """Bound on |sqrt(sum fma(x,x)/n - m^2) - two-pass std| for the one-pass
    form: the variance error (n+3) u (E[x^2] + m^2), halved through sqrt."""
    x = np.asarray(x, dtype=float)
    n = x.size
    e2 = math.fsum(x * x) / n
    s = np.std(x)
    return 4 * (n + 3) * U * e2 / max(s, 1e-300) + 4 * U * s


# ----------------------------------------------------------------------------
# sentinels
# ----------------------------------------------------------------------------

def test_divergent_sentinels():
    # chaotic Lorenz leaves |x| <= 5 within the record: flag 0
    ref = features(O.SEM_REFERENCE, 1.0, 1.0, 1.0, 28.0, 5.0, 20.0, maxv=5.0)
    cfg = features(O.SEM_CONFIG, 1.0, 1.0, 1.0, 28.0, 5.0, 20.0, maxv=5.0)
    assert ref == (999.0, 999.0, 0.0, 0)
    assert cfg == (0.0, 0.0, 0.0, 0)
    # overflow to inf (REFERENCE: isnan || isinf, basinsCUDA.cu:350)
    ref = features(O.SEM_REFERENCE, 1e200, 1e200, 1e200, 28.0, 0.0, 1.0, maxv=0.0)
    assert ref == (999.0, 999.0, 0.0, 0)


def test_peakless_record_sentinel():
    # rho = 0.5: the origin is a stable node (real eigenvalues -0.47, -10.5,
    # -8/3), so x has at most one maximum: no kept peak -> (1000, 1000)
    ref = features(O.SEM_REFERENCE, 5.0, 3.0, 4.0, 0.5, 0.0, 10.0)
    assert ref == (1000.0, 1000.0, 0.0, 1)
    # CONFIG: means of the record, std 0
    cfg = features(O.SEM_CONFIG, 5.0, 3.0, 4.0, 0.5, 0.0, 10.0)
    rec, _ = record(5.0, 3.0, 4.0, 0.5, 0, 1000, 1)
    assert cfg[3] == 1 and cfg[2] == 0.0
    assert abs(cfg[0] - math.fsum(rec[:, 0]) / 1000) <= mean_tol(rec[:, 0])
    assert abs(cfg[1] - math.fsum(rec[:, 2]) / 1000) <= mean_tol(rec[:, 2])


def test_one_peak_record_keeps_nothing():
    # a single maximum is dropped with the first peak (peakFinder keeps P-1)
    rec, _ = record(5.0, 3.0, 4.0, 0.5, 0, 1000, 1)
    assert len(strict_peaks(rec[:, 1])) <= 1
    ref = features(O.SEM_REFERENCE, 5.0, 3.0, 4.0, 0.5, 0.0, 10.0, wv=1)
    assert ref[:2] == (1000.0, 1000.0)


# ----------------------------------------------------------------------------
# equilibria: closed forms
# ----------------------------------------------------------------------------

RHO_EQ = 10.0
XS = math.sqrt(BETA * (RHO_EQ - 1.0))
ZS = RHO_EQ - 1.0


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_equilibrium_reference_takes_last_sample(sign):
    # start 1e-2 off C+- with no transient: the spiral (Re lambda ~ -0.59)
    # has settled to the floating-point equilibrium long before the end of
    # the 60 t.u. record, so the last step moves < 1e-9 (flag -1) and the
    # feature is the LAST sample, x*, while the first sample is x* + 1e-2
    x0 = sign * (XS + 1e-2)
    f0, f1, f2, flag = features(O.SEM_REFERENCE, x0, sign * XS, ZS, RHO_EQ, 0.0, 60.0)
    assert flag == -1
    assert f1 == -1.0
    assert abs(f0 - sign * XS) <= 1e-12
    assert abs(f0 - x0) > 1e-3


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_equilibrium_config_closed_form(sign):
    # after a 100 t.u. transient the whole record sits on C+-: (x*, z*, 0)
    f0, f1, f2, flag = features(O.SEM_CONFIG, sign * (XS + 1e-2), sign * XS, ZS, RHO_EQ, 100.0, 20.0)
    assert flag == -1
    assert abs(f0 - sign * XS) <= 1e-12
    assert abs(f1 - ZS) <= 1e-12
    assert f2 == 0.0


def test_equilibrium_recorded_variable_picks_feature():
    # writableVar = 2 records z: the REFERENCE sentinel is z* (last sample)
    f0, f1, _, flag = features(O.SEM_REFERENCE, XS + 1e-2, XS, ZS, RHO_EQ, 0.0, 60.0, wv=2)
    assert flag == -1 and f1 == -1.0 and abs(f0 - ZS) <= 1e-12


# ----------------------------------------------------------------------------
# Lorenz mirror map
# ----------------------------------------------------------------------------

MIRROR_ICS = [(-7.5, 3.25, 25.0), (12.0, -9.0, 25.0), (0.5, 0.75, 20.0)]


@pytest.mark.parametrize("x0,y0,z0", MIRROR_ICS)
@pytest.mark.parametrize("pre", [1, 3])
def test_mirror_map_recording_z(x0, y0, z0, pre):
    # z is invariant under the mirror: every feature is identical
    for sem in (O.SEM_REFERENCE, O.SEM_CONFIG):
        a = features(sem, x0, y0, z0, 24.5, 10.0, 30.0, pre=pre, wv=2)
        b = features(sem, -x0, -y0, z0, 24.5, 10.0, 30.0, pre=pre, wv=2)
        if sem == O.SEM_CONFIG:
            assert (b[0], b[1], b[2], b[3]) == (-a[0], a[1], a[2], a[3])
        else:
            assert a == b


@pytest.mark.parametrize("x0,y0,z0", MIRROR_ICS)
def test_mirror_map_recording_x(x0, y0, z0):
    # recording x: mean x flips sign exactly, mean z is unchanged, and the
    # mirrored std is the std of the original's troughs (peaks of -x)
    a = features(O.SEM_CONFIG, x0, y0, z0, 24.5, 10.0, 30.0, wv=0)
    b = features(O.SEM_CONFIG, -x0, -y0, z0, 24.5, 10.0, 30.0, wv=0)
    assert b[0] == -a[0] and b[1] == a[1] and b[3] == a[3]
    rec, _ = record(x0, y0, z0, 24.5, 1000, 3000, 1)
    troughs = -rec[strict_peaks(-rec[:, 0]), 0]
    kept = -troughs[1:]
    want = float(np.std(kept)) if kept.size >= 2 else 0.0
    assert abs(b[2] - want) <= std_tol(kept)


# ----------------------------------------------------------------------------
# independent evaluation on chaotic records
# ----------------------------------------------------------------------------

CHAOS_CASES = [
    # (sysid, values, ic, tT, tMax, pre, wv)
    (LOR, [SIG, 28.0, BETA], (1.0, 1.0, 1.0), 10.0, 30.0, 1, 0),
    (LOR, [SIG, 28.0, BETA], (-3.0, 4.0, 22.0), 5.0, 40.0, 2, 0),
    (LOR, [SIG, 24.5, BETA], (-12.0, 7.0, 25.0), 20.0, 30.0, 3, 2),
    (ROS, [0.2, 0.2, 5.7], (0.1, 0.1, 0.1), 50.0, 200.0, 2, 1),
]


def _case(sysid, values, ic, tT, tMax, pre, wv):
    S = int(round(tT / H))
    B = int(round(tMax / H / pre))
    rec, _ = record(ic[0], ic[1], ic[2], None, S, B, pre, sysid=sysid, values=values)
    idx = strict_peaks(rec[:, wv])
    assert idx.size >= 6, "case must hold several peaks"
    return rec, idx, B


@pytest.mark.parametrize("case", CHAOS_CASES)
def test_reference_features_independent(case):
    sysid, values, ic, tT, tMax, pre, wv = case
    rec, idx, B = _case(*case)
    got = features(O.SEM_REFERENCE, ic[0], ic[1], ic[2], None, tT, tMax, pre=pre, wv=wv,
                   sysid=sysid, values=values)
    assert got[3] == 1
    vals = rec[idx, wv]
    kept = vals[1:]  # the first peak is dropped
    P = kept.size
    dts = np.diff(idx).astype(float) * (H * pre)
    assert abs(got[0] - math.fsum(kept) / P) <= mean_tol(kept)
    assert abs(got[1] - math.fsum(dts) / P) <= mean_tol(dts)
    tele = (idx[-1] - idx[0]) * H * pre / P
    assert abs(got[1] - tele) <= mean_tol(dts) + 4 * U * tele
    # a mean interval in units of h instead of h * pre differs by far more
    if pre > 1:
        assert abs(got[1] - tele / pre) > 1e3 * mean_tol(dts)


@pytest.mark.parametrize("case", CHAOS_CASES)
def test_config_features_independent(case):
    sysid, values, ic, tT, tMax, pre, wv = case
    rec, idx, B = _case(*case)
    got = features(O.SEM_CONFIG, ic[0], ic[1], ic[2], None, tT, tMax, pre=pre, wv=wv,
                   sysid=sysid, values=values)
    assert got[3] == 1
    assert abs(got[0] - math.fsum(rec[:, 0]) / B) <= mean_tol(rec[:, 0])
    assert abs(got[1] - math.fsum(rec[:, 2]) / B) <= mean_tol(rec[:, 2])
    kept = rec[idx[1:], wv]
    assert abs(got[2] - float(np.std(kept))) <= std_tol(kept)
    # the plausible slips are all far outside the bounds
    assert abs(got[0] - math.fsum(rec[:-1, 0]) / (B - 1)) > 10 * mean_tol(rec[:, 0])
    assert abs(got[0] - math.fsum(rec[:, 0]) / (B - 1)) > 10 * mean_tol(rec[:, 0])
    assert abs(got[2] - float(np.std(rec[idx, wv]))) > 10 * std_tol(kept)
    assert abs(got[2] - float(np.std(kept, ddof=1))) > 10 * std_tol(kept)


def test_reference_interval_counts_prescaller():
    # the same trajectory sampled every step and every second step: the
    # strict peaks of the coarse record sit at even fine indices when the
    # maxima fall there, so compare the mean interval against its own
    # telescoped form instead: pre = 2 must give time units h * 2
    case = CHAOS_CASES[1]
    sysid, values, ic, tT, tMax, pre, wv = case
    rec, idx, B = _case(*case)
    got = features(O.SEM_REFERENCE, ic[0], ic[1], ic[2], None, tT, tMax, pre=pre, wv=wv,
                   sysid=sysid, values=values)
    P = idx.size - 1
    assert abs(got[1] - (idx[-1] - idx[0]) * (H * pre) / P) <= 1e-12 * got[1]
    assert abs(got[1] - (idx[-1] - idx[0]) * H / P) > 0.4 * got[1]

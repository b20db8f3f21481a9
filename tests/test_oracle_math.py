"""Pins for the oracle's portable transcendentals (orc_sin, orc_cos,
orc_exp, orc_atan in oracle/chaos_oracle.c), which the reference systems'
steps use on both the host and the device so that a step gives the same
bits on both.  Checked against the C library (numpy's float64 ufuncs) in
units in the last place, over the ranges the systems meet (Chameleon's
cos(18 x), BASINS' sin of bounded states, the DAMIR circuit's exp and atan
arguments) and beyond; special values map as the library's do."""
import math

import numpy as np
import pytest

import _oracle as O


def ulps(got, want):
    got, want = np.asarray(got), np.asarray(want)
    return np.abs(got - want) / np.spacing(np.maximum(np.abs(want), np.finfo(float).tiny))


def _args(lo, hi, n, seed):
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.uniform(lo, hi, n), np.linspace(lo, hi, n)])


@pytest.mark.parametrize("fn,ref", [("orc_sin", np.sin), ("orc_cos", np.cos)])
@pytest.mark.parametrize("lo,hi", [(-1e-3, 1e-3), (-4.0, 4.0), (-400.0, 400.0), (-2e5, 2e5)])
def test_trig_within_two_ulp(fn, ref, lo, hi):
    x = _args(lo, hi, 20000, 1)
    f = getattr(O.lib, fn)
    got = np.array([f(v) for v in x])
    assert ulps(got, ref(x)).max() <= 2.0


def test_trig_special_values():
    for f in (O.lib.orc_sin, O.lib.orc_cos):
        assert math.isnan(f(math.inf)) and math.isnan(f(-math.inf)) and math.isnan(f(math.nan))
    assert O.lib.orc_sin(0.0) == 0.0 and O.lib.orc_cos(0.0) == 1.0
    assert O.lib.orc_sin(-0.0) == 0.0
    for k in range(-50, 50):
        x = k * math.pi / 2
        assert abs(O.lib.orc_sin(x) - math.sin(x)) <= 1e-15 * max(1, abs(k))
        assert abs(O.lib.orc_cos(x) - math.cos(x)) <= 1e-15 * max(1, abs(k))


@pytest.mark.parametrize("lo,hi", [(-1.0, 1.0), (-50.0, 50.0), (-700.0, 700.0), (-745.0, -700.0)])
def test_exp_within_two_ulp(lo, hi):
    x = _args(lo, hi, 20000, 2)
    got = np.array([O.lib.orc_exp(v) for v in x])
    want = np.exp(x)
    normal = want >= np.finfo(float).tiny
    assert ulps(got[normal], want[normal]).max() <= 2.0
    # subnormal results: within one subnormal step
    assert np.abs(got[~normal] - want[~normal]).max() <= 2 * 5e-324 if (~normal).any() else True


def test_exp_special_values():
    assert O.lib.orc_exp(0.0) == 1.0
    assert O.lib.orc_exp(710.0) == math.inf and O.lib.orc_exp(1e300) == math.inf
    assert O.lib.orc_exp(-800.0) == 0.0 and O.lib.orc_exp(-math.inf) == 0.0
    assert math.isnan(O.lib.orc_exp(math.nan))
    assert O.lib.orc_exp(1.0) == pytest.approx(math.e, rel=3e-16)


@pytest.mark.parametrize("lo,hi", [(-1.0, 1.0), (-30.0, 30.0), (-1e6, 1e6)])
def test_atan_within_four_ulp(lo, hi):
    x = _args(lo, hi, 20000, 3)
    got = np.array([O.lib.orc_atan(v) for v in x])
    want = np.arctan(x)
    nz = np.abs(want) > 1e-300
    assert ulps(got[nz], want[nz]).max() <= 4.0


def test_atan_special_values():
    assert O.lib.orc_atan(0.0) == 0.0
    assert O.lib.orc_atan(math.inf) == pytest.approx(math.pi / 2, rel=2e-16)
    assert O.lib.orc_atan(-math.inf) == pytest.approx(-math.pi / 2, rel=2e-16)
    assert O.lib.orc_atan(1.0) == pytest.approx(math.pi / 4, rel=3e-16)
    assert math.isnan(O.lib.orc_atan(math.nan))

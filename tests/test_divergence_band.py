"""The sweep kernels' divergence fast path against the oracle's exact test.

A live sample v with |v| < L (L = divergence_fast_bound(check, maxValue),
chaoslibrary_amd/csrc/sweep.hip) skips the exact test on the device, so the
band is sound only if no such v is diverged under the reference's checks
(Bifurcation/src/bifurcationCUDA.cu:445-453: v != v || v == v + 1, then
(double)fabsf((float)v) > maxValue; Basins/src/basinsCUDA.cu:350-361:
isnan || isinf, then the same float-cast bound).  The float cast sends every
|v| above FLT_MAX (plus half an ulp) to inf, so for maxValue beyond FLT_MAX
the band must stop at FLT_MAX.  CPU only: the bound is host code of the
library, the exact test is the oracle's.
"""
import ctypes as C

import numpy as np
import pytest

import _oracle as O

FLT_MAX = float(np.finfo(np.float32).max)


@pytest.fixture(scope="module")
def band():
    import chaoslibrary_amd as ch
    f = getattr(ch.lib, "_ZN5chaos21divergence_fast_boundEid")
    f.restype = C.c_double
    f.argtypes = [C.c_int, C.c_double]
    return f


def _probes(L):
    """Values just inside the band: below L, down to a few ulps, both signs."""
    out = []
    if np.isfinite(L) and L > 0:
        v = L
        for _ in range(64):
            v = np.nextafter(v, 0.0)
            out += [v, -v]
        out += [L * 0.999, L * 0.5, -L * 0.5]
    out += [0.0, 1.0, -1.0, 1e10, 3.0e38, -3.0e38]
    return [x for x in out if abs(x) < L]


MAXVALUES = [0.0, 1.0, 1e3, 1e38, float(np.float32(1e38)), FLT_MAX, np.nextafter(FLT_MAX, np.inf),
             FLT_MAX * (1 + 2 ** -23), 3.5e38, 1e39, 1e300, float("inf"), -5.0]


@pytest.mark.parametrize("check", [O.CHK_BIF_REF, O.CHK_BASINS_REF, O.CHK_CONFIG])
@pytest.mark.parametrize("maxv", MAXVALUES)
def test_band_is_sound(band, check, maxv):
    L = band(check, maxv)
    for v in _probes(L):
        assert O.lib.orc_diverged(v, maxv, check) == 0, (check, maxv, v, L)


@pytest.mark.parametrize("check", [O.CHK_BIF_REF, O.CHK_BASINS_REF])
def test_float_cast_band_stops_at_flt_max(band, check):
    # maxValue far above FLT_MAX: a sample between FLT_MAX and the maxValue
    # band is diverged for the reference (inf > maxValue), so it must not be
    # inside the band
    maxv = 1e300
    L = band(check, maxv)
    assert L <= FLT_MAX
    for v in (FLT_MAX * (1 + 2 ** -23), 3.5e38, 1e39, 1e200):
        assert O.lib.orc_diverged(v, maxv, check) == 1
        assert not abs(v) < L


def test_config_band_is_maxvalue(band):
    assert band(O.CHK_CONFIG, 1e300) == 1e300
    assert band(O.CHK_CONFIG, 0.0) == float("inf")
    assert band(O.CHK_BIF_REF, 0.0) == 2.0 ** 53

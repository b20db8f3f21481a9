"""CPU checks of the LLE lane program the engine's persistent kernels run
(LleProg in chaoslibrary_amd/csrc/lanes.hpp, built for the host from the same
source by tests/native/lle_ticks_host.cpp).

* Tick-split invariance: running a point's ticks in arbitrary segments, with
  the lane state parked and reloaded in between (as the engine does between
  launches), gives the same exponent bit-for-bit as one unsplit run, with the
  perturbation generated in the loop (one-shot kernels) or stored with the
  job beforehand (engine), and equal to the oracle's orc_lle_point.
* Progress guard: the window loop advances min(NTp - local, left) ticks per
  trip, which is 0 (and negative after it) when NTp = 0 -- the value a
  kernel reads from a job-table slot whose LLE block was never filled in
  (a sweep job, a slot not yet uploaded).  Unguarded, that loop never ends
  on the device.  The guard refuses such ranges (and reversed ones) up
  front, and the engine reports the refusal through its error word.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import _oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "lle_ticks_host.cpp")
ROS, LOR = O.SYS["rossler_rk4"], O.SYS["lorenz_rk4"]
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)
_L = C.POINTER(C.c_longlong)


@pytest.fixture(scope="module")
def host(tmp_path_factory):
    hipcc = "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    so = str(tmp_path_factory.mktemp("lleticks") / "lle_ticks_host.so")
    subprocess.run([hipcc, "-x", "hip", "--cuda-host-only", "-std=c++17", "-O2", "-fPIC", "-shared",
                    "-ffp-contract=off", "-I", os.path.join(ROOT, "chaoslibrary_amd", "csrc"),
                    "-I", os.path.join(ROOT, "include"), SRC, "-o", so], check=True, capture_output=True)
    lib = C.CDLL(so)
    lib.host_lle_split.restype = C.c_double
    lib.host_lle_split.argtypes = [C.c_int, _D, _D, _D, _I, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                   C.c_double, C.c_double, C.c_uint64, C.c_int, C.c_int, _L, C.c_int, C.c_int, _I]
    lib.host_lle_guard.restype = C.c_int
    lib.host_lle_guard.argtypes = [C.c_int, C.c_int, C.c_longlong, C.c_longlong, _L]
    return lib


CASES = [
    # sys, values, ic, ranges, idx, nPts, i, j, S, NTp, W
    (ROS, [0.2, 0.2, 5.7], [0.1, 0.1, 0.1], [0.05, 0.35, 2.0, 10.0], [0, 2], 16, 5, 11, 500, 50, 20),
    (LOR, [10.0, 28.0, 8 / 3], [1.0, 1.0, 1.0], [20.0, 200.0, 0.5, 4.0], [1, 2], 8, 3, 6, 300, 25, 17),
    (ROS, [0.2, 0.2, 5.7], [0.1, 0.1, 0.1], [0.05, 0.35, 2.0, 10.0], [0, 2], 4, 0, 3, 0, 7, 9),
]


def _run(host, case, cuts, dir_stored):
    sysid, values, ic, ranges, idx, nPts, i, j, S, NTp, W = case
    v = np.asarray(values, float)
    x = np.asarray(ic, float)
    r = np.asarray(ranges, float)
    ix = np.asarray(idx, np.int32)
    c = np.asarray(cuts, np.int64)
    st = C.c_int(7)
    out = host.host_lle_split(sysid, O.dptr(v), O.dptr(x), O.dptr(r), O.iptr(ix), nPts, S, NTp, W, 1e-8,
                              W * NTp * 0.01, 0.01, 12345, i, j, c.ctypes.data_as(_L), c.size, dir_stored,
                              C.byref(st))
    return out, st.value


@pytest.mark.parametrize("case", CASES)
def test_split_runs_equal_unsplit_run(host, case):
    S, NTp, W = case[8], case[9], case[10]
    total = S + W * NTp
    want, st = _run(host, case, [], 0)
    assert st == 0 and np.isfinite(want)
    rng = np.random.default_rng(11)
    splits = [[1], [S] if S else [1], [S + 1], [S + NTp], [S + NTp - 1, S + NTp + 1], [total - 1],
              list(range(1, total, 3)), list(range(1, total, NTp))]
    for _ in range(12):
        k = int(rng.integers(1, 20))
        splits.append(sorted(set(int(x) for x in rng.integers(1, total, k))))
    for cuts in splits:
        cuts = [c for c in cuts if 0 < c < total]
        for dir_stored in (0, 1):
            got, st = _run(host, case, cuts, dir_stored)
            assert st == 0
            assert got == want, (cuts, dir_stored, got, want)


@pytest.mark.parametrize("case", CASES)
def test_host_program_equals_oracle(host, case):
    sysid, values, ic, ranges, idx, nPts, i, j, S, NTp, W = case
    a = np.asarray(values, float).copy()
    a[idx[0]] = O.lib.orc_grid_value(i, nPts, ranges[0], ranges[1])
    a[idx[1]] = O.lib.orc_grid_value(j, nPts, ranges[2], ranges[3])
    tMax = W * NTp * 0.01
    want = O.lle_point(sysid, O.SEM_CONFIG, tMax, NTp * 0.01, 0.01, 1e-8, ic, a, S * 0.01, 12345,
                       j * nPts + i)
    got, st = _run(host, case, [], 0)
    assert st == 0 and got == want


@pytest.mark.timeout(30)
@pytest.mark.parametrize("NTp,S,g0,g1", [(0, 0, 0, 100), (0, 0, 5, 9), (0, 50, 50, 80), (-3, 0, 0, 10)])
def test_zero_lle_block_is_refused(host, NTp, S, g0, g1):
    trip = C.c_longlong()
    assert host.host_lle_guard(NTp, S, g0, g1, C.byref(trip)) == 0
    # the loop the guard keeps the kernel out of would not advance
    assert trip.value <= 0


@pytest.mark.timeout(30)
def test_reversed_or_negative_range_is_refused(host):
    trip = C.c_longlong()
    assert host.host_lle_guard(50, 10, 40, 20, C.byref(trip)) == 0
    assert host.host_lle_guard(50, 10, -1, 20, C.byref(trip)) == 0
    assert host.host_lle_guard(50, -1, 0, 20, C.byref(trip)) == 0


def test_valid_range_advances(host):
    trip = C.c_longlong()
    assert host.host_lle_guard(50, 10, 0, 200, C.byref(trip)) == 1
    assert host.host_lle_guard(50, 10, 10, 11, C.byref(trip)) == 1
    assert host.host_lle_guard(50, 10, 7, 7, C.byref(trip)) == 1  # empty range: nothing to do

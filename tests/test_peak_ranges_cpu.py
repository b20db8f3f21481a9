"""CPU checks of how a peak-count sweep is cut into ranges of whole 64-point
waves by the device memory (chaoslibrary_amd/csrc/peak_ranges.hpp, used by
the 2-D bifurcation driver), built for the host by g++ from the same header.

exact_ranges: every range but the last ends where one more wave would
exceed the budget, every range fits it, the cuts are whole waves and cover
the grid, and a wave that alone exceeds the budget is reported.
first_pass_cuts: equal ranges of floor(budget / per-wave guess) waves, the
whole grid when that many cover it."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHUNK, CHUNK_BYTES = 256, 256 * 16 + 8  # the library's store: 256 entries of 16 B + 2 meta words


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    so = str(tmp_path_factory.mktemp("ranges") / "ranges_host.so")
    subprocess.run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-I", os.path.join(ROOT, "chaoslibrary_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "ranges_host.cpp"), "-o", so], check=True)
    L = C.CDLL(so)
    LL, ULL = C.c_longlong, C.c_ulonglong
    L.exact_ranges_host.restype = LL
    L.exact_ranges_host.argtypes = [C.POINTER(C.c_int), LL, ULL, C.c_int, ULL, C.POINTER(LL), LL, C.POINTER(ULL)]
    L.first_pass_cuts_host.restype = LL
    L.first_pass_cuts_host.argtypes = [LL, C.c_double, LL, C.POINTER(LL), LL]
    L.wave_pass_bytes_host.restype = ULL
    L.wave_pass_bytes_host.argtypes = [LL, LL, C.c_int, ULL]
    return L


def _exact(lib, counts, budget):
    counts = np.ascontiguousarray(counts, dtype=np.int32)
    n = len(counts)
    out = (C.c_longlong * (n // 64 + 3))()
    wb = C.c_ulonglong(0)
    k = lib.exact_ranges_host(counts.ctypes.data_as(C.POINTER(C.c_int)), n, budget, CHUNK, CHUNK_BYTES, out,
                              len(out), C.byref(wb))
    return (None, wb.value) if k < 0 else (list(out[:k]), None)


def _wave_bytes(lib, counts, w, n):
    e = min(n, w + 64)
    return lib.wave_pass_bytes_host(int(np.sum(counts[w:e], dtype=np.int64)), e - w, CHUNK, CHUNK_BYTES)


def test_wave_pass_bytes_closed_form(lib):
    # 0 peaks: one open chunk; 192 usable entries per chunk after a switch
    assert lib.wave_pass_bytes_host(0, 64, CHUNK, CHUNK_BYTES) == CHUNK_BYTES + 128 * 64
    assert lib.wave_pass_bytes_host(192, 64, CHUNK, CHUNK_BYTES) == 2 * CHUNK_BYTES + 32 * 192 + 128 * 64
    assert lib.wave_pass_bytes_host(193, 64, CHUNK, CHUNK_BYTES) == 3 * CHUNK_BYTES + 32 * 193 + 128 * 64
    assert lib.wave_pass_bytes_host(5, 7, CHUNK, CHUNK_BYTES) == 2 * CHUNK_BYTES + 32 * 5 + 128 * 7


@pytest.mark.parametrize("seed", range(6))
def test_exact_ranges_are_maximal_whole_waves_within_budget(lib, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 5000))
    counts = rng.integers(0, 1200, n)
    wb = [_wave_bytes(lib, counts, w, n) for w in range(0, n, 64)]
    budget = int(max(wb) * rng.uniform(1.0, 6.0))
    starts, _ = _exact(lib, counts, budget)
    assert starts[0] == 0 and starts[-1] == n
    assert all(a < b for a, b in zip(starts, starts[1:]))
    assert all(s % 64 == 0 for s in starts[:-1])
    for k in range(len(starts) - 1):
        a, b = starts[k], starts[k + 1]
        used = sum(wb[a // 64:(b + 63) // 64])
        assert used <= budget
        if k + 2 < len(starts):  # the next wave would not have fitted
            assert used + wb[b // 64] > budget


def test_exact_ranges_whole_grid_and_hand_case(lib):
    counts = np.full(640, 100)  # 10 waves of 6400 peaks each
    w = _wave_bytes(lib, counts, 0, 640)
    assert _exact(lib, counts, 10 * w)[0] == [0, 640]
    assert _exact(lib, counts, 3 * w)[0] == [0, 192, 384, 576, 640]
    assert _exact(lib, counts, 3 * w - 1)[0] == [0, 128, 256, 384, 512, 640]


def test_exact_ranges_reports_a_wave_over_budget(lib):
    counts = np.zeros(200, dtype=np.int64)
    counts[130] = 100000
    w = _wave_bytes(lib, counts, 128, 200)
    starts, need = _exact(lib, counts, w - 1)
    assert starts is None and need == w


def test_first_pass_cuts(lib):
    def cuts(n, guess, budget):
        out = (C.c_longlong * (n // 64 + 3))()
        k = lib.first_pass_cuts_host(n, guess, budget, out, len(out))
        return list(out[:k])
    assert cuts(1000, 10.0, 1000) == [0, 1000]             # 100 waves per range cover 16
    n = 65536 * 64  # the 2048^2 peak-count map
    c = cuts(n, 521.8, 11_627_000)  # a sixth of 286 GB in 4104-byte chunks, 1562.5 peaks per point guessed
    per = int(11_627_000 / 521.8)
    assert c == [0] + [64 * w for w in range(per, 65536, per)] + [n] and len(c) == 4
    assert cuts(130, 5.0, 1) == [0, 64, 128, 130]          # at least one wave per range

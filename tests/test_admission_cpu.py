"""CPU checks of the job server's admission rule (admit_fifo in
chaoslibrary_amd/csrc/admission.hpp, the function Engine::admit calls),
built for the host by g++ from the same header.

The rule: pending jobs enter in submission order while a slot is free, the
footprint fits the HBM budget beside what is still charged, and the tiles fit
the queue; the first job that does not fit stops admission (no overtaking).
A simulated serving loop on top of it (jobs finish after a fixed number of
rounds and release their charges) checks that every job is eventually
admitted, in submission order, and that no round ever exceeds the slot, HBM
or tile limits.  The round sizing of the same header (family_waves) is
checked here too.
"""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LL = C.POINTER(C.c_longlong)


@pytest.fixture(scope="module")
def native(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    so = str(tmp_path_factory.mktemp("admit") / "admit_host.so")
    subprocess.run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-I", os.path.join(ROOT, "chaoslibrary_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "admit_host.cpp"), "-o", so], check=True)
    lib = C.CDLL(so)
    lib.family_waves_host.restype = C.c_int
    lib.family_waves_host.argtypes = [C.c_longlong, C.c_int, C.c_longlong]
    return lib


@pytest.fixture(scope="module")
def admit(native):
    lib = native
    lib.admit_host.restype = C.c_int
    lib.admit_host.argtypes = [_LL, _LL, C.c_int, C.c_int, _LL, C.c_longlong, _LL, C.c_longlong]

    def run(fp, tl, free_slots, used, budget, active, max_tiles):
        n = len(fp)
        fp = (C.c_longlong * max(1, len(fp)))(*fp)
        tl = (C.c_longlong * max(1, len(tl)))(*tl)
        u, a = C.c_longlong(used), C.c_longlong(active)
        k = lib.admit_host(fp, tl, n, free_slots, C.byref(u), budget, C.byref(a), max_tiles)
        return k, u.value, a.value
    return run


def test_admits_fifo_prefix_within_budget(admit):
    k, used, active = admit([10, 20, 30, 5], [1, 2, 3, 1], 8, 0, 60, 0, 100)
    assert (k, used, active) == (3, 60, 6)


def test_head_of_line_blocks_smaller_followers(admit):
    # job 1 does not fit; job 2 would, but may not overtake it
    k, used, active = admit([10, 100, 5], [1, 1, 1], 8, 40, 100, 0, 100)
    assert (k, used, active) == (1, 50, 1)


def test_slots_limit(admit):
    k, used, active = admit([1] * 10, [1] * 10, 4, 0, 1000, 0, 1000)
    assert (k, used, active) == (4, 4, 4)


def test_tile_limit(admit):
    k, used, active = admit([1, 1, 1], [40, 40, 40], 8, 0, 1000, 30, 100)
    assert (k, used, active) == (1, 1, 70)


def test_nothing_pending_or_no_slot(admit):
    assert admit([], [], 8, 5, 10, 3, 10) == (0, 5, 3)
    assert admit([1], [1], 0, 5, 10, 3, 10) == (0, 5, 3)


def test_exact_fit_is_admitted(admit):
    assert admit([40], [7], 1, 60, 100, 3, 10) == (1, 100, 10)


def test_serving_loop_admits_every_job_in_order(admit):
    """A loop over rounds: admit, run every active job one round, retire
    finished jobs (their footprint released one round later, as the worker
    finishes them).  Every job is admitted exactly once, in order, and the
    limits hold in every round."""
    rng = np.random.default_rng(7)
    n = 400
    fp = rng.integers(1, 60, n).tolist()
    tl = rng.integers(1, 20, n).tolist()
    rounds_needed = rng.integers(1, 6, n).tolist()
    slots, budget, max_tiles = 16, 200, 64
    pending = list(range(n))
    active = {}      # job -> rounds left
    finishing = []   # jobs whose footprint is still charged
    used = act_tiles = 0
    order = []
    for _ in range(10000):
        if not pending and not active and not finishing:
            break
        free = slots - len(active)
        k, used, act_tiles = admit([fp[j] for j in pending], [tl[j] for j in pending], free, used, budget,
                                   act_tiles, max_tiles)
        for j in pending[:k]:
            active[j] = rounds_needed[j]
            order.append(j)
        pending = pending[k:]
        assert len(active) <= slots and used <= budget and act_tiles <= max_tiles
        assert used == sum(fp[j] for j in list(active) + finishing)
        for j in finishing:  # the worker releases last round's finished jobs
            used -= fp[j]
        finishing = []
        for j in list(active):
            active[j] -= 1
            if active[j] == 0:
                del active[j]
                act_tiles -= tl[j]  # slot and tiles free at once, HBM after the worker
                finishing.append(j)
    assert order == list(range(n))
    assert used == 0 and act_tiles == 0


def test_family_waves_round_sizing(native):
    """family_waves (admission.hpp), the wavefronts of a family kernel that
    take part in a round: no tiles -> 0; otherwise one per claim of `take`
    tiles, at most the grid's wavefronts."""
    fw = native.family_waves_host
    assert fw(0, 1, 4096) == 0
    assert fw(1, 1, 4096) == 1
    assert fw(1407, 1, 4096) == 1407   # a 300x300 job: 1407 64-point groups
    assert fw(1407, 2, 4096) == 704    # 128-point claims
    assert fw(1407, 4, 4096) == 352
    assert fw(4096, 1, 4096) == 4096
    assert fw(5625, 1, 4096) == 4096   # a 600x600 job: the whole grid
    for tiles in [1, 2, 63, 64, 65, 1000, 4095, 4096, 4097, 8191, 8192, 8193, 12289, 10 ** 6]:
        for take in [1, 2, 4]:
            for g in [1, 7, 64, 1024, 4096, 8192]:
                w = fw(tiles, take, g)
                claims = -(-tiles // take)
                assert w == min(claims, g) and w >= 1

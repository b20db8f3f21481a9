"""CPU checks of the served-stream loop (chaoslibrary_amd/csrc/serve_loop.hpp,
what chaos_stream_run runs) over a stand-in engine, built for the host by
g++ from the same header (tests/native/serve_host.cpp).

The stand-in completes a job on its "finishing thread" between a round and
the loop's next idle check.  The loop must count such completions before it
waits: without that, the stream's last job stayed uncounted and the loop
waited for it for ever (a GPU stream stopped at 1974 of 1975 jobs done).
The same stand-in under the loop without that count does not end, so the
stand-in reproduces the interleaving."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    so = str(tmp_path_factory.mktemp("serve") / "serve_host.so")
    subprocess.run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-I", os.path.join(ROOT, "chaoslibrary_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "serve_host.cpp"), "-o", so], check=True)
    L = C.CDLL(so)
    D, I = C.POINTER(C.c_double), C.POINTER(C.c_int)
    L.serve_stand_in.restype = C.c_int
    L.serve_stand_in.argtypes = [C.c_int, D, I, C.c_int, D, C.POINTER(C.c_longlong)]
    L.serve_stand_in_without_idle_count.restype = C.c_int
    L.serve_stand_in_without_idle_count.argtypes = [C.c_int, D, I]
    return L


def _serve(lib, arrival, rounds, fail_job=-1):
    arrival = np.ascontiguousarray(arrival, dtype=np.float64)
    rounds = np.ascontiguousarray(rounds, dtype=np.int32)
    done = np.zeros(len(arrival))
    rel = C.c_longlong(0)
    rc = lib.serve_stand_in(len(arrival), arrival.ctypes.data_as(C.POINTER(C.c_double)),
                            rounds.ctypes.data_as(C.POINTER(C.c_int)), fail_job,
                            done.ctypes.data_as(C.POINTER(C.c_double)), C.byref(rel))
    return rc, done, rel.value


@pytest.mark.parametrize("seed", range(5))
def test_every_job_is_served_and_counted(lib, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 400))
    arrival = np.cumsum(rng.exponential(5.0, n))
    rounds = rng.integers(1, 6, n)
    rc, done, released = _serve(lib, arrival, rounds)
    assert rc == 0 and released == n
    assert np.all(done >= arrival)  # a job is done only after it arrived


def test_last_job_completed_before_the_idle_check(lib):
    # one job: its round hands it to the finishing thread, which completes
    # it before the loop's next idle check
    rc, done, released = _serve(lib, [0.5], [1])
    assert rc == 0 and released == 1 and done[0] > 0.5
    bad = lib.serve_stand_in_without_idle_count(
        1, np.array([0.5]).ctypes.data_as(C.POINTER(C.c_double)), np.array([1], dtype=np.int32).ctypes.data_as(C.POINTER(C.c_int)))
    assert bad == 2


def test_a_failed_job_reaches_on_fail(lib):
    rc, _, _ = _serve(lib, [1.0, 2.0, 3.0], [1, 2, 1], fail_job=1)
    assert rc == 1

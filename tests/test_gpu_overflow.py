"""The peak store's overflow path (chaoslibrary_amd/csrc/bifurcation.hip:
first guess max(256, B/32) peaks per point, overflow flagged on the device,
sweep rerun with the exact per-wave need; the engine reruns the job one-shot).

Lorenz RK4 at h = 0.05 peaks about once every 15 samples, far above the
first guess of one per 32, so the first sweep must overflow.  Results are
compared with the oracle (bit-exact) and, for the engine, with the one-shot
entry point; the engine reports the rerun in chaos_job_info.reran."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LOR = dict(ic=[1.0, 1.0, 1.0], values=[10.0, 28.0, 8 / 3])
H = 0.05
T_MAX, T_T = 600.0, 10.0
B = int(T_MAX / H)


def _first_guess_overflows(stored, n):
    """The first pool holds ceil(n * max_pp / 192) + waves chunks of 256
    entries and every chunk a wave fills holds at most 256 of its peaks, so
    more than 256 * chunks stored peaks cannot have fitted."""
    max_pp = min(B / 2 + 1, max(256.0, B / 32))
    chunks = int(np.ceil(n * max_pp / 192)) + (n + 63) // 64
    return stored > 256 * chunks


def _same_bif1d(got, want):
    """counts, peaks and intervals everywhere; the first sample where it is
    defined (fixed points, counts == -1)."""
    (c, f, pk, iv), (oc, of, opk, oiv) = got, want
    assert np.array_equal(c, oc)
    assert np.array_equal(f[c == -1], of[c == -1])
    assert np.array_equal(pk, opk) and np.array_equal(iv, oiv)


@pytest.mark.parametrize("sem", [0, 1])
def test_bif1d_overflow_rerun_matches_oracle(chaos, oracle, sem):
    n = 96
    print(f"bif1d overflow sem={sem}", flush=True)
    st = chaos.Stats()
    got = chaos.bif1d(T_MAX, n, H, LOR["ic"], [24.0, 60.0], [1], 0, 0.0, T_T, LOR["values"], 1,
                      system="lorenz_rk4", semantics=sem, stats=st)
    want = oracle.bif1d(oracle.SYS["lorenz_rk4"], sem, T_MAX, n, H, LOR["ic"], [24.0, 60.0], [1], 0, 0.0, T_T,
                        LOR["values"], 1)
    assert _first_guess_overflows(st.peaks, n), st.peaks / n
    _same_bif1d(got, want)


def test_bif2d_overflow_rerun_matches_oracle(chaos, oracle):
    n = 8
    ranges = [40.0, 80.0, 2.0, 3.0]
    print("bif2d overflow", flush=True)
    st = chaos.Stats()
    got = chaos.bif2d(T_MAX, n, H, LOR["ic"], ranges, [1, 2], 0, 0.0, T_T, LOR["values"], 1, 0.01,
                      system="lorenz_rk4", semantics=1, stats=st)
    want = oracle.bif2d(oracle.SYS["lorenz_rk4"], 1, T_MAX, n, H, LOR["ic"], ranges, [1, 2], 0, 0.0, T_T,
                        LOR["values"], 1, 0.01)
    assert _first_guess_overflows(st.peaks, n * n), st.peaks / (n * n)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("kind", ["bif1d", "bif2d"])
def test_engine_overflow_reruns_job(chaos, oracle, kind):
    J = chaos.make_job
    if kind == "bif1d":
        job = J(chaos.JOB_BIF1D, "lorenz_rk4", 1, 96, LOR["ic"], LOR["values"], [24.0, 60.0], [1], h=H,
                tMax=T_MAX, transientTime=T_T)
    else:
        job = J(chaos.JOB_BIF2D, "lorenz_rk4", 1, 8, LOR["ic"], LOR["values"], [40.0, 80.0, 2.0, 3.0], [1, 2],
                h=H, tMax=T_MAX, transientTime=T_T, eps=0.01)
    print(f"engine overflow {kind}", flush=True)
    srv = chaos.Server(params=chaos.engine_params(steps_per_launch=4000))
    jid = srv.submit(job)
    srv.drain()
    info = srv.info(jid)
    assert info.status == 2 and info.reran == 1
    got = srv.result(jid, job)
    if kind == "bif1d":
        want = oracle.bif1d(oracle.SYS["lorenz_rk4"], 1, T_MAX, 96, H, LOR["ic"], [24.0, 60.0], [1], 0, 0.0, T_T,
                            LOR["values"], 1)
        _same_bif1d(got, want)
    else:
        want = oracle.bif2d(oracle.SYS["lorenz_rk4"], 1, T_MAX, 8, H, LOR["ic"], [40.0, 80.0, 2.0, 3.0], [1, 2], 0,
                            0.0, T_T, LOR["values"], 1, 0.01)
        assert np.array_equal(got, want)
    srv.close()


_RANGES_WANT = {}


@pytest.mark.parametrize("cap_mb", [8, 4])
@pytest.mark.parametrize("sem", [0, 1])
def test_bif2d_passes_in_ranges_match_oracle(chaos, oracle, monkeypatch, sem, cap_mb):
    """When the peak store of the whole grid does not fit the device, the
    grid runs in ranges of whole 64-point waves, each writing its part of
    the map.  CHAOS_PEAK_STORE_BYTES caps the store here: at 8 MB the first
    pass (a first guess of 375 peaks per point, about 0.5 MB per wave) covers
    the whole grid and its exact rerun (about 2.8 MB per wave) runs in ranges
    of two waves; at 4 MB the first pass itself is cut in two ranges, each
    rerun one wave at a time.  The map equals the oracle's and the uncapped
    call's; the ranges show as more launches."""
    n = 24
    ranges = [40.0, 80.0, 2.0, 3.0]
    args = (T_MAX, n, H, LOR["ic"], ranges, [1, 2], 0, 0.0, T_T, LOR["values"], 1, 0.01)
    print(f"bif2d passes in ranges sem={sem} cap={cap_mb} MB", flush=True)
    whole = chaos.Stats()
    got_whole = chaos.bif2d(*args, system="lorenz_rk4", semantics=sem, stats=whole)
    monkeypatch.setenv("CHAOS_PEAK_STORE_BYTES", str(cap_mb << 20))
    split = chaos.Stats()
    got = chaos.bif2d(*args, system="lorenz_rk4", semantics=sem, stats=split)
    if sem not in _RANGES_WANT:
        _RANGES_WANT[sem] = oracle.bif2d(oracle.SYS["lorenz_rk4"], sem, *args)
    want = _RANGES_WANT[sem]
    assert _first_guess_overflows(whole.peaks, n * n)
    assert split.peaks == whole.peaks
    assert split.launches > whole.launches
    assert np.array_equal(got, want) and np.array_equal(got_whole, want)

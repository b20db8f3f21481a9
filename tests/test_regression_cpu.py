"""Host-side checks of the regression harness (no GPU needed): the binary is
built, parses its options, fails cleanly without a device, and its suite
lines parse into the inputs the cpu rows re-run."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "chaoslibrary_amd", "chaos_regression")


def test_harness_built_and_parses_options():
    assert os.path.exists(EXE), "chaos_regression is not built (run __graft_entry__.build())"
    p = subprocess.run([EXE, "--help"], capture_output=True, text=True, timeout=60)
    assert p.returncode == 2 and "usage: chaos_regression" in p.stderr
    for bad in (["--libraries", "cuda"], ["--params", "other"], ["--resolutions", "1"], ["--bogus", "1"],
                ["--out"]):
        p = subprocess.run([EXE, *bad], capture_output=True, text=True, timeout=60)
        assert p.returncode == 2, bad


def test_harness_without_device_fails_cleanly(chaos, tmp_path):
    if chaos.device_count() > 0:
        pytest.skip("a device is present")
    p = subprocess.run([EXE, "--results", str(tmp_path), "--out", str(tmp_path)], capture_output=True, text=True,
                       timeout=60)
    assert p.returncode == 1 and "no HIP device" in p.stderr


def test_suite_line_parses():
    sys.path.insert(0, ROOT)
    import bench
    d = bench._parse_suite("SUITE LLE system=5 semantics=0 CT=400 TT=2000 NT=0.5 h=0.01 eps=9.9999999999999995e-07 "
                           "maxValue=10000 preScaller=1 writableVar=0 seed=12345 init=0.10000000000000001,0.1,0.1 "
                           "values=0.20000000000000001,0.2,5.7000000000000002 ranges=0.05,0.35,2,10 idx=0,2")
    assert d["name"] == "LLE" and d["system"] == 5 and d["seed"] == 12345
    assert d["eps"] == 1e-6 and d["init"] == [0.1, 0.1, 0.1] and d["values"] == [0.2, 0.2, 5.7]
    assert d["ranges"] == [0.05, 0.35, 2.0, 10.0] and d["idx"] == [0, 2] and d["CT"] == 400.0

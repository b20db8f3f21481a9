"""The batched engine's file writer (chaos_server_write_result) against the
one-shot file-writing entry points, and the src/main.cu-style regression
harness (chaoslibrary_amd/chaos_regression) end to end at small sizes."""
import csv
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "chaoslibrary_amd", "chaos_regression")

ROS = dict(ic=[0.1, 0.1, 0.1], values=[0.2, 0.2, 5.7])
LOR = dict(ic=[1.0, 1.0, 1.0], values=[10.0, 28.0, 8 / 3])


@pytest.mark.parametrize("sem", [0, 1])
@pytest.mark.parametrize("kind", ["lle", "bif1d", "bif2d", "basins"])
def test_server_write_result_matches_one_shot_files(chaos, tmp_path, kind, sem):
    print(f"write_result kind={kind} sem={sem}", flush=True)
    srv = chaos.Server()
    if kind == "lle":
        rg, ix = [0.05, 0.35, 2.0, 10.0], [0, 2]
        j = chaos.make_job(chaos.JOB_LLE2D, "rossler_rk4", sem, 12, ROS["ic"], ROS["values"], rg, ix, tMax=20.0,
                           NT=0.5, eps=1e-6, transientTime=5.0, seed=777)
        chaos.LLE2D(20.0, 0.5, 12, 0.01, 1e-6, ROS["ic"], 3, rg, ix, 0, 0.0, 5.0, ROS["values"], 3,
                    tmp_path / "one", options=chaos.make_options("rossler_rk4", sem, seed=777))
        files = [""]
    elif kind == "bif1d":
        rg, ix = [2.0, 6.0], [2]
        j = chaos.make_job(chaos.JOB_BIF1D, "rossler_rk4", sem, 40, ROS["ic"], ROS["values"], rg, ix, tMax=60.0,
                           transientTime=20.0, preScaller=2)
        chaos.bifurcation1D(60.0, 40, 0.01, 3, ROS["ic"], rg, ix, 0, 0.0, 20.0, ROS["values"], 3, 2,
                            tmp_path / "one", options=chaos.make_options("rossler_rk4", sem))
        files = [""]
    elif kind == "bif2d":
        rg, ix = [20.0, 200.0, 0.5, 4.0], [1, 2]
        j = chaos.make_job(chaos.JOB_BIF2D, "lorenz_rk4", sem, 10, LOR["ic"], LOR["values"], rg, ix, tMax=20.0,
                           eps=0.01, transientTime=10.0)
        chaos.bifurcation2D(20.0, 10, 0.01, 3, LOR["ic"], rg, ix, 0, 0.0, 10.0, LOR["values"], 3, 1, 0.01,
                            tmp_path / "one", options=chaos.make_options("lorenz_rk4", sem))
        files = [""]
    else:
        rg, ix = [-20.0, 20.0, -20.0, 20.0], [0, 1]
        vals = [10.0, 24.5, 8 / 3]
        j = chaos.make_job(chaos.JOB_BASINS2D, "lorenz_rk4", sem, 16, [0.0, 0.0, 25.0], vals, rg, ix, tMax=20.0,
                           eps=0.5, transientTime=10.0, min_pts=4)
        chaos.basinsOfAttraction_2(20.0, 16, 0.01, 3, [0.0, 0.0, 25.0], rg, ix, 0, 0.0, 10.0, vals, 3, 1, 0.5,
                                   tmp_path / "one", options=chaos.make_options("lorenz_rk4", sem, min_pts=4))
        files = ["", "_1.csv", "_2.csv", "_3.csv"] + (["_4.csv"] if sem == 1 else [])
    jid = srv.submit(j)
    srv.drain()
    srv.write_result(jid, str(tmp_path / "two"))
    srv.write_result(jid, str(tmp_path / "three"), output=chaos.OUT_BINARY)
    srv.release(jid)
    srv.close()
    for f in files:
        one = (tmp_path / ("one" + f)).read_bytes()
        assert one and (tmp_path / ("two" + f)).read_bytes() == one, f
    if kind != "bif1d":
        # the binary sidecars convert back to the same CSV text
        for f in files:
            assert not (tmp_path / ("three" + f)).exists()
            chaos.map_bin_to_csv(str(tmp_path / ("three" + f + ".bin")), str(tmp_path / ("conv" + f)))
            assert (tmp_path / ("conv" + f)).read_bytes() == (tmp_path / ("one" + f)).read_bytes()


def _rows(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _harness(tmp_path, *extra):
    assert os.path.exists(EXE), "chaos_regression is not built"
    cmd = [EXE, "--resolutions", "8,12", "--times", "4,6", "--fixed-resolution", "8", "--ct", "4",
           "--fixed-time", "3", "--out", str(tmp_path / "maps"), "--results", str(tmp_path / "res"), *extra]
    print("+", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    print(p.stdout[-3000:], p.stderr[-2000:], flush=True)
    return p


@pytest.mark.parametrize("sem", ["reference", "config"])
def test_regression_harness_rk4(tmp_path, sem):
    p = _harness(tmp_path, "--semantics", sem)
    assert p.returncode == 0
    summary = [ln for ln in p.stdout.splitlines() if ln.startswith("rows ")][0]
    # 3 suites x 2 libraries x (2 + 2 runs + 2 sweep rows)
    assert summary.startswith("rows 36, errors 0") and summary.endswith("mismatched 0")
    compared = int(summary.split("files compared ")[1].split(",")[0])
    # 4 runs per suite; basins runs write 4 (REFERENCE) or 5 (CONFIG) files
    assert compared == 4 + 4 + 4 * (4 if sem == "reference" else 5)
    for name, test, extra in (("bifurcation", "Bifurcation", "Dimension"), ("lle", "LLE", None),
                              ("basins", "Basins", None)):
        rows = _rows(tmp_path / "res" / f"{name}_performance_results.csv")
        assert len(rows) == 12
        head = ["Test", "Parameter", "Value", "Library", "ExecutionTime_ms"] + ([extra] if extra else [])
        assert list(rows[0].keys()) == head
        for r in rows:
            assert r["Test"] == test and float(r["ExecutionTime_ms"]) > 0
            assert r["Library"] in ("hip-per-job", "hip-batched")
            if extra:
                assert r["Dimension"] == "2D"
        res = sorted(int(r["Value"]) for r in rows if r["Parameter"] == "Resolution")
        assert res == [8, 8, 12, 12]
        tt = sorted(int(r["Value"]) for r in rows if r["Parameter"] == "ModelingTime")
        assert tt == [4, 4, 6, 6]
        assert sorted(r["Parameter"] for r in rows if r["Parameter"].endswith("Sweep")) == \
            ["ModelingTimeSweep"] * 2 + ["ResolutionSweep"] * 2
    runs = [ln.split() for ln in p.stdout.splitlines() if ln.startswith("RUN ")]
    assert len(runs) == 36 and all(float(r[6]) > 0 for r in runs)


def test_regression_harness_reference_params(tmp_path):
    # src/main.cu's own sets (Chameleon, BASINS): the engine runs RK4 only
    p = _harness(tmp_path, "--params", "reference")
    assert p.returncode == 0
    assert "hip-batched rows skipped" in p.stdout
    for name in ("bifurcation", "lle", "basins"):
        rows = _rows(tmp_path / "res" / f"{name}_performance_results.csv")
        assert len(rows) == 6 and {r["Library"] for r in rows} == {"hip-per-job"}
    assert (tmp_path / "maps" / "LLE_hip-per-job_Resolution_12.csv").read_text().startswith("0 5\n0 20\n")

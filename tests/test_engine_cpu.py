"""Host-side checks of the engine/server API (no GPU needed)."""
import pytest


def test_engine_params_default(chaos):
    p = chaos.engine_params()
    assert (p.block, p.waves_per_simd, p.steps_per_launch, p.use_graph) == (256, 4, 40000, 1)
    assert p.points_per_tile == 64
    assert p.max_jobs >= 1 and p.max_tiles >= 1 and 0 < p.hbm_fraction <= 1


def test_stream_config_default_matches_stream_spec(chaos):
    c = chaos.stream_config()
    assert c.lambda_ == 200.0 and c.duration_s == 60.0 and c.horizon_scale == 1.0
    assert (c.bif_min, c.bif_max) == (256, 8192)
    assert (c.lle_min, c.lle_max) == (64 * 64, 1024 * 1024)
    assert (c.basins_min, c.basins_max) == (128 * 128, 1024 * 1024)


def test_autotune_cache_roundtrip(chaos, tmp_path):
    f = tmp_path / "tune.txt"
    f.write_text("block 64\nwaves_per_simd 6\nsteps_per_launch 16000\nsteps_per_s 1.0e+12\n")
    p = chaos.autotune_load(str(f))
    assert (p.block, p.waves_per_simd, p.steps_per_launch) == (64, 6, 16000)
    with pytest.raises(chaos.ChaosError):
        chaos.autotune_load(str(tmp_path / "missing.txt"))


def test_server_without_device_fails_cleanly(chaos):
    if chaos.device_count() > 0:
        pytest.skip("a device is present")
    with pytest.raises(chaos.ChaosError) as e:
        chaos.Server()
    assert e.value.code in (-6, -2)


def test_job_struct_layout(chaos):
    j = chaos.make_job(chaos.JOB_LLE2D, "rossler_rk4", 1, 8, [0.1] * 3, [0.2, 0.2, 5.7], [0, 1, 2, 3], [0, 2])
    assert j.amountOfInitialConditions == 3 and j.amountOfValues == 3 and list(j.indicesOfMutVars) == [0, 2]


# ctypes mirror -> C struct name in include/chaoslibrary.h
MIRRORS = {"Options": "chaos_options", "Stats": "chaos_stats", "Job": "chaos_job",
           "EngineParams": "chaos_engine_params", "JobInfo": "chaos_job_info",
           "ServerStats": "chaos_server_stats", "StreamConfig": "chaos_stream_config",
           "StreamReport": "chaos_stream_report"}


def test_ctypes_mirrors_match_the_header(chaos, tmp_path):
    """Every ctypes structure has the size and field offsets gcc gives the
    header's struct (the header is the ABI)."""
    import ctypes as C
    import os
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "chaoslibrary.h"', "int main(void) {"]
    for py, cs in MIRRORS.items():
        lines.append(f'printf("{py} size %zu\\n", sizeof({cs}));')
        for f, _ in getattr(chaos, py)._fields_:
            cf = "lambda" if f == "lambda_" else f
            lines.append(f'printf("{py} {f} %zu\\n", offsetof({cs}, {cf}));')
    lines.append("return 0; }")
    src = tmp_path / "abi.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "abi"
    subprocess.run(["gcc", "-I", os.path.join(root, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for line in filter(None, out):
        py, f, v = line.split()
        cls = getattr(chaos, py)
        got = C.sizeof(cls) if f == "size" else getattr(cls, f).offset
        assert got == int(v), (py, f, got, v)

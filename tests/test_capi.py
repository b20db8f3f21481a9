"""CPU checks of the C-ABI library: it loads, exports every symbol that
include/*.h declares, and rejects bad arguments before touching a device."""
import ctypes as C
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(chaos_[A-Za-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    for must in ("chaos_lle2d", "chaos_LLE2D", "chaos_bif1d", "chaos_bifurcation1D", "chaos_bif2d",
                 "chaos_bifurcation2D", "chaos_basins2d", "chaos_basinsOfAttraction_2",
                 "chaos_integrate", "chaos_last_error"):
        assert must in names


def test_every_declared_symbol_is_exported(chaos):
    missing = [n for n in declared_symbols() if not hasattr(chaos.lib, n)]
    assert not missing, missing


def test_version_and_dims(chaos):
    assert chaos.lib.chaos_api_version() == 1
    assert chaos.system_dims("rossler_rk4") == (3, 3)
    assert chaos.system_dims("chameleon") == (3, 7)
    assert chaos.system_dims("basins_2") == (2, 2)
    with pytest.raises(chaos.ChaosError) as e:
        chaos.system_dims(42)
    assert e.value.code == -1


def test_invalid_arguments_rejected_without_device(chaos):
    with pytest.raises(chaos.ChaosError) as e:
        chaos.lle2d(10.0, 0.5, 0, 0.01, 1e-8, [0.1] * 3, [0, 1, 0, 1], [0, 2], 1.0, [0.2, 0.2, 5.7],
                    system="rossler_rk4")
    assert e.value.code == -1 and "nPts" in str(e.value)
    with pytest.raises(chaos.ChaosError):
        chaos.lle2d(10.0, 0.5, 4, 0.01, 1e-8, [0.1] * 3, [0, 1, 0, 1], [0, 5], 1.0, [0.2, 0.2, 5.7],
                    system="rossler_rk4")
    with pytest.raises(chaos.ChaosError):  # wrong parameter count for the system
        chaos.lle2d(10.0, 0.5, 4, 0.01, 1e-8, [0.1] * 3, [0, 1, 0, 1], [0, 2], 1.0, [0.2, 0.2],
                    system="rossler_rk4")


def test_device_count_does_not_crash(chaos):
    assert chaos.device_count() >= 0


def test_cpp_api_links(chaos, tmp_path):
    """Every C++ entry point of include/*HOST.h / LLEHost.h resolves in
    libchaos.so (a program taking their addresses links)."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    fns = {"bifurcationHOST.h": ["Bifurcation::bifurcation1D", "Bifurcation::bifurcation2D"],
           "LLEHost.h": ["LLE::LLE2D"], "basinsHOST.h": ["Basins::basinsOfAttraction_2"],
           "legacyHOST.h": ["old_library::TimeDomainCalculation", "old_library::bifurcation1DIC",
                            "old_library::bifurcation2DIC", "old_library::LLE1D", "old_library::LLE1DIC",
                            "old_library::LLE2DIC", "old_library::LS1D", "old_library::LS2D"]}
    src = "".join(f'#include "{h}"\n' for h in fns) + "#include <cstdio>\nint main() {\n"
    for names in fns.values():
        for n in names:
            src += f'    std::printf("%p\\n", (void*)&{n});\n'
    src += "    return 0;\n}\n"
    f = tmp_path / "api.cpp"
    f.write_text(src)
    lib = os.path.join(ROOT, "chaoslibrary_amd")
    r = subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(f), "-o", str(tmp_path / "api"),
                        "-L", lib, "-lchaos", "-Wl,--unresolved-symbols=ignore-in-shared-libs"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]


def test_map_sidecar_converter_cpu(chaos, tmp_path):
    """A hand-built map sidecar converts to the reference text layout."""
    import struct
    import numpy as np
    vals = np.array([0.5, float("nan"), 1.25e-7, -3.0, 2.0, 7.0], dtype=np.float64)
    hdr = struct.pack("<8sIIqiiiiii4d", b"CHAOSMAP", 1, 0, vals.size, 3, 6, 1, 2, 6, 0, 0.05, 0.35, 2.0, 10.0)
    assert len(hdr) == 80
    b = tmp_path / "m.csv.bin"
    b.write_bytes(hdr + vals.tobytes())
    chaos.map_bin_to_csv(b, tmp_path / "m.csv")
    assert (tmp_path / "m.csv").read_text() == "0.05 0.35\n2 10\n0.5, 0, 1.25e-07, \n-3, 2, 7"
    ints = np.array([1, -1, 0, 3], dtype=np.int32)
    hdr = struct.pack("<8sIIqiiiiii4d", b"CHAOSMAP", 1, 1, ints.size, 2, 6, 0, 0, 6, 0, 0, 0, 0, 0)
    b.write_bytes(hdr + ints.tobytes())
    chaos.map_bin_to_csv(b, tmp_path / "i.csv")
    assert (tmp_path / "i.csv").read_text() == "1, -1, \n0, 3"
    b.write_bytes(b"NOTAMAP!" + bytes(72))
    with pytest.raises(chaos.ChaosError):
        chaos.map_bin_to_csv(b, tmp_path / "x.csv")
    b.write_bytes(hdr + ints.tobytes()[:-4])            # truncated
    with pytest.raises(chaos.ChaosError):
        chaos.map_bin_to_csv(b, tmp_path / "x.csv")

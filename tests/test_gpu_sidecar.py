"""Map files as binary sidecars (chaos_options.output): every map an entry
point writes can also be written as <path>.bin, and the converter turns the
sidecar into exactly the CSV bytes the entry point writes itself."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROS = ([0.1, 0.1, 0.1], [0.2, 0.2, 5.7])
LOR = ([1.0, 1.0, 1.0], [10.0, 28.0, 8 / 3])


def both(chaos, system, sem=0):
    return chaos.make_options(system, sem, output=chaos.OUT_CSV | chaos.OUT_BINARY)


def same_after_conversion(chaos, tmp_path, csv_files):
    for f in csv_files:
        out = tmp_path / ("conv_" + f.name)
        chaos.map_bin_to_csv(str(f) + ".bin", out)
        assert out.read_bytes() == f.read_bytes(), f.name


def test_sidecars_reproduce_csv(chaos, tmp_path):
    ic, vals = ROS
    f = tmp_path / "lle.csv"
    print("sidecar LLE2D", flush=True)
    chaos.LLE2D(20.0, 0.5, 12, 0.01, 1e-8, ic, 3, [0.05, 0.35, 2.0, 10.0], [0, 2], 0, 0.0, 5.0, vals, 3, str(f),
                options=both(chaos, "rossler_rk4", 1))
    g = tmp_path / "bif.csv"
    chaos.bifurcation2D(20.0, 10, 0.01, 3, LOR[0], [20.0, 200.0, 0.5, 4.0], [1, 2], 0, 0.0, 10.0, LOR[1], 3, 1,
                        1e-2, str(g), options=both(chaos, "lorenz_rk4", 1))
    b = tmp_path / "basins.csv"
    chaos.basinsOfAttraction_2(10.0, 12, 0.01, 3, [0.0, 0.0, 25.0], [-20.0, 20.0, -20.0, 20.0], [0, 1], 0, 0.0,
                               10.0, [10.0, 24.5, 8 / 3], 3, 1, 0.5, str(b), options=both(chaos, "lorenz_rk4", 1))
    ls = tmp_path / "ls"
    chaos.LS2D(10.0, 0.5, 4, 0.01, 1e-8, ic, 3, [0.1, 0.3, 4.0, 6.0], [0, 2], 0, 0.0, 5.0, vals, 3, str(ls),
               options=both(chaos, "rossler_rk4"))
    lic = tmp_path / "lle2dic.csv"
    chaos.LLE2DIC(10.0, 0.5, 5, 0.01, 1e-8, LOR[0], 3, [-5.0, 5.0, -5.0, 5.0], [0, 1], 0, 0.0, 2.0, LOR[1], 3,
                  str(lic), options=both(chaos, "lorenz_rk4"))
    files = [f, g, b, tmp_path / "basins.csv_1.csv", tmp_path / "basins.csv_2.csv", tmp_path / "basins.csv_3.csv",
             tmp_path / "basins.csv_4.csv", tmp_path / "ls1.csv", tmp_path / "ls2.csv", tmp_path / "ls3.csv", lic]
    same_after_conversion(chaos, tmp_path, files)


def test_binary_only_writes_no_csv(chaos, tmp_path):
    ic, vals = ROS
    f = tmp_path / "lle.csv"
    o = chaos.make_options("rossler_rk4", 1, output=chaos.OUT_BINARY)
    chaos.LLE2D(10.0, 0.5, 8, 0.01, 1e-8, ic, 3, [0.05, 0.35, 2.0, 10.0], [0, 2], 0, 0.0, 2.0, vals, 3, str(f),
                options=o)
    assert not f.exists() and (tmp_path / "lle.csv.bin").exists()
    raw = np.frombuffer((tmp_path / "lle.csv.bin").read_bytes()[80:], dtype=np.float64)
    want = chaos.lle2d(10.0, 0.5, 8, 0.01, 1e-8, ic, [0.05, 0.35, 2.0, 10.0], [0, 2], 2.0, vals,
                       system="rossler_rk4", semantics=1)
    assert np.array_equal(raw, want.reshape(-1))

"""GPU parity of the per-step maps through the C-ABI (chaos_integrate)
against the oracle: bit-exact over 10^4 steps for every system, including
the reference maps that call cos / sin / exp / atan (both sides evaluate the
library's portable routines, operation for operation)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def grid_params(O, n, ranges, fixed, idx, na):
    rng = np.random.default_rng(1234)
    P = np.tile(np.asarray(fixed, dtype=float), (n, 1))
    ia = rng.integers(0, 2048, n)
    ib = rng.integers(0, 2048, n)
    for k in range(n):
        P[k, idx[0]] = O.lib.orc_grid_value(int(ia[k]), 2048, ranges[0], ranges[1])
        P[k, idx[1]] = O.lib.orc_grid_value(int(ib[k]), 2048, ranges[2], ranges[3])
    return P


@pytest.mark.parametrize("system,fixed,ranges,idx,ic", [
    ("rossler_rk4", (0.2, 0.2, 5.7), (0.05, 0.35, 2.0, 10.0), (0, 2), (0.1, 0.1, 0.1)),
    ("lorenz_rk4", (10.0, 28.0, 8 / 3), (20.0, 200.0, 0.5, 4.0), (1, 2), (1.0, 1.0, 1.0)),
    ("rossler_cd", (0.0, 0.2, 0.2, 5.7), (0.05, 0.35, 2.0, 10.0), (1, 3), (0.1, 0.1, 0.1)),
])
def test_trajectories_bit_exact_1e4_steps(chaos, oracle, system, fixed, ranges, idx, ic):
    sid = oracle.SYS[system]
    nx, na = oracle.dims(sid)
    n = 1000
    P = grid_params(oracle, n, ranges, fixed, idx, na)
    X = np.tile(np.asarray(ic, dtype=float), (n, 1))
    print(f"integrate {system} n={n} steps=10000", flush=True)
    got = chaos.integrate(system, P, X, 10000, 0.01)
    want = oracle.integrate(sid, P, X, 10000, 0.01)
    same = (got.view(np.uint64) == want.view(np.uint64))
    assert same.all(), f"{(~same.all(1)).sum()} of {n} trajectories differ"


@pytest.mark.parametrize("system,params,ic,h", [
    ("chameleon", (0.5, 3.0, 0.0, 2.0, 1.0, 18.0, 1.0), (3.0, 3.0, 0.0), 0.01),
    ("basins", (0.5, 0.1, 1.4, 15.552, 2.0), (0.3, -0.2, 0.1), 0.01),
    ("basins_2", (0.5, 0.1), (0.3, -0.2), 0.01),
    ("damir", (1e-5, 1e5), (0.1, 0.5, 0.0), 1e-10),
])
def test_transcendental_maps_bit_exact_1e4_steps(chaos, oracle, system, params, ic, h):
    sid = oracle.SYS[system]
    n = 256
    rng = np.random.default_rng(5)
    P = np.tile(np.asarray(params, dtype=float), (n, 1))
    X = np.asarray(ic, dtype=float) + rng.uniform(-0.05, 0.05, (n, len(ic)))
    if system == "damir":
        X[:, 2] = 0.0
    print(f"integrate {system} n={n} steps=10000", flush=True)
    got = chaos.integrate(system, P, X, 10000, h)
    want = oracle.integrate(sid, P, X, 10000, h)
    same = (got.view(np.uint64) == want.view(np.uint64))
    assert same.all(), f"{(~same.all(1)).sum()} of {n} trajectories differ"


def test_zero_steps_and_empty(chaos):
    X = np.array([[0.1, 0.2, 0.3]])
    assert np.array_equal(chaos.integrate("rossler_rk4", [[0.2, 0.2, 5.7]], X, 0, 0.01), X)

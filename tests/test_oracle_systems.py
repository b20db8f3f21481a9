"""Pins for the oracle's per-step maps (oracle/chaos_oracle.c).

The RK4 steps are checked against a textbook RK4 written from the ODEs (not
from the oracle's fma placement), against scipy's DOP853 (order-4 error
decay), and against exact invariants (Lorenz origin, Lorenz mirror symmetry).
The reference composition schemes are checked for their convergence order
against scipy on the ODE they discretise.
"""
import numpy as np
import pytest
from scipy.integrate import solve_ivp

import _oracle as O

ROS = O.SYS["rossler_rk4"]
LOR = O.SYS["lorenz_rk4"]


def ode_rossler(t, x, a, b, c):
    return [-x[1] - x[2], x[0] + a * x[1], b + x[2] * (x[0] - c)]


def ode_lorenz(t, x, s, r, b):
    return [s * (x[1] - x[0]), x[0] * (r - x[2]) - x[1], x[0] * x[1] - b * x[2]]


def textbook_rk4(f, x, h, p):
    x = np.asarray(x, dtype=float)
    k1 = np.array(f(0, x, *p))
    k2 = np.array(f(0, x + h / 2 * k1, *p))
    k3 = np.array(f(0, x + h / 2 * k2, *p))
    k4 = np.array(f(0, x + h * k3, *p))
    return x + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)


@pytest.mark.parametrize("sysid,f,params", [
    (ROS, ode_rossler, (0.2, 0.2, 5.7)),
    (LOR, ode_lorenz, (10.0, 28.0, 8.0 / 3.0)),
])
def test_rk4_matches_textbook(sysid, f, params):
    rng = np.random.default_rng(7)
    for _ in range(200):
        x = rng.uniform(-15, 15, 3)
        got = O.steps(sysid, x, params, 0.01, 1)
        want = textbook_rk4(f, x, 0.01, params)
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("sysid,f,params,x0", [
    (ROS, ode_rossler, (0.2, 0.2, 5.7), (0.1, 0.1, 0.1)),
    (LOR, ode_lorenz, (10.0, 28.0, 8.0 / 3.0), (1.0, 1.0, 1.0)),
])
def test_rk4_order_four(sysid, f, params, x0):
    T = 0.5
    ref = solve_ivp(f, (0, T), x0, args=params, method="DOP853", rtol=1e-13, atol=1e-13).y[:, -1]
    errs = []
    for h in (0.01, 0.005, 0.0025):
        x = O.steps(sysid, x0, params, h, int(round(T / h)))
        errs.append(np.linalg.norm(x - ref))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 12 < r1 < 20 and 12 < r2 < 20, (errs, r1, r2)


def test_lorenz_origin_is_exact_fixed_point():
    x = O.steps(LOR, [0.0, 0.0, 0.0], (10, 28, 8 / 3), 0.01, 10000)
    assert np.all(x == 0.0)


def test_lorenz_mirror_symmetry_bit_exact():
    rng = np.random.default_rng(3)
    for _ in range(50):
        x = rng.uniform(-20, 20, 3)
        a = O.steps(LOR, x, (10, 28, 8 / 3), 0.01, 500)
        b = O.steps(LOR, [-x[0], -x[1], x[2]], (10, 28, 8 / 3), 0.01, 500)
        assert a[0] == -b[0] and a[1] == -b[1] and a[2] == b[2]


def test_rossler_rk4_equilibrium_nearly_fixed():
    a, b, c = 0.2, 0.2, 5.7
    xe = (c - np.sqrt(c * c - 4 * a * b)) / 2
    x0 = np.array([xe, -xe / a, xe / a])
    x = O.steps(ROS, x0, (a, b, c), 0.01, 10)
    np.testing.assert_allclose(x, x0, atol=1e-12)


def _order(sysid, f, params, x0, T, hs, to_ode=lambda p: p):
    ref = solve_ivp(f, (0, T), x0, args=to_ode(params), method="DOP853", rtol=1e-12, atol=1e-12).y[:, -1]
    errs = [np.linalg.norm(O.steps(sysid, x0, params, h, int(round(T / h))) - ref) for h in hs]
    return [errs[i] / errs[i + 1] for i in range(len(errs) - 1)], errs


def test_rossler_cd_is_second_order():
    # systems/systems.cuh:41-63 with symmetry parameter a[0]=0 is a symmetric
    # composition of the Rossler flow (a[1]=a, a[2]=b, a[3]=c).
    ratios, errs = _order(O.SYS["rossler_cd"], ode_rossler, (0.0, 0.2, 0.2, 5.7), (1.0, 1.0, 0.1),
                          2.0, (0.02, 0.01, 0.005), to_ode=lambda p: p[1:])
    assert all(3.3 < r < 4.7 for r in ratios), (ratios, errs)


def ode_chameleon(t, x, a0, a1, a2, a3, a4, a5, a6):
    return [-a6 * x[1], a6 * x[0] + a1 * x[2], a2 - a3 * x[2] + a4 * np.cos(a5 * x[1])]


def test_chameleon_is_second_order():
    # systems/systems.cuh:11-39 with a[0]=0.5: explicit half step then the
    # reversed semi-implicit half step (a symmetric composition).
    p = (0.5, 3.0, 0.0, 2.0, 1.0, 18.0, 1.0)
    ratios, errs = _order(O.SYS["chameleon"], ode_chameleon, p, (3.0, 3.0, 0.0), 1.0,
                          (0.004, 0.002, 0.001))
    assert all(3.3 < r < 4.7 for r in ratios), (ratios, errs)


def ode_basins(t, x, a0, b, *unused):
    return [np.sin(x[1]) - b * x[0], np.sin(x[2]) - b * x[1], np.sin(x[0]) - b * x[2]]


def _order_per_step(sysid, f, params, x0, T, hs, per_step, method="DOP853", norm=np.linalg.norm):
    """Error ratios of the map applied T / (per_step * h) times against the
    ODE flow at T (a step of the map advances the flow by per_step * h)."""
    ref = solve_ivp(f, (0, T), x0, args=params, method=method, rtol=1e-12, atol=1e-14).y[:, -1]
    errs = [norm(O.steps(sysid, list(x0) + [0.0] * (O.dims(sysid)[0] - len(x0)), params, h,
                         int(round(T / (per_step * h))))[:len(x0)] - ref) for h in hs]
    return [errs[i] / errs[i + 1] for i in range(len(errs) - 1)], errs


def test_basins_map_converges_to_its_ode():
    # systems/systems.cuh:67-83: a Gauss-Seidel explicit sweep of length h
    # followed by a reversed semi-implicit sweep of length h2 = h*(1-a[0])
    # (rounded to float) of x'=sin(y)-b x, y'=sin(z)-b y, z'=sin(x)-b z; a
    # step advances the flow by h + h2.  a[0] = 0 makes the two sweeps
    # adjoint (second order); otherwise first order.  Powers of two keep h2
    # exact in float.
    hs = (2.0 ** -7, 2.0 ** -8, 2.0 ** -9, 2.0 ** -10)
    r2, e2 = _order_per_step(O.SYS["basins"], ode_basins, (0.0, 0.2, 1.4, 15.552, 2.0), (0.3, -1.0, 2.0), 2.0,
                             hs, 2.0)
    assert all(3.9 < r < 4.1 for r in r2), (r2, e2)
    r1, e1 = _order_per_step(O.SYS["basins"], ode_basins, (0.5, 0.2, 1.4, 15.552, 2.0), (0.3, -1.0, 2.0), 3.0,
                             hs, 1.5)
    assert all(1.9 < r < 2.1 for r in r1), (r1, e1)


def ode_basins2(t, x, a0, a1):
    return [-x[1], a1 * x[0] + np.sin(x[1])]


@pytest.mark.parametrize("a0", [0.5, 0.3])
def test_basins2_map_converges_to_its_ode(a0):
    # systems/systems.cuh:85-101: stages of length h*a[0] and h*(1-a[0]) of
    # x'=-y, y'=a[1] x + sin(y); first order (the float-rounded y between the
    # stages stays far below the discretisation error at these h).
    r, e = _order_per_step(O.SYS["basins_2"], ode_basins2, (a0, -1.5), (1.0, 0.5), 2.0,
                           (2.0 ** -4, 2.0 ** -5, 2.0 ** -6, 2.0 ** -7), 1.0)
    assert all(1.9 < x < 2.1 for x in r), (r, e)


def ode_damir(t, y, A, w):
    # the circuit of systems/systems.cuh:103-178 written from its equations:
    # capacitor voltage Vc and memristor state XSV, sinusoidal drive A sin(w t)
    Vc, XSV = y
    Vd, Vm = Vc - 0.04, Vc
    Id = (1.15e-7 * (np.exp(15 * Vd) - np.exp(-15 * Vd)) + 2.1e-5 / 0.09 * Vd * np.exp(-(Vd - 0.09) / 0.09)
          - 3e-6 * (np.arctan(26 * (Vd - 0.14)) + np.arctan(26 * (Vd + 0.14))))
    kT = 2.0 * 1.380649e-23 / 1.602176634e-19
    f1 = 1 / (1 + np.exp(-(Vm - 0.28) * (Vm + 0.12) / kT))
    f2 = 1 / (1 + np.exp(-(Vm + 0.006) * (Vm - 0.14) / kT))
    g = XSV / 1434.0 + (1 - XSV) / 1e6
    return [(A * np.sin(w * t) - Vm * g - Id) / 22e-9, (f1 * (1 - XSV) - (1 - f2) * XSV) / 1e-7]


def test_damir_map_is_euler_of_its_ode():
    # explicit Euler with the time in X[2]: first-order convergence of both
    # the voltage and the memristor state (mid-transition at T = 2 TAU)
    r, e = _order_per_step(O.SYS["damir"], ode_damir, (2e-4, 5e6), (0.4, 0.2), 2e-7,
                           (4e-10, 2e-10, 1e-10, 5e-11), 1.0, method="Radau",
                           norm=lambda d: np.abs(d) / np.array([1e-6, 1e-4]))
    for ratios in np.array(r).T:
        assert all(1.95 < x < 2.05 for x in ratios), (r, e)


def test_damir_time_channel_advances_by_h():
    x = O.steps(O.SYS["damir"], [0.1, 0.5, 0.0], (1e-5, 1e5), 1e-10, 1000)
    assert abs(x[2] - 1000 * 1e-10) < 1e-20 and np.all(np.isfinite(x))


def test_integrate_batch_equals_single_steps():
    rng = np.random.default_rng(11)
    params = np.column_stack([rng.uniform(0.05, 0.35, 32), np.full(32, 0.2), rng.uniform(2, 10, 32)])
    states = np.tile([0.1, 0.1, 0.1], (32, 1))
    out = O.integrate(ROS, params, states, 1000, 0.01)
    for i in range(32):
        assert np.array_equal(out[i], O.steps(ROS, states[i], params[i], 0.01, 1000))


def test_dims():
    assert O.dims(O.SYS["chameleon"]) == (3, 7)
    assert O.dims(O.SYS["basins_2"]) == (2, 2)
    assert O.dims(ROS) == (3, 3)

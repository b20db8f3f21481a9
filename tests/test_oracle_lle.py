"""Pins for the oracle's Benettin LLE (LLE/src/LLECUDA.cu:56-166 restated).

* Linear closed form: at the Lorenz origin with rho < 1 the perturbation obeys
  the linear map M = R(hJ) (RK4 stability polynomial of the Jacobian), so the
  Benettin sum equals sum_k log|M^NT u_k| computed here with numpy from the
  same initial direction -- a dropped window, a wrong renormalisation or a
  wrong divisor fails it.
* REFERENCE semantics runs two extra windows (all/NTp + 2).
* Published exponents: Lorenz(10,28,8/3) ~ 0.906, Rossler(0.2,0.2,5.7) ~ 0.071,
  and ~0 on a periodic Rossler orbit.
* The seeded direction is a unit vector; SplitMix64 reproduces its published
  test vector.
"""
import ctypes as C
import math

import numpy as np
import pytest

import _oracle as O

ROS = O.SYS["rossler_rk4"]
LOR = O.SYS["lorenz_rk4"]


def test_splitmix64_published_vector():
    s = C.c_uint64(1234567)
    got = [O.lib.orc_splitmix64(C.byref(s)) for _ in range(5)]
    assert got == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                   4593380528125082431, 16408922859458223821]


def test_unit_vector_norm_and_spread():
    vs = np.array([O.unit_vector(12345, p) for p in range(4000)])
    norms = np.sqrt((vs ** 2).sum(1))
    assert np.all(np.abs(norms - 1) < 4e-16)
    assert np.all(np.abs(vs.mean(0)) < 0.05)
    np.testing.assert_allclose((vs ** 2).mean(0), 1 / 3, atol=0.03)
    assert not np.array_equal(vs[0], vs[1])
    assert np.array_equal(O.unit_vector(12345, 17), O.unit_vector(12345, 17))
    assert not np.array_equal(O.unit_vector(1, 17), O.unit_vector(2, 17))


def _linear_benettin(J, h, NTp, W, u, T):
    hJ = h * J
    M = np.eye(3) + hJ + hJ @ hJ / 2 + hJ @ hJ @ hJ / 6 + hJ @ hJ @ hJ @ hJ / 24
    P = np.linalg.matrix_power(M, NTp)
    s = 0.0
    for _ in range(W):
        u = P @ u
        n = np.linalg.norm(u)
        s += math.log(n)
        u = u / n
    return s / T


@pytest.mark.parametrize("sem,extra", [(O.SEM_CONFIG, 0), (O.SEM_REFERENCE, 2)])
def test_lle_linear_closed_form(sem, extra):
    sigma, rho, beta = 10.0, 0.5, 8.0 / 3.0
    h, NT, T = 0.01, 0.5, 2000.0
    J = np.array([[-sigma, sigma, 0.0], [rho, -1.0, 0.0], [0.0, 0.0, -beta]])
    u = O.unit_vector(99, 0)
    want = _linear_benettin(J, h, 50, 4000 + extra, u, T)
    got = O.lle_point(LOR, sem, T, NT, h, 1e-8, [0.0, 0.0, 0.0], [sigma, rho, beta], 0.0, 99, 0)
    assert abs(got - want) < 1e-7, (got, want)
    # and the asymptote is ln R(h lambda1)/h
    lam1 = (-(sigma + 1) + math.sqrt((sigma + 1) ** 2 - 4 * sigma * (1 - rho))) / 2
    z = h * lam1
    asym = math.log(1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24) / h
    assert abs(got * T / (T + extra * NT) - asym) < 5e-3


def test_lle_published_values():
    lor = O.lle_point(LOR, O.SEM_CONFIG, 2000.0, 0.5, 0.01, 1e-8, [1, 1, 1], [10, 28, 8 / 3], 50.0, 1, 0)
    ros = O.lle_point(ROS, O.SEM_CONFIG, 2000.0, 0.5, 0.01, 1e-8, [0.1, 0.1, 0.1], [0.2, 0.2, 5.7], 500.0, 1, 0)
    per = O.lle_point(ROS, O.SEM_CONFIG, 1000.0, 0.5, 0.01, 1e-8, [0.1, 0.1, 0.1], [0.2, 0.2, 2.5], 500.0, 1, 0)
    assert abs(lor - 0.906) < 0.03, lor
    assert abs(ros - 0.071) < 0.012, ros
    assert abs(per) < 0.005, per


def test_lle2d_layout_rows_are_second_parameter():
    ranges = [0.05, 0.35, 2.0, 10.0]
    out = O.lle2d(ROS, O.SEM_CONFIG, 10.0, 0.5, 4, 0.01, 1e-8, [0.1, 0.1, 0.1], ranges, [0, 2], 5.0,
                  [0.2, 0.2, 5.7], 12345)
    for j in range(4):
        for i in range(4):
            a = O.lib.orc_grid_value(i, 4, 0.05, 0.35)
            c = O.lib.orc_grid_value(j, 4, 2.0, 10.0)
            want = O.lle_point(ROS, O.SEM_CONFIG, 10.0, 0.5, 0.01, 1e-8, [0.1, 0.1, 0.1], [a, 0.2, c], 5.0,
                               12345, j * 4 + i)
            assert out[j, i] == want


def test_grid_value_endpoints():
    assert O.lib.orc_grid_value(0, 2048, 0.05, 0.35) == 0.05
    assert abs(O.lib.orc_grid_value(2047, 2048, 0.05, 0.35) - 0.35) < 1e-15
    assert O.lib.orc_grid_value(0, 1, 3.0, 9.0) == 3.0
    for k in range(0, 2048, 97):
        assert O.lib.orc_grid_value_pow(k + 2048 * 5, 2048, -20, 20, 0) == O.lib.orc_grid_value(k, 2048, -20, 20)
        assert O.lib.orc_grid_value_pow(k * 2048 + 5, 2048, -20, 20, 1) == O.lib.orc_grid_value(k, 2048, -20, 20)

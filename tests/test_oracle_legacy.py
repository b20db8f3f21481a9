"""Pins of the oracle's legacy functions (debri/, namespace old_library):
Lyapunov-spectrum Gram-Schmidt, the two legacy LLE normalisation forms, the
early stop of the legacy integration loop, the Lyapunov spectrum and the
time-domain recorder -- against closed forms of a linear map, numpy QR, the
LLE module's oracle and exact invariants."""
import numpy as np
import pytest

import _oracle as O

LOR = O.SYS["lorenz_rk4"]


def rk4_matrix(J, h):
    hJ = h * J
    I = np.eye(len(J))
    return I + hJ + hJ @ hJ / 2 + hJ @ hJ @ hJ / 6 + hJ @ hJ @ hJ @ hJ / 24


def lorenz_jacobian_origin(s, r, b):
    return np.array([[-s, s, 0.0], [r, -1.0, 0.0], [0.0, 0.0, -b]])


def test_gram_schmidt_matches_qr():
    rng = np.random.default_rng(0)
    for n in (2, 3, 4):
        A = rng.normal(size=(n, n))
        B, den = O.gram_schmidt(A)
        assert np.allclose(B @ B.T, np.eye(n), atol=1e-12)
        Q, R = np.linalg.qr(A.T)          # columns of Q = orthonormalised rows of A
        sgn = np.sign(np.diag(R))
        assert np.allclose(B, (Q * sgn).T, atol=1e-12)
        assert np.allclose(den, np.abs(np.diag(R)), rtol=1e-12)
    # a zero row stays zero with norm 0 (the reference's 0/0 guard)
    B, den = O.gram_schmidt(np.array([[1.0, 0.0], [2.0, 0.0]]))
    assert np.allclose(B[0], [1, 0]) and np.all(B[1] == 0) and den[1] == 0


@pytest.mark.parametrize("form", [0, 1])
def test_legacy_lle_linear_closed_form(form):
    # rho = 0.5: the origin is an exact, stable fixed point, so the
    # perturbation follows the linear RK4 map R(hJ) exactly to O(eps^2)
    s, r, b, h, NT, tMax, eps = 10.0, 0.5, 8 / 3, 0.01, 0.5, 10.0, 1e-8
    got = O.lib.orc_legacy_lle_point(LOR, 0, form, tMax, NT, h, eps, O.dptr(np.zeros(3)), 3,
                                     O.dptr(np.array([s, r, b])), 0.0, 0.0, 7, 11)
    v = np.zeros(3)
    O.lib.orc_unit_vector(7, 11, 3, O.dptr(v))
    M = np.linalg.matrix_power(rk4_matrix(lorenz_jacobian_origin(s, r, b), h), int(NT / h))
    y, acc = eps * v, 0.0
    for _ in range(int(tMax / NT)):
        y = M @ y
        d = np.linalg.norm(y) / eps
        acc += np.log(d)
        y = y / d
    # O(eps) relative: the quadratic terms of Lorenz at |y| ~ eps
    assert got == pytest.approx(acc / tMax, rel=1e-7)
    # over a long run Benettin converges to the top exponent of the linear map
    lam = np.max(np.log(np.abs(np.linalg.eigvals(rk4_matrix(lorenz_jacobian_origin(s, r, b), h))))) / h
    long = O.lib.orc_legacy_lle_point(LOR, 0, form, 400.0, NT, h, eps, O.dptr(np.zeros(3)), 3,
                                      O.dptr(np.array([s, r, b])), 0.0, 0.0, 7, 11)
    assert long == pytest.approx(lam, abs=0.01)


def test_legacy_forms_agree_and_match_lle_module():
    ic, vals = np.array([1.0, 1.0, 1.0]), np.array([10.0, 28.0, 8 / 3])
    args = (20.0, 0.5, 0.01, 1e-8, O.dptr(ic), 3, O.dptr(vals), 0.0, 5.0, 3, 17)
    f0 = O.lib.orc_legacy_lle_point(LOR, 0, 0, *args)
    f1 = O.lib.orc_legacy_lle_point(LOR, 0, 1, *args)
    # same windows (int(20/0.5) = 40 = llround(20/h)/llround(NT/h)), same
    # direction and transient as the LLE module in CONFIG semantics
    mod = O.lle_point(LOR, 1, 20.0, 0.5, 0.01, 1e-8, ic, vals, 5.0, 3, 17)
    assert abs(f0 - f1) < 1e-9 and abs(f1 - mod) < 1e-9
    assert 0.5 < f1 < 1.4                          # Lorenz(10, 28, 8/3) top exponent ~0.9


def test_legacy_early_stop_equals_one_step_windows():
    # with maxValue below |x| the legacy loop stops after the first step of
    # every window: the sum of logs equals that of 1-step windows without
    # the bound (same step count per window, same transient of 1 step)
    ic, vals = np.array([1.0, 1.0, 1.0]), np.array([10.0, 28.0, 8 / 3])
    h = 0.01
    a = O.lib.orc_legacy_lle_point(LOR, 0, 0, 20.0, 0.5, h, 1e-8, O.dptr(ic), 3, O.dptr(vals), 1e-3, h, 5, 2)
    b = O.lib.orc_legacy_lle_point(LOR, 0, 0, 0.4, h, h, 1e-8, O.dptr(ic), 3, O.dptr(vals), 0.0, h, 5, 2)
    assert a * 20.0 == pytest.approx(b * 0.4, rel=1e-15)
    c = O.lib.orc_legacy_lle_point(LOR, 0, 0, 20.0, 0.5, h, 1e-8, O.dptr(ic), 3, O.dptr(vals), 0.0, h, 5, 2)
    assert c != a


def test_legacy_ls_linear_spectrum():
    s, r, b, h, NT, tMax, eps = 10.0, 0.5, 8 / 3, 0.01, 0.5, 20.0, 1e-8
    out = np.zeros(3)
    assert O.lib.orc_legacy_ls_point(LOR, 0, tMax, NT, h, eps, O.dptr(np.zeros(3)), 3,
                                     O.dptr(np.array([s, r, b])), 0.0, 0.0, 9, 4, O.dptr(out)) == 0
    R = rk4_matrix(lorenz_jacobian_origin(s, r, b), h)
    M = np.linalg.matrix_power(R, int(NT / h))
    Z = np.zeros((3, 3))
    for j in range(3):
        O.lib.orc_unit_vector(9, 4 * 4 + j, 3, O.dptr(Z[j]))
    Q, _ = np.linalg.qr(Z.T)
    Y = Q.T * np.sign(np.diag(np.linalg.qr(Z.T)[1]))[:, None]   # GS rows
    acc = np.zeros(3)
    for _ in range(int(tMax / NT)):
        D = (M @ (eps * Y).T).T
        q, rr = np.linalg.qr(D.T)
        acc += np.log(np.abs(np.diag(rr)) / eps)
        Y = (q * np.sign(np.diag(rr))).T
    assert np.allclose(out, acc / tMax, rtol=1e-7, atol=1e-9)
    lam = np.sort(np.log(np.abs(np.linalg.eigvals(R))) / h)[::-1]
    assert np.allclose(out, lam, atol=0.1)
    # volume contraction: the sum is ln|det R^NT| per unit time (exact per window)
    assert out.sum() == pytest.approx(np.log(abs(np.linalg.det(M))) / NT * (int(tMax / NT) * NT / tMax),
                                      rel=1e-8)


def test_legacy_ls_top_exponent_matches_lle():
    ic, vals = np.array([1.0, 1.0, 1.0]), np.array([10.0, 28.0, 8 / 3])
    out = np.zeros(3)
    O.lib.orc_legacy_ls_point(LOR, 0, 200.0, 0.5, 0.01, 1e-8, O.dptr(ic), 3, O.dptr(vals), 0.0, 20.0, 1, 0,
                              O.dptr(out))
    assert out[0] == pytest.approx(0.906, abs=0.08)
    assert abs(out[1]) < 0.05
    assert out.sum() == pytest.approx(-(10 + 1 + 8 / 3), abs=0.05)   # trace of the Lorenz Jacobian


def test_time_domain_samples_are_the_trajectory():
    vals, ic = [10.0, 28.0, 8 / 3], [1.0, 2.0, 3.0]
    h, pre, tT, tMax = 0.01, 3, 0.25, 0.6
    out, flags = O.time_domain(LOR, 0, tMax, 4, h, ic, [20.0, 30.0], [1], 0, 0.0, tT, vals, pre)
    S, B = int(tT / h), int(tMax / h / pre)
    assert out.shape == (4, B) and np.all(flags == 1)
    for p in range(4):
        a = np.array(vals)
        a[1] = O.lib.orc_grid_value(p, 4, 20.0, 30.0)
        X = np.array(ic)
        O.lib.orc_steps(LOR, O.dptr(X), O.dptr(a), h, S * pre)     # the transient runs preScaller steps
        for t in range(B):
            assert out[p, t] == X[0]
            O.lib.orc_steps(LOR, O.dptr(X), O.dptr(a), h, pre)
    # a bound below |x| stops the transient: nothing recorded
    out, flags = O.time_domain(LOR, 0, tMax, 3, h, ic, [20.0, 30.0], [1], 0, 0.5, tT, vals, pre)
    assert np.all(flags == 0) and not out.any()


def test_legacy_bif1d_ic_equals_parameter_sweep_on_one_point():
    # a degenerate IC range puts every point at the same start: with
    # preScaller 1 and no divergence or equilibrium, the legacy IC sweep and
    # the Bifurcation module's parameter sweep record the same samples
    ic, vals = [1.0, 1.0, 1.0], [10.0, 28.0, 8 / 3]
    counts, peaks = O.legacy_bif1d_ic(LOR, 0, 20.0, 3, 0.01, ic, [1.0, 1.0], [0], 0, 0.0, 5.0, vals, 1)
    c2, _, p2, _ = O.bif1d(LOR, 0, 20.0, 2, 0.01, ic, [28.0, 28.0], [1], 0, 0.0, 5.0, vals, 1)
    assert np.all(counts == c2[0]) and c2[0] > 5
    assert np.array_equal(peaks[:counts[0]], p2[:c2[0]])


def test_legacy_ic_transient_fixed_point_still_records():
    # rho = 0.5, start at the origin: an exact equilibrium.  The legacy IC
    # sweep records it and reports -1; the Bifurcation module's parameter
    # sweep drops a fixed-point transient to 0 (bifurcationCUDA.cu:334-342)
    vals = [10.0, 0.5, 8 / 3]
    counts, peaks = O.legacy_bif1d_ic(LOR, 0, 5.0, 2, 0.01, [0.0, 0.0, 0.0], [0.0, 0.0], [0], 0, 0.0, 1.0, vals, 1)
    assert list(counts) == [-1, -1] and peaks.size == 0
    c2, _, _, _ = O.bif1d(LOR, 0, 5.0, 2, 0.01, [0.0, 0.0, 0.0], [0.5, 0.5], [1], 0, 0.0, 1.0, vals, 1)
    assert list(c2) == [0, 0]
    m = O.legacy_bif2d_ic(LOR, 0, 5.0, 2, 0.01, [0.0, 0.0, 0.0], [0.0, 0.0, 0.0, 0.0], [0, 1], 0, 0.0, 1.0,
                          vals, 1, 0.1)
    assert list(m) == [-1] * 4

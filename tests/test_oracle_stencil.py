"""Pins of the oracle's stencil kernels (pmgc/matvecd.cc, pmgc/gsd.cc)
against dense brute force, symmetry and fixed-point properties."""
import numpy as np
import pytest

from tests import helpers as H


@pytest.mark.parametrize("numdia", [7, 27])
@pytest.mark.parametrize("n", [5, 7])
def test_matvec_equals_dense(orc, n, numdia):
    rng = np.random.default_rng(n * 100 + numdia)
    ac, cc = H.random_op(n, numdia, rng)
    A, _ = H.dense_A(n, numdia, ac, cc)
    x = H.from_int(n, rng.standard_normal((n - 2) ** 3))
    y = orc.matvec(n, numdia, ac, cc, x)
    np.testing.assert_allclose(H.to_int(n, y), A @ H.to_int(n, x), rtol=1e-13, atol=1e-13)
    # residual r = f - A x
    f = H.from_int(n, rng.standard_normal((n - 2) ** 3))
    r = orc.mresid(n, numdia, ac, cc, f, x)
    np.testing.assert_allclose(H.to_int(n, r), H.to_int(n, f) - A @ H.to_int(n, x),
                               rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("numdia", [7, 27])
def test_matvec_symmetric(orc, numdia):
    n = 9
    rng = np.random.default_rng(3)
    ac, cc = H.random_op(n, numdia, rng)
    x = H.from_int(n, rng.standard_normal((n - 2) ** 3))
    y = H.from_int(n, rng.standard_normal((n - 2) ** 3))
    ax = orc.matvec(n, numdia, ac, cc, x)
    ay = orc.matvec(n, numdia, ac, cc, y)
    assert abs(np.dot(y, ax) - np.dot(x, ay)) < 1e-12 * abs(np.dot(y, ax))


def test_laplacian_extreme_eigenvalues(orc):
    """Unscaled 3-D Laplacian eigenvalues quoted in pmgc/powerd.h:63-72."""
    n = 9
    N = n ** 3
    ac = np.zeros(4 * N)
    for (i, j, k) in H.interior(n):
        p = H.IX(n, i, j, k)
        ac[p] = 6.0
        if i < n - 1:
            ac[N + p] = 1.0
        if j < n - 1:
            ac[2 * N + p] = 1.0
        if k < n - 1:
            ac[3 * N + p] = 1.0
    cc = np.zeros(N)
    A = np.column_stack([H.to_int(n, orc.matvec(n, 7, ac, cc, H.from_int(n, e)))
                         for e in np.eye((n - 2) ** 3)])
    ev = np.linalg.eigvalsh(A)
    lmin = 6 - 6 * np.cos(np.pi / (n - 1))
    lmax = 6 - 6 * np.cos((n - 2) * np.pi / (n - 1))
    assert abs(ev[0] - lmin) < 1e-12 and abs(ev[-1] - lmax) < 1e-12


@pytest.mark.parametrize("numdia", [7, 27])
def test_gs_fixed_point_and_serial_order(orc, numdia):
    """The exact discrete solution is a fixed point of the smoother; one sweep
    equals a serial loop over colours in the reference's colour order
    (pmgc/gsd.cc:85-93, 161-169)."""
    n = 7
    rng = np.random.default_rng(11)
    ac, cc = H.random_op(n, numdia, rng)
    A, pts = H.dense_A(n, numdia, ac, cc)
    xs = rng.standard_normal(len(pts))
    f = H.from_int(n, A @ xs)
    out = orc.gsrb(n, numdia, ac, cc, f, H.from_int(n, xs), 3, 0)
    np.testing.assert_allclose(H.to_int(n, out), xs, rtol=1e-12, atol=1e-12)
    # serial reference: Gauss-Seidel point updates ordered by colour
    x0 = rng.standard_normal(len(pts))
    for iadj in (0, 1):
        x = x0.copy()
        if numdia == 7:
            cols = [0, 1] if iadj == 0 else [1, 0]
            colour = [(i + j + k) % 2 for (i, j, k) in pts]
        else:
            cols = list(range(8)) if iadj == 0 else list(range(7, -1, -1))
            colour = [(i % 2) + 2 * (j % 2) + 4 * (k % 2) for (i, j, k) in pts]
        fi = H.to_int(n, f)
        for c in cols:
            for a in range(len(pts)):
                if colour[a] == c:
                    x[a] = (fi[a] - (A[a] @ x - A[a, a] * x[a])) / A[a, a]
        got = orc.gsrb(n, numdia, ac, cc, f, H.from_int(n, x0), 1, iadj)
        np.testing.assert_allclose(H.to_int(n, got), x, rtol=1e-12, atol=1e-12)


def test_cghs_solves_small_system(orc):
    n = 7
    rng = np.random.default_rng(5)
    ac, cc = H.random_op(n, 27, rng)
    A, pts = H.dense_A(n, 27, ac, cc)
    b = rng.standard_normal(len(pts))
    x, it = orc.cghs(n, 27, ac, cc, H.from_int(n, b), 100, 0.0)
    assert it <= 100
    np.testing.assert_allclose(H.to_int(n, x), np.linalg.solve(A, b), rtol=1e-9, atol=1e-11)


def test_ion_terms_closed_form(orc):
    """1:1 salt with sconc 1/2: c_vec = coef*sinh(u), dc_vec = coef*cosh(u)
    (pmgc/mypdec.cc:136-212, 348-414), clamped at |u| = 85."""
    orc.mypdefinit([1.0, -1.0], [0.5, 0.5])
    u = np.array([-100.0, -3.0, -0.5, 0.0, 0.25, 2.0, 84.0, 90.0])
    coef = np.array([1.0, 2.0, 0.5, 1.0, 3.0, 1.0, 1e-3, 1e-3])
    uc = np.clip(u, -85, 85)
    np.testing.assert_allclose(orc.c_vec(coef, u), coef * np.sinh(uc), rtol=1e-13)
    np.testing.assert_allclose(orc.c_vec(coef, u, deriv=True), coef * np.cosh(uc), rtol=1e-13)
    z = orc.c_vec(np.zeros_like(u), u)
    assert np.all(z == 0.0)

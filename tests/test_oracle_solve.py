"""Pins of the oracle's fine operator and full solves: closed-form harmonic
solutions (Laplacian_pmgc.cc:145-147, Laplacian_3D_GS.cc:225),
manufactured solutions for the linear and Newton drivers."""
import numpy as np
import pytest

from tests import helpers as H


def uniform_inputs(n, h, g, eps=1.0, kappa2=0.0, fcf=None):
    c = np.arange(n) * h
    inp = {"xf": c.copy(), "yf": c.copy(), "zf": c.copy()}
    N = n ** 3
    for k in ("a1cf", "a2cf", "a3cf"):
        inp[k] = np.full(N, eps)
    inp["ccf"] = np.full(N, kappa2)
    inp["fcf"] = np.zeros(N) if fcf is None else fcf
    inp["tcf"] = np.zeros(N)
    gx, gy, gz = np.zeros(4 * n * n), np.zeros(4 * n * n), np.zeros(4 * n * n)
    for b in range(n):
        for a in range(n):
            gx[b * n + a] = g(c[0], c[a], c[b])
            gx[n * n + b * n + a] = g(c[-1], c[a], c[b])
            gy[b * n + a] = g(c[a], c[0], c[b])
            gy[n * n + b * n + a] = g(c[a], c[-1], c[b])
            gz[b * n + a] = g(c[a], c[b], c[0])
            gz[n * n + b * n + a] = g(c[a], c[b], c[-1])
    inp["gxcf"], inp["gycf"], inp["gzcf"] = gx, gy, gz
    return inp


def test_buildA_uniform_closed_form(orc):
    """Uniform grid: oE = h a1cf, oC = h * (6 faces), cc = h^3 ccf,
    fc = h^3 fcf + h * face * g on boundary-adjacent points
    (pmgc/buildAd.cc:145-208)."""
    n, h = 7, 0.4
    rng = np.random.default_rng(0)
    inp = uniform_inputs(n, h, lambda x, y, z: 1.0 + x - 2 * y + 3 * z)
    for k in ("a1cf", "a2cf", "a3cf", "ccf", "fcf"):
        inp[k] = rng.uniform(0.5, 2.0, n ** 3)
    ac, cc, fc, ipc = orc.buildA_fv(n, inp)
    N = n ** 3
    a1, a2, a3 = inp["a1cf"], inp["a2cf"], inp["a3cf"]
    g = lambda i, j, k: 1.0 + (i - 1) * h - 2 * (j - 1) * h + 3 * (k - 1) * h
    for (i, j, k) in H.interior(n):
        p = H.IX(n, i, j, k)
        faces = [a1[p], a1[H.IX(n, i - 1, j, k)], a2[p], a2[H.IX(n, i, j - 1, k)],
                 a3[p], a3[H.IX(n, i, j, k - 1)]]
        assert ac[p] == pytest.approx(h * sum(faces), rel=1e-14)
        assert ac[N + p] == (pytest.approx(h * a1[p], rel=1e-14) if i < n - 1 else 0.0)
        assert ac[2 * N + p] == (pytest.approx(h * a2[p], rel=1e-14) if j < n - 1 else 0.0)
        assert ac[3 * N + p] == (pytest.approx(h * a3[p], rel=1e-14) if k < n - 1 else 0.0)
        assert cc[p] == pytest.approx(h ** 3 * inp["ccf"][p], rel=1e-14)
        f = h ** 3 * inp["fcf"][p]
        if i == n - 1: f += h * a1[p] * g(n, j, k)
        if j == n - 1: f += h * a2[p] * g(i, n, k)
        if k == n - 1: f += h * a3[p] * g(i, j, n)
        if i == 2: f += h * faces[1] * g(1, j, k)
        if j == 2: f += h * faces[3] * g(i, 1, k)
        if k == 2: f += h * faces[5] * g(i, j, 1)
        assert fc[p] == pytest.approx(f, rel=1e-13)
    assert list(ipc[9:12]) == [0, 7, 1]


@pytest.mark.parametrize("g", [lambda x, y, z: z,
                               lambda x, y, z: -3 * x * x - 4 * y * y + 7 * z * z])
def test_harmonic_closed_form(orc, g):
    n, nlev, h = 17, 3, 0.25
    inp = uniform_inputs(n, h, g)
    ierr, u, iw, rw = orc.solve(inp, n, nlev, errtol=1e-13, itmax=100)
    assert ierr == 0
    c = np.arange(n) * h
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    exact = g(X, Y, Z).transpose(2, 1, 0).ravel()  # k slowest, i fastest
    assert np.max(np.abs(u - exact)) < 1e-9 * max(1, np.max(np.abs(exact)))


def test_manufactured_variable_coefficients(orc):
    n, nlev, h = 17, 3, 0.5
    rng = np.random.default_rng(9)
    inp = uniform_inputs(n, h, lambda x, y, z: 0.0)
    N = n ** 3
    for k in ("a1cf", "a2cf", "a3cf"):
        inp[k] = rng.choice([2.0, 78.54], N)
    inp["ccf"] = rng.uniform(0.0, 2.0, N)
    ac, cc, _, _ = orc.buildA_fv(n, inp)
    us = H.from_int(n, rng.standard_normal((n - 2) ** 3))
    Au = orc.matvec(n, 7, ac, cc, us)
    inp["fcf"] = Au / h ** 3
    ierr, u, _, _ = orc.solve(inp, n, nlev, errtol=1e-12, itmax=200)
    assert ierr == 0
    assert np.max(np.abs(u - us)) < 1e-8


def test_newton_manufactured(orc):
    """f := A u* + c_vec(u*) is solved back to u* by the Newton driver."""
    n, nlev, h = 17, 3, 0.5
    rng = np.random.default_rng(21)
    orc.mypdefinit([1.0, -1.0], [0.5, 0.5])
    inp = uniform_inputs(n, h, lambda x, y, z: 0.0)
    N = n ** 3
    for k in ("a1cf", "a2cf", "a3cf"):
        inp[k] = rng.choice([2.0, 78.54], N)
    inp["ccf"] = rng.uniform(0.0, 5.0, N)
    ac, cc, _, _ = orc.buildA_fv(n, inp)
    us = H.from_int(n, 3.0 * rng.standard_normal((n - 2) ** 3))
    rhs = orc.matvec(n, 7, ac, np.zeros(N), us) + orc.c_vec(cc, us)
    inp["fcf"] = H.from_int(n, H.to_int(n, rhs)) / h ** 3
    ierr, u, _, _ = orc.solve(inp, n, nlev, nonlin=True, errtol=1e-12, itmax=50)
    assert ierr == 0
    assert np.max(np.abs(u - us)) < 1e-8


def test_vcycle_contracts_on_pb_problem(orc, mg):
    n, nlev = 33, 4
    inp = orc.pbe_synth(n, 0.5, 40, 6.0, 0.15, 0, 3)
    ierr, u, _, _ = orc.solve(inp, n, nlev)
    it, ie, rel, hist = orc.last_stats()
    assert ierr == 0 and rel <= 1e-6 and 2 <= it <= 30
    rho = hist[1:] / hist[:-1]
    assert np.all(rho < 0.5)


def _newton_inputs(orc, n, h, seed):
    rng = np.random.default_rng(seed)
    orc.mypdefinit([1.0, -1.0], [0.5, 0.5])
    inp = uniform_inputs(n, h, lambda x, y, z: 0.0)
    N = n ** 3
    for k in ("a1cf", "a2cf", "a3cf"):
        inp[k] = rng.choice([2.0, 78.54], N)
    inp["ccf"] = rng.uniform(0.0, 5.0, N)
    ac, cc, _, _ = orc.buildA_fv(n, inp)
    us = H.from_int(n, 3.0 * rng.standard_normal((n - 2) ** 3))
    rhs = orc.matvec(n, 7, ac, np.zeros(N), us) + orc.c_vec(cc, us)
    inp["fcf"] = H.from_int(n, H.to_int(n, rhs)) / h ** 3
    return inp, us


def test_nested_newton_manufactured(orc):
    """mgkey 1 (Vfnewton, newtond.cc:60-138): the nested iteration
    recovers the manufactured solution as the plain Newton does."""
    n, nlev, h = 17, 3, 0.5
    inp, us = _newton_inputs(orc, n, h, 21)
    ierr, u, _, _ = orc.solve(inp, n, nlev, nonlin=True, errtol=1e-12, itmax=50, mgkey=1)
    assert ierr == 0
    assert np.max(np.abs(u - us)) < 1e-8


def test_nested_newton_starts_from_coarse_solution(orc):
    """The finest Newton of the nested iteration starts from the prolongated
    coarse-level solutions: its first step leaves a smaller nonlinear
    residual than the plain Newton's first step from zero, it needs no more
    iterations, and both reach the same potential."""
    n, nlev = 33, 3
    orc.mypdefinit([1.0, -1.0], [0.5, 0.5])
    inp = orc.pbe_synth(n, 0.5, 40, 6.0, 0.3, 0, 5)
    e0, u0, _, _ = orc.solve(inp, n, nlev, nonlin=True, errtol=1e-8)
    it0, _, _, h0 = orc.last_stats()
    e1, u1, _, _ = orc.solve(inp, n, nlev, nonlin=True, errtol=1e-8, mgkey=1)
    it1, _, _, h1 = orc.last_stats()
    assert e0 == e1 == 0
    assert h1[0] < h0[0] and it1 <= it0
    assert np.max(np.abs(u1 - u0)) <= 1e-7 * np.max(np.abs(u0))


def test_newton_starts_from_callers_u(orc, mg):
    """Vnewdriv2 zeroes only the boundary of u (VfboundPMG00,
    newdrvd.cc:263): the interior given is the starting iterate, so the
    converged potential given back converges in one Newton step."""
    n, nlev = 17, 3
    orc.mypdefinit([1.0, -1.0], [0.5, 0.5])
    inp = orc.pbe_synth(n, 0.5, 12, 3.0, 0.3, 0, 6)
    _, u0, _, _ = orc.solve(inp, n, nlev, nonlin=True, errtol=1e-10)
    it0 = orc.last_stats()[0]
    iparm, rparm = mg.make_params(n, nlev, nonlin=1, errtol=1e-10)
    sz = orc.mgsz(2, 0, 0, n, n, n, nlev)[1]
    a = {k: np.ascontiguousarray(inp[k], np.float64).copy() for k in orc.PBE_KEYS}
    u = u0.copy()
    e = orc.newdriv(iparm, rparm, np.zeros(int(sz[11]), np.int32), np.zeros(int(iparm[0])),
                    u, *(a[k] for k in orc.PBE_KEYS))
    it = orc.last_stats()[0]
    assert e == 0 and it0 > 2 and it == 1

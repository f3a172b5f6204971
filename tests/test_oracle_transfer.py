"""Pins of prolongation / restriction (pmgc/buildPd.cc, pmgc/matvecd.cc)."""
import numpy as np
import pytest

from tests import helpers as H


def laplace_like(n, eps=None, h=1.0, kappa=0.0):
    """7-point box operator of -div(eps grad u) + kappa u with Dirichlet
    folding (the shape VbuildA_fv produces), built directly."""
    N = n ** 3
    ac = np.zeros(4 * N)
    cc = np.zeros(N)
    if eps is None:
        eps = lambda i, j, k, ax: 1.0
    for (i, j, k) in H.interior(n):
        p = H.IX(n, i, j, k)
        e = [eps(i, j, k, 0), eps(i - 1, j, k, 0), eps(i, j, k, 1), eps(i, j - 1, k, 1),
             eps(i, j, k, 2), eps(i, j, k - 1, 2)]
        ac[p] = h * sum(e)
        ac[N + p] = h * e[0] if i < n - 1 else 0.0
        ac[2 * N + p] = h * e[2] if j < n - 1 else 0.0
        ac[3 * N + p] = h * e[4] if k < n - 1 else 0.0
        cc[p] = kappa * h ** 3
    return ac, cc


def weights(pc, nc, ii, jj, kk):
    Nc = nc ** 3
    c = H.IX(nc, ii, jj, kk)
    return {H.PPLANE[pl]: pc[pl * Nc + c] for pl in range(27)}


def test_trilinear_constant_coefficients(orc):
    """Operator-based P of a constant-coefficient 7-point operator is the
    trilinear stencil 1/2, 1/4, 1/8 (buildPd.cc:272-275, 415-418)."""
    nf, nc = 17, 9
    ac, _ = laplace_like(nf)
    for mgprol in (0, 1):
        pc = orc.buildP(nf, nc, mgprol, 7, ac)
        for (ii, jj, kk) in H.interior(nc):
            w = weights(pc, nc, ii, jj, kk)
            for d, v in w.items():
                m = sum(1 for t in d if t)
                assert abs(v - 0.5 ** m) < 1e-14, (ii, jj, kk, d, v)


def test_galerkin_level_weights_trilinear_in_interior(orc):
    """The 27-point rule gives 1/2, 1/4, 1/8 for the zero-row-sum Galerkin
    operator of a constant Laplacian, away from the boundary."""
    nf, nc, ncc = 33, 17, 9
    ac, _ = laplace_like(nf)
    pc = orc.buildP(nf, nc, 1, 7, ac)
    acc = orc.buildG(nf, nc, 7, pc, ac)
    pc2 = orc.buildP(nc, ncc, 1, 27, acc)
    for (ii, jj, kk) in H.interior(ncc):
        if min(ii, jj, kk) < 3 or max(ii, jj, kk) > ncc - 2:
            continue
        for d, v in weights(pc2, ncc, ii, jj, kk).items():
            m = sum(1 for t in d if t)
            assert abs(v - 0.5 ** m) < 1e-13, (ii, jj, kk, d, v)


def test_layered_medium_flux_weights(orc):
    """eps varying only along y: the y-face weight is the 1-D flux-continuous
    interpolation a_j / (a_j + a_{j+1}) for both stencils."""
    nf, nc = 17, 9
    rng = np.random.default_rng(1)
    ay = rng.uniform(0.5, 80.0, nf + 2)
    eps = lambda i, j, k, ax: ay[j] if ax == 1 else 1.0
    ac, _ = laplace_like(nf, eps)
    N = nf ** 3
    pc = orc.buildP(nf, nc, 1, 7, ac)
    for (ii, jj, kk) in H.interior(nc):
        j = 2 * jj - 1
        w = weights(pc, nc, ii, jj, kk)
        assert abs(w[(0, 1, 0)] - ay[j] / (ay[j] + ay[j + 1])) < 1e-14
        assert abs(w[(0, -1, 0)] - ay[j - 1] / (ay[j - 1] + ay[j - 2])) < 1e-14
        assert abs(w[(1, 0, 0)] - 0.5) < 1e-14


def test_explicit_reference_weights(orc):
    """Selected weights written out as in pmgc/buildPd.cc (op7 :272-275
    oPN, :316-321 oPNE, :415-420 dPNE; op27 :690-697 oPN, :738-746 oPNE,
    :893-901 dPNE) at random coefficients."""
    nf, nc = 9, 5
    rng = np.random.default_rng(7)
    for numdia in (7, 27):
        ac, _ = H.random_op(nf, numdia, rng)
        N = nf ** 3
        P = lambda pl, i, j, k: ac[pl * N + H.IX(nf, i, j, k)]
        oC, oE, oN, uC = [lambda i, j, k, q=q: P(q, i, j, k) for q in range(4)]
        (oNE, oNW, uE, uW, uN, uS, uNE, uNW, uSE, uSW) = [
            (lambda i, j, k, q=q: P(q, i, j, k) if numdia == 27 else 0.0) for q in range(4, 14)]
        pc = orc.buildP(nf, nc, 1, numdia, ac)
        for (ii, jj, kk) in H.interior(nc):
            i, j, k = 2 * ii - 1, 2 * jj - 1, 2 * kk - 1
            w = weights(pc, nc, ii, jj, kk)
            oPN = ((uNE(i - 1, j, k - 1) + uN(i, j, k - 1) + uNW(i + 1, j, k - 1)
                    + oNE(i - 1, j, k) + oN(i, j, k) + oNW(i + 1, j, k)
                    + uSW(i, j + 1, k) + uS(i, j + 1, k) + uSE(i, j + 1, k))
                   / (oC(i, j + 1, k) - oE(i - 1, j + 1, k) - oE(i, j + 1, k)
                      - uC(i, j + 1, k - 1) - uE(i - 1, j + 1, k - 1) - uW(i + 1, j + 1, k - 1)
                      - uC(i, j + 1, k) - uW(i, j + 1, k) - uE(i, j + 1, k)))
            oPE = w[(1, 0, 0)]
            oPNE = ((uNE(i, j, k - 1) + oNE(i, j, k) + uSW(i + 1, j + 1, k)
                     + (uN(i + 1, j, k - 1) + oN(i + 1, j, k) + uS(i + 1, j + 1, k)) * oPE
                     + (uE(i, j + 1, k - 1) + oE(i, j + 1, k) + uW(i + 1, j + 1, k)) * oPN)
                    / (oC(i + 1, j + 1, k) - uC(i + 1, j + 1, k - 1) - uC(i + 1, j + 1, k)))
            dPC, dPN, dPE = w[(0, 0, -1)], w[(0, 1, -1)], w[(1, 0, -1)]
            dPNE = ((uSW(i + 1, j + 1, k - 1) + uW(i + 1, j + 1, k - 1) * oPN
                     + uS(i + 1, j + 1, k - 1) * oPE + uC(i + 1, j + 1, k - 1) * oPNE
                     + oNE(i, j, k - 1) * dPC + oE(i, j + 1, k - 1) * dPN
                     + oN(i + 1, j, k - 1) * dPE) / oC(i + 1, j + 1, k - 1))
            assert abs(w[(0, 1, 0)] - oPN) < 1e-14 * max(1, abs(oPN))
            assert abs(w[(1, 1, 0)] - oPNE) < 1e-13 * max(1, abs(oPNE))
            assert abs(w[(1, 1, -1)] - dPNE) < 1e-13 * max(1, abs(dPNE))


def test_restrict_is_transpose_of_interp(orc):
    nf, nc = 17, 9
    rng = np.random.default_rng(2)
    ac, _ = H.random_op(nf, 7, rng)
    pc = orc.buildP(nf, nc, 1, 7, ac)
    r = H.from_int(nf, rng.standard_normal((nf - 2) ** 3))
    e = H.from_int(nc, rng.standard_normal((nc - 2) ** 3))
    lhs = np.dot(orc.restrc(nf, nc, r, pc), e)
    rhs = np.dot(r, orc.interp(nc, nf, e, pc))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_interp_weight_positions(orc):
    """Interpolating a unit coarse vector e_C puts weight P_C(d) at C_f + d
    (P plane order pmgc/matvecd.cc:809-816); the coincident fine point is a
    plain copy (matvecd.cc:898), independent of oPC."""
    nf, nc = 9, 5
    rng = np.random.default_rng(4)
    pc = np.zeros(27 * nc ** 3)
    for (ii, jj, kk) in H.interior(nc):
        for pl in range(27):
            pc[pl * nc ** 3 + H.IX(nc, ii, jj, kk)] = rng.uniform(0.1, 1.0)
    for C in [(2, 2, 2), (3, 2, 4), (4, 4, 4)]:
        e = np.zeros(nc ** 3)
        e[H.IX(nc, *C)] = 1.0
        f = orc.interp(nc, nf, e, pc)
        w = weights(pc, nc, *C)
        for d, v in w.items():
            q = tuple(2 * c - 1 + t for c, t in zip(C, d))
            expect = 1.0 if d == (0, 0, 0) else v
            assert f[H.IX(nf, *q)] == pytest.approx(expect, abs=0, rel=1e-15)
        assert np.count_nonzero(f) == 27


def test_trilinear_reproduces_linear_functions(orc):
    nf, nc = 17, 9
    pc = orc.buildP(nf, nc, 0, 7, np.zeros(4 * nf ** 3))
    g = lambda i, j, k: 1.5 * i - 0.7 * j + 0.3 * k + 2.0
    e = np.zeros(nc ** 3)
    for (ii, jj, kk) in H.interior(nc):
        e[H.IX(nc, ii, jj, kk)] = g(2 * ii - 1, 2 * jj - 1, 2 * kk - 1)
    f = orc.interp(nc, nf, e, pc)
    # fine points whose interpolation stencil lies on interior coarse points
    for (i, j, k) in H.interior(nf):
        if min(i, j, k) >= 3 and max(i, j, k) <= nf - 2:
            assert abs(f[H.IX(nf, i, j, k)] - g(i, j, k)) < 1e-12

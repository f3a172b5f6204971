"""Galerkin coarse operator (pmgc/buildGd.cc) against dense P^T A P."""
import numpy as np
import pytest

from tests import helpers as H


def dense_P(orc, nf, nc, pc):
    cols = []
    for e in np.eye((nc - 2) ** 3):
        cols.append(H.to_int(nf, orc.interp(nc, nf, H.from_int(nc, e), pc)))
    return np.column_stack(cols)


@pytest.mark.parametrize("numdia,nf", [(7, 9), (7, 17), (27, 9), (27, 17)])
def test_rap_equals_dense(orc, numdia, nf):
    nc = (nf - 1) // 2 + 1
    rng = np.random.default_rng(nf + numdia)
    ac, cc = H.random_op(nf, numdia, rng)
    pc = orc.buildP(nf, nc, 1, numdia, ac)
    acc = orc.buildG(nf, nc, numdia, pc, ac)
    A, _ = H.dense_A(nf, numdia, ac, np.zeros(nf ** 3))
    P = dense_P(orc, nf, nc, pc)
    RAP = P.T @ A @ P
    Ac, _ = H.dense_A(nc, 27, acc, np.zeros(nc ** 3))
    scale = np.max(np.abs(RAP))
    assert np.max(np.abs(Ac - RAP)) <= 1e-12 * scale
    # stored couplings to the coarse boundary are exactly zero
    Nc = nc ** 3
    for (ii, jj, kk) in H.interior(nc):
        for pl in range(1, 14):
            ex, ey, ez = H.UPPER27[pl]
            q = (ii + ex, jj + ey, kk + ez)
            if min(q) == 1 or max(q) == nc:
                assert acc[pl * Nc + H.IX(nc, ii, jj, kk)] == 0.0


def test_rap_zero_row_sum_interior(orc):
    nf, nc = 33, 17
    from tests.test_oracle_transfer import laplace_like
    ac, _ = laplace_like(nf)
    pc = orc.buildP(nf, nc, 1, 7, ac)
    acc = orc.buildG(nf, nc, 7, pc, ac)
    A, pts = H.dense_A(nc, 27, acc, np.zeros(nc ** 3))
    for a, (i, j, k) in enumerate(pts):
        if min(i, j, k) >= 3 and max(i, j, k) <= nc - 2:
            assert abs(A[a].sum()) < 1e-12 * A[a, a]

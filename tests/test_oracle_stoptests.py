"""Pins the oracle's stopping tests istop 2-5 (pmgc/mgcsd.cc:126-163,
483-527) and the algebraic right-hand side of istop 4 and 5 (Vbuildalg,
mgdrvd.cc:424-433, mgsubd.cc:891-950) against their definitions, evaluated
independently from the oracle's iterates after 1, 2 and 3 cycles:

* 2: ||u_{k-1} - u_k||_1 / sqrt(nx ny nz), u_0 = tcf (successive iterates);
* 3: ||tcf - u_k||_2 / ||tcf||_2 (error against a given solution);
* 4: the same norm, with the right-hand side replaced by A tcf on every
  level, so that the solve converges to tcf;
* 5: sqrt(e . A e) / sqrt(tcf . A tcf), e = tcf - u_k (energy norm), with
  the same right-hand side.

Norms are over interior points; A is the dense matrix of the fine operator
assembled from its planes (tests/helpers.py)."""
import numpy as np
import pytest

from tests import helpers as H

N, NLEV = 17, 3


@pytest.fixture(scope="module")
def problem(orc):
    inp = orc.pbe_synth(N, 0.5, 30, 3.5, 0.15, 0, 4)
    rng = np.random.default_rng(17)
    m = np.zeros((N, N, N))
    m[1:-1, 1:-1, 1:-1] = 1
    tcf = (rng.standard_normal((N, N, N)) * m).ravel()
    inp = dict(inp)
    inp["tcf"] = tcf
    ac, cc, fc, _ = orc.buildA_fv(N, inp)
    A, pts = H.dense_A(N, 7, ac, cc)
    return inp, tcf, A


def _runs(orc, inp, istop, cycles=3):
    out, hist = [], None
    for k in range(1, cycles + 1):
        e, u, iw, rw = orc.solve(inp, N, NLEV, istop=istop, itmax=k, errtol=1e-300)
        it, ie, _, h = orc.last_stats()
        assert e == 1 and it == k
        out.append(H.to_int(N, u))
        hist = h[:cycles]
    return out, hist, rw


def test_istop2_successive_iterates(orc, problem):
    inp, tcf, _ = problem
    us, hist, _ = _runs(orc, inp, 2)
    prev = [H.to_int(N, tcf)] + us[:-1]
    want = [np.abs(a - b).sum() / np.sqrt(N ** 3) for a, b in zip(prev, us)]
    assert np.allclose(hist, want, rtol=1e-12, atol=0)


def test_istop3_error_norm(orc, problem):
    inp, tcf, _ = problem
    us, hist, _ = _runs(orc, inp, 3)
    t = H.to_int(N, tcf)
    want = [np.linalg.norm(t - u) / np.linalg.norm(t) for u in us]
    assert np.allclose(hist, want, rtol=1e-12, atol=0)


@pytest.mark.parametrize("istop", [4, 5])
def test_istop45_algebraic_rhs(orc, problem, istop):
    inp, tcf, A = problem
    us, hist, rw = _runs(orc, inp, istop)
    t = H.to_int(N, tcf)
    if istop == 4:
        want = [np.linalg.norm(t - u) / np.linalg.norm(t) for u in us]
    else:
        want = [np.sqrt((t - u) @ A @ (t - u)) / np.sqrt(t @ A @ t) for u in us]
    assert np.allclose(hist, want, rtol=1e-10, atol=0)
    # the fine right-hand side in rwork is A tcf (Vbuildalg), so the solve
    # converges to tcf
    import multigrid_amd as mg
    sz = mg.Vmgsz(2, 0, 0, N, N, N, NLEV)
    fc0 = sz["n_rpc"] + sz["narr"]
    f = H.to_int(N, rw[fc0:fc0 + N ** 3])
    assert np.max(np.abs(f - A @ t)) <= 1e-12 * np.max(np.abs(f))
    e, u, _, _ = orc.solve(inp, N, NLEV, istop=istop, errtol=1e-10)
    assert e == 0
    assert np.max(np.abs(H.to_int(N, u) - t)) <= 1e-8 * np.max(np.abs(t))
    assert hist[0] > hist[1] > hist[2]

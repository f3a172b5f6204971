"""Stopping tests iparm(11) = istop 0-5 (pmgc/mgcsd.cc:126-163, 483-527)
and the per-cycle report (Vprtstp, mgsubd.cc:486-536) on the GPU against
the oracle, whose stopping tests are pinned to their definitions
(test_oracle_stoptests.py): same exit status and cycle count, potential
within 1e-10, the relative measure of every cycle within 1e-9 of the
history's scale (late residuals are rounding-dominated), and for
istop 4 and 5 the algebraic right-hand side A tcf in rwork (Vbuildalg)
within 1e-12.  The mixed-precision solve takes the same tests through a
resident context (pmg_ctx_set_tru)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, NLEV = 33, 3


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def problem(orc):
    inp = orc.pbe_synth(N, 0.5, 40, 6.0, 0.15, 0, 21)
    rng = np.random.default_rng(33)
    m = np.zeros((N, N, N))
    m[1:-1, 1:-1, 1:-1] = 1
    inp = dict(inp)
    inp["tcf"] = (rng.standard_normal((N, N, N)) * m).ravel()
    return inp


def _kw(istop):
    # istop 3 measures against a tcf that is not the solution: it levels off
    # above errtol, so both runs stop at itmax
    return dict(istop=istop, itmax=5) if istop == 3 else dict(istop=istop)


@pytest.mark.parametrize("istop", [0, 1, 2, 3, 4, 5])
def test_mgdriv_stop_tests(mg, orc, problem, istop):
    print(f"istop {istop}", flush=True)
    iparm, rparm = mg.make_params(N, NLEV, **_kw(istop))
    sz = mg.Vmgsz(2, 0, 0, N, N, N, NLEV)
    iw = np.zeros(sz["iintot"], np.int32)
    rw = np.zeros(int(iparm[0]))
    u = np.zeros(N ** 3)
    a = {k: problem[k].copy() for k in mg.PBE_KEYS}
    e1 = mg.Vmgdriv(iparm, rparm, iw, rw, u, *(a[k] for k in mg.PBE_KEYS))
    it1, ie1, rr1 = mg.last_stats()
    h1 = mg.last_history()
    e0, u0, iw0, rw0 = orc.solve(problem, N, NLEV, **_kw(istop))
    it0, ie0, rr0, h0 = orc.last_stats()
    print(f"  cycles gpu {it1} oracle {it0}; history gpu {h1[:3]} oracle {h0[:3]}", flush=True)
    assert e1 == e0 and ie1 == ie0 and it1 == it0 == len(h1) == len(h0)
    assert _rel(u1 := u, u0) <= 1e-10
    assert np.max(np.abs(h1 - h0)) <= 1e-9 * np.max(h0)
    assert abs(rr1 - rr0) <= 1e-9 * np.max(h0)
    assert _rel(rw, rw0) <= 1e-12
    if istop in (4, 5):  # tcf is the solution of the algebraic system
        inner = np.zeros((N, N, N), bool)
        inner[1:-1, 1:-1, 1:-1] = True
        # a relative 2-norm or energy-norm error of 1e-6 bounds the
        # max-norm error only loosely
        assert _rel(u1.reshape(N, N, N)[inner], problem["tcf"].reshape(N, N, N)[inner]) < 1e-4


@pytest.mark.parametrize("istop", [2, 3, 4, 5])
def test_ctx_mixed_stop_tests(mg, orc, problem, istop):
    """The FP32 cycle under the FP64 stopping measures: within 1e-6 of the
    FP64 oracle; the FP64 context solve equals the drop-in's history."""
    iparm, rparm = mg.make_params(N, NLEV, **_kw(istop))
    ctx = mg.Context(iparm, rparm, problem)
    with pytest.raises(mg.PMGError):
        ctx.solve(mg.PMG_FP64)  # no comparison solution yet
    ctx.set_tru(problem["tcf"])
    ctx.solve(mg.PMG_FP64)
    h64 = ctx.history()
    ctx.set_tru(problem["tcf"])  # istop 2 overwrote it with the iterates
    ctx.rebuild()                # istop 4, 5 replaced fc
    ctx.solve(mg.PMG_MIXED)
    umx, smx, hmx = ctx.u(), ctx.stats(), ctx.history()
    ctx.close()
    _, u0, _, _ = orc.solve(problem, N, NLEV, **_kw(istop))
    it0, _, _, h0 = orc.last_stats()
    print(f"istop {istop}: fp64 {len(h64)} mixed {smx['iters']} oracle {it0} cycles, "
          f"mixed err {_rel(umx, u0):.2e}", flush=True)
    assert len(h64) == it0 and np.max(np.abs(h64 - h0)) <= 1e-9 * np.max(h0)
    assert len(hmx) == smx["iters"] and hmx[-1] == pytest.approx(smx["relres"], rel=1e-12)
    # converged solves: the contract's 1e-6; istop 3 stops both at 5 cycles,
    # where the FP32 cycle's iterate is compared before convergence
    assert _rel(umx, u0) <= (1e-4 if istop == 3 else 1e-6)


def test_per_cycle_report(mg, orc, problem, capfd):
    """iinfo > 0: one Vprtstp line per cycle, with the relative residual and
    the contraction number of that cycle."""
    iparm, rparm = mg.make_params(N, NLEV, iinfo=1)
    sz = mg.Vmgsz(2, 0, 0, N, N, N, NLEV)
    a = {k: problem[k].copy() for k in mg.PBE_KEYS}
    capfd.readouterr()
    mg.Vmgdriv(iparm, rparm, np.zeros(sz["iintot"], np.int32), np.zeros(int(iparm[0])),
               np.zeros(N ** 3), *(a[k] for k in mg.PBE_KEYS))
    lines = [l for l in capfd.readouterr().out.splitlines() if l.startswith("iteration =")]
    h = mg.last_history()
    assert len(lines) == len(h) == mg.last_stats()[0]
    for q, l in enumerate(lines):
        f = l.split()
        assert int(f[2]) == q + 1
        assert float(f[6]) == pytest.approx(h[q], rel=1e-6)
        if q:
            assert float(f[10]) == pytest.approx(h[q] / h[q - 1], rel=1e-6)

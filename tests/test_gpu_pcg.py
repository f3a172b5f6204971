"""GPU parity of the class-style solver (PBE_pmgc_class.h:1070-1133 with
CG_Solver.h:45-242): CG preconditioned by the class V-cycle on the PMG
hierarchy, against the oracle's restatement.  FP64: potential 1e-10
relative max-norm, the same CG iteration count, iwork identical and rwork
to 1e-12; mixed (FP32 V-cycle preconditioner under the FP64 CG): potential
within 1e-6 of the FP64 oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _drive(mg, inp, n, nlev, num_iters=0, **kw):
    iparm, rparm = mg.make_params(n, nlev, **kw)
    sz = mg.Vmgsz(2, 0, 0, n, n, n, nlev)
    iw = np.zeros(sz["iintot"], np.int32)
    rw = np.zeros(int(iparm[0]))
    u = np.zeros(n ** 3)
    a = {k: inp[k].copy() for k in ("xf", "yf", "zf", "gxcf", "gycf", "gzcf", "a1cf",
                                     "a2cf", "a3cf", "ccf", "fcf", "tcf")}
    ierr = mg.Vpcgdriv(iparm, rparm, iw, rw, u, a["xf"], a["yf"], a["zf"], a["gxcf"],
                       a["gycf"], a["gzcf"], a["a1cf"], a["a2cf"], a["a3cf"], a["ccf"],
                       a["fcf"], a["tcf"], num_iters=num_iters)
    return ierr, u, iw, rw


@pytest.mark.parametrize("n,nlev,natoms,R", [(33, 3, 40, 6.0), (65, 4, 200, 14.0)])
def test_pcgdriv_parity(mg, orc, n, nlev, natoms, R):
    print(f"pcg parity n={n} nlev={nlev}", flush=True)
    c, inp = orc.config_inputs("pb65", n=n, natoms=natoms, sphere_r=R)
    e1, u1, iw1, rw1 = _drive(mg, inp, n, nlev)
    it1, ie1, rr1 = mg.last_stats()
    e0, u0, iw0, rw0 = orc.pcg(inp, n, nlev)
    it0, ie0, rr0, _ = orc.last_stats()
    print(f"  iters gpu={it1} oracle={it0} relres gpu={rr1:.3e} oracle={rr0:.3e} "
          f"potential err={_rel(u1, u0):.2e}", flush=True)
    assert e1 == e0 == 0 and it1 == it0
    assert rr1 <= 1e-6
    assert _rel(u1, u0) < 1e-10
    assert np.array_equal(iw1, iw0)
    assert _rel(rw1, rw0) < 1e-12


def test_pcg_counting_criterion(mg, orc):
    """num_iters > 0 runs exactly that many CG iterations, as the
    reference's counting_criterion (CG_Solver.h:38-43)."""
    print("pcg counting n=33", flush=True)
    n, nlev = 33, 3
    c, inp = orc.config_inputs("pb65", n=n, natoms=40, sphere_r=6.0)
    e1, u1, _, _ = _drive(mg, inp, n, nlev, num_iters=3)
    it1 = mg.last_stats()[0]
    e0, u0, _, _ = orc.pcg(inp, n, nlev, num_iters=3)
    print(f"  iters gpu={it1} err={_rel(u1, u0):.2e}", flush=True)
    assert it1 == 3 and orc.last_stats()[0] == 3
    assert _rel(u1, u0) < 1e-10


def test_ctx_pcg_fp64_and_mixed(mg, orc):
    print("ctx pcg fp64 + mixed n=65", flush=True)
    n, nlev = 65, 4
    c, inp = orc.config_inputs("pb65", n=n, natoms=200, sphere_r=14.0)
    iparm, rparm = mg.make_params(n, nlev)
    ctx = mg.Context(iparm, rparm, inp)
    ctx.pcg(mg.PMG_FP64)
    u64, s64 = ctx.u(), ctx.stats()
    ctx.pcg(mg.PMG_MIXED)
    umx, smx = ctx.u(), ctx.stats()
    ctx.solve(mg.PMG_FP64)  # the PMG cycle on the same context still matches
    upmg = ctx.u()
    ctx.close()
    _, u0, _, _ = orc.pcg(inp, n, nlev)
    it0 = orc.last_stats()[0]
    _, up0, _, _ = orc.solve(inp, n, nlev)
    print(f"  fp64 iters={s64['iters']} mixed iters={smx['iters']} oracle={it0} "
          f"fp64 err={_rel(u64, u0):.2e} mixed err={_rel(umx, u0):.2e}", flush=True)
    assert s64["iters"] == it0 and _rel(u64, u0) < 1e-10
    assert smx["relres"] <= 1e-6 and smx["ierror"] == 0
    assert _rel(umx, u0) < 1e-6
    assert _rel(upmg, up0) < 1e-10


def test_pcg_needs_two_levels(mg, orc):
    print("pcg nlev=1 refused", flush=True)
    n = 9
    c, inp = orc.config_inputs("pb65", n=n, natoms=6, sphere_r=2.5, h=1.0)
    with pytest.raises(mg.PMGError) as ex:
        _drive(mg, inp, n, 1)
    assert ex.value.code == mg.PMG_EBADPARM

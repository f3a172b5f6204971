"""GPU parity of the banded coarsest solve (iparm(21) = mgsolv = 1):
Vbuildband + Vdpbfa at setup and Vdpbsl per V-cycle (pmgc/buildBd.cc,
pmgc/mlinpckd.cc, pmgc/mgcsd.cc:391-407) against the CPU oracle.

Tolerances: the factored band in rwork 1e-12 relative to its largest entry
(the GPU factors right-looking, the oracle left-looking as Vdpbfa: same
factor, different rounding); potential 1e-10 relative max-norm with the
same V-cycle count; iwork identical.  The band matrix needs to be positive
definite for the reference to proceed, which with the reference's mirrored
diagonal-direction planes holds for a uniform dielectric, or a coarsest
level of 5^3 or smaller, on these molecules."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    s = max(np.max(np.abs(b)), 1e-300)
    return np.max(np.abs(a - b)) / s


def _inputs(orc, mg, n, natoms, R, eps_in, ionic=0.15, seed=3):
    return orc.pbe_assemble(n, 0.5, ionic, mg.pbe_molecule(natoms, R, 0, seed), eps_in=eps_in)


def _drive(mg, fn, inp, n, nlev, nonlin=False, **kw):
    iparm, rparm = mg.make_params(n, nlev, nonlin=1 if nonlin else 0, mgsolv=1, **kw)
    sz = mg.Vmgsz(2, 0, 1, n, n, n, nlev)
    iw = np.zeros(sz["iintot"], np.int32)
    rw = np.zeros(int(iparm[0]))
    u = np.zeros(n ** 3)
    a = {k: inp[k].copy() for k in ("xf", "yf", "zf", "gxcf", "gycf", "gzcf", "a1cf",
                                     "a2cf", "a3cf", "ccf", "fcf", "tcf")}
    ierr = fn(iparm, rparm, iw, rw, u, a["xf"], a["yf"], a["zf"], a["gxcf"], a["gycf"],
              a["gzcf"], a["a1cf"], a["a2cf"], a["a3cf"], a["ccf"], a["fcf"], a["tcf"])
    return ierr, u, iw, rw


def _band(mg, n, nlev, iw, rw, newton=False):
    sz = mg.Vmgsz(2, 0, 1, n, n, n, nlev)
    niz = sz["n_iz"]
    iz = iw[:niz].reshape(nlev + 1, 50)
    ipcb = iw[niz + iz[nlev, 4] - 1: niz + iz[nlev, 4] - 1 + 4]
    acb = sz["n_rpc"] + 2 * sz["narr"] + (2 * sz["nf"] if newton else 0) + 27 * sz["narrc"]
    o = acb + iz[nlev, 6] - 1
    return ipcb.copy(), rw[o: o + int(ipcb[0]) * int(ipcb[2])]


# (n, nlev, natoms, R, eps_in): coarsest 9^3 (one-slot window), 5^3, 4^3
# (colliding band offsets), 11^3 and 12^3 (two-slot windows)
CASES = [(33, 3, 40, 6.0, 78.54), (33, 4, 40, 6.0, 2.0), (25, 4, 20, 4.0, 2.0),
         (41, 3, 60, 8.0, 78.54), (45, 3, 80, 9.0, 78.54)]


@pytest.mark.parametrize("n,nlev,natoms,R,eps_in", CASES)
def test_mgdriv_band_parity(mg, orc, n, nlev, natoms, R, eps_in):
    print(f"mgdriv mgsolv=1 n={n} nlev={nlev} eps_in={eps_in}", flush=True)
    inp = _inputs(orc, mg, n, natoms, R, eps_in)
    e1, u1, iw1, rw1 = _drive(mg, mg.Vmgdriv, inp, n, nlev)
    it1 = mg.last_stats()[0]
    e0, u0, iw0, rw0 = orc.solve(inp, n, nlev, mgsolv=1)
    it0 = orc.last_stats()[0]
    ipc1, b1 = _band(mg, n, nlev, iw1, rw1)
    ipc0, b0 = _band(mg, n, nlev, iw0, rw0)
    print(f"  iters gpu={it1} oracle={it0} potential err={_rel(u1, u0):.2e} "
          f"band err={_rel(b1, b0):.2e} ipc={list(ipc1)}", flush=True)
    assert e1 == e0 == 0 and it1 == it0
    assert np.array_equal(iw1, iw0) and list(ipc1) == list(ipc0)
    assert ipc1[3] == 1
    assert _rel(b1, b0) < 1e-12
    assert _rel(u1, u0) < 1e-10


@pytest.mark.parametrize("ionic", [0.0, 0.15])
def test_one_level_direct_solve(mg, orc, ionic):
    """nlev = 1: the fine 7-point system solved directly through its band
    (mgcsd.cc:195-212); cc is left out of the band as in the reference."""
    print(f"mgsolv=1 nlev=1 n=9 ionic={ionic}", flush=True)
    n = 9
    inp = orc.pbe_assemble(n, 1.0, ionic, mg.pbe_molecule(6, 2.5, 0, 3))
    e1, u1, iw1, rw1 = _drive(mg, mg.Vmgdriv, inp, n, 1)
    e0, u0, iw0, rw0 = orc.solve(inp, n, 1, mgsolv=1)
    print(f"  potential err={_rel(u1, u0):.2e}", flush=True)
    assert e1 == e0 == 0
    assert _rel(u1, u0) < 1e-12


def test_indefinite_band_reports_enotspd(mg, orc):
    """Dielectric jump at 33^3 / 3 levels: Vdpbfa meets a non-positive pivot
    (the reference prints and exits, mgsubd.cc:255-260): PMG_ENOTSPD from
    the driver and from a resident context."""
    print("mgsolv=1 indefinite band", flush=True)
    n, nlev = 33, 3
    inp = _inputs(orc, mg, n, 40, 6.0, 2.0, seed=7)
    assert orc.solve(inp, n, nlev, mgsolv=1)[0] == 3
    with pytest.raises(mg.PMGError) as ex:
        _drive(mg, mg.Vmgdriv, inp, n, nlev)
    assert ex.value.code == mg.PMG_ENOTSPD
    iparm, rparm = mg.make_params(n, nlev, mgsolv=1)
    with pytest.raises(mg.PMGError) as ex:
        mg.Context(iparm, rparm, inp)
    assert ex.value.code == mg.PMG_ENOTSPD


def test_ctx_fp64_and_mixed_band(mg, orc):
    print("ctx mgsolv=1 fp64 + mixed n=65 nlev=4", flush=True)
    n, nlev = 65, 4
    inp = _inputs(orc, mg, n, 200, 14.0, 78.54)
    iparm, rparm = mg.make_params(n, nlev, mgsolv=1)
    ctx = mg.Context(iparm, rparm, inp)
    ctx.solve(mg.PMG_FP64)
    u64, s64 = ctx.u(), ctx.stats()
    ctx.solve(mg.PMG_MIXED)
    umx, smx = ctx.u(), ctx.stats()
    ctx.close()
    _, u0, _, _ = orc.solve(inp, n, nlev, mgsolv=1)
    it0 = orc.last_stats()[0]
    print(f"  fp64 iters={s64['iters']} mixed iters={smx['iters']} oracle={it0} "
          f"fp64 err={_rel(u64, u0):.2e} mixed err={_rel(umx, u0):.2e}", flush=True)
    assert s64["iters"] == it0 and _rel(u64, u0) < 1e-10
    assert smx["relres"] <= 1e-6 and smx["ierror"] == 0
    assert _rel(umx, u0) < 1e-6


def test_newdriv_band_parity(mg, orc):
    print("newdriv mgsolv=1 n=33", flush=True)
    n, nlev = 33, 3
    ions = ([1.0, -1.0], [0.5, 0.5])
    mg.Vmypdefinitnpbe(2, *ions)
    orc.mypdefinit(*ions)
    inp = _inputs(orc, mg, n, 40, 6.0, 78.54, ionic=0.3, seed=5)
    e1, u1, iw1, rw1 = _drive(mg, mg.Vnewdriv, inp, n, nlev, nonlin=True, errtol=1e-8)
    it1 = mg.last_stats()[0]
    e0, u0, iw0, rw0 = orc.solve(inp, n, nlev, nonlin=True, errtol=1e-8, mgsolv=1)
    it0 = orc.last_stats()[0]
    ipc1, b1 = _band(mg, n, nlev, iw1, rw1, newton=True)
    ipc0, b0 = _band(mg, n, nlev, iw0, rw0, newton=True)
    print(f"  newton iters gpu={it1} oracle={it0} err={_rel(u1, u0):.2e} "
          f"band err={_rel(b1, b0):.2e}", flush=True)
    assert e1 == e0 == 0 and it1 == it0
    assert _rel(b1, b0) < 1e-12
    assert _rel(u1, u0) < 1e-9


def test_band_width_limit(mg, orc):
    """A 13^3 coarsest level (half-bandwidth 133) is beyond the band
    factor's window of 128: refused with PMG_EBADPARM before any work."""
    print("mgsolv=1 13^3 coarsest refused", flush=True)
    n = 49
    inp = _inputs(orc, mg, n, 80, 9.0, 78.54)
    with pytest.raises(mg.PMGError) as ex:
        _drive(mg, mg.Vmgdriv, inp, n, 3)
    assert ex.value.code == mg.PMG_EBADPARM

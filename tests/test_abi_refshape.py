"""The reference-shaped entry points: pmg_Vmgsz (pmgc/mgdrvd.h:236-255),
pmg_Vpackmg (pmgc/mgsubd.h:273), pmg_Vmypdefinitlpbe / npbe
(pmgc/mypdec.h:89-121) take the reference's pointer lists, so a call site
written for pmgc binds them unchanged.  The GPU test drives complete solves
through those calls and pmg_mgdriv / pmg_newdriv only, as a maintainer's
ctypes (or C) binding would."""
import ctypes

import numpy as np
import pytest

C_INT = ctypes.c_int
_I = ctypes.POINTER(ctypes.c_int)
_D = ctypes.POINTER(ctypes.c_double)


def _ref_vmgsz(L, mgcoar, mgdisc, mgsolv, nx, ny, nz, nlev):
    ins = [C_INT(v) for v in (mgcoar, mgdisc, mgsolv, nx, ny, nz, nlev)]
    outs = [C_INT(-7) for _ in range(12)]
    rc = L.pmg_Vmgsz(*(ctypes.byref(v) for v in ins), *(ctypes.byref(v) for v in outs))
    return rc, [v.value for v in outs]


def _ref_vpackmg(L, nrwk, niwk, nx, ny, nz, nlev, nu1=2, nu2=2, mgkey=0, itmax=200,
                 istop=1, ipcon=0, nonlin=0, mgsmoo=1, mgprol=1, mgcoar=2, mgsolv=0,
                 mgdisc=0, iinfo=0, errtol=1e-6, ipkey=0, omegal=0.8, omegan=0.9,
                 irite=8, iperf=0):
    iparm = np.full(22, -1, np.int32)
    rparm = np.zeros(10)
    ints = [C_INT(v) for v in (niwk, nx, ny, nz, nlev, nu1, nu2, mgkey, itmax, istop,
                               ipcon, nonlin, mgsmoo, mgprol, mgcoar, mgsolv, mgdisc, iinfo)]
    nr = ctypes.c_size_t(nrwk)
    et, ol, on = ctypes.c_double(errtol), ctypes.c_double(omegal), ctypes.c_double(omegan)
    ik, ir, ip_ = C_INT(ipkey), C_INT(irite), C_INT(iperf)
    rc = L.pmg_Vpackmg(iparm.ctypes.data_as(_I), rparm.ctypes.data_as(_D), ctypes.byref(nr),
                       *(ctypes.byref(v) for v in ints), ctypes.byref(et), ctypes.byref(ik),
                       ctypes.byref(ol), ctypes.byref(on), ctypes.byref(ir),
                       ctypes.byref(ip_))
    return rc, iparm, rparm


def test_vmgsz_pointer_form(mg):
    L = mg.lib()
    for (n, nlev, mgsolv) in [(65, 4, 0), (161, 5, 0), (289, 6, 0), (33, 3, 1), (97, 5, 1)]:
        rc, out = _ref_vmgsz(L, 2, 0, mgsolv, n, n, n, nlev)
        s = mg.Vmgsz(2, 0, mgsolv, n, n, n, nlev)
        assert rc == 0
        assert out == [s[k] for k in ("nxc", "nyc", "nzc", "nf", "nc", "narr", "narrc",
                                      "n_rpc", "n_iz", "n_ipc", "iretot", "iintot")]
    # the reference prints "invalid mgdisc parameter" (mgdrvd.cc:561)
    assert _ref_vmgsz(L, 2, 5, 0, 33, 33, 33, 3)[0] == mg.PMG_EBADPARM
    assert _ref_vmgsz(L, 3, 0, 0, 33, 33, 33, 3)[0] == mg.PMG_EBADPARM
    assert _ref_vmgsz(L, 2, 0, 4, 33, 33, 33, 3)[0] == mg.PMG_EBADPARM


def test_vpackmg_pointer_form(mg):
    L = mg.lib()
    rc, iparm, rparm = _ref_vpackmg(L, 1234567, 890, 33, 25, 17, 3, nu1=1, nu2=3, itmax=77,
                                    istop=0, mgsolv=1, errtol=2.5e-7, iinfo=2)
    ip2, rp2 = mg.Vpackmg(1234567, 890, 33, 25, 17, 3, nu1=1, nu2=3, itmax=77, istop=0,
                          mgsolv=1, errtol=2.5e-7, iinfo=2)
    assert rc == 0
    assert np.array_equal(iparm, ip2) and np.array_equal(rparm, rp2)
    # VAT(iparm, k) positions of mgsubd.cc:549-570
    assert list(iparm[[0, 1, 2, 3, 4, 5, 6, 7, 9, 10, 11, 20]]) == \
        [1234567, 890, 33, 25, 17, 3, 1, 3, 77, 0, 2, 1]
    assert rparm[0] == 2.5e-7 and rparm[8] == 0.8 and rparm[9] == 0.9
    # nrwk is size_t in the reference; iparm(1) holds an int
    assert _ref_vpackmg(L, 1 << 33, 10, 9, 9, 9, 2)[0] == mg.PMG_EBADPARM


def _solve_ref_shaped(mg, inp, n, nlev, nonlin, errtol):
    """A complete solve through the reference-shaped calls only."""
    L = mg.lib()
    rc, sz = _ref_vmgsz(L, 2, 0, 0, n, n, n, nlev)
    assert rc == 0
    nf, iretot, iintot = sz[3], sz[10], sz[11]
    nrwk = iretot + (2 * nf if nonlin else 0)  # newdrvd.cc:123
    rc, iparm, rparm = _ref_vpackmg(L, nrwk, iintot, n, n, n, nlev, nonlin=nonlin,
                                    errtol=errtol)
    assert rc == 0
    iwork = np.zeros(iintot, np.int32)
    rwork = np.zeros(nrwk)
    u = np.zeros(n ** 3)
    a = {k: inp[k].copy() for k in mg.PBE_KEYS}
    fn = L.pmg_newdriv if nonlin else L.pmg_mgdriv
    rc = fn(iparm.ctypes.data_as(_I), rparm.ctypes.data_as(_D),
            iwork.ctypes.data_as(_I), rwork.ctypes.data_as(_D), u.ctypes.data_as(_D),
            *(a[k].ctypes.data_as(_D) for k in mg.PBE_KEYS))
    return rc, u, iwork, rwork


@pytest.mark.gpu
def test_solves_through_reference_shaped_calls(mg, orc):
    n, nlev = 33, 3
    inp = orc.pbe_synth(n, 0.5, 40, 6.0, 0.3, 0, 5)
    print("reference-shaped linear solve", flush=True)
    rc, u, iw, rw = _solve_ref_shaped(mg, inp, n, nlev, 0, 1e-6)
    e0, u0, iw0, rw0 = orc.solve(inp, n, nlev)
    assert rc == e0 == 0
    assert np.max(np.abs(u - u0)) <= 1e-10 * np.max(np.abs(u0))
    assert np.array_equal(iw, iw0)
    assert np.max(np.abs(rw - rw0)) <= 1e-12 * np.max(np.abs(rw0))
    print("reference-shaped Newton solve", flush=True)
    nion = C_INT(2)
    charge = np.array([1.0, -1.0])
    sconc = np.array([0.5, 0.5])
    assert mg.lib().pmg_Vmypdefinitnpbe(ctypes.byref(nion), charge.ctypes.data_as(_D),
                                        sconc.ctypes.data_as(_D)) == 0
    orc.mypdefinit(charge, sconc)
    rc, u, _, _ = _solve_ref_shaped(mg, inp, n, nlev, 1, 1e-8)
    it = mg.last_stats()[0]
    e0, u0, _, _ = orc.solve(inp, n, nlev, nonlin=True, errtol=1e-8)
    assert rc == e0 == 0 and it == orc.last_stats()[0]
    assert np.max(np.abs(u - u0)) <= 1e-9 * np.max(np.abs(u0))
    # the linear ion table call (Vmypdefinitlpbe) sets the same table
    assert mg.lib().pmg_Vmypdefinitlpbe(ctypes.byref(nion), charge.ctypes.data_as(_D),
                                        sconc.ctypes.data_as(_D)) == 0
    rc2, u2, _, _ = _solve_ref_shaped(mg, inp, n, nlev, 1, 1e-8)
    assert rc2 == 0 and np.array_equal(u2, u)

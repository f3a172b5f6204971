"""ctypes binding of the CPU oracle (oracle/libpmg_oracle.so).

Test infrastructure only: the oracle is the checker for the HIP library,
never part of the product path.  Function names follow the reference
routines each oracle function restates (see oracle/pmg_oracle.h).
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OLIB = os.path.join(ROOT, "oracle", "libpmg_oracle.so")
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_o = None


def olib():
    global _o
    if _o is None:
        if not os.path.exists(OLIB):
            subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
        L = ctypes.CDLL(OLIB)
        i, d, dp, ip = ctypes.c_int, ctypes.c_double, _D, _I
        sig = {
            "pmgo_mgsz": (i, [i] * 7 + [ip]),
            "pmgo_maxlev": (i, [i, i, i]),
            "pmgo_buildstr": (None, [i, i, i, i, ip]),
            "pmgo_buildA_fv": (None, [i, i, i, i, ip] + [dp] * 14),
            "pmgo_buildP": (None, [i] * 8 + [dp, dp]),
            "pmgo_buildG": (None, [i] * 7 + [dp, dp, dp]),
            "pmgo_matvec": (None, [i, i, i, i, dp, dp, dp, dp]),
            "pmgo_mresid": (None, [i, i, i, i, dp, dp, dp, dp, dp]),
            "pmgo_nmresid": (None, [i, i, i, i, dp, dp, dp, dp, dp, dp]),
            "pmgo_gsrb": (None, [i, i, i, i, dp, dp, dp, dp, i, i]),
            "pmgo_restrc": (None, [i] * 6 + [dp, dp, dp]),
            "pmgo_interp": (None, [i] * 6 + [dp, dp, dp]),
            "pmgo_extrac": (None, [i] * 6 + [dp, dp]),
            "pmgo_xnrm1": (d, [i, i, i, dp]),
            "pmgo_xnrm2": (d, [i, i, i, dp]),
            "pmgo_xdot": (d, [i, i, i, dp, dp]),
            "pmgo_cghs": (i, [i, i, i, i, dp, dp, dp, dp, dp, dp, dp, i, d]),
            "pmgo_buildband": (i, [i, i, i, i, dp, dp, ip]),
            "pmgo_dpbfa": (None, [dp, i, i, i, ip]),
            "pmgo_dpbsl": (None, [dp, i, i, i, dp]),
            "pmgo_bandsolve": (None, [i, i, i, dp, ip, dp, dp, dp]),
            "pmgo_pcgdriv": (i, [ip, dp, ip, dp, dp] + [dp] * 12 + [i]),
            "pmgo_cg_class": (i, [i] * 4 + [dp] * 7 + [d, i]),
            "pmgo_mypdefinit": (None, [i, dp, dp]),
            "pmgo_c_vec": (None, [dp, dp, dp, i, i, i]),
            "pmgo_dc_vec": (None, [dp, dp, dp, i, i, i]),
            "pmgo_mgdriv": (i, [ip, dp, ip, dp, dp] + [dp] * 12),
            "pmgo_newdriv": (i, [ip, dp, ip, dp, dp] + [dp] * 12),
            "pmgo_last_stats": (i, [ip, ip, dp, dp, i]),
            "pbo_constants": (None, [d, dp, dp]),
            "pbo_assemble": (None, [i, d, d, i, dp] + [dp] * 12),
            "pbo_assemble_ex": (None, [i, d, d, d, d, d, i, dp] + [dp] * 12),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _o = L
    return _o


def dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def ip(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_I)


def mgsz(mgcoar, mgdisc, mgsolv, nx, ny, nz, nlev):
    out = np.zeros(12, np.int32)
    rc = olib().pmgo_mgsz(mgcoar, mgdisc, mgsolv, nx, ny, nz, nlev, ip(out))
    return rc, out


def buildA_fv(n, inp):
    N = n ** 3
    ac, cc, fc = np.zeros(4 * N), np.zeros(N), np.zeros(N)
    ipc = np.zeros(100, np.int32)
    olib().pmgo_buildA_fv(n, n, n, 0, ip(ipc), dp(ac), dp(cc), dp(fc),
                          *(dp(inp[k]) for k in ("xf", "yf", "zf", "gxcf", "gycf",
                                                 "gzcf", "a1cf", "a2cf", "a3cf",
                                                 "ccf", "fcf")))
    return ac, cc, fc, ipc


def buildP(nf, nc, mgprol, numdia, acf):
    pc = np.zeros(27 * nc ** 3)
    olib().pmgo_buildP(nf, nf, nf, nc, nc, nc, mgprol, numdia, dp(acf), dp(pc))
    return pc


def buildG(nf, nc, numdia, pc, acf):
    acc = np.zeros(14 * nc ** 3)
    olib().pmgo_buildG(nf, nf, nf, nc, nc, nc, numdia, dp(pc), dp(acf), dp(acc))
    return acc


def matvec(n, numdia, ac, cc, x):
    y = np.zeros(n ** 3)
    olib().pmgo_matvec(n, n, n, numdia, dp(ac), dp(cc), dp(x), dp(y))
    return y


def mresid(n, numdia, ac, cc, fc, x):
    r = np.zeros(n ** 3)
    olib().pmgo_mresid(n, n, n, numdia, dp(ac), dp(cc), dp(fc), dp(x), dp(r))
    return r


def nmresid(n, numdia, ac, cc, fc, x):
    r, w = np.zeros(n ** 3), np.zeros(n ** 3)
    olib().pmgo_nmresid(n, n, n, numdia, dp(ac), dp(cc), dp(fc), dp(x), dp(r), dp(w))
    return r


def gsrb(n, numdia, ac, cc, fc, x, itmax, iadjoint, dims=None):
    """dims = (nx, ny, nz) for a non-cubic grid (default n^3)."""
    nx, ny, nz = dims if dims else (n, n, n)
    x = x.copy()
    olib().pmgo_gsrb(nx, ny, nz, numdia, dp(ac), dp(cc), dp(fc), dp(x), itmax, iadjoint)
    return x


def restrc(nf, nc, xin, pc):
    out = np.zeros(nc ** 3)
    olib().pmgo_restrc(nf, nf, nf, nc, nc, nc, dp(xin), dp(out), dp(pc))
    return out


def interp(nc, nf, xin, pc):
    out = np.zeros(nf ** 3)
    olib().pmgo_interp(nc, nc, nc, nf, nf, nf, dp(xin), dp(out), dp(pc))
    return out


def cghs(n, numdia, ac, cc, fc, itmax=100, errtol=0.0):
    N = n ** 3
    x, p, ap, r = np.zeros(N), np.zeros(N), np.zeros(N), np.zeros(N)
    it = olib().pmgo_cghs(n, n, n, numdia, dp(ac), dp(cc), dp(fc), dp(x), dp(p),
                          dp(ap), dp(r), itmax, errtol)
    return x, it


def band_dims(n, numdia):
    """(n_unknowns, m, lda) of the band of an n^3 level (buildBd.cc:264-285)."""
    s = n - 2
    m = s * s if numdia == 7 else s * s + s + 1
    return s ** 3, m, m + 1


def buildband(n, numdia, ac):
    """Factored band of an n^3 level; returns (info, abd (lda, nu) view, ipcb)."""
    nu, m, lda = band_dims(n, numdia)
    abd = np.zeros(lda * nu)
    ipcb = np.zeros(4, np.int32)
    info = olib().pmgo_buildband(n, n, n, numdia, dp(np.ascontiguousarray(ac, np.float64)),
                                 dp(abd), ip(ipcb))
    return info, abd, ipcb


def dpbsl(abd, lda, nu, m, b):
    x = np.ascontiguousarray(b, np.float64).copy()
    olib().pmgo_dpbsl(dp(abd), lda, nu, m, dp(x))
    return x


def bandsolve(n, abd, ipcb, rhs):
    x = np.full(n ** 3, 7.0)
    wk = np.zeros(int(ipcb[0]))
    olib().pmgo_bandsolve(n, n, n, dp(abd), ip(ipcb), dp(np.ascontiguousarray(rhs)),
                          dp(x), dp(wk))
    return x


def mypdefinit(charge, sconc):
    charge = np.ascontiguousarray(charge, np.float64)
    sconc = np.ascontiguousarray(sconc, np.float64)
    olib().pmgo_mypdefinit(len(charge), dp(charge), dp(sconc))


def c_vec(coef, u, deriv=False):
    out = np.zeros_like(u)
    n = len(u)
    f = olib().pmgo_dc_vec if deriv else olib().pmgo_c_vec
    f(dp(coef), dp(u), dp(out), n, 1, 1)
    return out


def _drv(fn, iparm, rparm, iwork, rwork, u, xf, yf, zf, gxcf, gycf, gzcf,
         a1cf, a2cf, a3cf, ccf, fcf, tcf):
    return fn(ip(iparm), dp(rparm), ip(iwork), dp(rwork), dp(u), dp(xf), dp(yf),
              dp(zf), dp(gxcf), dp(gycf), dp(gzcf), dp(a1cf), dp(a2cf), dp(a3cf),
              dp(ccf), dp(fcf), dp(tcf))


def mgdriv(*args):
    return _drv(olib().pmgo_mgdriv, *args)


def newdriv(*args):
    return _drv(olib().pmgo_newdriv, *args)


def last_stats(cap=4096):
    it, ie, rr = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    hist = np.zeros(cap)
    m = olib().pmgo_last_stats(ctypes.byref(it), ctypes.byref(ie), ctypes.byref(rr),
                               dp(hist), cap)
    return it.value, ie.value, rr.value, hist[:m].copy()


def solve(inp, n, nlev, driver=None, nonlin=False, **kw):
    """Run a driver (oracle by default) on inputs; returns (ierror, u, iwork, rwork)."""
    import multigrid_amd as mg

    iparm, rparm = mg.make_params(n, nlev, nonlin=1 if nonlin else 0, **kw)
    sz = mgsz(2, 0, 0, n, n, n, nlev)[1]
    iw = np.zeros(int(sz[11]), np.int32)
    rw = np.zeros(int(iparm[0]))
    u = np.zeros(n ** 3)
    a = {k: np.ascontiguousarray(inp[k], np.float64).copy() for k in
         ("xf", "yf", "zf", "gxcf", "gycf", "gzcf", "a1cf", "a2cf", "a3cf",
          "ccf", "fcf", "tcf")}
    fn = driver or (newdriv if nonlin else mgdriv)
    ierr = fn(iparm, rparm, iw, rw, u, a["xf"], a["yf"], a["zf"], a["gxcf"],
              a["gycf"], a["gzcf"], a["a1cf"], a["a2cf"], a["a3cf"], a["ccf"],
              a["fcf"], a["tcf"])
    return ierr, u, iw, rw


PBE_KEYS = ("xf", "yf", "zf", "gxcf", "gycf", "gzcf", "a1cf", "a2cf", "a3cf",
            "ccf", "fcf", "tcf")


def pcg(inp, n, nlev, num_iters=0, **kw):
    """Class-style PCG (oracle) on inputs; returns (ierror, u, iwork, rwork)."""
    import multigrid_amd as mg

    iparm, rparm = mg.make_params(n, nlev, **kw)
    sz = mgsz(2, 0, int(iparm[20]), n, n, n, nlev)[1]
    iw = np.zeros(int(sz[11]), np.int32)
    rw = np.zeros(int(iparm[0]))
    u = np.zeros(n ** 3)
    a = {k: np.ascontiguousarray(inp[k], np.float64).copy() for k in PBE_KEYS}
    ierr = olib().pmgo_pcgdriv(ip(iparm), dp(rparm), ip(iw), dp(rw), dp(u),
                               *(dp(a[k]) for k in PBE_KEYS), int(num_iters))
    return ierr, u, iw, rw


def cg_class(n, numdia, ac, cc, f, x0, thresh=1e-8, itmax=100000):
    x = np.ascontiguousarray(x0, np.float64).copy()
    wk = [np.zeros(n ** 3) for _ in range(3)]
    it = olib().pmgo_cg_class(n, n, n, numdia, dp(ac), dp(cc), dp(np.ascontiguousarray(f)),
                              dp(x), dp(wk[0]), dp(wk[1]), dp(wk[2]), thresh, itmax)
    return it, x


def pbe_constants(ionic):
    lb, k2 = ctypes.c_double(), ctypes.c_double()
    olib().pbo_constants(ionic, ctypes.byref(lb), ctypes.byref(k2))
    return lb.value, k2.value


def pbe_assemble(n, h, ionic, atoms, eps_in=2.0, eps_out=78.54, ion_shell=2.0):
    """Oracle assembly of the fine-level inputs from an (natoms, 5) array."""
    at = np.ascontiguousarray(atoms, np.float64).reshape(-1)
    out = {k: np.zeros(n) for k in ("xf", "yf", "zf")}
    for k in ("gxcf", "gycf", "gzcf"):
        out[k] = np.zeros(4 * n * n)
    for k in ("a1cf", "a2cf", "a3cf", "ccf", "fcf", "tcf"):
        out[k] = np.zeros(n * n * n)
    olib().pbo_assemble_ex(n, h, ionic, eps_in, eps_out, ion_shell, at.size // 5, dp(at),
                           *(dp(out[k]) for k in PBE_KEYS))
    out["atoms"] = at.reshape(-1, 5)
    return out


def pbe_synth(n, h, natoms, sphere_r, ionic=0.15, charge_mode=0, seed=20261202):
    """Synthetic molecule from the library's (host-only) generator, assembled
    by the oracle: the inputs CPU-only tests and the CPU baseline use."""
    import multigrid_amd as mg
    return pbe_assemble(n, h, ionic, mg.pbe_molecule(natoms, sphere_r, charge_mode, seed))


def config_inputs(name, **over):
    import multigrid_amd as mg
    c = dict(mg.CONFIGS[name])
    c.update(over)
    return c, pbe_synth(c["n"], c["h"], c["natoms"], c["sphere_r"], c["ionic"],
                        c["charge_mode"], c["seed"])

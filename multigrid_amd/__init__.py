"""multigrid_amd -- MI355X-native PMG (APBS pmgc) multigrid solver.

Python host mirror of the reference's PMG entry points
(pmgc/mgdrvd.h:79-255, pmgc/newdrvd.h, pmgc/mgsubd.cc:538-576,
pmgc/mypdec.cc:72-106) over the C ABI of libpmg_mi355x.so
(include/pmg_mi355x.h).  Arrays are numpy, 1-based Fortran order
flattened (i fastest), exactly as the reference's pointers.

Names follow the reference: Vmgsz, Vpackmg, Vmgdriv, Vnewdriv,
Vmypdefinitlpbe / Vmypdefinitnpbe.  Errors that the reference reports with
printf/exit raise PMGError here; itmax exhaustion returns ierror = 1 as the
reference's Vmvcs does (pmgc/mgcsd.cc:529).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libpmg_mi355x.so")

(PMG_OK, PMG_ITMAX, PMG_EBADPARM, PMG_ENOTSPD, PMG_ENOMEM, PMG_EDEVICE, PMG_EBUSY,
 PMG_ESCRATCH) = range(8)
PMG_FP64, PMG_MIXED = 0, 1
_ERRS = {
    PMG_EBADPARM: "invalid or unsupported iparm/rparm or array shape",
    PMG_ENOTSPD: "coarse operator breakdown",
    PMG_ENOMEM: "device memory allocation failed",
    PMG_EDEVICE: "no usable GPU or HIP runtime error",
    PMG_EBUSY: "admission refused",
    PMG_ESCRATCH: "beyond the fixed scratch of an allocation-free call",
}


class PMGError(RuntimeError):
    def __init__(self, code, where=""):
        self.code = code
        super().__init__(f"{where}: {_ERRS.get(code, 'error')} (code {code})")


_lib = None
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def lib():
    """Load the HIP library; fails loudly if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ImportError(
                f"{LIBPATH} is missing: build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'`")
        L = ctypes.CDLL(LIBPATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    dp, ip, i, d = _D, _I, ctypes.c_int, ctypes.c_double
    fp = ctypes.POINTER(ctypes.c_float)
    vp = ctypes.c_void_p
    sig = {
        "pmg_mgsz": (i, [i, i, i, i, i, i, i, ip]),
        "pmg_Vmgsz": (i, [ip] * 19),
        "pmg_Vpackmg": (i, [ip, dp, ctypes.POINTER(ctypes.c_size_t)] + [ip] * 18
                        + [dp, ip, dp, dp, ip, ip]),
        "pmg_Vmypdefinitlpbe": (i, [ip, dp, dp]),
        "pmg_Vmypdefinitnpbe": (i, [ip, dp, dp]),
        "pmg_maxlev": (i, [i, i, i]),
        "pmg_mgdriv": (i, [ip, dp, ip, dp, dp] + [dp] * 12),
        "pmg_newdriv": (i, [ip, dp, ip, dp, dp] + [dp] * 12),
        "pmg_pcgdriv": (i, [ip, dp, ip, dp, dp] + [dp] * 12 + [i]),
        "pmg_ctx_pcg": (i, [vp, i, i]),
        "pmg_mypdefinit": (i, [i, dp, dp]),
        "pmg_last_stats": (i, [ip, ip, dp]),
        "pmg_last_history": (i, [dp, i]),
        "pmg_ctx_history": (i, [vp, dp, i]),
        "pmg_ctx_set_tru": (i, [vp, dp]),
        "pmg_ctx_create": (i, [ctypes.POINTER(vp), ip, dp] + [dp] * 11),
        "pmg_ctx_solve": (i, [vp, i]),
        "pmg_ctx_solve_group": (i, [ctypes.POINTER(vp), i, i]),
        "pmg_ctx_get_u": (i, [vp, dp]),
        "pmg_ctx_stats": (i, [vp, ip, ip, dp, dp, dp]),
        "pmg_ctx_rebuild": (i, [vp]),
        "pmg_ctx_bytes": (ctypes.c_size_t, [vp]),
        "pmg_ctx_destroy": (None, [vp]),
        "pmg_ctx_get_level": (i, [vp, i, dp, dp, dp, dp]),
        "pmg_ctx_ktime_enable": (i, [vp, i]),
        "pmg_ctx_ktime": (i, [vp, dp, ctypes.POINTER(ctypes.c_longlong)]),
        "pmg_ctx_set_graphs": (i, [vp, i]),
        "pmg_matvec": (i, [i, i, i, i, dp, dp, dp, dp]),
        "pmg_mresid": (i, [i, i, i, i, dp, dp, dp, dp, dp]),
        "pmg_gsrb": (i, [i, i, i, i, dp, dp, dp, dp, i, i]),
        "pmg_restrc": (i, [i, i, i, i, i, i, dp, dp, dp]),
        "pmg_interp": (i, [i, i, i, i, i, i, dp, dp, dp]),
        "pmg_correct": (i, [i] * 7 + [dp] * 6),
        "pmg_gsrb_zero": (i, [i, i, i, i, dp, dp, dp, dp, i, i]),
        "pmg_mresid_f32": (i, [i, i, i, i, fp, fp, fp, fp, fp]),
        "pmg_gsrb_f32": (i, [i, i, i, i, fp, fp, fp, fp, i, i, i]),
        "pmg_restrc_f32": (i, [i] * 6 + [fp, fp, fp]),
        "pmg_interp_f32": (i, [i] * 6 + [fp, fp, fp]),
        "pmg_correct_f32": (i, [i] * 7 + [fp] * 6),
        "pmg_mresid_mixed": (i, [i, i, i, dp, dp, dp, dp, fp, dp]),
        "pmg_addcorr_f32": (i, [i, i, i, fp, dp]),
        "pmg_mresid_mixed_noalloc": (i, [i, i, i] + [vp] * 7),
        "pmg_buildP": (i, [i, i, i, i, i, i, i, i, dp, dp]),
        "pmg_buildG": (i, [i, i, i, i, i, i, i, dp, dp, dp]),
        "pmg_pbe_synth": (i, [i, d, i, d, d, i, ctypes.c_uint64] + [dp] * 13),
        "pmg_pbe_molecule": (i, [i, d, i, ctypes.c_uint64, dp]),
        "pmg_pbe_assemble": (i, [i, d, d, i, dp] + [dp] * 12),
        "pmg_pbe_assemble_ex": (i, [i, d, d, d, d, d, i, dp] + [dp] * 12),
        "pmg_dx_write": (i, [ctypes.c_char_p, i, i, i, d, d, d, d, dp, ctypes.c_char_p]),
        "pmg_dx_read": (i, [ctypes.c_char_p, ip, dp, dp, dp, ctypes.c_size_t]),
        "pmg_ctx_create_pbe": (i, [ctypes.POINTER(vp), ip, dp, d, d, i, dp]),
        "pmg_ctx_load_pbe": (i, [vp, d, d, i, dp]),
        "pmg_server_create": (i, [ctypes.POINTER(vp), ctypes.c_size_t, i]),
        "pmg_server_submit": (i, [vp, ip, dp, i, d, d, i, dp, dp,
                                  ctypes.POINTER(ctypes.c_int64)]),
        "pmg_server_wait": (i, [vp, ctypes.c_int64, i, ip, dp, dp]),
        "pmg_server_stats": (i, [vp, ctypes.POINTER(ctypes.c_longlong), dp, dp, dp,
                                 ctypes.POINTER(ctypes.c_size_t), ip]),
        "pmg_server_reset_stats": (i, [vp]),
        "pmg_server_create_grouped": (i, [ctypes.POINTER(vp), ctypes.c_size_t, i, i]),
        "pmg_server_destroy": (None, [vp]),
        "pmg_version": (ctypes.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        if hasattr(L, name):
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args


def _dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def _ip(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_I)


def _check(rc, where):
    if rc >= PMG_EBADPARM:
        raise PMGError(rc, where)
    return rc


# ---------------------------------------------------------------------------
# Sizing and control arrays
# ---------------------------------------------------------------------------
def Vmgsz(mgcoar, mgdisc, mgsolv, nx, ny, nz, nlev):
    """Workspace sizes (pmgc/mgdrvd.cc:505-601) as a dict."""
    out = np.zeros(12, np.int32)
    _check(lib().pmg_mgsz(mgcoar, mgdisc, mgsolv, nx, ny, nz, nlev, _ip(out)), "Vmgsz")
    keys = ["nxc", "nyc", "nzc", "nf", "nc", "narr", "narrc", "n_rpc", "n_iz",
            "n_ipc", "iretot", "iintot"]
    return dict(zip(keys, (int(v) for v in out)))


def Vmaxlev(nx, ny, nz):
    """Deepest admissible level count (pmgc/mgsubd.cc:456-484)."""
    return int(lib().pmg_maxlev(nx, ny, nz))


def Vpackmg(nrwk, niwk, nx, ny, nz, nlev, nu1=2, nu2=2, mgkey=0, itmax=200,
            istop=1, ipcon=0, nonlin=0, mgsmoo=1, mgprol=1, mgcoar=2,
            mgsolv=0, mgdisc=0, iinfo=0, errtol=1e-6, ipkey=0, omegal=0.8,
            omegan=0.9, irite=8, iperf=0):
    """Pack iparm (22 ints) and rparm (10 doubles) (pmgc/mgsubd.cc:538-576)."""
    iparm = np.zeros(22, np.int32)
    rparm = np.zeros(10, np.float64)
    vals = [nrwk, niwk, nx, ny, nz, nlev, nu1, nu2, mgkey, itmax, istop, iinfo,
            irite, ipkey, ipcon, nonlin, mgprol, mgcoar, mgdisc, mgsmoo, mgsolv,
            iperf]
    iparm[:] = vals
    rparm[0], rparm[8], rparm[9] = errtol, omegal, omegan
    return iparm, rparm


def make_params(n, nlev, nonlin=0, **kw):
    """iparm/rparm for an n^3 grid with workspace sizes filled in."""
    sz = Vmgsz(kw.get("mgcoar", 2), kw.get("mgdisc", 0), kw.get("mgsolv", 0), n, n, n, nlev)
    nrwk = sz["iretot"] + (2 * sz["nf"] if nonlin else 0)
    return Vpackmg(nrwk, sz["iintot"], n, n, n, nlev, nonlin=nonlin, **kw)


# ---------------------------------------------------------------------------
# Drop-in drivers
# ---------------------------------------------------------------------------
def _driver(fn, name, iparm, rparm, iwork, rwork, u, xf, yf, zf, gxcf, gycf,
            gzcf, a1cf, a2cf, a3cf, ccf, fcf, tcf):
    rc = fn(_ip(iparm), _dp(rparm), _ip(iwork), _dp(rwork), _dp(u), _dp(xf),
            _dp(yf), _dp(zf), _dp(gxcf), _dp(gycf), _dp(gzcf), _dp(a1cf),
            _dp(a2cf), _dp(a3cf), _dp(ccf), _dp(fcf), _dp(tcf))
    return _check(rc, name)


def Vmgdriv(iparm, rparm, iwork, rwork, u, xf, yf, zf, gxcf, gycf, gzcf,
            a1cf, a2cf, a3cf, ccf, fcf, tcf):
    """Linear PMG solve (pmgc/mgdrvd.cc:60-171) on the GPU, FP64.
    Returns ierror (0 converged, 1 itmax reached)."""
    return _driver(lib().pmg_mgdriv, "Vmgdriv", iparm, rparm, iwork, rwork, u,
                   xf, yf, zf, gxcf, gycf, gzcf, a1cf, a2cf, a3cf, ccf, fcf, tcf)


def Vnewdriv(iparm, rparm, iwork, rwork, u, xf, yf, zf, gxcf, gycf, gzcf,
             a1cf, a2cf, a3cf, ccf, fcf, tcf):
    """Nonlinear PB Newton solve (pmgc/newdrvd.cc:60-300) on the GPU, FP64.
    The interior of u is the starting iterate (its boundary is zeroed);
    iparm[8] = mgkey 1 runs the nested iteration (Vfnewton,
    newtond.cc:60-138) from the coarsest level up."""
    return _driver(lib().pmg_newdriv, "Vnewdriv", iparm, rparm, iwork, rwork,
                   u, xf, yf, zf, gxcf, gycf, gzcf, a1cf, a2cf, a3cf, ccf, fcf, tcf)


def Vpcgdriv(iparm, rparm, iwork, rwork, u, xf, yf, zf, gxcf, gycf, gzcf,
             a1cf, a2cf, a3cf, ccf, fcf, tcf, num_iters=0):
    """The reference's class-style solver (PBE_pmgc_class.h:1070-1133,
    CG_Solver.h:45-242): CG preconditioned by the class V-cycle, on the
    Vmgdriv hierarchy.  num_iters > 0 runs that many CG iterations; 0 runs
    to rparm[0] (relative L1 residual) or iparm[9] iterations."""
    rc = lib().pmg_pcgdriv(_ip(iparm), _dp(rparm), _ip(iwork), _dp(rwork), _dp(u),
                           _dp(xf), _dp(yf), _dp(zf), _dp(gxcf), _dp(gycf), _dp(gzcf),
                           _dp(a1cf), _dp(a2cf), _dp(a3cf), _dp(ccf), _dp(fcf), _dp(tcf),
                           int(num_iters))
    return _check(rc, "Vpcgdriv")


def Vmypdefinitlpbe(nion, charge, sconc):
    """Ion table (pmgc/mypdec.cc:72-88)."""
    charge = np.ascontiguousarray(charge, np.float64)
    sconc = np.ascontiguousarray(sconc, np.float64)
    return _check(lib().pmg_mypdefinit(int(nion), _dp(charge), _dp(sconc)),
                  "Vmypdefinitlpbe")


Vmypdefinitnpbe = Vmypdefinitlpbe


def last_stats():
    it, ie, rr = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    lib().pmg_last_stats(ctypes.byref(it), ctypes.byref(ie), ctypes.byref(rr))
    return it.value, ie.value, rr.value


def last_history(cap=4096):
    """Relative stopping measure after each V-cycle of the last Vmgdriv call
    on this thread (the reference's Vprtstp report, mgsubd.cc:486-536)."""
    out = np.zeros(cap)
    m = lib().pmg_last_history(_dp(out), cap)
    return out[:min(m, cap)].copy()


# ---------------------------------------------------------------------------
# Resident-hierarchy context
# ---------------------------------------------------------------------------
def solve_group(contexts, mode=PMG_FP64):
    """Solve several resident contexts together (pmg_ctx_solve_group): their
    iterations run in lockstep, each kernel launched once for all of them;
    every result is bit-identical to that context's own solve."""
    arr = (ctypes.c_void_p * len(contexts))(*[c._h.value for c in contexts])
    return _check(lib().pmg_ctx_solve_group(arr, len(contexts), mode),
                  "pmg_ctx_solve_group")


class Context:
    """A solve whose inputs and hierarchy stay resident in HBM."""

    def __init__(self, iparm, rparm, inp=None, pbe=None):
        """inp: dict of host input arrays (pmg_ctx_create), or
        pbe=(h, ionic_strength, atoms): inputs assembled on the GPU from an
        (natoms, 5) atom array (pmg_ctx_create_pbe)."""
        self._h = ctypes.c_void_p()
        L = lib()
        self.n = int(iparm[2]), int(iparm[3]), int(iparm[4])
        self.nlev = int(iparm[5])
        if pbe is not None:
            h, ionic, atoms = pbe
            at = np.ascontiguousarray(atoms, np.float64).reshape(-1)
            rc = L.pmg_ctx_create_pbe(ctypes.byref(self._h), _ip(iparm), _dp(rparm),
                                      h, ionic, at.size // 5, _dp(at))
            _check(rc, "pmg_ctx_create_pbe")
            return
        rc = L.pmg_ctx_create(ctypes.byref(self._h), _ip(iparm), _dp(rparm),
                              _dp(inp["xf"]), _dp(inp["yf"]), _dp(inp["zf"]),
                              _dp(inp["gxcf"]), _dp(inp["gycf"]), _dp(inp["gzcf"]),
                              _dp(inp["a1cf"]), _dp(inp["a2cf"]), _dp(inp["a3cf"]),
                              _dp(inp["ccf"]), _dp(inp["fcf"]))
        _check(rc, "pmg_ctx_create")

    def load_pbe(self, h, ionic_strength, atoms):
        """Re-assemble the inputs for another molecule and rebuild."""
        at = np.ascontiguousarray(atoms, np.float64).reshape(-1)
        return _check(lib().pmg_ctx_load_pbe(self._h, h, ionic_strength, at.size // 5,
                                             _dp(at)), "pmg_ctx_load_pbe")

    def solve(self, mode=PMG_FP64):
        return _check(lib().pmg_ctx_solve(self._h, mode), "pmg_ctx_solve")

    def pcg(self, mode=PMG_FP64, num_iters=0):
        """The class-style PCG (pmg_ctx_pcg) on the resident hierarchy."""
        return _check(lib().pmg_ctx_pcg(self._h, mode, int(num_iters)), "pmg_ctx_pcg")

    def u(self):
        out = np.empty(self.n[0] * self.n[1] * self.n[2], np.float64)
        _check(lib().pmg_ctx_get_u(self._h, _dp(out)), "pmg_ctx_get_u")
        return out

    def stats(self):
        it, ie = ctypes.c_int(), ctypes.c_int()
        rr, su, so = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        lib().pmg_ctx_stats(self._h, ctypes.byref(it), ctypes.byref(ie),
                            ctypes.byref(rr), ctypes.byref(su), ctypes.byref(so))
        return dict(iters=it.value, ierror=ie.value, relres=rr.value,
                    setup_ms=su.value, solve_ms=so.value)

    def rebuild(self):
        return _check(lib().pmg_ctx_rebuild(self._h), "pmg_ctx_rebuild")

    def set_tru(self, tcf):
        """The comparison solution of the stopping tests istop 2-5."""
        t = np.ascontiguousarray(tcf, np.float64)
        return _check(lib().pmg_ctx_set_tru(self._h, _dp(t)), "pmg_ctx_set_tru")

    def history(self, cap=4096):
        """Relative stopping measure after each iteration of the last solve."""
        out = np.zeros(cap)
        m = lib().pmg_ctx_history(self._h, _dp(out), cap)
        return out[:min(m, cap)].copy()

    def nbytes(self):
        return int(lib().pmg_ctx_bytes(self._h))

    def level(self, lev):
        """FP64 operators of level lev (1-based): (ac, cc, fc, pc)."""
        n = self.n
        dims = [n]
        for _ in range(1, self.nlev):
            dims.append(tuple((v - 1) // 2 + 1 for v in dims[-1]))
        d = dims[lev - 1]
        nl = d[0] * d[1] * d[2]
        ac = np.zeros((4 if lev == 1 else 14) * nl)
        cc, fc = np.zeros(nl), np.zeros(nl)
        pc = None
        if lev < self.nlev:
            c = dims[lev]
            pc = np.zeros(27 * c[0] * c[1] * c[2])
        _check(lib().pmg_ctx_get_level(self._h, lev, _dp(ac), _dp(cc), _dp(fc),
                                       _dp(pc)), "pmg_ctx_get_level")
        return ac, cc, fc, pc

    def ktime_enable(self, max_events=8192):
        return _check(lib().pmg_ctx_ktime_enable(self._h, max_events), "ktime")

    def set_graphs(self, on):
        return _check(lib().pmg_ctx_set_graphs(self._h, 1 if on else 0), "set_graphs")

    def ktime(self):
        ms, cnt = ctypes.c_double(), ctypes.c_longlong()
        lib().pmg_ctx_ktime(self._h, ctypes.byref(ms), ctypes.byref(cnt))
        return ms.value, cnt.value

    def close(self):
        if self._h:
            lib().pmg_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Serving loop
# ---------------------------------------------------------------------------
class Server:
    """Continuous-batching PB solve server (pmg_server_*)."""

    def __init__(self, mem_budget=0, max_active=8, groups=0):
        """groups > 0: the solves' iterations run grouped
        (pmg_server_create_grouped); 0: each on its own stream."""
        self._h = ctypes.c_void_p()
        _check(lib().pmg_server_create_grouped(ctypes.byref(self._h), mem_budget,
                                               max_active, int(groups)),
               "pmg_server_create_grouped")
        self._out = {}

    def submit(self, iparm, rparm, mode, h, ionic_strength, atoms, want_u=True):
        """Queue a solve; returns the ticket.  The potential (if want_u)
        comes back from wait()."""
        at = np.ascontiguousarray(atoms, np.float64).reshape(-1)
        n = int(iparm[2]) * int(iparm[3]) * int(iparm[4])
        u = np.empty(n) if want_u else None
        t = ctypes.c_int64()
        rc = lib().pmg_server_submit(self._h, _ip(iparm), _dp(rparm), mode, h,
                                     ionic_strength, at.size // 5, _dp(at), _dp(u),
                                     ctypes.byref(t))
        _check(rc, "pmg_server_submit")
        self._out[t.value] = u
        return t.value

    def wait(self, ticket, timeout_ms=-1):
        it, rr, lat = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        rc = lib().pmg_server_wait(self._h, ticket, timeout_ms, ctypes.byref(it),
                                   ctypes.byref(rr), ctypes.byref(lat))
        _check(rc, "pmg_server_wait")
        return dict(rc=rc, iters=it.value, relres=rr.value, latency_ms=lat.value,
                    u=self._out.pop(ticket))

    def stats(self):
        c, mem, mx = ctypes.c_longlong(), ctypes.c_size_t(), ctypes.c_int()
        p50, p99, sps = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        lib().pmg_server_stats(self._h, ctypes.byref(c), ctypes.byref(p50),
                               ctypes.byref(p99), ctypes.byref(sps), ctypes.byref(mem),
                               ctypes.byref(mx))
        return dict(completed=c.value, p50_ms=p50.value, p99_ms=p99.value,
                    solves_per_s=sps.value, mem_used=mem.value, max_concurrent=mx.value)

    def reset_stats(self):
        """Start the statistics afresh (no request may be in flight); the
        pooled contexts stay."""
        _check(lib().pmg_server_reset_stats(self._h), "pmg_server_reset_stats")

    def close(self):
        if self._h:
            lib().pmg_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# Solve shapes of the serving workload: grid size -> levels (coarsest 5^3
# to 11^3), all at h = 0.5 A.
SERVE_LEVELS = {33: 3, 65: 4, 97: 5, 129: 5, 161: 5}


# ---------------------------------------------------------------------------
# Poisson-Boltzmann inputs: synthetic molecules (host) and their assembly
# into grid inputs (GPU)
# ---------------------------------------------------------------------------
def pbe_molecule(natoms, sphere_r, charge_mode=0, seed=20261202):
    """Seeded synthetic molecule, (natoms, 5) = x, y, z, radius, charge."""
    atoms = np.zeros(5 * max(natoms, 1))
    _check(lib().pmg_pbe_molecule(natoms, sphere_r, charge_mode, ctypes.c_uint64(seed),
                                  _dp(atoms)), "pbe_molecule")
    return atoms[: 5 * natoms].reshape(natoms, 5)


def _pbe_arrays(n):
    out = {k: np.zeros(n) for k in ("xf", "yf", "zf")}
    for k in ("gxcf", "gycf", "gzcf"):
        out[k] = np.zeros(4 * n * n)
    for k in ("a1cf", "a2cf", "a3cf", "ccf", "fcf", "tcf"):
        out[k] = np.zeros(n * n * n)
    return out


PBE_KEYS = ("xf", "yf", "zf", "gxcf", "gycf", "gzcf", "a1cf", "a2cf", "a3cf",
            "ccf", "fcf", "tcf")


def pbe_assemble(n, h, ionic_strength, atoms, eps_in=2.0, eps_out=78.54, ion_shell=2.0):
    """Fine-level Vmgdriv inputs assembled on the GPU from an atom array."""
    at = np.ascontiguousarray(atoms, np.float64).reshape(-1)
    out = _pbe_arrays(n)
    rc = lib().pmg_pbe_assemble_ex(n, h, ionic_strength, eps_in, eps_out, ion_shell,
                                   at.size // 5, _dp(at), *(_dp(out[k]) for k in PBE_KEYS))
    _check(rc, "pbe_assemble")
    out["atoms"] = at.reshape(-1, 5)
    return out


def pbe_synth(n, h, natoms, sphere_r, ionic_strength=0.15, charge_mode=0, seed=20261202):
    """Fine-level inputs for a seeded synthetic molecule (GPU assembly)."""
    return pbe_assemble(n, h, ionic_strength, pbe_molecule(natoms, sphere_r, charge_mode, seed))


# ---------------------------------------------------------------------------
# Molecule and field files (PQR in, OpenDX in/out)
# ---------------------------------------------------------------------------
def read_pqr(path):
    """Atoms of a PQR file as (natoms, 5) = x, y, z, radius, charge (the
    reference reads the same fields, import_molecular_structure.cc)."""
    rows = []
    for line in open(path):
        if line.startswith(("ATOM", "HETATM")):
            f = line.split()
            x, y, z, q, r = (float(v) for v in f[-5:])
            rows.append((x, y, z, r, q))
    return np.array(rows, np.float64).reshape(-1, 5)


def dx_write(path, u, dims, origin, h, title="POTENTIAL (kT/e)"):
    """Write an n^3 field (PMG layout, x fastest) as OpenDX."""
    u = np.ascontiguousarray(u, np.float64).reshape(-1)
    nx, ny, nz = dims
    assert u.size == nx * ny * nz
    _check(lib().pmg_dx_write(str(path).encode(), nx, ny, nz, origin[0], origin[1],
                              origin[2], h, _dp(u), title.encode()), "dx_write")


def dx_read(path):
    """Read an OpenDX regular-grid field: (u in PMG layout, counts, origin,
    delta)."""
    c = np.zeros(3, np.int32)
    o, dl = np.zeros(3), np.zeros(3)
    _check(lib().pmg_dx_read(str(path).encode(), _ip(c), _dp(o), _dp(dl), None, 0),
           "dx_read")
    u = np.zeros(int(c[0]) * int(c[1]) * int(c[2]))
    _check(lib().pmg_dx_read(str(path).encode(), _ip(c), _dp(o), _dp(dl), _dp(u), u.size),
           "dx_read")
    return u, tuple(int(v) for v in c), o, dl


# Benchmark configurations (sizes and molecule parameters of the five
# workloads this solver is measured on).
CONFIGS = {
    "pb65": dict(n=65, h=0.5, nlev=4, natoms=200, sphere_r=14.0, ionic=0.15, charge_mode=0, seed=20261202),
    "pb161": dict(n=161, h=0.4, nlev=5, natoms=1000, sphere_r=25.0, ionic=0.15, charge_mode=0, seed=20261203),
    "pb289": dict(n=289, h=0.35, nlev=6, natoms=5000, sphere_r=45.0, ionic=0.15, charge_mode=1, seed=20261204),
    "newton129": dict(n=129, h=0.5, nlev=5, natoms=800, sphere_r=20.0, ionic=0.30, charge_mode=0, seed=20261205),
}


def config_inputs(name, **over):
    c = dict(CONFIGS[name])
    c.update(over)
    inp = pbe_synth(c["n"], c["h"], c["natoms"], c["sphere_r"], c["ionic"],
                    c["charge_mode"], c["seed"])
    return c, inp

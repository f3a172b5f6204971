"""GPU parity on the benchmark workloads themselves, at full size, through
the C ABI against the CPU oracle on the same inputs:

* 161^3 (h 0.4, 5 levels, 1000 atoms): FP64 drop-in potential <= 1e-10
  relative max-norm, the same V-cycle count, every coarse operator
  (VbuildG, buildGd.cc:437-3274), P, cc and fc of every level <= 1e-12;
  the mixed-precision potential <= 1e-6 of the oracle.
* 289^3 (h 0.35, 6 levels, 5000 atoms, charges over {-1,-0.5,0,0.5,1}):
  the same FP64 bars and the coarse operators; the mixed potential <= 1e-6
  of the oracle, and the L1 relative residual of both solutions recomputed
  on the host <= 1e-6 (the stop test of mgcsd.cc:483-527).
* 129^3 nonlinear (sinh) PBE, 800 atoms, 0.3 M, Newton to 1e-8: the same
  Newton iteration count as the oracle and the potential <= 1e-9, for FP64
  and for FP32 inner V-cycles.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    s = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(a - b))) / s


def _drv(mg, fn, inp, iparm, rparm):
    sz = mg.Vmgsz(2, 0, int(iparm[20]), int(iparm[2]), int(iparm[3]), int(iparm[4]),
                  int(iparm[5]))
    iw = np.zeros(sz["iintot"], np.int32)
    rw = np.zeros(int(iparm[0]))
    u = np.zeros(int(iparm[2]) * int(iparm[3]) * int(iparm[4]))
    a = {k: inp[k].copy() for k in mg.PBE_KEYS}
    ierr = fn(iparm.copy(), rparm.copy(), iw, rw, u, *(a[k] for k in mg.PBE_KEYS))
    return ierr, u, iw, rw


def _levels(n, nlev):
    out = [n]
    for _ in range(1, nlev):
        out.append((out[-1] - 1) // 2 + 1)
    return out


def _rwork_sections(mg, n, nlev, newton=False):
    """(name, level, slice) of every operator array in rwork, in the layout
    Vmgdriv leaves it (mgdrvd.cc:150-166): rpc | cc | fc | [Newton 2 nf] |
    pc | ac, each multi-level array level by level."""
    sz = mg.Vmgsz(2, 0, 0, n, n, n, nlev)
    dims = _levels(n, nlev)
    pts = [d ** 3 for d in dims]
    cc0 = sz["n_rpc"]
    fc0 = cc0 + sz["narr"]
    pc0 = fc0 + sz["narr"] + (2 * sz["nf"] if newton else 0)
    ac0 = pc0 + 27 * sz["narrc"]
    out = []
    off = 0
    for lev in range(nlev):
        out.append(("cc", lev + 1, slice(cc0 + off, cc0 + off + pts[lev])))
        out.append(("fc", lev + 1, slice(fc0 + off, fc0 + off + pts[lev])))
        off += pts[lev]
    off = 0
    for lev in range(1, nlev):
        out.append(("pc", lev + 1, slice(pc0 + off, pc0 + off + 27 * pts[lev])))
        off += 27 * pts[lev]
    off = 0
    for lev in range(nlev):
        m = (4 if lev == 0 else 14) * pts[lev]
        out.append(("ac", lev + 1, slice(ac0 + off, ac0 + off + m)))
        off += m
    return out


def _check_operators(mg, orc, n, nlev, rw1, rw0, tol=1e-12):
    """Every array the setup leaves in rwork, level by level.  ac (the fine
    operator and the Galerkin products of VbuildG) and pc (VbuildP): max-norm
    error <= tol relative to the level's largest entry.  cc and fc of the
    coarse levels are restrictions (Vrestrc, mgsubd.cc:349-373) of a
    source with charges of both signs, so partial sums cancel: their error
    is bounded point by point by tol times the same restriction applied to
    the absolute values (the magnitude of the terms summed), which the
    order in which a restriction adds its 27 terms cannot exceed."""
    secs = {(name, lev): sl for name, lev, sl in _rwork_sections(mg, n, nlev)}
    dims = _levels(n, nlev)
    bad, worst = [], {}
    for (name, lev), sl in secs.items():
        a, b = rw1[sl], rw0[sl]
        if not np.any(b):
            if np.any(a):
                bad.append(f"{name} level {lev}: nonzero where the oracle is zero")
            continue
        if name in ("cc", "fc"):
            continue
        e = _rel(a, b)
        worst[name, lev] = e
        if e > tol:
            bad.append(f"{name} level {lev}: rel err {e:.2e}")
    for name in ("cc", "fc"):
        mag = np.abs(rw0[secs[name, 1]])
        e = _rel(rw1[secs[name, 1]], rw0[secs[name, 1]])
        worst[name, 1] = e
        if e > tol:
            bad.append(f"{name} level 1: rel err {e:.2e}")
        for lev in range(2, nlev + 1):
            pcl = np.abs(rw0[secs["pc", lev]])
            mag = orc.restrc(dims[lev - 2], dims[lev - 1], np.ascontiguousarray(mag), pcl)
            a, b = rw1[secs[name, lev]], rw0[secs[name, lev]]
            over = np.abs(a - b) - tol * mag
            worst[name, lev] = float(np.max(np.abs(a - b) / np.maximum(mag, 1e-300)))
            if np.any(over > 0):
                bad.append(f"{name} level {lev}: {int(np.sum(over > 0))} points beyond "
                           f"{tol:g} x the restricted magnitudes")
    print("  operator errors: " + ", ".join(f"{k[0]}{k[1]} {v:.1e}"
                                              for k, v in sorted(worst.items())), flush=True)
    assert not bad, bad
    return max(v for (nm, _), v in worst.items() if nm in ("ac", "pc"))


def _host_relres(orc, n, inp, u):
    """Relative L1 residual of the interior unknowns of u, recomputed by the
    oracle's VbuildA_fv + Vmresid (Dirichlet values folded into fc)."""
    ac, cc, fc, _ = orc.buildA_fv(n, inp)
    ub = u.copy().reshape(n, n, n)
    ub[0, :, :] = ub[-1, :, :] = 0
    ub[:, 0, :] = ub[:, -1, :] = 0
    ub[:, :, 0] = ub[:, :, -1] = 0
    r = orc.mresid(n, 7, ac, cc, fc, ub.ravel())
    return float(np.abs(r).sum() / np.abs(fc).sum())


def _linear_config(mg, orc, name):
    c = mg.CONFIGS[name]
    n, nlev = c["n"], c["nlev"]
    atoms = mg.pbe_molecule(c["natoms"], c["sphere_r"], c["charge_mode"], c["seed"])
    inp = orc.pbe_assemble(n, c["h"], c["ionic"], atoms)
    iparm, rparm = mg.make_params(n, nlev)
    print(f"{name}: {n}^3 nlev={nlev} FP64 drop-in on the GPU", flush=True)
    e1, u1, iw1, rw1 = _drv(mg, mg.Vmgdriv, inp, iparm, rparm)
    it1, _, rr1 = mg.last_stats()
    print(f"{name}: mixed precision on a resident context", flush=True)
    ctx = mg.Context(iparm, rparm, inp)
    ctx.solve(mg.PMG_MIXED)
    umx, smx = ctx.u(), ctx.stats()
    ctx.close()
    print(f"{name}: oracle", flush=True)
    e0, u0, iw0, rw0 = orc.solve(inp, n, nlev)
    it0, _, rr0, _ = orc.last_stats()
    return dict(n=n, nlev=nlev, inp=inp, e1=e1, u1=u1, iw1=iw1, rw1=rw1, it1=it1, rr1=rr1,
                umx=umx, smx=smx, e0=e0, u0=u0, iw0=iw0, rw0=rw0, it0=it0, rr0=rr0)


def test_pb161_fp64_and_mixed_vs_oracle(mg, orc):
    r = _linear_config(mg, orc, "pb161")
    n, nlev = r["n"], r["nlev"]
    worst = _check_operators(mg, orc, n, nlev, r["rw1"], r["rw0"])
    print(f"  V-cycles gpu={r['it1']} oracle={r['it0']} potential err "
          f"{_rel(r['u1'], r['u0']):.2e} operators worst {worst:.2e} mixed iters "
          f"{r['smx']['iters']} err {_rel(r['umx'], r['u0']):.2e}", flush=True)
    assert r["e1"] == r["e0"] == 0
    assert r["it1"] == r["it0"]
    assert _rel(r["u1"], r["u0"]) <= 1e-10
    assert np.array_equal(r["iw1"], r["iw0"])
    # the final residual is rounding-dominated (r = f - A u of a converged u
    # cancels to ~1e-7 of f), so its L1 norm agrees only to ~1e-8 relative
    assert abs(r["rr1"] - r["rr0"]) <= 1e-6 * r["rr0"]
    assert r["smx"]["ierror"] == 0 and r["smx"]["relres"] <= 1e-6
    assert _rel(r["umx"], r["u0"]) <= 1e-6


def test_pb289_fp64_and_mixed_vs_oracle(mg, orc):
    r = _linear_config(mg, orc, "pb289")
    n, nlev = r["n"], r["nlev"]
    worst = _check_operators(mg, orc, n, nlev, r["rw1"], r["rw0"])
    host64 = _host_relres(orc, n, r["inp"], r["u1"])
    hostmx = _host_relres(orc, n, r["inp"], r["umx"])
    print(f"  V-cycles gpu={r['it1']} oracle={r['it0']} potential err "
          f"{_rel(r['u1'], r['u0']):.2e} operators worst {worst:.2e}; mixed iters "
          f"{r['smx']['iters']} err {_rel(r['umx'], r['u0']):.2e}; host relres fp64 "
          f"{host64:.2e} mixed {hostmx:.2e}", flush=True)
    assert r["e1"] == r["e0"] == 0
    assert r["it1"] == r["it0"]
    assert _rel(r["u1"], r["u0"]) <= 1e-10
    assert np.array_equal(r["iw1"], r["iw0"])
    assert r["smx"]["ierror"] == 0
    assert _rel(r["umx"], r["u0"]) <= 1e-6
    assert host64 <= 1e-6 * 1.0001 and hostmx <= 1e-6 * 1.0001
    assert abs(host64 - r["rr1"]) <= 1e-9 and abs(hostmx - r["smx"]["relres"]) <= 1e-9


def test_newton129_vs_oracle(mg, orc):
    c = mg.CONFIGS["newton129"]
    n, nlev = c["n"], c["nlev"]
    ions = ([1.0, -1.0], [0.5, 0.5])
    mg.Vmypdefinitnpbe(2, *ions)
    orc.mypdefinit(*ions)
    atoms = mg.pbe_molecule(c["natoms"], c["sphere_r"], c["charge_mode"], c["seed"])
    inp = orc.pbe_assemble(n, c["h"], c["ionic"], atoms)
    iparm, rparm = mg.make_params(n, nlev, nonlin=1, errtol=1e-8)
    print(f"newton129: {n}^3 Newton FP64 drop-in on the GPU", flush=True)
    e1, u1, _, _ = _drv(mg, mg.Vnewdriv, inp, iparm, rparm)
    it1 = mg.last_stats()[0]
    print("newton129: Newton with FP32 inner V-cycles", flush=True)
    ctx = mg.Context(iparm, rparm, inp)
    ctx.solve(mg.PMG_MIXED)
    umx, smx = ctx.u(), ctx.stats()
    ctx.close()
    print("newton129: oracle", flush=True)
    e0, u0, _, _ = orc.solve(inp, n, nlev, nonlin=True, errtol=1e-8)
    it0 = orc.last_stats()[0]
    print(f"  Newton iters gpu={it1} mixed={smx['iters']} oracle={it0} err fp64 "
          f"{_rel(u1, u0):.2e} mixed {_rel(umx, u0):.2e}", flush=True)
    assert e1 == e0 == 0 and smx["ierror"] == 0
    assert it1 == it0 and smx["iters"] == it0
    assert _rel(u1, u0) <= 1e-9
    assert _rel(umx, u0) <= 1e-9


def test_newton129_nested_vs_oracle(mg, orc):
    """The 129^3 Newton workload with the nested iteration (mgkey 1, Vfnewton
    newtond.cc:60-138): the same finest-level Newton iterations and
    per-iteration residuals as the oracle, the potential within 1e-9, in the
    drop-in and through a resident context with FP32 inner cycles."""
    c = mg.CONFIGS["newton129"]
    n, nlev = c["n"], c["nlev"]
    ions = ([1.0, -1.0], [0.5, 0.5])
    mg.Vmypdefinitnpbe(2, *ions)
    orc.mypdefinit(*ions)
    atoms = mg.pbe_molecule(c["natoms"], c["sphere_r"], c["charge_mode"], c["seed"])
    inp = orc.pbe_assemble(n, c["h"], c["ionic"], atoms)
    iparm, rparm = mg.make_params(n, nlev, nonlin=1, errtol=1e-8, mgkey=1)
    print(f"newton129 nested: {n}^3 Newton from the coarsest level, FP64 drop-in", flush=True)
    e1, u1, _, _ = _drv(mg, mg.Vnewdriv, inp, iparm, rparm)
    it1, h1 = mg.last_stats()[0], mg.last_history()
    print("newton129 nested: FP32 inner V-cycles", flush=True)
    ctx = mg.Context(iparm, rparm, inp)
    ctx.solve(mg.PMG_MIXED)
    umx, smx = ctx.u(), ctx.stats()
    ctx.close()
    print("newton129 nested: oracle", flush=True)
    e0, u0, _, _ = orc.solve(inp, n, nlev, nonlin=True, errtol=1e-8, mgkey=1)
    it0, _, _, h0 = orc.last_stats()
    print(f"  Newton iters gpu={it1} mixed={smx['iters']} oracle={it0} err fp64 "
          f"{_rel(u1, u0):.2e} mixed {_rel(umx, u0):.2e}", flush=True)
    assert e1 == e0 == 0 and smx["ierror"] == 0
    assert it1 == it0 == len(h1) == len(h0)
    assert np.allclose(h1, h0, rtol=1e-6, atol=1e-9 * h0[0])
    assert _rel(u1, u0) <= 1e-9
    assert _rel(umx, u0) <= 1e-9


def test_grouped_workloads_equal_alone(mg, orc):
    """The 65^3 and 161^3 workloads and a 129^3 linear solve of the Newton
    workload's molecule solved together with grouped iterations
    (pmg_ctx_solve_group): each potential, iteration count and history is
    that of its own solve, bit for bit, in both precisions."""
    specs = []
    for name in ("pb65", "pb161", "newton129"):
        c = mg.CONFIGS[name]
        atoms = mg.pbe_molecule(c["natoms"], c["sphere_r"], c["charge_mode"], c["seed"])
        specs.append((c["n"], c["nlev"], (c["h"], c["ionic"], atoms)))
    for mode in (mg.PMG_MIXED, mg.PMG_FP64):
        ctxs = [mg.Context(*mg.make_params(n, nlev), pbe=mol) for n, nlev, mol in specs]
        alone = []
        for c in ctxs:
            c.solve(mode)
            alone.append((c.u(), c.stats()["iters"], c.history()))
        print(f"grouped workloads mode={mode}", flush=True)
        mg.solve_group(ctxs, mode)
        for (n, _, _), c, (u0, it0, h0) in zip(specs, ctxs, alone):
            assert c.stats()["iters"] == it0 and np.array_equal(c.history(), h0)
            assert np.array_equal(c.u(), u0), f"{n}^3 grouped potential differs"
        for c in ctxs:
            c.close()

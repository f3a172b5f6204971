"""GPU parity: every HIP kernel and driver against the CPU oracle, through
the C ABI.  Tolerances: single kernels 1e-13 relative (FMA contraction is
the only difference); potential 1e-10 relative max-norm with identical
V-cycle count; Galerkin coarse operators 1e-12 relative."""
import ctypes

import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu


def _rel(a, b):
    s = max(np.max(np.abs(b)), 1e-300)
    return np.max(np.abs(a - b)) / s


def _call(mg, name, *args):
    rc = getattr(mg.lib(), name)(*args)
    assert rc == 0, f"{name} returned {rc}"


dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@pytest.mark.parametrize("numdia", [7, 27])
@pytest.mark.parametrize("n", [9, 17])
def test_matvec_resid(mg, orc, n, numdia):
    print(f"matvec/resid n={n} numdia={numdia}", flush=True)
    rng = np.random.default_rng(n + numdia)
    ac, cc = H.random_op(n, numdia, rng)
    x = H.from_int(n, rng.standard_normal((n - 2) ** 3))
    f = H.from_int(n, rng.standard_normal((n - 2) ** 3))
    y = np.zeros(n ** 3)
    _call(mg, "pmg_matvec", n, n, n, numdia, dp(ac), dp(cc), dp(x), dp(y))
    assert _rel(y, orc.matvec(n, numdia, ac, cc, x)) < 1e-13
    r = np.zeros(n ** 3)
    _call(mg, "pmg_mresid", n, n, n, numdia, dp(ac), dp(cc), dp(f), dp(x), dp(r))
    assert _rel(r, orc.mresid(n, numdia, ac, cc, f, x)) < 1e-13


@pytest.mark.parametrize("dims", [(126, 40, 1000), (131, 70, 300), (129, 41, 9), (126, 40, 60)])
def test_mresid7_shapes(mg, orc, dims):
    """The 7-point residual on non-cubic grids: the k-march (k chunks of 24
    and 18 planes on 256 compute units) over several tiles in i and j with
    ragged last tiles and ragged last chunks (126 x 40 x 1000,
    131 x 70 x 300), and the column kernel on a grid of 127 interior columns
    (129 x 41 x 9) and on one whose chunks would be short (126 x 40 x 60)."""
    print(f"mresid7 dims={dims}", flush=True)
    nx, ny, nz = dims
    N = nx * ny * nz
    rng = np.random.default_rng(sum(dims))
    ac = rng.uniform(0.1, 1.0, 4 * N)
    ac[:N] += 6.0
    cc = rng.uniform(0.0, 1.0, N)
    mask = np.zeros((nz, ny, nx))
    mask[1:-1, 1:-1, 1:-1] = 1
    f = (rng.standard_normal((nz, ny, nx)) * mask).ravel()
    x = (rng.standard_normal((nz, ny, nx)) * mask).ravel()
    r = np.zeros(N)
    _call(mg, "pmg_mresid", nx, ny, nz, 7, dp(ac), dp(cc), dp(f), dp(x), dp(r))
    ref = np.zeros(N)
    orc.olib().pmgo_mresid(nx, ny, nz, 7, dp(ac), dp(cc), dp(f), dp(x), dp(ref))
    assert _rel(r, ref) < 1e-13


@pytest.mark.parametrize("numdia", [7, 27])
@pytest.mark.parametrize("iadj", [0, 1])
def test_gsrb(mg, orc, numdia, iadj):
    print(f"gsrb numdia={numdia} iadjoint={iadj}", flush=True)
    n = 17
    rng = np.random.default_rng(numdia * 10 + iadj)
    ac, cc = H.random_op(n, numdia, rng)
    f = H.from_int(n, rng.standard_normal((n - 2) ** 3))
    x0 = H.from_int(n, rng.standard_normal((n - 2) ** 3))
    x = x0.copy()
    _call(mg, "pmg_gsrb", n, n, n, numdia, dp(ac), dp(cc), dp(f), dp(x), 3, iadj)
    assert _rel(x, orc.gsrb(n, numdia, ac, cc, f, x0, 3, iadj)) < 1e-12


@pytest.mark.parametrize("numdia,n", [(7, 97), (27, 97), (27, 131), (27, 41), (7, 19),
                                      (7, 22), (7, 23)])
def test_gsrb_large(mg, orc, numdia, n):
    """Several workgroups and k-chunks for the fused 7-point sweep (97^3);
    the row-parallel 27-point sweep at 97^3 and with two point pairs per
    lane for rows longer than 129 points (131^3); for 27 points a pair of sweeps
    runs as two paired-class launches per sweep (strips of 7 rows at 97^3
    and 131^3, of 3 rows at 41^3), an odd third sweep as the four class
    launches.  The fused 7-point sweep marches its k
    chunk five planes per loop trip: 97^3 takes chunks of 6 planes (one trip
    and one plane over, the last chunk of 5 one trip), 19^3, 22^3, 23^3
    chunks of 4 planes (no full trip) ending on a chunk of 1, 4 and 1
    planes."""
    print(f"gsrb large n={n} numdia={numdia}", flush=True)
    rng = np.random.default_rng(n + numdia)
    ac, cc = H.random_op_vec(n, numdia, rng)
    mask = np.zeros((n, n, n)); mask[1:-1, 1:-1, 1:-1] = 1
    f = (rng.standard_normal((n, n, n)) * mask).ravel()
    x0 = (rng.standard_normal((n, n, n)) * mask).ravel()
    for iadj, sweeps in ((0, 2), (1, 3)):
        x = x0.copy()
        _call(mg, "pmg_gsrb", n, n, n, numdia, dp(ac), dp(cc), dp(f), dp(x), sweeps, iadj)
        assert _rel(x, orc.gsrb(n, numdia, ac, cc, f, x0, sweeps, iadj)) < 1e-12


@pytest.mark.parametrize("numdia", [7, 27])
@pytest.mark.parametrize("dims", [(33, 25, 19), (23, 41, 17)])
def test_gsrb_noncubic(mg, orc, numdia, dims):
    """Non-cubic grids: tile, strip and plane counts differ per axis for the
    fused 7-point sweep and the paired 27-point launches (strips of 3 rows
    at these sizes), two sweeps forward and three backward.  Positive random
    couplings everywhere (the sweep formula, not the discretisation, is
    under test)."""
    print(f"gsrb non-cubic dims={dims} numdia={numdia}", flush=True)
    nx, ny, nz = dims
    N = nx * ny * nz
    rng = np.random.default_rng(nx + 3 * ny + 7 * nz + numdia)
    nplanes = 4 if numdia == 7 else 14
    ac = rng.uniform(0.1, 1.0, nplanes * N)
    ac[:N] += 2.0 * (nplanes - 1)  # oC dominates the couplings
    cc = rng.uniform(0.0, 1.0, N)
    mask = np.zeros((nz, ny, nx)); mask[1:-1, 1:-1, 1:-1] = 1
    f = (rng.standard_normal((nz, ny, nx)) * mask).ravel()
    x0 = (rng.standard_normal((nz, ny, nx)) * mask).ravel()
    for iadj, sweeps in ((0, 2), (1, 3)):
        x = x0.copy()
        _call(mg, "pmg_gsrb", nx, ny, nz, numdia, dp(ac), dp(cc), dp(f), dp(x), sweeps, iadj)
        ref = orc.gsrb(0, numdia, ac, cc, f, x0, sweeps, iadj, dims=dims)
        assert _rel(x, ref) < 1e-12, (iadj, sweeps)


@pytest.mark.parametrize("mgprol", [0, 1])
@pytest.mark.parametrize("numdia", [7, 27])
def test_buildP_buildG_noncubic(mg, orc, numdia, mgprol):
    """One non-cubic coarsening step, 33x25x19 -> 17x13x10: the prolongation
    and Galerkin kernels give each lane one of the consecutive interior
    points of a coarse plane, across row ends, with 165 points per plane
    (several workgroups per plane, a partial last one).  Positive random
    couplings everywhere, diagonal dominant."""
    print(f"buildP/buildG non-cubic numdia={numdia} mgprol={mgprol}", flush=True)
    df, dc = (33, 25, 19), (17, 13, 10)
    Nf, Nc = df[0] * df[1] * df[2], dc[0] * dc[1] * dc[2]
    rng = np.random.default_rng(11 * numdia + mgprol)
    nplanes = 4 if numdia == 7 else 14
    ac = rng.uniform(0.1, 1.0, nplanes * Nf)
    ac[:Nf] += 2.0 * (nplanes - 1)
    pc, pco = np.zeros(27 * Nc), np.zeros(27 * Nc)
    _call(mg, "pmg_buildP", *df, *dc, mgprol, numdia, dp(ac), dp(pc))
    orc.olib().pmgo_buildP(*df, *dc, mgprol, numdia, dp(ac), dp(pco))
    assert np.abs(pco).max() > 0 and _rel(pc, pco) < 1e-13
    acc, acco = np.zeros(14 * Nc), np.zeros(14 * Nc)
    _call(mg, "pmg_buildG", *df, *dc, numdia, dp(pco), dp(ac), dp(acc))
    orc.olib().pmgo_buildG(*df, *dc, numdia, dp(pco), dp(ac), dp(acco))
    assert np.abs(acco).max() > 0 and _rel(acc, acco) < 1e-12


@pytest.mark.parametrize("df", [(5, 5, 5), (7, 5, 9), (9, 17, 5), (131, 7, 7)])
@pytest.mark.parametrize("numdia", [7, 27])
def test_buildG_shapes_bitwise(mg, orc, numdia, df):
    """The Galerkin product (w = A P_C, one wavefront per fine plane, partial
    sums added through LDS) on the smallest and most uneven coarsenings: one
    coarse interior point, fewer coarse points per plane than lanes, 128
    coarse points per plane (two full workgroups).  Same grouping and
    roundings as the oracle, so the entries are equal to the bit."""
    print(f"buildG shapes numdia={numdia} df={df}", flush=True)
    dc = tuple((v - 1) // 2 + 1 for v in df)
    Nf, Nc = df[0] * df[1] * df[2], dc[0] * dc[1] * dc[2]
    rng = np.random.default_rng(sum(df) + numdia)
    nplanes = 4 if numdia == 7 else 14
    ac = rng.uniform(0.1, 1.0, nplanes * Nf)
    ac[:Nf] += 2.0 * (nplanes - 1)
    pc = np.zeros(27 * Nc)
    orc.olib().pmgo_buildP(*df, *dc, 1, numdia, dp(ac), dp(pc))
    acc, acco = np.zeros(14 * Nc), np.zeros(14 * Nc)
    _call(mg, "pmg_buildG", *df, *dc, numdia, dp(pc), dp(ac), dp(acc))
    orc.olib().pmgo_buildG(*df, *dc, numdia, dp(pc), dp(ac), dp(acco))
    assert np.abs(acco).max() > 0
    assert np.array_equal(acc, acco)


@pytest.mark.parametrize("mgprol", [0, 1])
@pytest.mark.parametrize("numdia", [7, 27])
def test_buildP_buildG_transfer(mg, orc, numdia, mgprol):
    print(f"buildP/buildG numdia={numdia} mgprol={mgprol}", flush=True)
    nf, nc = 17, 9
    rng = np.random.default_rng(numdia + 3 * mgprol)
    ac, _ = H.random_op(nf, numdia, rng)
    pc = np.zeros(27 * nc ** 3)
    _call(mg, "pmg_buildP", nf, nf, nf, nc, nc, nc, mgprol, numdia, dp(ac), dp(pc))
    pco = orc.buildP(nf, nc, mgprol, numdia, ac)
    assert _rel(pc, pco) < 1e-13
    acc = np.zeros(14 * nc ** 3)
    _call(mg, "pmg_buildG", nf, nf, nf, nc, nc, nc, numdia, dp(pco), dp(ac), dp(acc))
    assert _rel(acc, orc.buildG(nf, nc, numdia, pco, ac)) < 1e-12
    r = H.from_int(nf, rng.standard_normal((nf - 2) ** 3))
    xc = np.zeros(nc ** 3)
    _call(mg, "pmg_restrc", nf, nf, nf, nc, nc, nc, dp(r), dp(xc), dp(pco))
    assert _rel(xc, orc.restrc(nf, nc, r, pco)) < 1e-13
    e = H.from_int(nc, rng.standard_normal((nc - 2) ** 3))
    xf = np.zeros(nf ** 3)
    _call(mg, "pmg_interp", nc, nc, nc, nf, nf, nf, dp(e), dp(xf), dp(pco))
    assert _rel(xf, orc.interp(nc, nf, e, pco)) < 1e-13


def _run_driver(mg, fn, inp, n, nlev, nonlin=False, **kw):
    iparm, rparm = mg.make_params(n, nlev, nonlin=1 if nonlin else 0, **kw)
    sz = mg.Vmgsz(2, 0, 0, n, n, n, nlev)
    iw = np.zeros(sz["iintot"], np.int32)
    rw = np.zeros(int(iparm[0]))
    u = np.zeros(n ** 3)
    a = {k: inp[k].copy() for k in ("xf", "yf", "zf", "gxcf", "gycf", "gzcf", "a1cf",
                                     "a2cf", "a3cf", "ccf", "fcf", "tcf")}
    ierr = fn(iparm, rparm, iw, rw, u, a["xf"], a["yf"], a["zf"], a["gxcf"], a["gycf"],
              a["gzcf"], a["a1cf"], a["a2cf"], a["a3cf"], a["ccf"], a["fcf"], a["tcf"])
    return ierr, u, iw, rw


@pytest.mark.parametrize("n,nlev,natoms,R,seed", [(33, 3, 40, 6.0, 1), (65, 4, 200, 14.0, 20261202)])
def test_mgdriv_parity(mg, orc, n, nlev, natoms, R, seed):
    print(f"mgdriv parity n={n} nlev={nlev}", flush=True)
    inp = orc.pbe_synth(n, 0.5, natoms, R, 0.15, 0, seed)
    e1, u1, iw1, rw1 = _run_driver(mg, mg.Vmgdriv, inp, n, nlev)
    it1 = mg.last_stats()[0]
    e0, u0, iw0, rw0 = orc.solve(inp, n, nlev)
    it0 = orc.last_stats()[0]
    print(f"  iters gpu={it1} oracle={it0} potential rel err={_rel(u1, u0):.2e}", flush=True)
    assert e1 == e0 == 0
    assert it1 == it0
    assert _rel(u1, u0) < 1e-10
    assert np.array_equal(iw1, iw0)
    # coarse operators (Galerkin) in the rwork layout: ac of levels >= 2
    sz = mg.Vmgsz(2, 0, 0, n, n, n, nlev)
    k_ac = sz["n_rpc"] + 2 * sz["narr"] + 27 * sz["narrc"]
    fine = 4 * sz["nf"]
    assert _rel(rw1[k_ac + fine:k_ac + fine + 14 * sz["narrc"]],
                rw0[k_ac + fine:k_ac + fine + 14 * sz["narrc"]]) < 1e-12
    assert _rel(rw1, rw0) < 1e-12


def test_ctx_fp64_and_mixed(mg, orc):
    print("ctx fp64 + mixed n=65", flush=True)
    n, nlev = 65, 4
    inp = orc.pbe_synth(n, 0.5, 200, 14.0, 0.15, 0, 20261202)
    iparm, rparm = mg.make_params(n, nlev)
    ctx = mg.Context(iparm, rparm, inp)
    ctx.solve(mg.PMG_FP64)
    u64, s64 = ctx.u(), ctx.stats()
    ctx.solve(mg.PMG_MIXED)  # first FP32 use: operators converted after setup
    umx, smx = ctx.u(), ctx.stats()
    ctx.rebuild()  # FP32 copies now written by the setup kernels themselves
    ctx.solve(mg.PMG_MIXED)
    umx2 = ctx.u()
    ctx.close()
    assert np.array_equal(umx2, umx), "FP32 operators from setup differ"
    _, u0, _, _ = orc.solve(inp, n, nlev)
    it0 = orc.last_stats()[0]
    print(f"  fp64 iters={s64['iters']} mixed iters={smx['iters']} oracle={it0} "
          f"mixed err={_rel(umx, u0):.2e}", flush=True)
    assert s64["iters"] == it0 and _rel(u64, u0) < 1e-10
    assert smx["relres"] <= 1e-6 and smx["ierror"] == 0
    assert _rel(umx, u0) < 1e-6
    # the FP32 cycle (with its coarsest CG stopped at a relative 1e-3)
    # contracts as the FP64 cycle does: as many outer iterations, and the
    # potentials agree to the solver tolerance
    assert smx["iters"] == s64["iters"] and _rel(umx, u64) < 1e-6


def test_newdriv_parity(mg, orc):
    print("newdriv parity n=33", flush=True)
    n, nlev = 33, 3
    ions = ([1.0, -1.0], [0.5, 0.5])
    mg.Vmypdefinitnpbe(2, *ions)
    orc.mypdefinit(*ions)
    inp = orc.pbe_synth(n, 0.5, 40, 6.0, 0.3, 0, 5)
    e1, u1, _, _ = _run_driver(mg, mg.Vnewdriv, inp, n, nlev, nonlin=True, errtol=1e-8)
    it1 = mg.last_stats()[0]
    e0, u0, _, _ = orc.solve(inp, n, nlev, nonlin=True, errtol=1e-8)
    it0 = orc.last_stats()[0]
    print(f"  newton iters gpu={it1} oracle={it0} err={_rel(u1, u0):.2e}", flush=True)
    assert e1 == e0 == 0 and it1 == it0
    assert _rel(u1, u0) < 1e-9


def _newton_problem(mg, orc, n, seed):
    ions = ([1.0, -1.0], [0.5, 0.5])
    mg.Vmypdefinitnpbe(2, *ions)
    orc.mypdefinit(*ions)
    return orc.pbe_synth(n, 0.5, 40, 6.0, 0.3, 0, seed)


@pytest.mark.parametrize("n,nlev", [(33, 3), (65, 4)])
def test_nested_newton_parity(mg, orc, n, nlev):
    """mgkey 1 (Vfnewton, newtond.cc:60-138): Newton on each level from the
    coarsest up; the same Newton iteration count and per-iteration relative
    residuals as the oracle, the potential within 1e-9."""
    print(f"nested newton parity n={n} nlev={nlev}", flush=True)
    inp = _newton_problem(mg, orc, n, 5)
    e1, u1, _, _ = _run_driver(mg, mg.Vnewdriv, inp, n, nlev, nonlin=True, errtol=1e-8, mgkey=1)
    it1, h1 = mg.last_stats()[0], mg.last_history()
    e0, u0, _, _ = orc.solve(inp, n, nlev, nonlin=True, errtol=1e-8, mgkey=1)
    it0, _, _, h0 = orc.last_stats()
    print(f"  nested newton iters gpu={it1} oracle={it0} err={_rel(u1, u0):.2e}", flush=True)
    assert e1 == e0 == 0 and it1 == it0 == len(h1) == len(h0)
    assert np.allclose(h1, h0, rtol=1e-6, atol=1e-9 * h0[0])
    assert _rel(u1, u0) < 1e-9


def test_newton_from_callers_u(mg, orc):
    """Vnewdriv starts from the interior of the caller's u (boundary zeroed,
    VfboundPMG00 newdrvd.cc:263): a perturbed converged potential takes the
    oracle's Newton iterations from that start, to the same potential."""
    print("newton from caller's u n=33", flush=True)
    n, nlev = 33, 3
    inp = _newton_problem(mg, orc, n, 6)
    _, u0, _, _ = orc.solve(inp, n, nlev, nonlin=True, errtol=1e-10)
    it_zero = orc.last_stats()[0]
    rng = np.random.default_rng(7)
    start = u0 * (1.0 + 1e-3 * rng.standard_normal(u0.size))
    iparm, rparm = mg.make_params(n, nlev, nonlin=1, errtol=1e-10)
    sz = mg.Vmgsz(2, 0, 0, n, n, n, nlev)
    runs = []
    for fn in (mg.Vnewdriv, orc.newdriv):
        a = {k: inp[k].copy() for k in mg.PBE_KEYS}
        u = start.copy()
        e = fn(iparm.copy(), rparm.copy(), np.zeros(sz["iintot"], np.int32),
               np.zeros(int(iparm[0])), u, *(a[k] for k in mg.PBE_KEYS))
        runs.append((e, u, (mg.last_stats() if fn is mg.Vnewdriv else orc.last_stats())[0]))
    (e1, u1, it1), (e0, u0b, it0) = runs
    print(f"  iters gpu={it1} oracle={it0} (from zero {it_zero})", flush=True)
    assert e1 == e0 == 0 and it1 == it0 < it_zero
    assert _rel(u1, u0b) < 1e-9


def test_ctx_nested_newton_mixed(mg, orc):
    """The resident context's Newton (nested, mgkey 1) in FP64 and with FP32
    inner cycles: FP64 as the drop-in, mixed within 1e-6."""
    print("ctx nested newton fp64 + mixed n=33", flush=True)
    n, nlev = 33, 3
    inp = _newton_problem(mg, orc, n, 5)
    iparm, rparm = mg.make_params(n, nlev, nonlin=1, errtol=1e-8, mgkey=1)
    ctx = mg.Context(iparm, rparm, inp)
    ctx.solve(mg.PMG_FP64)
    u64, s64 = ctx.u(), ctx.stats()
    ctx.solve(mg.PMG_MIXED)
    umx, smx = ctx.u(), ctx.stats()
    ctx.close()
    _, u0, _, _ = orc.solve(inp, n, nlev, nonlin=True, errtol=1e-8, mgkey=1)
    it0 = orc.last_stats()[0]
    print(f"  fp64 iters={s64['iters']} mixed iters={smx['iters']} oracle={it0} "
          f"mixed err={_rel(umx, u0):.2e}", flush=True)
    assert s64["iters"] == it0 and _rel(u64, u0) < 1e-9
    assert smx["ierror"] == 0 and _rel(umx, u0) < 1e-6


def test_error_paths(mg, orc):
    print("error paths", flush=True)
    n, nlev = 17, 3
    inp = orc.pbe_synth(n, 0.5, 8, 2.0, 0.15, 0, 1)
    # itmax exhausted -> ierror 1
    e, _, _, _ = _run_driver(mg, mg.Vmgdriv, inp, n, nlev, itmax=1, errtol=1e-14)
    assert e == 1 and mg.last_stats()[1] == 1
    # unsupported smoother -> PMGError(EBADPARM)
    with pytest.raises(mg.PMGError) as ex:
        _run_driver(mg, mg.Vmgdriv, inp, n, nlev, mgsmoo=0)
    assert ex.value.code == mg.PMG_EBADPARM
    # zero right-hand side: rsden = 1 branch, u = 0 (pmgc/mgcsd.cc:154-157)
    z = {k: (np.zeros_like(v) if k in ("fcf", "gxcf", "gycf", "gzcf") else v) for k, v in inp.items()}
    e, u, _, _ = _run_driver(mg, mg.Vmgdriv, z, n, nlev)
    assert e == 0 and np.all(u == 0.0)


def test_full_size_pb161_properties(mg, orc):
    """161^3 / 5 levels: the FP64 GPU solution's L1 relative residual,
    recomputed by the oracle on the host, meets 1e-6; the mixed-precision
    solution agrees with it to 1e-6."""
    print("pb161 161^3 full-size properties", flush=True)
    c = mg.CONFIGS["pb161"]
    n, nlev = c["n"], c["nlev"]
    atoms = mg.pbe_molecule(c["natoms"], c["sphere_r"], c["charge_mode"], c["seed"])
    inp = orc.pbe_assemble(n, c["h"], c["ionic"], atoms)
    iparm, rparm = mg.make_params(n, nlev)
    # inputs assembled on the GPU from the atom list
    ctx = mg.Context(iparm, rparm, pbe=(c["h"], c["ionic"], atoms))
    ctx.solve(mg.PMG_FP64)
    u64, s64 = ctx.u(), ctx.stats()
    ctx.solve(mg.PMG_MIXED)
    umx, smx = ctx.u(), ctx.stats()
    ctx.close()
    ac, cc, fc, _ = orc.buildA_fv(n, inp)
    # residual of the interior unknowns (Dirichlet values are folded into fc)
    ub = u64.copy().reshape(n, n, n)
    ub[0, :, :] = ub[-1, :, :] = 0
    ub[:, 0, :] = ub[:, -1, :] = 0
    ub[:, :, 0] = ub[:, :, -1] = 0
    r = orc.mresid(n, 7, ac, cc, fc, ub.ravel())
    rel = np.abs(r).sum() / np.abs(fc).sum()
    print(f"  fp64 iters={s64['iters']} relres={s64['relres']:.2e} host={rel:.2e} "
          f"mixed iters={smx['iters']} diff={_rel(umx, u64):.2e}", flush=True)
    assert s64["ierror"] == 0 and rel <= 1e-6 * 1.0001
    assert abs(rel - s64["relres"]) <= 1e-9
    assert _rel(umx, u64) < 1e-6


def test_repeatable_bit_for_bit(mg, orc):
    """Every kernel is deterministic: repeated solves on one context and on
    a second context give identical bits (a race in LDS or in a reduction
    shows up here)."""
    print("repeatability 65^3 / 97^3", flush=True)
    for n, nlev in ((65, 4), (97, 5)):
        atoms = mg.pbe_molecule(150, 10.0, 1, 31 + n)
        iparm, rparm = mg.make_params(n, nlev)
        ref = {}
        for rep in range(2):
            ctx = mg.Context(iparm, rparm, pbe=(0.5, 0.15, atoms))
            for _ in range(3):
                for mode in (mg.PMG_FP64, mg.PMG_MIXED):
                    ctx.solve(mode)
                    u = ctx.u()
                    if mode not in ref:
                        ref[mode] = u
                    assert np.array_equal(u, ref[mode]), (n, rep, mode)
            ctx.close()

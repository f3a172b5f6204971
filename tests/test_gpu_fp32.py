"""Kernel-level parity of the mixed-precision inner cycle: the FP32
instances of the smoother (including the zero-start first sweep), the
residual, restriction, prolongation and the damped coarse-grid correction,
plus the FP64-to-FP32 outer residual and the FP64 += FP32 update, each
called through the C ABI on operators and vectors rounded to FP32 and
compared with the FP64 oracle applied to the same rounded values.

Tolerances are FP32 rounding bounds: a sum of m terms carries at most
about m * 2^-24 of the sum of the terms' magnitudes, so single-pass kernels
are checked point by point against 32 * 2^-24 * (the same operation on the
absolute values); three smoothing sweeps against 1e-5 of the largest
value.  An index, colour or weight mistake makes errors of order one.
The FP64 coarse-grid correction (pmg_correct) and the zero-start FP64
sweep (pmg_gsrb_zero) are checked at the FP64 bars of the other kernels."""
import ctypes

import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu

EPS32 = 2.0 ** -24
_F = ctypes.POINTER(ctypes.c_float)
_D = ctypes.POINTER(ctypes.c_double)


def fp(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_F)


def dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def _call(mg, name, *args):
    rc = getattr(mg.lib(), name)(*args)
    assert rc == 0, f"{name} returned {rc}"


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def _back(a32):
    return np.ascontiguousarray(a32, np.float64)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _op(dims, numdia, seed):
    """Random diagonally dominant symmetric operator with couplings into the
    boundary zero (as the reference builds them), rounded to FP32, plus
    random interior vectors f, x."""
    nx, ny, nz = dims
    rng = np.random.default_rng(seed)
    if nx == ny == nz:
        ac, cc = H.random_op_vec(nx, numdia, rng)
    else:
        N = nx * ny * nz
        nplanes = 4 if numdia == 7 else 14
        ac = rng.uniform(0.1, 1.0, nplanes * N)
        ac[:N] += 2.0 * (nplanes - 1)
        cc = rng.uniform(0.0, 1.0, N)
    mask = np.zeros((nz, ny, nx))
    mask[1:-1, 1:-1, 1:-1] = 1
    f = (rng.standard_normal((nz, ny, nx)) * mask).ravel()
    x = (rng.standard_normal((nz, ny, nx)) * mask).ravel()
    return _f32(ac), _f32(cc), _f32(f), _f32(x)


def _abs_matvec(orc, dims, numdia, ac, cc, x):
    """|A| |x| with the oracle's matvec: off-diagonal planes negated."""
    nx, ny, nz = dims
    N = nx * ny * nz
    a = np.abs(_back(ac))
    a[N:] = -a[N:]
    y = np.zeros(N)
    orc.olib().pmgo_matvec(nx, ny, nz, numdia, dp(a), dp(np.abs(_back(cc))),
                           dp(np.abs(_back(x))), dp(y))
    return y


# (145, 37, 21), 27 points: rows longer than 129 points (two point pairs per
# lane, all four points' operands loaded before any update).  (126, 40, 1000)
# and (131, 70, 300), 7 points: the residual's march (k chunks of 24 and 18
# planes on 256 compute units) over several tiles in i and j, with ragged
# last tiles and ragged last chunks.
CASES = [((17, 17, 17), 7), ((17, 17, 17), 27), ((97, 97, 97), 7), ((97, 97, 97), 27),
         ((33, 25, 19), 7), ((33, 25, 19), 27), ((145, 37, 21), 27),
         ((126, 40, 1000), 7), ((131, 70, 300), 7)]


@pytest.mark.parametrize("dims,numdia", CASES)
def test_mresid_f32(mg, orc, dims, numdia):
    print(f"mresid f32 dims={dims} numdia={numdia}", flush=True)
    nx, ny, nz = dims
    ac, cc, f, x = _op(dims, numdia, sum(dims) + numdia)
    r = np.zeros(nx * ny * nz, np.float32)
    _call(mg, "pmg_mresid_f32", nx, ny, nz, numdia, fp(ac), fp(cc), fp(f), fp(x), fp(r))
    ref = np.zeros(nx * ny * nz)
    orc.olib().pmgo_mresid(nx, ny, nz, numdia, dp(_back(ac)), dp(_back(cc)), dp(_back(f)),
                           dp(_back(x)), dp(ref))
    mag = np.abs(_back(f)) + _abs_matvec(orc, dims, numdia, ac, cc, x)
    err = np.abs(_back(r) - ref)
    assert np.all(err <= 32 * EPS32 * mag + 1e-30), float(np.max(err / (mag + 1e-30)))


@pytest.mark.parametrize("dims,numdia", CASES)
@pytest.mark.parametrize("zero_start", [0, 1])
def test_gsrb_f32(mg, orc, dims, numdia, zero_start):
    print(f"gsrb f32 dims={dims} numdia={numdia} zero_start={zero_start}", flush=True)
    nx, ny, nz = dims
    ac, cc, f, x0 = _op(dims, numdia, 3 * sum(dims) + numdia)
    for iadj, sweeps in ((0, 2), (1, 3)):
        x = x0.copy()
        if zero_start:
            x[x != 0] = np.float32(np.nan)  # never read from the interior
        _call(mg, "pmg_gsrb_f32", nx, ny, nz, numdia, fp(ac), fp(cc), fp(f), fp(x), sweeps,
              iadj, zero_start)
        start = np.zeros(nx * ny * nz) if zero_start else _back(x0)
        ref = orc.gsrb(0, numdia, _back(ac), _back(cc), _back(f), start, sweeps, iadj,
                       dims=dims)
        assert np.all(np.isfinite(x))
        assert _rel(_back(x), ref) <= 1e-5, (iadj, sweeps, _rel(_back(x), ref))


@pytest.mark.parametrize("dims,numdia", CASES)
def test_gsrb_zero_start_fp64(mg, orc, dims, numdia):
    """The FP64 zero-start sweeps (the coarse levels' first smoothing) equal
    the oracle's sweeps from x = 0 to the FP64 sweep bar."""
    print(f"gsrb zero-start fp64 dims={dims} numdia={numdia}", flush=True)
    nx, ny, nz = dims
    ac, cc, f, _ = (_back(a) for a in _op(dims, numdia, 5 * sum(dims) + numdia))
    for iadj, sweeps in ((0, 2), (1, 3)):
        x = np.zeros(nx * ny * nz)
        x[_interior(dims).ravel() > 0] = np.nan  # never read from the interior
        _call(mg, "pmg_gsrb_zero", nx, ny, nz, numdia, dp(ac), dp(cc), dp(f), dp(x), sweeps,
              iadj)
        ref = orc.gsrb(0, numdia, ac, cc, f, np.zeros(nx * ny * nz), sweeps, iadj, dims=dims)
        assert _rel(x, ref) <= 1e-12, (iadj, sweeps)


TRANSFER = [(17, 9, 7), (17, 9, 27), (97, 49, 7), (97, 49, 27)]


def _transfer_op(orc, nf, nc, numdia, seed):
    rng = np.random.default_rng(seed)
    ac, _ = H.random_op_vec(nf, numdia, rng)
    pc = orc.buildP(nf, nc, 1, numdia, ac)
    return rng, pc


@pytest.mark.parametrize("nf,nc,numdia", TRANSFER)
def test_restrc_interp_f32(mg, orc, nf, nc, numdia):
    print(f"restrc/interp f32 {nf}->{nc} numdia={numdia}", flush=True)
    rng, pc = _transfer_op(orc, nf, nc, numdia, nf + nc + numdia)
    pc32 = _f32(pc)
    r = _f32((rng.standard_normal((nf, nf, nf)) * _interior(nf)).ravel())
    xc = np.zeros(nc ** 3, np.float32)
    _call(mg, "pmg_restrc_f32", nf, nf, nf, nc, nc, nc, fp(r), fp(xc), fp(pc32))
    ref = orc.restrc(nf, nc, _back(r), _back(pc32))
    mag = orc.restrc(nf, nc, np.abs(_back(r)), np.abs(_back(pc32)))
    assert np.all(np.abs(_back(xc) - ref) <= 32 * EPS32 * mag + 1e-30)
    e = _f32((rng.standard_normal((nc, nc, nc)) * _interior(nc)).ravel())
    xf = np.zeros(nf ** 3, np.float32)
    _call(mg, "pmg_interp_f32", nc, nc, nc, nf, nf, nf, fp(e), fp(xf), fp(pc32))
    ref = orc.interp(nc, nf, _back(e), _back(pc32))
    mag = orc.interp(nc, nf, np.abs(_back(e)), np.abs(_back(pc32)))
    assert np.all(np.abs(_back(xf) - ref) <= 32 * EPS32 * mag + 1e-30)


def _interior(n):
    """Interior mask of an n^3 grid, or of dims (nx, ny, nz), as [k, j, i]."""
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    m = np.zeros((nz, ny, nx))
    m[1:-1, 1:-1, 1:-1] = 1
    return m


def _correct_ref(orc, nf, nc, numdia, acc, ccc, xc, w0c, pc, xf):
    """xf + xdamp P xc with xdamp = (xc . w0c) / (xc . A_c xc) computed from
    its defining formula (mgcsd.cc:429-452, dots over the interior)."""
    y = orc.matvec(nc, numdia, acc, ccc, xc)
    m = _interior(nc).ravel() > 0
    num, den = float(np.dot(xc[m], w0c[m])), float(np.dot(xc[m], y[m]))
    xdamp = num / den if den != 0.0 else 0.0
    out = xf.copy()
    mf = _interior(nf).ravel() > 0
    out[mf] += xdamp * orc.interp(nc, nf, xc, pc)[mf]
    return out, xdamp


@pytest.mark.parametrize("nf,nc,numdia", TRANSFER)
def test_coarse_correction(mg, orc, nf, nc, numdia):
    """The damped coarse-grid correction in FP64 (1e-13) and FP32 (1e-5).
    The coarse operator is a Galerkin 27-point one (what the cycle's coarse
    levels hold); the coarse solution is a perturbed solve so xdamp is near
    but not at 1."""
    print(f"coarse correction {nf}->{nc} numdia={numdia}", flush=True)
    rng = np.random.default_rng(7 * nf + numdia)
    acf, _ = H.random_op_vec(nf, numdia, rng)
    pc = orc.buildP(nf, nc, 1, numdia, acf)
    acc = orc.buildG(nf, nc, numdia, pc, acf)
    ccc = np.zeros(nc ** 3)
    w0c = (rng.standard_normal((nc, nc, nc)) * _interior(nc)).ravel()
    xc = (rng.standard_normal((nc, nc, nc)) * _interior(nc)).ravel()
    xf = (rng.standard_normal((nf, nf, nf)) * _interior(nf)).ravel()
    ref, xdamp = _correct_ref(orc, nf, nc, 27, acc, ccc, xc, w0c, pc, xf)
    out = xf.copy()
    _call(mg, "pmg_correct", nc, nc, nc, nf, nf, nf, 27, dp(acc), dp(ccc), dp(xc), dp(w0c),
          dp(pc), dp(out))
    print(f"  xdamp {xdamp:.6f} fp64 err {_rel(out, ref):.2e}", flush=True)
    assert _rel(out, ref) <= 1e-13
    a32 = [_f32(v) for v in (acc, ccc, xc, w0c, pc, xf)]
    ref32, _ = _correct_ref(orc, nf, nc, 27, *(_back(v) for v in a32))
    out32 = a32[5].copy()
    _call(mg, "pmg_correct_f32", nc, nc, nc, nf, nf, nf, 27, *(fp(v) for v in a32[:5]),
          fp(out32))
    assert _rel(_back(out32), ref32) <= 1e-5
    # a zero coarse solution gives a zero correction (xdamp 0/0 -> 0)
    z = np.zeros(nc ** 3)
    out = xf.copy()
    _call(mg, "pmg_correct", nc, nc, nc, nf, nf, nf, 27, dp(acc), dp(ccc), dp(z), dp(w0c),
          dp(pc), dp(out))
    assert np.array_equal(out, xf)


@pytest.mark.parametrize("n", [17, 97, 161])
def test_mixed_outer_residual_and_update(mg, orc, n):
    """r = f - A u in FP64 stored as FP32 equals the oracle's FP64 residual
    rounded to FP32 (within one FP32 rounding), with its FP64 L1 norm; the
    update u += e adds the FP32 correction exactly."""
    print(f"mixed outer residual / update n={n}", flush=True)
    rng = np.random.default_rng(n)
    ac, cc = H.random_op_vec(n, 7, rng)
    m = _interior(n).ravel()
    f = rng.standard_normal(n ** 3) * m
    u = rng.standard_normal(n ** 3)  # boundary values are read as Dirichlet data
    r = np.zeros(n ** 3, np.float32)
    l1 = ctypes.c_double()
    _call(mg, "pmg_mresid_mixed", n, n, n, dp(ac), dp(cc), dp(f), dp(u), fp(r),
          ctypes.byref(l1))
    ref = orc.mresid(n, 7, ac, cc, f, u)
    # one FP32 rounding of the FP64 residual, which itself may differ from
    # the oracle's by FP64 roundings of the terms (multiply-add contraction)
    mag = np.abs(f) + _abs_matvec(orc, (n, n, n), 7, ac, cc, u)
    assert np.all(np.abs(_back(r) - ref) <= EPS32 * np.abs(ref) + 1e-13 * mag)
    l1ref = orc.olib().pmgo_xnrm1(n, n, n, dp(ref))
    assert abs(l1.value - l1ref) <= 1e-13 * l1ref
    e = _f32(rng.standard_normal(n ** 3) * m)
    uu = u.copy()
    _call(mg, "pmg_addcorr_f32", n, n, n, fp(e), dp(uu))
    want = u.copy()
    want[m > 0] += _back(e)[m > 0]
    assert np.array_equal(uu, want)

"""Pins of the oracle's banded coarsest solve (mgsolv = 1): Vbuildband
(pmgc/buildBd.cc:62-304) and Vdpbfa / Vdpbsl (pmgc/mlinpckd.cc:167-235)
against dense numpy linear algebra, and of the V-cycle and drivers that use
it (pmgc/mgcsd.cc:195-212, 391-407; pmgc/mgsubd.cc:249-262)."""
import numpy as np
import pytest

from tests import helpers as H


def band_upper(abd, lda, nu, m):
    """Dense upper triangle stored in LINPACK band form: A(i, j), i <= j, at
    abd(m + 1 + i - j, j) (1-based, column-major)."""
    cols = abd.reshape(nu, lda)
    R = np.zeros((nu, nu))
    for j in range(nu):
        for i in range(max(0, j - m), j + 1):
            R[i, j] = cols[j, m + i - j]
    return R


def test_dpbfa_dpbsl_on_random_banded_spd(orc):
    """Vdpbfa / Vdpbsl alone: a random SPD matrix of half-bandwidth m packed
    into band form factors to numpy's Cholesky factor and solves as numpy."""
    rng = np.random.default_rng(5)
    nu, m = 40, 6
    L = np.tril(rng.uniform(-1, 1, (nu, nu)))
    L = np.triu(L, -m)
    L[np.arange(nu), np.arange(nu)] = rng.uniform(1.0, 2.0, nu)
    A = L @ L.T  # half-bandwidth m
    lda = m + 1
    abd = np.zeros(nu * lda)
    cols = abd.reshape(nu, lda)
    for j in range(nu):
        for i in range(max(0, j - m), j + 1):
            cols[j, m + i - j] = A[i, j]
    info = np.zeros(1, np.int32)
    orc.olib().pmgo_dpbfa(orc.dp(abd), lda, nu, m, orc.ip(info))
    assert info[0] == 0
    R = band_upper(abd, lda, nu, m)
    np.testing.assert_allclose(R, np.linalg.cholesky(A).T, rtol=0, atol=1e-12 * np.abs(R).max())
    b = rng.standard_normal(nu)
    x = orc.dpbsl(abd, lda, nu, m, b)
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=0, atol=1e-11 * np.abs(x).max())


# Plane index of the symmetric 27-point storage (pmgc/matvecd.cc:185-189).
PL = dict(oC=0, oE=1, oN=2, uC=3, oNE=4, oNW=5, uE=6, uW=7, uN=8, uS=9,
          uNE=10, uNW=11, uSE=12, uSW=13)
# The band takes the coupling of point (i,j,k) with a lower neighbour from a
# plane at (i, j-1, k) or (i, j, k-1) in the mirrored direction
# (buildBd.cc:192-245), where the operator applied by Vmatvec27
# (matvecd.cc:137-175) holds it at the neighbour: e.g. (i+1, j-1, k) couples
# through oNW(i+1, j-1, k) there and through oNE(i, j-1, k) in the band.  The
# banded matrix is therefore the operator of these shifted planes:
# target plane <- (source plane, shift applied as out[p] = src[p - s]).
BAND_PLANES = {"oNW": ("oNE", (1, 0, 0)), "oNE": ("oNW", (-1, 0, 0)),
               "uW": ("uE", (1, 0, 0)), "uE": ("uW", (-1, 0, 0)),
               "uS": ("uN", (0, 1, 0)), "uN": ("uS", (0, -1, 0)),
               "uSW": ("uNE", (1, 1, 0)), "uSE": ("uNW", (-1, 1, 0)),
               "uNW": ("uSE", (1, -1, 0)), "uNE": ("uSW", (-1, -1, 0))}


def band_operator(n, numdia, ac):
    """Dense matrix the reference's band holds for the planes ac (zero
    couplings to boundary points assumed, as the Galerkin operators have)."""
    N = n ** 3
    planes = ac.reshape(-1, n, n, n)
    out = planes.copy()
    if numdia == 27:
        for tgt, (src, (ex, ey, ez)) in BAND_PLANES.items():
            out[PL[tgt]] = H._shift_from(planes[PL[src]], ex, ey, ez)
    A, _ = H.dense_A(n, numdia, out.reshape(-1), np.zeros(N))
    return A


@pytest.mark.parametrize("n,numdia", [(5, 27), (6, 27), (4, 7), (6, 7)])
def test_band_factor_is_cholesky_of_banded_operator(orc, n, numdia):
    """The factored band is the Cholesky factor R (A = R^T R, positive
    diagonal) of the banded matrix: the diagonal is oC alone (cc is not part
    of it, buildBd.cc:104, 180) and the diagonal-direction couplings come
    from the mirrored planes."""
    rng = np.random.default_rng(n * 100 + numdia)
    ac, cc = H.random_op(n, numdia, rng)
    info, abd, ipcb = orc.buildband(n, numdia, ac)
    nu, m, lda = orc.band_dims(n, numdia)
    assert info == 0
    assert list(ipcb) == [nu, m, lda, 1]
    A = band_operator(n, numdia, ac)
    R = band_upper(abd, lda, nu, m)
    assert np.all(np.diag(R) > 0)
    np.testing.assert_allclose(R.T @ R, A, rtol=0, atol=1e-13 * np.abs(A).max())
    np.testing.assert_allclose(R, np.linalg.cholesky(A).T, rtol=0,
                               atol=1e-13 * np.abs(R).max())


def test_band_offsets_collide_with_two_interior_columns(orc):
    """With nx - 2 == 2 several of the 13 band offsets of buildBd.cc:182-245
    coincide (east = north-east = up-north-east = 1, ...) and the later
    writes win: the banded matrix then differs from the interior operator
    and, for an operator whose boundary couplings vanish, loses the
    east coupling of every point with i = nx - 1.  The oracle keeps this,
    as a drop-in must."""
    n = 4
    rng = np.random.default_rng(21)
    ac, _ = H.random_op(n, 27, rng)
    info, abd, ipcb = orc.buildband(n, 27, ac)
    nu, m, lda = orc.band_dims(n, 27)
    assert info == 0
    R = band_upper(abd, lda, nu, m)
    B = R.T @ R
    A, pts = H.dense_A(n, 27, ac, np.zeros(n ** 3))
    np.testing.assert_allclose(np.diag(B), np.diag(A), rtol=1e-13)
    for a, (i, j, k) in enumerate(pts):
        if i == 3:  # east (lower) neighbour at a - 1 is (2, j, k)
            assert A[a - 1, a] != 0.0
            assert abs(B[a - 1, a]) < 1e-13 * np.abs(A).max()


def test_band_equals_operator_for_mirror_symmetric_planes(orc):
    """Where each diagonal-direction plane is constant and equal to its
    mirror (oNE = oNW, uE = uW, uN = uS, uNE = uSW, uNW = uSE), as for a
    Galerkin-coarsened constant Laplacian, the band holds exactly the
    operator of Vmatvec27 (with cc removed)."""
    n = 6
    rng = np.random.default_rng(9)
    ac, _ = H.random_op(n, 27, rng)
    planes = ac.reshape(14, n, n, n)
    mask = np.zeros((n, n, n), bool)
    mask[1:-1, 1:-1, 1:-1] = True
    for a, b, v in (("oNE", "oNW", 0.11), ("uE", "uW", 0.07), ("uN", "uS", 0.05),
                    ("uNE", "uSW", 0.03), ("uNW", "uSE", 0.02)):
        for name in (a, b):
            pl = PL[name]
            ex, ey, ez = H.UPPER27[pl]
            nb = H._shift_from(mask.astype(float), -ex, -ey, -ez) > 0
            planes[pl] = v * (mask & nb)
    info, abd, ipcb = orc.buildband(n, 27, ac)
    nu, m, lda = orc.band_dims(n, 27)
    R = band_upper(abd, lda, nu, m)
    A, _ = H.dense_A(n, 27, ac, np.zeros(n ** 3))
    np.testing.assert_allclose(R.T @ R, A, rtol=0, atol=1e-13 * np.abs(A).max())


@pytest.mark.parametrize("n,numdia", [(6, 27), (6, 7)])
def test_dpbsl_matches_dense_solve(orc, n, numdia):
    rng = np.random.default_rng(7 + numdia)
    ac, _ = H.random_op(n, numdia, rng)
    info, abd, ipcb = orc.buildband(n, numdia, ac)
    nu, m, lda = orc.band_dims(n, numdia)
    A = band_operator(n, numdia, ac)
    b = rng.standard_normal(nu)
    x = orc.dpbsl(abd, lda, nu, m, b)
    ref = np.linalg.solve(A, b)
    assert np.abs(x - ref).max() <= 1e-12 * np.abs(ref).max()


def test_bandsolve_packs_interior_and_zeroes_boundary(orc):
    n = 6
    rng = np.random.default_rng(3)
    ac, _ = H.random_op(n, 27, rng)
    info, abd, ipcb = orc.buildband(n, 27, ac)
    A = band_operator(n, 27, ac)
    rhs = rng.standard_normal(n ** 3)
    x = orc.bandsolve(n, abd, ipcb, rhs)
    ref = np.linalg.solve(A, H.to_int(n, rhs))
    np.testing.assert_allclose(H.to_int(n, x), ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    xb = x.reshape(n, n, n)
    for face in (xb[0], xb[-1], xb[:, 0], xb[:, -1], xb[:, :, 0], xb[:, :, -1]):
        assert np.all(face == 0.0)


def test_dpbfa_reports_first_nonpositive_pivot(orc):
    """info = the first column whose leading principal minor is not
    positive definite (mlinpckd.cc:197-201)."""
    n = 6
    rng = np.random.default_rng(11)
    ac, _ = H.random_op(n, 27, rng)
    bad = 10  # 0-based interior index whose diagonal is made negative
    i, j, k = H.interior(n)[bad]
    ac[H.IX(n, i, j, k)] = -1.0
    info, _, ipcb = orc.buildband(n, 27, ac)
    A = band_operator(n, 27, ac)
    first = next(s for s in range(1, A.shape[0] + 1)
                 if np.any(np.linalg.eigvalsh(A[:s, :s]) <= 0))
    assert info == first == bad + 1
    assert ipcb[3] == 1


def _band_region(orc, n, nlev, iw, rw):
    sz = orc.mgsz(2, 0, 1, n, n, n, nlev)[1]
    nrpc, narr, narrc, niz = int(sz[7]), int(sz[5]), int(sz[6]), int(sz[8])
    iz = iw[:niz].reshape(nlev + 1, 50)
    acb = nrpc + 2 * narr + 27 * narrc
    ipcb = iw[niz + iz[nlev, 4] - 1: niz + iz[nlev, 4] - 1 + 4]
    band = rw[acb + iz[nlev, 6] - 1: acb + iz[nlev, 6] - 1 + int(ipcb[0]) * int(ipcb[2])]
    coarse = rw[acb + iz[nlev - 1, 6] - 1:]
    return ipcb, band, coarse


def test_driver_nlev1_direct_solve_is_exact_without_salt(orc):
    """One level, mgsolv 1: the driver solves the fine system directly
    (mgcsd.cc:195-212); with no ions cc = 0, so the result is the exact
    discrete solution of the 7-point operator."""
    n = 9
    c, inp = orc.config_inputs("pb65", n=n, h=1.0, natoms=6, sphere_r=2.5, ionic=0.0)
    ierr, u, iw, rw = orc.solve(inp, n, 1, mgsolv=1)
    assert ierr == 0
    it, ie, rr, hist = orc.last_stats()
    assert it == 1 and rr < 1e-13


def test_driver_nlev1_band_omits_cc(orc):
    """With ions the band still carries oC alone, so the one-level direct
    solve returns (A - diag(cc))^{-1} fc on the interior."""
    n = 9
    c, inp = orc.config_inputs("pb65", n=n, h=1.0, natoms=6, sphere_r=2.5)
    ierr, u, iw, rw = orc.solve(inp, n, 1, mgsolv=1)
    sz = orc.mgsz(2, 0, 1, n, n, n, 1)[1]
    nrpc, narr = int(sz[7]), int(sz[5])
    N = n ** 3
    cc = rw[nrpc: nrpc + N]
    fc = rw[nrpc + narr: nrpc + narr + N]
    assert np.abs(cc).max() > 0
    acb = nrpc + 2 * narr
    ac = rw[acb: acb + 4 * N]
    A, _ = H.dense_A(n, 7, ac, np.zeros(N))
    ref = np.linalg.solve(A, H.to_int(n, fc))
    np.testing.assert_allclose(H.to_int(n, u), ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def test_driver_multilevel_band_layout_and_convergence(orc):
    """Multilevel V-cycles with the banded coarsest solve converge; the
    factored band sits at ac(iz(7, nlev+1)) with ipc(iz(5, nlev+1)) =
    {n, m, lda, 1} (mgsubd.cc:249-262, mgcsd.cc:398-401) and is the factor of
    the coarsest Galerkin operator left in rwork; the potential agrees with
    the CG coarsest solve to the stopping tolerance."""
    import multigrid_amd as mg
    n, nlev = 33, 3
    # uniform dielectric: the band of the coarsest operator is SPD
    inp = orc.pbe_assemble(n, 0.5, 0.15, mg.pbe_molecule(40, 6.0, 0, 7), eps_in=78.54)
    ierr1, u1, iw1, rw1 = orc.solve(inp, n, nlev, mgsolv=1)
    it1 = orc.last_stats()[0]
    ierr0, u0, iw0, rw0 = orc.solve(inp, n, nlev, mgsolv=0)
    it0 = orc.last_stats()[0]
    assert ierr1 == 0 and ierr0 == 0
    assert abs(it1 - it0) <= 2
    nc = 9
    ipcb, band, coarse = _band_region(orc, n, nlev, iw1, rw1)
    nu, m, lda = orc.band_dims(nc, 27)
    assert list(ipcb) == [nu, m, lda, 1]
    info, ref_band, _ = orc.buildband(nc, 27, coarse[:14 * nc ** 3].copy())
    assert info == 0
    np.testing.assert_array_equal(band, ref_band)
    assert np.abs(u1 - u0).max() <= 1e-4 * np.abs(u0).max()


def test_driver_reports_indefinite_band(orc):
    """With a dielectric jump (2 inside, 78.54 outside) the mirrored-plane
    band of the 9^3 coarsest level is indefinite: Vdpbfa reports a pivot and
    the reference stops (mgsubd.cc:255-260); the oracle returns 3."""
    import multigrid_amd as mg
    n = 33
    inp = orc.pbe_assemble(n, 0.5, 0.15, mg.pbe_molecule(40, 6.0, 0, 7))
    ierr, *_ = orc.solve(inp, n, 3, mgsolv=1)
    assert ierr == 3


def test_mgsolv_out_of_range_rejected(orc):
    n = 17
    c, inp = orc.config_inputs("pb65", n=n, natoms=10, sphere_r=3.0)
    import multigrid_amd as mg
    iparm, rparm = mg.make_params(n, 2)
    iparm[20] = 2
    sz = orc.mgsz(2, 0, 0, n, n, n, 2)[1]
    iw = np.zeros(int(sz[11]), np.int32)
    rw = np.zeros(int(iparm[0]))
    u = np.zeros(n ** 3)
    a = {k: np.ascontiguousarray(inp[k], np.float64).copy() for k in orc.PBE_KEYS}
    assert orc.mgdriv(iparm, rparm, iw, rw, u, *(a[k] for k in orc.PBE_KEYS)) == 2


def test_band_refused_with_one_interior_column(orc):
    """A 3-point coarsest grid (one interior point per direction) would put
    band offsets on the diagonal and past the band column: refused."""
    n = 17
    c, inp = orc.config_inputs("pb65", n=n, natoms=10, sphere_r=3.0)
    ierr, *_ = orc.solve(inp, n, 4, mgsolv=1)
    assert ierr == 2

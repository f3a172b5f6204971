"""Independent numpy restatements used to pin the oracle (dense matrices,
closed forms).  Indices are 1-based as in the reference (pmgc/vmatrix.h)."""
import numpy as np

# upper directions of the symmetric 27-point storage, plane -> (ex, ey, ez)
# (the neighbour each plane multiplies in pmgc/matvecd.cc:137-175)
UPPER27 = {1: (1, 0, 0), 2: (0, 1, 0), 3: (0, 0, 1), 4: (1, 1, 0), 5: (-1, 1, 0),
           6: (1, 0, 1), 7: (-1, 0, 1), 8: (0, 1, 1), 9: (0, -1, 1), 10: (1, 1, 1),
           11: (-1, 1, 1), 12: (1, -1, 1), 13: (-1, -1, 1)}
# P plane -> offset d (pmgc/matvecd.cc:809-816)
PPLANE = {}
for zb, dz in ((0, 0), (1, 1), (2, -1)):
    for q, (dx, dy) in enumerate([(0, 0), (0, 1), (0, -1), (1, 0), (-1, 0), (1, 1),
                                  (-1, 1), (1, -1), (-1, -1)]):
        PPLANE[9 * zb + q] = (dx, dy, dz)


def IX(n, i, j, k):
    return (k - 1) * n * n + (j - 1) * n + (i - 1)


def interior(n):
    return [(i, j, k) for k in range(2, n) for j in range(2, n) for i in range(2, n)]


def dense_A(n, numdia, ac, cc):
    """Dense matrix over interior points assembled from the symmetric planes."""
    N = n ** 3
    pts = interior(n)
    pos = {IX(n, *p): a for a, p in enumerate(pts)}
    m = len(pts)
    A = np.zeros((m, m))
    planes = [1, 2, 3] if numdia == 7 else list(range(1, 14))
    for a, (i, j, k) in enumerate(pts):
        p = IX(n, i, j, k)
        A[a, a] = ac[p] + cc[p]
        for pl in planes:
            ex, ey, ez = UPPER27[pl]
            q = IX(n, i + ex, j + ey, k + ez)
            if q in pos:
                A[a, pos[q]] -= ac[pl * N + p]
                A[pos[q], a] -= ac[pl * N + p]
    return A, pts


def to_int(n, v):
    return np.array([v[IX(n, *p)] for p in interior(n)])


def from_int(n, w):
    out = np.zeros(n ** 3)
    for a, p in enumerate(interior(n)):
        out[IX(n, *p)] = w[a]
    return out


def random_op(n, numdia, rng, zero_boundary=True):
    """Random symmetric positive couplings; planes zero where they would
    couple an interior point to the boundary (as the reference builds)."""
    N = n ** 3
    nplanes = 4 if numdia == 7 else 14
    ac = np.zeros(nplanes * N)
    cc = np.zeros(N)
    for (i, j, k) in interior(n):
        p = IX(n, i, j, k)
        for pl in range(1, nplanes):
            ex, ey, ez = UPPER27[pl]
            ii, jj, kk = i + ex, j + ey, k + ez
            if 2 <= ii <= n - 1 and 2 <= jj <= n - 1 and 2 <= kk <= n - 1:
                ac[pl * N + p] = rng.uniform(0.1, 1.0) * (1.0 if pl <= 3 else 0.3)
        cc[p] = rng.uniform(0.0, 0.5)
    # diagonal: row sum of couplings + margin (diagonally dominant)
    A, pts = dense_A(n, numdia, ac, np.zeros(N))
    offsum = -(A - np.diag(np.diag(A))).sum(axis=1)
    for a, (i, j, k) in enumerate(pts):
        ac[IX(n, i, j, k)] = offsum[a] + rng.uniform(0.5, 1.5)
    return ac, cc


def _interior_mask(n):
    m = np.zeros((n, n, n), bool)
    m[1:-1, 1:-1, 1:-1] = True
    return m


def _shift_from(a, ex, ey, ez):
    """out[k, j, i] = a[k - ez, j - ey, i - ex] (zero outside)."""
    n = a.shape[0]
    out = np.zeros_like(a)
    sl = lambda d: (slice(d, n), slice(0, n - d)) if d >= 0 else (slice(0, n + d), slice(-d, n))
    (ko, ki), (jo, ji), (io, ii) = sl(ez), sl(ey), sl(ex)
    out[ko, jo, io] = a[ki, ji, ii]
    return out


def random_op_vec(n, numdia, rng):
    """Vectorised random_op for large grids (arrays indexed [k, j, i])."""
    nplanes = 4 if numdia == 7 else 14
    mask = _interior_mask(n)
    ac = np.zeros((nplanes, n, n, n))
    for pl in range(1, nplanes):
        ex, ey, ez = UPPER27[pl]
        nb = _shift_from(mask.astype(float), -ex, -ey, -ez) > 0  # p+e interior
        ac[pl] = rng.uniform(0.1, 1.0, (n, n, n)) * (1.0 if pl <= 3 else 0.3) * (mask & nb)
    diag = np.zeros((n, n, n))
    for pl in range(1, nplanes):
        ex, ey, ez = UPPER27[pl]
        diag += ac[pl] + _shift_from(ac[pl], ex, ey, ez)
    ac[0] = (diag + rng.uniform(0.5, 1.5, (n, n, n))) * mask
    cc = rng.uniform(0.0, 0.5, (n, n, n)) * mask
    return ac.reshape(-1), cc.reshape(-1)

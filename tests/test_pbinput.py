"""Synthetic PB inputs: the library's seeded molecule generator (host) and
the oracle's assembly of grid inputs from it, pinned by the physical closed
forms they are built from (charge conservation, the dielectric and kappa
values, the exclusion shell, the Debye-Hueckel sum evaluated here
independently)."""
import math

import numpy as np

KB, QE, EPS0, NA, T = 1.380649e-23, 1.602176634e-19, 8.8541878128e-12, 6.02214076e23, 298.15
LB = QE * QE / (4 * math.pi * EPS0 * KB * T) * 1e10


def test_deterministic(mg, orc):
    assert np.array_equal(mg.pbe_molecule(30, 3.0, 0, 5), mg.pbe_molecule(30, 3.0, 0, 5))
    a = orc.pbe_synth(17, 0.5, 30, 3.0, 0.15, 0, 5)
    b = orc.pbe_synth(17, 0.5, 30, 3.0, 0.15, 0, 5)
    c = orc.pbe_synth(17, 0.5, 30, 3.0, 0.15, 0, 6)
    for k in a:
        assert np.array_equal(a[k], b[k])
    assert not np.array_equal(a["fcf"], c["fcf"])


def test_charges_and_source(orc):
    n, h = 33, 0.5
    for mode in (0, 1):
        d = orc.pbe_synth(n, h, 41, 5.0, 0.15, mode, 9)
        q = d["atoms"][:, 4]
        if mode == 0:
            assert sorted(set(q)) == [-1.0, 0.0, 1.0] and q.sum() == 0 and (q == 0).sum() == 1
        else:
            assert set(q) <= {-1.0, -0.5, 0.0, 0.5, 1.0}
        # trilinear spreading conserves charge: sum fcf h^3 = 4 pi lB sum q
        assert abs(d["fcf"].sum() * h ** 3 - 4 * math.pi * LB * q.sum()) < 1e-9 * 4 * math.pi * LB * 41
        r = np.linalg.norm(d["atoms"][:, :3], axis=1)
        assert np.all(r <= 5.0) and np.all((d["atoms"][:, 3] >= 1.2) & (d["atoms"][:, 3] <= 2.0))


def test_dielectric_and_kappa(orc):
    n, h, I = 33, 0.5, 0.15
    d = orc.pbe_synth(n, h, 20, 4.0, I, 0, 3)
    hm = 2 * 2.0 * 78.54 / (2.0 + 78.54)
    for k in ("a1cf", "a2cf", "a3cf"):
        vals = set(np.round(d[k], 12))
        assert vals <= {2.0, 78.54, round(hm, 12)} and 2.0 in vals
    kap2 = 2 * I * 1000 * NA * QE * QE / (EPS0 * 78.54 * KB * T) * 1e-20
    assert abs(1 / math.sqrt(kap2) - 7.85) < 0.05  # Debye length at 0.15 M
    assert set(np.round(d["ccf"] / (78.54 * kap2), 12)) == {0.0, 1.0}
    # node inside an atom sphere -> excluded from the ion term
    at = d["atoms"]
    c = -(n - 1) * h / 2 + np.arange(n) * h
    i, j, k = (int(round((at[0, m] - c[0]) / h)) for m in range(3))
    p = k * n * n + j * n + i
    assert d["ccf"][p] == 0.0


def test_debye_hueckel_boundary(orc):
    n, h, I = 17, 0.5, 0.15
    d = orc.pbe_synth(n, h, 10, 2.0, I, 0, 4)
    at = d["atoms"]
    kap = math.sqrt(2 * I * 1000 * NA * QE * QE / (EPS0 * 78.54 * KB * T) * 1e-20)
    c = -(n - 1) * h / 2 + np.arange(n) * h

    def dh(x, y, z):
        s = 0.0
        for (ax, ay, az, ar, aq) in at:
            dd = math.sqrt((x - ax) ** 2 + (y - ay) ** 2 + (z - az) ** 2)
            s += LB * aq / 78.54 * math.exp(-kap * (dd - ar)) / ((1 + kap * ar) * dd)
        return s

    for (a, b) in [(0, 0), (3, 7), (16, 16)]:
        assert abs(d["gxcf"][b * n + a] - dh(c[0], c[a], c[b])) < 1e-12
        assert abs(d["gycf"][n * n + b * n + a] - dh(c[a], c[-1], c[b])) < 1e-12
        assert abs(d["gzcf"][b * n + a] - dh(c[a], c[b], c[0])) < 1e-12


def test_spreading_first_moments(mg, orc):
    """Trilinear spreading reproduces linear functions: the source's first
    moments equal the charges' (a swapped lower/upper weight fails this)."""
    n, h = 25, 0.45
    atoms = mg.pbe_molecule(60, 4.0, 1, 17)
    d = orc.pbe_assemble(n, h, 0.15, atoms)
    x = d["xf"]
    f = d["fcf"].reshape(n, n, n) * h ** 3 / (4 * math.pi * LB)
    q = atoms[:, 4]
    assert abs(f.sum() - q.sum()) < 1e-10
    for axis, col in ((2, 0), (1, 1), (0, 2)):
        shape = [1, 1, 1]
        shape[axis] = n
        m = (f * x.reshape(shape)).sum()
        assert abs(m - (q * atoms[:, col]).sum()) < 1e-9


def test_single_atom_marks(mg, orc):
    """One atom: nodes inside its sphere carry eps 2 on both adjacent faces,
    the ion-exclusion shell is exactly the nodes within r + 2 A."""
    n, h, r = 21, 0.5, 1.6
    atoms = np.array([[0.13, -0.21, 0.37, r, 1.0]])
    d = orc.pbe_assemble(n, h, 0.15, atoms)
    x = d["xf"]
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")  # X[i,j,k]
    d2 = ((X - 0.13) ** 2 + (Y + 0.21) ** 2 + (Z - 0.37) ** 2).transpose(2, 1, 0).ravel()
    excl = d["ccf"] == 0.0
    assert np.array_equal(excl, d2 <= (r + 2) ** 2)
    a1 = d["a1cf"].reshape(n, n, n)
    ins = (d2 <= r * r).reshape(n, n, n)
    both = ins[:, :, :-1] & ins[:, :, 1:]
    assert np.all(a1[:, :, :-1][both] == 2.0)
    out = ~ins[:, :, :-1] & ~ins[:, :, 1:]
    assert np.allclose(a1[:, :, :-1][out], 78.54, rtol=0, atol=1e-12)

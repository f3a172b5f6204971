"""OpenDX / PQR I/O and the end-to-end physics check against APBS: the
reference holds an APBS 3.4.1 potential for one atom (17^3 grid, h = 1 A,
pdie 4, sdie 78.4, no salt, multiple-Debye-Hueckel boundary, spline charge);
the oracle's assembly and PMG solve with the same parameters must reproduce
it on the boundary (both are the analytic Debye-Hueckel sum) and, away from
the atom, to the accuracy the different dielectric discretisations allow.
APBS's staggered molecular-surface dielectric differs from the node-based
harmonic faces here, so the bars loosen towards the atom."""
import os

import numpy as np
import pytest

from tests.conftest import ROOT

GOLD = os.path.join(ROOT, "tests", "golden")


def apbs_case(mg):
    u, c, o, dl = mg.dx_read(os.path.join(GOLD, "apbs_single_atom_pot.dx"))
    at = mg.read_pqr(os.path.join(GOLD, "apbs_single_atom.pqr"))
    n, h = c[0], dl[0]
    a = at.copy()
    a[:, :3] -= o + (n - 1) / 2 * h  # grid centred on the atom (gcent mol 1)
    return u, n, h, a


def check_against_apbs(u, ua, n):
    U, A = u.reshape(n, n, n), ua.reshape(n, n, n)
    rel = np.abs(U - A) / np.abs(A)
    b = np.zeros((n, n, n), bool)
    b[0] = b[-1] = True
    b[:, 0] = b[:, -1] = True
    b[:, :, 0] = b[:, :, -1] = True
    x = np.arange(n) - (n - 1) // 2
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    assert rel[b].max() < 1e-4
    assert rel[(r >= 3) & ~b].max() < 0.08
    assert rel[(r >= 6) & ~b].max() < 0.005
    m = (r >= 2) & ~b
    assert np.corrcoef(U[m], A[m])[0, 1] > 0.998
    c = (n - 1) // 2
    assert abs(U[c, c, c] / A[c, c, c] - 1) < 0.1


def test_pqr_and_dx_header(mg):
    u, c, o, dl = mg.dx_read(os.path.join(GOLD, "apbs_single_atom_pot.dx"))
    at = mg.read_pqr(os.path.join(GOLD, "apbs_single_atom.pqr"))
    assert c == (17, 17, 17) and np.allclose(dl, 1.0) and u.size == 17 ** 3
    assert np.allclose(at, [[-8.544, 3.578, 14.046, 1.824, 0.1305]])
    assert np.allclose(o + 8.0, at[0, :3], atol=1e-9)  # centred on the atom
    # the file lists x slowest, z fastest: its first values run along z
    with open(os.path.join(GOLD, "apbs_single_atom_pot.dx")) as f:
        first = [l for l in f if l[:1].isdigit() or l[:1] == "-"][0].split()
    assert np.allclose([u[0], u[17 * 17], u[2 * 17 * 17]], [float(v) for v in first])


def test_dx_round_trip(mg, tmp_path):
    rng = np.random.default_rng(3)
    u = rng.standard_normal(7 * 5 * 6)
    path = tmp_path / "f.dx"
    mg.dx_write(path, u, (7, 5, 6), (-1.0, 2.0, 0.5), 0.25)
    v, c, o, dl = mg.dx_read(path)
    assert c == (7, 5, 6) and np.allclose(o, (-1.0, 2.0, 0.5)) and np.allclose(dl, 0.25)
    assert np.allclose(v, u, rtol=1e-6, atol=0)


def test_oracle_matches_apbs_single_atom(mg, orc):
    ua, n, h, a = apbs_case(mg)
    inp = orc.pbe_assemble(n, h, 0.0, a, eps_in=4.0, eps_out=78.4, ion_shell=0.0)
    e, u, _, _ = orc.solve(inp, n, 3)
    assert e == 0
    check_against_apbs(u, ua, n)

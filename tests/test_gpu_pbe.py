"""GPU Poisson-Boltzmann input assembly (pmg_pbe_assemble and the
atom-list contexts) against the oracle's restatement of the reference's
assembly (PBE_pmgc_class.h:191-344).

Bar: every grid array that involves no transcendental function (marks,
dielectric faces, kappa map, charge source, coordinates) is bit-identical
to the oracle; the Debye-Hueckel faces, which call exp(), agree to 1e-13
of their largest magnitude."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EXACT = ("xf", "yf", "zf", "a1cf", "a2cf", "a3cf", "ccf", "fcf", "tcf")
FACES = ("gxcf", "gycf", "gzcf")


def _clustered(n, h, seed):
    """Atoms packed into a few grid cells, so nodes receive charge from many
    atoms and the summation order is exercised; some atoms lie outside the
    grid and must be dropped."""
    rng = np.random.default_rng(seed)
    c0 = -(n - 1) * h / 2
    cells = [(3, 4, 5), (3, 4, 6), (4, 4, 5), (n - 2, n - 2, n - 2), (0, 0, 0)]
    at = []
    for (i, j, k) in cells:
        for _ in range(9):
            p = c0 + (np.array([i, j, k]) + rng.random(3)) * h
            at.append([*p, 1.2 + 0.8 * rng.random(), rng.choice([-1.0, -0.5, 0.5, 1.0])])
    at.append([c0 - 3 * h, 0.0, 0.0, 1.5, 1.0])          # outside: dropped
    at.append([0.0, -c0 + 0.5 * h, 0.0, 1.5, -1.0])       # outside: dropped
    at.append([c0 + (n - 1) * h, 0.1, 0.2, 1.4, 1.0])     # on the last plane
    return np.array(at)


CASES = [
    ("tiny", 17, 0.5, 0.15, lambda mg: mg.pbe_molecule(8, 2.0, 0, 1)),
    ("pb65-like", 65, 0.5, 0.15, lambda mg: mg.pbe_molecule(200, 14.0, 0, 20261202)),
    ("mode1", 33, 0.35, 0.3, lambda mg: mg.pbe_molecule(120, 4.5, 1, 11)),
    ("clustered", 21, 0.4, 0.15, lambda mg: _clustered(21, 0.4, 5)),
    ("no atoms", 9, 0.5, 0.15, lambda mg: np.zeros((0, 5))),
    ("no salt", 17, 0.5, 0.0, lambda mg: mg.pbe_molecule(10, 2.5, 0, 4)),
]


@pytest.mark.parametrize("name,n,h,ionic,make", CASES, ids=[c[0] for c in CASES])
def test_assemble_parity(mg, orc, name, n, h, ionic, make):
    atoms = make(mg)
    print(f"pbe assemble {name} n={n} natoms={len(atoms)}", flush=True)
    g = mg.pbe_assemble(n, h, ionic, atoms)
    o = orc.pbe_assemble(n, h, ionic, atoms)
    for k in EXACT:
        assert np.array_equal(g[k], o[k]), k
    for k in FACES:
        scale = max(np.max(np.abs(o[k])), 1e-300)
        assert np.max(np.abs(g[k] - o[k])) <= 1e-13 * scale, k
    if name == "clustered":
        assert np.count_nonzero(o["fcf"]) > 0


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def test_ctx_from_atoms_and_reload(mg, orc):
    """A context assembled from atoms solves like one given the oracle's
    host arrays; re-loading another molecule into it gives bit for bit the
    result of a fresh context."""
    print("ctx from atoms n=65", flush=True)
    n, nlev, h = 65, 4, 0.5
    iparm, rparm = mg.make_params(n, nlev)
    a1 = mg.pbe_molecule(200, 14.0, 0, 20261202)
    a2 = mg.pbe_molecule(150, 12.0, 1, 99)
    ref = mg.Context(iparm, rparm, orc.pbe_assemble(n, h, 0.15, a1))
    ref.solve(mg.PMG_FP64)
    u_ref, s_ref = ref.u(), ref.stats()
    ref.close()
    ctx = mg.Context(iparm, rparm, pbe=(h, 0.15, a1))
    ctx.solve(mg.PMG_FP64)
    u1, s1 = ctx.u(), ctx.stats()
    assert s1["iters"] == s_ref["iters"] and _rel(u1, u_ref) < 1e-12
    ctx.load_pbe(h, 0.15, a2)
    ctx.solve(mg.PMG_MIXED)
    u2 = ctx.u()
    ctx.close()
    fresh = mg.Context(iparm, rparm, pbe=(h, 0.15, a2))
    fresh.solve(mg.PMG_MIXED)
    u2f = fresh.u()
    fresh.close()
    print(f"  iters={s1['iters']} err vs host inputs={_rel(u1, u_ref):.2e}", flush=True)
    assert np.array_equal(u2, u2f)


def test_apbs_single_atom(mg, orc):
    """GPU assembly and solve of the APBS single-atom case: the oracle's
    result to 1e-10 and APBS's within the physics bars of test_dx_apbs."""
    from tests.test_dx_apbs import apbs_case, check_against_apbs
    ua, n, h, a = apbs_case(mg)
    print(f"apbs single atom n={n}", flush=True)
    g = mg.pbe_assemble(n, h, 0.0, a, eps_in=4.0, eps_out=78.4, ion_shell=0.0)
    iparm, rparm = mg.make_params(n, 3)
    ctx = mg.Context(iparm, rparm, g)
    ctx.solve(mg.PMG_FP64)
    u = ctx.u()
    ctx.close()
    o = orc.pbe_assemble(n, h, 0.0, a, eps_in=4.0, eps_out=78.4, ion_shell=0.0)
    _, u0, _, _ = orc.solve(o, n, 3)
    assert _rel(u, u0) < 1e-10
    check_against_apbs(u, ua, n)

"""Grouped iterations (pmg_ctx_solve_group, the kernels of the serving
loop): several resident solves whose V-cycles run in lockstep, each kernel
launched once for all of them.  Every context must give bit for bit the
potential, iteration count and per-cycle history of its own pmg_ctx_solve,
in both precisions, for a mix of grid sizes, level counts, coarsest solvers
and stopping tests; the FP64 results also match the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (n, nlev, natoms, R, seed, extra make_params keywords)
CASES = [
    (33, 3, 40, 6.0, 1, {}),
    (65, 4, 150, 12.0, 2, {}),
    (97, 5, 300, 16.0, 3, {}),
    (65, 4, 90, 10.0, 4, dict(istop=0, errtol=1e-3)),
    # banded coarsest solve; a uniform dielectric keeps the reference's band
    # positive definite
    (33, 3, 40, 6.0, 3, dict(mgsolv=1)),
]


@pytest.fixture(scope="module")
def problems(orc, mg):
    out = []
    for n, nlev, natoms, R, seed, kw in CASES:
        if kw.get("mgsolv") == 1:
            inp = orc.pbe_assemble(n, 0.5, 0.15, mg.pbe_molecule(natoms, R, 0, seed),
                                   eps_in=78.54)
        else:
            inp = orc.pbe_synth(n, 0.5, natoms, R, 0.15, 0, seed)
        out.append((n, nlev, inp, kw))
    return out


def _contexts(mg, problems):
    return [mg.Context(*mg.make_params(n, nlev, **kw), inp) for n, nlev, inp, kw in problems]


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_group_equals_alone(mg, orc, problems, mode):
    m = mg.PMG_FP64 if mode == "fp64" else mg.PMG_MIXED
    ctxs = _contexts(mg, problems)
    alone = []
    for (n, nlev, _, kw), c in zip(problems, ctxs):
        print(f"{mode}: alone n={n} nlev={nlev} {kw}", flush=True)
        c.solve(m)
        alone.append((c.u(), c.stats()["iters"], c.history()))
    print(f"{mode}: grouped, {len(ctxs)} contexts", flush=True)
    mg.solve_group(ctxs, m)
    for (n, nlev, inp, kw), c, (u0, it0, h0) in zip(problems, ctxs, alone):
        u1, st, h1 = c.u(), c.stats(), c.history()
        print(f"  n={n}: iters alone {it0} grouped {st['iters']}", flush=True)
        assert st["iters"] == it0 and np.array_equal(h1, h0)
        assert np.array_equal(u1, u0), f"n={n}: grouped potential differs"
        if m == mg.PMG_FP64 and not kw:
            _, uo, _, _ = orc.solve(inp, n, nlev)
            assert np.max(np.abs(u1 - uo)) <= 1e-10 * np.max(np.abs(uo))
    for c in ctxs:
        c.close()


def test_group_repeat_and_subsets(mg, problems):
    """A context solved in different groups (alone, in a pair, in the full
    group) and again after another group: always the same bits."""
    ctxs = _contexts(mg, problems[:3])
    print("subsets", flush=True)
    ctxs[2].solve(mg.PMG_MIXED)
    ref = ctxs[2].u()
    mg.solve_group(ctxs[1:], mg.PMG_MIXED)
    assert np.array_equal(ctxs[2].u(), ref)
    mg.solve_group([ctxs[2]], mg.PMG_MIXED)
    assert np.array_equal(ctxs[2].u(), ref)
    mg.solve_group(ctxs, mg.PMG_MIXED)
    assert np.array_equal(ctxs[2].u(), ref)
    for c in ctxs:
        c.close()


def test_group_refuses_what_it_does_not_group(mg, problems):
    n, nlev, inp, _ = problems[0]
    a = mg.Context(*mg.make_params(n, nlev), inp)
    b = mg.Context(*mg.make_params(n, nlev, istop=2), inp)
    c = mg.Context(*mg.make_params(n, nlev, nu1=1, nu2=1), inp)
    for bad in ([a, b], [a, c], [a, a]):
        with pytest.raises(mg.PMGError) as ex:
            mg.solve_group(bad, mg.PMG_MIXED)
        assert ex.value.code == mg.PMG_EBADPARM
    for x in (a, b, c):
        x.close()

"""Serving loop (pmg_server_*): results in a mix of concurrent requests of
uneven sizes are bit for bit those of the same solves run alone; admission
respects the device-memory budget; the mixed-precision potentials meet the
FP64 oracle to 1e-6."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests(mg):
    """(n, nlev, mode, atoms) of a small uneven mix: three grid sizes, both
    precisions, distinct molecules."""
    out = []
    specs = [(33, 40, 6.0, 1), (65, 150, 12.0, 2), (49, 80, 9.0, 3), (33, 30, 5.0, 4),
             (65, 120, 11.0, 5), (49, 60, 8.0, 6)]
    for q, (n, na, R, seed) in enumerate(specs):
        nlev = {33: 3, 49: 4, 65: 4}[n]
        mode = mg.PMG_MIXED if q % 2 == 0 else mg.PMG_FP64
        out.append((n, nlev, mode, mg.pbe_molecule(na, R, q % 2, seed)))
    return out


def test_same_alone_and_in_mix(mg):
    reqs = _requests(mg)
    print(f"serve: {len(reqs)} requests alone, then all at once", flush=True)
    alone = []
    for (n, nlev, mode, atoms) in reqs:
        srv = mg.Server(0, 1)
        iparm, rparm = mg.make_params(n, nlev)
        t = srv.submit(iparm, rparm, mode, 0.5, 0.15, atoms)
        alone.append(srv.wait(t))
        srv.close()
    srv = mg.Server(0, 4)
    tickets = []
    for (n, nlev, mode, atoms) in reqs:
        iparm, rparm = mg.make_params(n, nlev)
        tickets.append(srv.submit(iparm, rparm, mode, 0.5, 0.15, atoms))
    mix = [srv.wait(t) for t in tickets]
    st = srv.stats()
    srv.close()
    print(f"  stats {st}", flush=True)
    assert st["completed"] == len(reqs) and st["max_concurrent"] >= 2
    assert st["p50_ms"] <= st["p99_ms"]
    for a, b in zip(alone, mix):
        assert a["rc"] == b["rc"] == 0
        assert a["iters"] == b["iters"] and a["relres"] == b["relres"] <= 1e-6
        assert np.array_equal(a["u"], b["u"])


def test_matches_context_solve_and_oracle(mg, orc):
    """A served solve equals the resident-context solve of the same molecule
    bit for bit, and (mixed precision) the FP64 oracle to 1e-6."""
    n, nlev = 49, 4
    atoms = mg.pbe_molecule(70, 8.0, 0, 11)
    iparm, rparm = mg.make_params(n, nlev)
    srv = mg.Server(0, 2)
    r = srv.wait(srv.submit(iparm, rparm, mg.PMG_MIXED, 0.5, 0.15, atoms))
    srv.close()
    ctx = mg.Context(iparm, rparm, pbe=(0.5, 0.15, atoms))
    ctx.solve(mg.PMG_MIXED)
    u_ctx = ctx.u()
    ctx.close()
    assert np.array_equal(r["u"], u_ctx)
    _, u0, _, _ = orc.solve(orc.pbe_assemble(n, 0.5, 0.15, atoms), n, nlev)
    err = np.max(np.abs(r["u"] - u0)) / np.max(np.abs(u0))
    print(f"serve vs oracle rel err {err:.2e}", flush=True)
    assert err < 1e-6


def test_budget_admission(mg):
    """With room for one context the requests run one at a time and all
    complete; a solve that can never fit is refused at submission."""
    n, nlev = 33, 3
    iparm, rparm = mg.make_params(n, nlev)
    atoms = [mg.pbe_molecule(30 + 5 * q, 5.0, 0, 100 + q) for q in range(4)]
    probe = mg.Server(0, 1)
    t = probe.submit(iparm, rparm, mg.PMG_MIXED, 0.5, 0.15, atoms[0])
    ref = probe.wait(t)
    one = probe.stats()["mem_used"]
    probe.close()
    srv = mg.Server(one + one // 2, 4)
    ts = [srv.submit(iparm, rparm, mg.PMG_MIXED, 0.5, 0.15, a) for a in atoms]
    res = [srv.wait(t) for t in ts]
    st = srv.stats()
    srv.close()
    assert st["max_concurrent"] == 1 and st["completed"] == 4
    assert np.array_equal(res[0]["u"], ref["u"])
    small = mg.Server(one // 4, 2)
    with pytest.raises(mg.PMGError) as e:
        small.submit(iparm, rparm, mg.PMG_MIXED, 0.5, 0.15, atoms[0])
    assert e.value.code == mg.PMG_EBUSY
    small.close()

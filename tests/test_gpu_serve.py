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


@pytest.mark.parametrize("groups", [1, 2])
def test_grouped_server_same_results(mg, groups):
    """Grouped iterations (pmg_server_create_grouped): every request of the
    uneven mix, both precisions in flight at once, gives the bits, the
    iteration count and the residual of the ungrouped server.  The grouped
    mixed iterations add the FP32 correction to u in a pass of their own,
    the ungrouped ones in the last fine post-smoothing sweep,
    so this also checks that the two updates give the same bits."""
    reqs = _requests(mg) * 2
    res = {}
    for g in (0, groups):
        print(f"serve: {len(reqs)} requests, groups={g}", flush=True)
        srv = mg.Server(0, 8, groups=g)
        tickets = []
        for (n, nlev, mode, atoms) in reqs:
            iparm, rparm = mg.make_params(n, nlev)
            tickets.append(srv.submit(iparm, rparm, mode, 0.5, 0.15, atoms))
        res[g] = [srv.wait(t) for t in tickets]
        st = srv.stats()
        srv.close()
        assert st["completed"] == len(reqs)
    for a, b in zip(res[0], res[groups]):
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


def _ctx_solve(mg, iparm, rparm, mode, atoms, h=0.5, ionic=0.15):
    ctx = mg.Context(iparm, rparm, pbe=(h, ionic, atoms))
    rc = ctx.solve(mode)
    u, st = ctx.u(), ctx.stats()
    ctx.close()
    return rc, u, st


def test_budget_admission(mg):
    """With room for one context the requests run one at a time, reusing one
    pooled context (setup kernels rewrite its operators in place, its
    captured graph is replayed): every served result equals a fresh
    resident-context solve of the same molecule bit for bit.  A solve that
    can never fit is refused at submission."""
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
    for a, r in zip(atoms, res):
        rc, u, cst = _ctx_solve(mg, iparm, rparm, mg.PMG_MIXED, a)
        assert r["rc"] == rc == 0
        assert r["iters"] == cst["iters"] and r["relres"] == cst["relres"]
        assert np.array_equal(r["u"], u)
    small = mg.Server(one // 4, 2)
    with pytest.raises(mg.PMGError) as e:
        small.submit(iparm, rparm, mg.PMG_MIXED, 0.5, 0.15, atoms[0])
    assert e.value.code == mg.PMG_EBUSY
    small.close()


def test_pooled_context_takes_each_requests_stopping_rule(mg):
    """One pooled context serves requests of one shape with different
    errtol / istop / itmax (iparm(10), iparm(11), rparm(1),
    pmgc/mgcsd.cc:133-163, 483-527): each result (status, iterations,
    residual, potential) is that of a fresh context with the same
    parameters."""
    n, nlev = 33, 3
    atoms = mg.pbe_molecule(40, 6.0, 0, 77)
    variants = [dict(), dict(errtol=1e-9), dict(itmax=2, errtol=1e-12),
                dict(istop=0, errtol=1e-3), dict(errtol=1e-4)]
    for mode in (mg.PMG_FP64, mg.PMG_MIXED):
        srv = mg.Server(0, 1)
        served = []
        for kw in variants:
            iparm, rparm = mg.make_params(n, nlev, **kw)
            served.append(srv.wait(srv.submit(iparm, rparm, mode, 0.5, 0.15, atoms)))
        st = srv.stats()
        srv.close()
        print(f"mode {mode}: iters {[r['iters'] for r in served]} "
              f"relres {[r['relres'] for r in served]}", flush=True)
        assert st["completed"] == len(variants)
        for kw, r in zip(variants, served):
            iparm, rparm = mg.make_params(n, nlev, **kw)
            rc, u, cst = _ctx_solve(mg, iparm, rparm, mode, atoms)
            assert r["rc"] == rc, kw
            assert r["iters"] == cst["iters"] and r["relres"] == cst["relres"], kw
            assert np.array_equal(r["u"], u), kw
        its = [r["iters"] for r in served]
        assert its[1] > its[0] > its[4] and its[2] == 2 and served[2]["rc"] == mg.PMG_ITMAX


def test_large_request_not_starved(mg):
    """Tight budget: a 65^3 request cannot share the device with any 33^3
    context.  Queued behind one 33^3 request and ahead of many more, it is
    passed over at most max_active times and then holds the queue until the
    running solves finish and their idle contexts are evicted for it; it
    completes long before the stream of small requests drains."""
    p33 = mg.make_params(33, 3)
    p65 = mg.make_params(65, 4)
    small = [mg.pbe_molecule(30, 5.0, 0, 300 + q) for q in range(60)]
    big = mg.pbe_molecule(150, 12.0, 0, 299)
    probe = mg.Server(0, 1)
    probe.wait(probe.submit(*p33, mg.PMG_MIXED, 0.5, 0.15, small[0]))
    b33 = probe.stats()["mem_used"]
    probe.close()
    probe = mg.Server(0, 1)
    ref = probe.wait(probe.submit(*p65, mg.PMG_MIXED, 0.5, 0.15, big))
    b65 = probe.stats()["mem_used"]
    probe.close()
    assert b65 > 2 * b33
    srv = mg.Server(b65 + b33 - 1, 4)
    ts = [srv.submit(*p33, mg.PMG_MIXED, 0.5, 0.15, small[0])]
    tb = srv.submit(*p65, mg.PMG_MIXED, 0.5, 0.15, big)
    ts += [srv.submit(*p33, mg.PMG_MIXED, 0.5, 0.15, a) for a in small[1:]]
    rb = srv.wait(tb)
    lat = [srv.wait(t)["latency_ms"] for t in ts]
    srv.close()
    print(f"large request latency {rb['latency_ms']:.1f} ms; small p50 "
          f"{np.median(lat):.1f} max {max(lat):.1f} ms", flush=True)
    assert rb["rc"] == 0 and np.array_equal(rb["u"], ref["u"])
    assert rb["latency_ms"] < np.median(lat)


def test_config_sizes_alone_and_in_mix(mg, orc):
    """The serving workload's own shapes: one request each of 97^3, 129^3
    and 161^3 (h 0.5 A) with 1000-3000 atoms, in both precisions.  Each
    result is bit for bit the same alone and in the mix of all six; the
    FP64 potentials are within 1e-10 of the FP64 oracle with its V-cycle
    count, the mixed-precision ones within 1e-6."""
    specs = [(97, 1000, 11), (129, 2000, 12), (161, 3000, 13)]
    reqs = []
    for n, na, seed in specs:
        half = (n - 1) * 0.5 / 2
        R = min(14.0 * (na / 200.0) ** (1.0 / 3.0), 0.75 * half - 2.0)
        atoms = mg.pbe_molecule(na, R, 0, seed)
        for mode in (mg.PMG_FP64, mg.PMG_MIXED):
            reqs.append((n, mg.SERVE_LEVELS[n], mode, atoms))
    print(f"serve at workload sizes: {len(reqs)} requests alone", flush=True)
    alone = []
    for (n, nlev, mode, atoms) in reqs:
        srv = mg.Server(0, 1)
        alone.append(srv.wait(srv.submit(*mg.make_params(n, nlev), mode, 0.5, 0.15, atoms)))
        srv.close()
    print("  all at once", flush=True)
    srv = mg.Server(0, 6)
    ts = [srv.submit(*mg.make_params(n, nlev), mode, 0.5, 0.15, atoms)
          for (n, nlev, mode, atoms) in reqs]
    mix = [srv.wait(t) for t in ts]
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(reqs) and st["max_concurrent"] >= 2
    for a, b in zip(alone, mix):
        assert a["rc"] == b["rc"] == 0
        assert a["iters"] == b["iters"] and a["relres"] == b["relres"] <= 1e-6
        assert np.array_equal(a["u"], b["u"])
    for q, (n, nlev, mode, atoms) in enumerate(reqs):
        if mode != mg.PMG_FP64:
            continue
        print(f"  oracle {n}^3 {len(atoms)} atoms", flush=True)
        _, u0, _, _ = orc.solve(orc.pbe_assemble(n, 0.5, 0.15, atoms), n, nlev)
        it0 = orc.last_stats()[0]
        e64 = np.max(np.abs(mix[q]["u"] - u0)) / np.max(np.abs(u0))
        emx = np.max(np.abs(mix[q + 1]["u"] - u0)) / np.max(np.abs(u0))
        print(f"  {n}^3: iters fp64 {mix[q]['iters']} mixed {mix[q + 1]['iters']} "
              f"oracle {it0}; err fp64 {e64:.2e} mixed {emx:.2e}", flush=True)
        assert mix[q]["iters"] == it0 and e64 <= 1e-10
        assert emx <= 1e-6

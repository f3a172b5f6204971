"""Benchmark: PB solves/sec + p99 latency at 1e-6 relative residual,
mixed precision vs FP64 on one MI355X.

Headline workload: the serving stream of linearised Poisson-Boltzmann
requests (grid sizes uniform over {65, 97, 129, 161}^3 at h = 0.5 A,
100-3000 synthetic atoms each, 0.15 M salt, V(2,2) Galerkin cycles to 1e-6
L1 relative residual, contexts within a 200 GB device-memory budget).  One
step = one batch of --batch requests of the stream; the timed region
submits the --steps batches at once and waits for the last result, so
`value` is the serving loop's mixed-precision capacity in solves/s (each
request: GPU input assembly from its atom list, operator build, V-cycles).
The FP64 capacity, and p50/p99 latency of both precisions under Poisson
arrivals at 0.8x the FP64 capacity, sit beside it.

Further legs: the 161^3 and 289^3 single solves (p50/p99 over >= 10
timed solves each), the 129^3 Newton solve (median of 5 timed solves),
the banded coarsest solve,
the class-style PCG; the roofline of the fine FP32 sweep at 289^3; the CPU
oracle on the host cores.

Prints one JSON line (see the contract at the end of main()).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="timed steps: batches of --batch requests of the serving stream")
    ap.add_argument("--warmup", type=int, default=3,
                    help="untimed warm-up batches (fill the context pool of every shape)")
    ap.add_argument("--batch", type=int, default=100, help="requests per step")
    ap.add_argument("--serve-active", type=int, default=8,
                    help="solves in flight at once in the serving loop")
    ap.add_argument("--serve-groups", type=int, default=2,
                    help="also measure the stream's capacity with grouped iterations in "
                         "this many groups (0: skip)")
    ap.add_argument("--single-steps", type=int, default=10,
                    help="timed solves of each single-solve leg (161^3, 289^3)")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the serving stream (profiling runs of the other legs; "
                         "no headline value then)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the single-solve, Newton, band and PCG legs")
    ap.add_argument("--legs", default="single,large,newton,band,pcg",
                    help="extra legs: single (161^3), large (289^3, also the roofline), "
                         "newton (129^3), band (161^3 with the banded coarsest solve), "
                         "pcg (161^3 with the class-style PCG)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first run a plain device copy of this many MiB (a reference "
                         "kernel of known traffic for hardware-counter runs)")
    return ap.parse_args()


def timed_steps(ctx, mol, mode, steps, warmup):
    for _ in range(warmup):
        ctx.load_pbe(*mol)
        ctx.solve(mode)
    lat = []
    iters = []
    t0 = time.perf_counter()
    for _ in range(steps):
        s = time.perf_counter()
        ctx.load_pbe(*mol)
        ctx.solve(mode)
        lat.append((time.perf_counter() - s) * 1e3)
        iters.append(ctx.stats()["iters"])
    wall = time.perf_counter() - t0
    return wall, np.array(lat), iters


def sweep_bytes(pts, sweeps, nu=(2, 2)):
    """Algorithmic bytes of `sweeps` timed FP32 fine sweeps of a mixed solve
    (V(nu1, nu2) cycles).  A sweep reads the diagonal oC + cc (precomputed
    once per setup), fc, oE, oN, uC and x and writes x: 7 arrays x 4 B per
    interior point.  The first pre-smoothing sweep of each correction cycle
    starts from x = 0 and does not read x: 6 arrays.  The last
    post-smoothing sweep adds its result to the FP64 iterate u instead of
    writing x: 6 arrays x 4 B + u read and written (2 x 8 B)."""
    cycles = sweeps // (nu[0] + nu[1])
    return 4 * pts * (7 * sweeps - cycles) + 12 * pts * cycles


def measured_traffic(n):
    """HBM bytes per launch of the fine FP32 sweep of an n^3 level, as
    measured with rocprofv3 FETCH_SIZE / WRITE_SIZE counters and recorded in
    profiles/pb<n>_hbm_traffic.md (None when no such measurement exists)."""
    path = os.path.join(ROOT, "profiles", f"pb{n}_hbm_traffic.md")
    try:
        for line in open(path):
            if line.startswith(f"traffic_bytes_per_launch k_gs7_sweep<float> n={n}:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _omp_threads(n):
    """Set the OpenMP thread count of the (already loaded) oracle."""
    import ctypes
    for name in ("libgomp.so.1", "libgomp.so"):
        try:
            ctypes.CDLL(name).omp_set_num_threads(int(n))
            return True
        except OSError:
            continue
    return False


def cpu_baseline(mg, stream_sizes, threads):
    """The CPU oracle (C restatement of the reference pmgc, OpenMP over the
    host cores) on this host: one request of each serving size (assembly +
    setup + V-cycles to 1e-6, the same work a served request does), the
    implied CPU capacity for the stream's size mix, a 1-thread solve of the
    65^3 workload, and the 161^3 / 129^3 Newton / 289^3 workloads.  Returns
    the baseline and the oracle's potentials for the error columns."""
    from tests import oracle as orc

    orc.olib()
    nthreads = int(threads or os.environ.get("OMP_NUM_THREADS", 0) or os.cpu_count() or 1)
    _omp_threads(nthreads)
    per_size = {}
    for n, natoms, seed in ((65, 1550, 1), (97, 1550, 2), (129, 1550, 3), (161, 1550, 4)):
        half = (n - 1) * 0.5 / 2
        R = min(14.0 * (natoms / 200.0) ** (1.0 / 3.0), 0.75 * half - 2.0)
        atoms = mg.pbe_molecule(natoms, R, 0, 77 + seed)
        t0 = time.perf_counter()
        inp = orc.pbe_assemble(n, 0.5, 0.15, atoms)
        t1 = time.perf_counter()
        orc.solve(inp, n, mg.SERVE_LEVELS[n])
        t2 = time.perf_counter()
        per_size[n] = {"assembly_s": t1 - t0, "solve_s": t2 - t1,
                       "request_s": t2 - t0, "vcycles": orc.last_stats()[0]}
    # the stream draws its sizes uniformly: mean request time over its mix
    mean_s = float(np.mean([per_size[n]["request_s"] for n in stream_sizes]))
    out = {"value": 1.0 / mean_s, "unit": "solves/s", "cores": nthreads, "kind": "port",
           "cpu_model": cpu_model(),
           "sample": "one 1550-atom request of each serving size (65/97/129/161^3, "
                     "h 0.5): oracle input assembly + Vmgdriv (setup + V-cycles to "
                     "1e-6); value = 1 / mean request time over the stream's size mix",
           "per_size": per_size}
    us = {}
    for name in ("pb161", "newton129", "pb289"):
        c = mg.CONFIGS[name]
        atoms = mg.pbe_molecule(c["natoms"], c["sphere_r"], c["charge_mode"], c["seed"])
        t0 = time.perf_counter()
        inp = orc.pbe_assemble(c["n"], c["h"], c["ionic"], atoms)
        t1 = time.perf_counter()
        if name == "newton129":
            orc.mypdefinit([1.0, -1.0], [0.5, 0.5])
            _, u, _, _ = orc.solve(inp, c["n"], c["nlev"], nonlin=True, errtol=1e-8)
        else:
            _, u, _, _ = orc.solve(inp, c["n"], c["nlev"])
        t2 = time.perf_counter()
        out[name] = {"assembly_s": t1 - t0, "solve_s": t2 - t1,
                     "iters": orc.last_stats()[0]}
        us[name] = (u, orc.last_stats()[0])
    c = mg.CONFIGS["pb65"]
    atoms = mg.pbe_molecule(c["natoms"], c["sphere_r"], c["charge_mode"], c["seed"])
    inp = orc.pbe_assemble(c["n"], c["h"], c["ionic"], atoms)
    _omp_threads(1)
    t0 = time.perf_counter()
    orc.solve(inp, c["n"], c["nlev"])
    t1 = time.perf_counter()
    _omp_threads(nthreads)
    out["pb65_one_thread"] = {"solve_s": t1 - t0, "vcycles": orc.last_stats()[0]}
    return out, us


def serve_requests(mg, count, seed=20261206):
    """The serving workload: grid sizes uniform over {65, 97, 129, 161}^3
    (h = 0.5 A), atom counts U[100, 3000] in a ball whose radius keeps the
    atom density of the 65^3 workload (14 A for 200 atoms), capped to fit
    the box; charges +-1, 0.15 M salt; V(2,2) cycles to 1e-6."""
    rng = np.random.default_rng(seed)
    sizes = [65, 97, 129, 161]
    out = []
    for q in range(count):
        n = sizes[int(rng.integers(0, 4))]
        na = int(rng.integers(100, 3001))
        half = (n - 1) * 0.5 / 2
        R = min(14.0 * (na / 200.0) ** (1.0 / 3.0), 0.75 * half - 2.0)
        atoms = mg.pbe_molecule(na, R, 0, seed * 7919 + q)
        out.append((n, mg.SERVE_LEVELS[n], atoms))
    return out


def serve_stream(mg, srv, reqs, mode, rate):
    """Submit the requests with Poisson arrivals at `rate` per second (all
    at once when rate is None) to the running server and wait for all;
    returns the server stats of this stream and the wall time from the
    first submission to the last result."""
    srv.reset_stats()
    params = {n: mg.make_params(n, nlev) for n, nlev in set((r[0], r[1]) for r in reqs)}
    rng = np.random.default_rng(7)
    tickets = []
    t0 = time.perf_counter()
    t_next = t0
    for (n, nlev, atoms) in reqs:
        if rate:
            t_next += rng.exponential(1.0 / rate)
            d = t_next - time.perf_counter()
            if d > 0:
                time.sleep(d)
        iparm, rparm = params[n]
        tickets.append(srv.submit(iparm, rparm, mode, 0.5, 0.15, atoms, want_u=False))
    res = [srv.wait(t) for t in tickets]
    wall = time.perf_counter() - t0
    st = srv.stats()
    st["wall_s"] = wall
    st["mean_iters"] = float(np.mean([r["iters"] for r in res]))
    st["max_relres"] = float(max(r["relres"] for r in res))
    st["all_ok"] = all(r["rc"] == 0 for r in res)
    print(f"serving: {len(reqs)} requests mode={mode} rate={rate} -> "
          f"{st['solves_per_s']:.1f} solves/s p50={st['p50_ms']:.2f} ms "
          f"p99={st['p99_ms']:.2f} ms wall={wall:.2f} s", file=sys.stderr, flush=True)
    return st


KEEP = ("solves_per_s", "p50_ms", "p99_ms", "max_concurrent", "mean_iters", "max_relres",
        "wall_s")


def serving_leg(mg, steps, warmup, batch, active, groups=2):
    """The headline: the stream's --steps batches submitted at once, mixed
    precision (timed: `value`) and FP64 (capacity beside it); then both
    precisions under Poisson arrivals at 0.8x the FP64 capacity for
    p50/p99 latency."""
    budget = 200 * (1 << 30)
    reqs = serve_requests(mg, steps * batch)
    warm = serve_requests(mg, max(warmup, 1) * batch, seed=20261207)
    # one server for every stream: the warm-up fills its context pool with
    # every shape and precision, as a running service holds it
    srv = mg.Server(budget, active)
    for mode in (mg.PMG_FP64, mg.PMG_MIXED):
        serve_stream(mg, srv, warm, mode, None)
    cap64 = serve_stream(mg, srv, reqs, mg.PMG_FP64, None)
    capmx = serve_stream(mg, srv, reqs, mg.PMG_MIXED, None)
    rate = 0.8 * cap64["solves_per_s"]
    mixed = serve_stream(mg, srv, reqs, mg.PMG_MIXED, rate)
    fp64 = serve_stream(mg, srv, reqs, mg.PMG_FP64, rate)
    srv.close()
    # the same stream with grouped iterations (two groups, each kernel of an
    # iteration launched once for up to 8 requests)
    grouped = {}
    if groups:
        gsrv = mg.Server(budget, active, groups=groups)
        for mode in (mg.PMG_FP64, mg.PMG_MIXED):
            serve_stream(mg, gsrv, warm, mode, None)
        for key, mode in (("fp64", mg.PMG_FP64), ("mixed", mg.PMG_MIXED)):
            st = serve_stream(mg, gsrv, reqs, mode, None)
            grouped[key] = {k: st[k] for k in KEEP}
        gsrv.close()
    return {
        "requests": len(reqs),
        "arrival_rate_per_s": rate,
        "max_active": active,
        "budget_bytes": budget,
        "results_on_device": True,
        "capacity": {"mixed": {k: capmx[k] for k in KEEP}, "fp64": {k: cap64[k] for k in KEEP}},
        "at_rate": {"mixed": {k: mixed[k] for k in KEEP}, "fp64": {k: fp64[k] for k in KEEP}},
        "mixed_vs_fp64_capacity": capmx["solves_per_s"] / cap64["solves_per_s"],
        "grouped_capacity": dict(groups=groups, **grouped) if groups else None,
        "all_ok": capmx["all_ok"] and cap64["all_ok"] and mixed["all_ok"] and fp64["all_ok"],
    }, capmx


def rel_err(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def single_leg(mg, name, steps, oracle=None):
    """One workload solved again and again on a resident context (each step:
    GPU input assembly from the atom list, hierarchy build, V-cycles to
    1e-6), FP64 and mixed precision, p50/p99 over `steps` timed solves; the
    fine FP32 sweep timed with HIP events on the context stream over further
    mixed solves; errors against the FP64 GPU solve and the oracle."""
    cfg = mg.CONFIGS[name]
    n, nlev = cfg["n"], cfg["nlev"]
    atoms = mg.pbe_molecule(cfg["natoms"], cfg["sphere_r"], cfg["charge_mode"], cfg["seed"])
    mol = (cfg["h"], cfg["ionic"], atoms)
    iparm, rparm = mg.make_params(n, nlev)
    ctx = mg.Context(iparm, rparm, pbe=mol)
    out = {"workload": f"linearised PBE {n}^3 h={cfg['h']} nlev={nlev} {cfg['natoms']} atoms, "
                       "V(2,2) Galerkin, stop 1e-6"}
    us = {}
    for key, mode in (("fp64", mg.PMG_FP64), ("mixed", mg.PMG_MIXED)):
        wall, lat, its = timed_steps(ctx, mol, mode, steps, 2)
        st = ctx.stats()
        out[key] = {"solves_per_s": steps / wall, "p50_ms": float(np.percentile(lat, 50)),
                    "p99_ms": float(np.percentile(lat, 99)), "vcycles": its[-1],
                    "relres": st["relres"], "setup_ms": st["setup_ms"],
                    "solve_ms": st["solve_ms"]}
        us[key] = ctx.u()
    out["mixed_vs_fp64_speedup"] = out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"]
    ctx.ktime_enable(16384)
    ksteps = 4
    for _ in range(ksteps):
        ctx.solve(mg.PMG_MIXED)
    kms, sweeps = ctx.ktime()
    ctx.close()
    pts = (n - 2) ** 3
    avg_s = (kms / 1e3) / max(sweeps, 1)
    achieved = sweep_bytes(pts, sweeps) / max(sweeps, 1) / avg_s / 1e9 if sweeps else 0.0
    out["fp32_sweep"] = {"gbs": achieved, "avg_launch_us": avg_s * 1e6, "launches": sweeps,
                         "bytes_per_launch": sweep_bytes(pts, sweeps) / max(sweeps, 1)}
    out["mixed_rel_err_vs_fp64_gpu"] = rel_err(us["mixed"], us["fp64"])
    if oracle is not None:
        u0, it0 = oracle
        out["oracle_vcycles"] = it0
        out["fp64_rel_err_vs_oracle"] = rel_err(us["fp64"], u0)
        out["mixed_rel_err_vs_oracle"] = rel_err(us["mixed"], u0)
    return out


def newton_leg(mg, oracle=None, steps=5):
    """The 129^3 nonlinear PBE (sinh ionic term, 0.3 M 1:1 salt, 800 atoms):
    Newton with FP64 and with mixed-precision inner V-cycles to 1e-8 (median
    and largest of `steps` timed solves after one untimed one), and (with
    the CPU leg) the error against the FP64 oracle's Newton."""
    cfg = mg.CONFIGS["newton129"]
    n, nlev = cfg["n"], cfg["nlev"]
    atoms = mg.pbe_molecule(cfg["natoms"], cfg["sphere_r"], cfg["charge_mode"], cfg["seed"])
    ions = ([1.0, -1.0], [0.5, 0.5])
    mg.Vmypdefinitnpbe(2, *ions)
    iparm, rparm = mg.make_params(n, nlev, nonlin=1, errtol=1e-8)
    ctx = mg.Context(iparm, rparm, pbe=(cfg["h"], cfg["ionic"], atoms))
    out = {"workload": f"nonlinear PBE {n}^3 h={cfg['h']} nlev={nlev} {cfg['natoms']} atoms, "
                       f"{cfg['ionic']} M, Newton to 1e-8"}
    us = {}
    for name, mode in (("fp64", mg.PMG_FP64), ("mixed", mg.PMG_MIXED)):
        ctx.solve(mode)  # warm
        lat = []
        for _ in range(steps):
            t0 = time.perf_counter()
            ctx.solve(mode)
            lat.append((time.perf_counter() - t0) * 1e3)
        st = ctx.stats()
        out[name] = {"ms": float(np.median(lat)), "max_ms": max(lat), "timed_solves": steps,
                     "newton_iters": st["iters"], "relres": st["relres"]}
        us[name] = ctx.u()
    ctx.close()
    out["mixed_rel_err_vs_fp64_gpu"] = rel_err(us["mixed"], us["fp64"])
    if oracle is not None:
        u0, it0 = oracle
        out["oracle_newton_iters"] = it0
        out["fp64_rel_err_vs_oracle"] = rel_err(us["fp64"], u0)
        out["mixed_rel_err_vs_oracle"] = rel_err(us["mixed"], u0)
    return out


def band_leg(mg, cpu, steps=5):
    """The 161^3 workload with the banded coarsest solve (iparm(21) = 1:
    Vbuildband + Vdpbfa at setup, Vdpbsl per V-cycle) instead of 100 CG
    iterations: FP64 and mixed solves (assembly + build + V-cycles) and,
    with the CPU leg, the error against the oracle's mgsolv = 1 solve."""
    cfg = mg.CONFIGS["pb161"]
    n, nlev = cfg["n"], cfg["nlev"]
    atoms = mg.pbe_molecule(cfg["natoms"], cfg["sphere_r"], cfg["charge_mode"], cfg["seed"])
    mol = (cfg["h"], cfg["ionic"], atoms)
    iparm, rparm = mg.make_params(n, nlev, mgsolv=1)
    ctx = mg.Context(iparm, rparm, pbe=mol)
    out = {"workload": f"linearised PBE {n}^3 h={cfg['h']} nlev={nlev} {cfg['natoms']} atoms, "
                       "banded coarsest solve (mgsolv=1)"}
    us = {}
    for name, mode in (("fp64", mg.PMG_FP64), ("mixed", mg.PMG_MIXED)):
        wall, lat, its = timed_steps(ctx, mol, mode, steps, 1)
        st = ctx.stats()
        out[name] = {"solves_per_s": steps / wall, "p50_ms": float(np.percentile(lat, 50)),
                     "vcycles": its[-1], "relres": st["relres"], "setup_ms": st["setup_ms"],
                     "solve_ms": st["solve_ms"]}
        us[name] = ctx.u()
    ctx.close()
    out["mixed_vs_fp64_speedup"] = out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"]
    out["mixed_rel_err_vs_fp64_gpu"] = rel_err(us["mixed"], us["fp64"])
    if cpu:
        from tests import oracle as orc
        inp = orc.pbe_assemble(n, cfg["h"], cfg["ionic"], atoms)
        _, u0, _, _ = orc.solve(inp, n, nlev, mgsolv=1)
        out["oracle_vcycles"] = orc.last_stats()[0]
        out["fp64_rel_err_vs_oracle"] = rel_err(us["fp64"], u0)
        out["mixed_rel_err_vs_oracle"] = rel_err(us["mixed"], u0)
    return out


def pcg_leg(mg, cpu, steps=5):
    """The 161^3 workload solved by the class-style solver (CG preconditioned
    by the class V-cycle, PBE_pmgc_class.h:1070-1133 + CG_Solver.h:45-242) on
    the same hierarchy, to the same 1e-6 relative L1 residual: FP64 and with
    the FP32 V-cycle preconditioner; with the CPU leg, the error against the
    oracle's class-style solve."""
    cfg = mg.CONFIGS["pb161"]
    n, nlev = cfg["n"], cfg["nlev"]
    atoms = mg.pbe_molecule(cfg["natoms"], cfg["sphere_r"], cfg["charge_mode"], cfg["seed"])
    mol = (cfg["h"], cfg["ionic"], atoms)
    iparm, rparm = mg.make_params(n, nlev)
    ctx = mg.Context(iparm, rparm, pbe=mol)
    out = {"workload": f"linearised PBE {n}^3 h={cfg['h']} nlev={nlev} {cfg['natoms']} atoms, "
                       "class-style PCG to 1e-6"}
    us = {}
    for name, mode in (("fp64", mg.PMG_FP64), ("mixed", mg.PMG_MIXED)):
        ctx.load_pbe(*mol)
        ctx.pcg(mode)  # warm
        lat = []
        t0 = time.perf_counter()
        for _ in range(steps):
            s = time.perf_counter()
            ctx.load_pbe(*mol)
            ctx.pcg(mode)
            lat.append((time.perf_counter() - s) * 1e3)
        wall = time.perf_counter() - t0
        st = ctx.stats()
        out[name] = {"solves_per_s": steps / wall, "p50_ms": float(np.percentile(lat, 50)),
                     "cg_iters": st["iters"], "relres": st["relres"],
                     "setup_ms": st["setup_ms"], "solve_ms": st["solve_ms"]}
        us[name] = ctx.u()
    ctx.close()
    out["mixed_vs_fp64_speedup"] = out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"]
    out["mixed_rel_err_vs_fp64_gpu"] = rel_err(us["mixed"], us["fp64"])
    if cpu:
        from tests import oracle as orc
        inp = orc.pbe_assemble(n, cfg["h"], cfg["ionic"], atoms)
        _, u0, _, _ = orc.pcg(inp, n, nlev)
        out["oracle_cg_iters"] = orc.last_stats()[0]
        out["fp64_rel_err_vs_oracle"] = rel_err(us["fp64"], u0)
        out["mixed_rel_err_vs_oracle"] = rel_err(us["mixed"], u0)
    return out


def copy_probe(mg, mib):
    """Three device-to-device copies of `mib` MiB through the HIP runtime the
    library itself uses (a second runtime, such as torch's, in the same
    process breaks the library's calls): the reference kernel of known
    traffic that scales the hardware counters."""
    import ctypes

    mg.lib()  # the library's HIP runtime is the one loaded
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    nb = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if h.hipMalloc(ctypes.byref(a), nb) or h.hipMalloc(ctypes.byref(b), nb):
        raise RuntimeError("copy probe: hipMalloc failed")
    try:
        if h.hipMemset(a, 1, nb):
            raise RuntimeError("copy probe: hipMemset failed")
        for _ in range(3):
            if h.hipMemcpy(b, a, nb, 3):  # hipMemcpyDeviceToDevice
                raise RuntimeError("copy probe: hipMemcpy failed")
        h.hipDeviceSynchronize()
    finally:
        h.hipFree(a)
        h.hipFree(b)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the solver runs on a single MI355X).", file=sys.stderr)
        sys.exit(2)
    import multigrid_amd as mg

    if args.copy_probe_mib:
        copy_probe(mg, args.copy_probe_mib)

    # headline: the serving stream (the timed region is its mixed capacity run)
    serving, wall, value = None, None, None
    if not args.no_serve:
        serving, capmx = serving_leg(mg, args.steps, args.warmup, args.batch,
                                     args.serve_active, args.serve_groups)
        wall = capmx["wall_s"]
        value = args.steps * args.batch / wall

    cpu, oracle_u = None, {}
    if not args.no_cpu:
        cpu, oracle_u = cpu_baseline(mg, (65, 97, 129, 161), 0)
        print(f"cpu: {cpu}", file=sys.stderr, flush=True)
    extra = {}
    legs = set() if args.no_extra else set(args.legs.split(","))
    if "single" in legs:
        extra["single_161"] = single_leg(mg, "pb161", args.single_steps, oracle_u.get("pb161"))
    if "large" in legs:
        extra["large_289"] = single_leg(mg, "pb289", args.single_steps, oracle_u.get("pb289"))
    if "newton" in legs:
        extra["newton_129"] = newton_leg(mg, oracle_u.get("newton129"))
    if "band" in legs:
        extra["band_161"] = band_leg(mg, not args.no_cpu)
    if "pcg" in legs:
        extra["pcg_161"] = pcg_leg(mg, not args.no_cpu)
    for k, v in extra.items():
        print(f"{k}: {v}", file=sys.stderr, flush=True)

    # roofline of the dominant kernel, the fine FP32 red-black sweep, on the
    # largest grid (289^3); the 161^3 figure when that leg was skipped
    rl = None
    for key, n in (("large_289", 289), ("single_161", 161)):
        if key in extra:
            sw = extra[key]["fp32_sweep"]
            rl = {"bound": "hbm", "achieved": sw["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": sw["gbs"] / HBM_PEAK_GBS, "traffic": measured_traffic(n),
                  "kernel": f"k_gs7_sweep<float> (fine fused red-black sweep, {n}^3)",
                  "avg_launch_us": sw["avg_launch_us"], "timed_launches": sw["launches"],
                  "algorithmic_bytes_per_launch": sw["bytes_per_launch"]}
            break
    nothing = {"mixed": {"p50_ms": None, "p99_ms": None}, "fp64": {"p50_ms": None,
                                                                  "p99_ms": None}}
    at = serving["at_rate"] if serving else nothing
    out = {
        "metric": "PB solves/sec + p99 latency at 1e-6 rel residual: mixed-precision vs FP64 GPU",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3 if wall else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-in-f64",
        "data": "synthetic",
        "config": {"workload": f"serving stream of {args.steps * args.batch} linearised PB "
                               "requests, n uniform over {65,97,129,161}^3, h=0.5, 100-3000 "
                               "atoms, stop 1e-6, 200 GB budget; value = mixed-precision "
                               "capacity (all requests submitted at once)",
                   "requests_per_step": args.batch, "max_active": args.serve_active},
        "p50_ms": at["mixed"]["p50_ms"],
        "p99_ms": at["mixed"]["p99_ms"],
        "fp64": {"value": serving["capacity"]["fp64"]["solves_per_s"] if serving else None,
                 "p50_ms": at["fp64"]["p50_ms"], "p99_ms": at["fp64"]["p99_ms"]},
        "latency_at_rate_per_s": serving["arrival_rate_per_s"] if serving else None,
        "mixed_vs_fp64_capacity": serving["mixed_vs_fp64_capacity"] if serving else None,
        "roofline": rl,
        "cpu_baseline": cpu,
        "serving": serving,
        **extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

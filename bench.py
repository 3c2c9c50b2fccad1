"""Benchmark of the explicit residual path on one MI355X.

Workload (default): the annular rotor with Hall body force, 193 x 49 x 129 nodes, JST,
RK41A, CFL 2, local time step, plain FP64 (synthetic grid and fields, see
cturbobfm_amd/cases.py).  A step is one explicit iteration: time step, radial-equilibrium
profile, and four fused residual + RK stages over all nodes.

Prints one JSON line:
  value       cell-stage updates/s = nodes * stages * steps / wall time of the timed steps
              (the production path: captured graphs, no per-kernel events; the same
              number of steps with HIP events around every stage kernel gives the kernel
              times, their wall time is ms_per_step_with_kernel_events)
  roofline    the fused stage kernel: algorithmic bytes per launch (208 B per node for 3-D
              FP64, 172 B with the FP32 face records of --precision mixed) / its average
              launch time from HIP events on the case stream, against 8 TB/s; traffic
              from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary when present
  cpu_baseline the CPU restatement (oracle/liboracle.so, OpenMP) timed on a bounded sample
              of the same workload on this host
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def body_force_nodes(case) -> int:
    """Blade-row nodes (bladePresent > 0) of the workload, reported in the config."""
    f = case["fields"]
    if case["config"].get("BFM_ACTIVE", "no") != "yes" or "bladePresent" not in f:
        return 0
    return int(np.count_nonzero(f["bladePresent"] > 0))


def traffic_from_profiles(kernel_prefix: str):
    """HBM bytes per launch of the stage kernel from the committed counter summary."""
    path = os.path.join(ROOT, "profiles", "stage_kernel_rotor3d_193x49x129_counters.md")
    if not os.path.exists(path):
        return None
    txt = open(path).read()
    m = re.search(r"hbm_bytes_per_launch:\s*([0-9.eE+]+)", txt)
    return float(m.group(1)) if m else None


def cpu_baseline(case, seconds: float = 15.0):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the test oracle; timed here only as the reported CPU baseline
    threads = int(os.environ.get("OMP_NUM_THREADS", str(os.cpu_count() or 1)))
    oc = O.OracleCase(case)
    stages = oc.params.n_stages
    oc.run(1, threads=threads)        # warm-up (page-in, thread start)
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < seconds:
        oc.run(1, threads=threads)
        n += 1
    dt = time.perf_counter() - t0
    nodes = case["ni"] * case["nj"] * case["nk"]
    return {"value": nodes * stages * n / dt, "unit": "cell-stage updates/s", "cores": threads,
            "kind": "port", "sample": "%d iterations (%d stages each) of the same %dx%dx%d case, %.1f s"
            % (n, stages, case["ni"], case["nj"], case["nk"], dt)}


def bench_batch(args):
    """Batched throttle sweep: rotor operating points on four grid sizes (outlet pressure
    drawn uniformly from 90-110 kPa, seeded), FP32 flux / FP64 state, all advanced in one
    batch (one launch per kernel, captured graph per iteration).  Compared with the plain
    path: each case alone in FP64, one after another, for the same iterations."""
    from cturbobfm_amd import Batch, Solver, SolverError, cases
    rng = np.random.default_rng(args.seed)
    sizes = [(97, 25, 65), (129, 33, 97), (193, 49, 129), (257, 65, 193)]
    per = args.batch_cases // len(sizes)
    pts = [(sz, float(rng.uniform(90000.0, 110000.0))) for sz in sizes for _ in range(per)]
    its = args.warmup + args.steps + 10
    t_setup = time.perf_counter()
    b = Batch()
    members = []
    for q, (sz, p) in enumerate(pts):
        s = Solver.from_case(cases.rotor3d(*sz, n_iterations=its, outlet_pressure=p), {"GPU_PRECISION": "mixed"})
        b.admit(s)
        members.append(s)
        print("admitted %d/%d: rotor %dx%dx%d p_out %.0f" % (q + 1, len(pts), *sz, p), file=sys.stderr, flush=True)
    t_setup = time.perf_counter() - t_setup
    nodes = sum(s.N for s in members)
    b.advance(args.warmup)
    t0 = time.perf_counter()
    b.advance(args.steps)
    wall = time.perf_counter() - t0
    # a member that stopped (converged or diverged) skips its work from then on, so the
    # rate counts every member's nodes only when none stopped in the run
    stopped = diverged = 0
    for s in members:
        try:
            stopped += bool(s.status()[2])
        except SolverError:
            diverged += 1
    if stopped or diverged:
        print("warning: %d members stopped and %d diverged; the rate below overstates the work done"
              % (stopped, diverged), file=sys.stderr, flush=True)
    tiling = sorted(set(s.tuning().split("->")[-1].strip() for s in members))
    b.close()
    for s in members:
        s.close()
    # plain path: one case of each size alone in FP64 (the same work per case of a size)
    seq = 0.0
    for sz in sizes:
        s = Solver.from_case(cases.rotor3d(*sz, n_iterations=its, outlet_pressure=95000.0), {"GPU_PRECISION": "fp64"})
        s.advance(args.warmup)
        s.sync()
        t = time.perf_counter()
        s.advance(args.steps)
        s.sync()
        seq += (time.perf_counter() - t) * per
        s.close()
    value = nodes * 4 * args.steps / wall
    return {
        "metric": "batched cell-stage updates/s (FP32 flux / FP64 state) vs plain FP64 one case at a time",
        "value": value, "unit": "cell-stage updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64-state/f32-flux", "data": "synthetic rotor grids and blade fields, free-stream start",
        "config": {"workload": "throttle sweep: %d rotor operating points, %d per size on %s, outlet 90-110 kPa "
                               "(seed %d), JST, RK41A, one batch" % (len(pts), per,
                                                                    ", ".join("%dx%dx%d" % z for z in sizes), args.seed),
                   "nodes": nodes, "cases": len(pts), "stopped_cases": stopped, "diverged_cases": diverged,
                   "stage_tiling": tiling,
                   "setup_s": t_setup},
        "plain_fp64_sequential_s": seq, "batched_s": wall, "speedup_vs_plain": seq / wall,
    }


def serve_leg(args, n_cases):
    """Served solves: a seeded Poisson stream (mean 2 cases/s) of 2-D bump and 3-D rotor
    cases, admitted by device-memory footprint into running batches (FP32 flux / FP64
    state) and retired on a density-residual drop; against the plain FP64 path serving
    the same stream one case at a time (its measured service times in a FIFO queue)."""
    import tempfile
    from cturbobfm_amd import serve
    tmp = tempfile.mkdtemp(prefix="tbfm_serve_")
    lut = serve.stream_table(tmp)
    specs = serve.mixed_stream(n_cases, seed=args.seed, precision="mixed", real_fraction=0.05,
                               lut_path=lut)
    arrivals = serve.poisson_arrivals(len(specs), rate=args.rate, seed=args.seed)
    progress = lambda m: print(m, file=sys.stderr, flush=True)
    srv = serve.Server(drop=args.drop, max_iter=args.max_iter, chunk=args.chunk, keep_turbo=True, log=progress)
    t0 = time.perf_counter()
    res = srv.run(specs, arrivals)
    wall = time.perf_counter() - t0
    served = serve.summary(res, wall)
    alone = serve.sequential(specs, drop=args.drop, max_iter=args.max_iter, precision="fp64", keep_turbo=True,
                             log=progress)
    # converged mass flow and total-pressure ratio, served (FP32 flux) against plain FP64
    dm, dpr = [], []
    for r, a in zip(res, alone):
        if r.turbo is not None and a.turbo is not None and r.converged and a.converged:
            dm.append(abs(r.turbo[6] - a.turbo[6]) / abs(a.turbo[6]))
            dpr.append(abs(r.turbo[7] - a.turbo[7]) / abs(a.turbo[7]))
    # the plain path serving the same arrivals: one case at a time, first come first served
    order = np.argsort(arrivals)
    free = 0.0
    ttc = np.zeros(len(specs))
    for q in order:
        service = alone[q].finished - alone[q].arrival
        start = max(free, arrivals[q])
        free = start + service
        ttc[q] = free - arrivals[q]
    plain_wall = free
    plain = {"p50_s": float(np.percentile(ttc, 50)), "p99_s": float(np.percentile(ttc, 99)),
             "cases_per_hour": len(specs) / plain_wall * 3600.0, "wall_s": plain_wall,
             "converged": int(sum(r.converged for r in alone)), "iterations": int(sum(r.iterations for r in alone))}
    # capacity: the same requests all arriving at once, so that neither path waits for
    # arrivals (the Poisson stream's cases/hour are bounded by its rate); the plain path's
    # capacity is its summed service time
    srv = serve.Server(drop=args.drop, max_iter=args.max_iter, chunk=args.chunk, log=progress)
    t0 = time.perf_counter()
    res_cap = srv.run(specs, np.zeros(len(specs)))
    wall_cap = time.perf_counter() - t0
    busy = float(sum(a.finished - a.arrival for a in alone))
    served_cap = serve.summary(res_cap, wall_cap)
    capacity = {"served_cases_per_hour": served_cap["cases_per_hour"], "served_wall_s": wall_cap,
                "served_converged": served_cap["converged"],
                "plain_fp64_cases_per_hour": len(specs) / busy * 3600.0, "plain_fp64_busy_s": busy,
                "cases_per_hour_ratio": served_cap["cases_per_hour"] / (len(specs) / busy * 3600.0)}
    return specs, served, plain, dm, dpr, res, capacity


def equal_iteration_diffs(specs, res):
    """Mass flow and total-pressure ratio of each served FP32-flux result against the
    plain FP64 path run from the same start for exactly the same number of iterations:
    the difference of precision alone (the converged comparison also holds the
    difference of two convergence levels).  Table-fluid requests compute in FP64 on both
    paths and are left out."""
    from cturbobfm_amd import Solver
    dm, dpr = [], []
    for sp, r in zip(specs, res):
        if r.turbo is None or not r.converged or "table-fluid" in sp.name:
            continue
        ov = dict(sp.overrides)
        ov["GPU_PRECISION"] = "fp64"
        s = Solver.from_case(sp.build(), ov)
        s.set_convergence(True, 1.0e3)   # a drop no run reaches: exactly r.iterations
        while s.iterations < r.iterations:
            s.advance(min(500, r.iterations - s.iterations))
        t = s.turbo()
        s.close()
        dm.append(abs(r.turbo[6] - t[6]) / abs(t[6]))
        dpr.append(abs(r.turbo[7] - t[7]) / abs(t[7]))
    return dm, dpr


def limit_cycle_diffs(specs, n=12000, window=4000):
    """Mass flow and total-pressure ratio averaged over the stalled limit cycle (the last
    `window` of n iterations; the cases stall about 3.5 orders down, where the inlet state
    moves in bisection-sized steps and the instantaneous figures wander), FP32 flux
    against FP64, for the first ideal-gas request of each grid size in the stream."""
    from cturbobfm_amd import Solver
    seen, rows = set(), []
    for sp in specs:
        if "table-fluid" in sp.name or (sp.three_d, sp.nodes) in seen:
            continue
        seen.add((sp.three_d, sp.nodes))
        means = {}
        for prec in ("fp64", "mixed"):
            case = sp.build()
            case["config"]["TURBO_OUTPUT"] = "yes"
            case["config"]["TURBO_DIRECTION"] = "i"
            ov = dict(sp.overrides)
            ov["GPU_PRECISION"] = prec
            s = Solver.from_case(case, ov)
            s.set_convergence(True, 1.0e3)   # a drop no run reaches: exactly n iterations
            while s.iterations < n:
                s.advance(min(3000, n - s.iterations))
            h = s.turbo_history(1, n)[-window:]
            s.close()
            means[prec] = (h[:, 1].mean(), h[:, 2].mean())
        (ma, pa), (mb, pb) = means["fp64"], means["mixed"]
        rows.append({"case": sp.name.split(" p_out")[0].split(" M_out")[0], "mass_flow_rel_diff": abs(mb - ma) / abs(ma),
                     "prtt_rel_diff": abs(pb - pa) / abs(pa)})
    return {"iterations": n, "window": window, "cases": rows,
            "mass_flow_max": max(r["mass_flow_rel_diff"] for r in rows) if rows else None,
            "prtt_max": max(r["prtt_rel_diff"] for r in rows) if rows else None}


def bench_serve(args):
    specs, served, plain, dm, dpr, res, capacity = serve_leg(args, args.serve_cases)
    em, epr = equal_iteration_diffs(specs, res)
    lc = limit_cycle_diffs(specs)
    return {
        "metric": "served cases/hour; p50/p99 time-to-converge vs plain FP64",
        "value": served["cases_per_hour"], "unit": "cases/hour", "n_gpus": 1, "steps": len(specs), "warmup": 0,
        "ms_per_step": served["wall_s"] / len(specs) * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64-state/f32-flux",
        "data": "synthetic bump and rotor cases (cturbobfm_amd/serve.py mixed_stream, seed %d)" % args.seed,
        "config": {"workload": "%d cases, Poisson arrivals %.1f/s, retire at %.1f orders of density drop "
                               "(cap %d iterations)" % (len(specs), args.rate, args.drop, args.max_iter),
                   "chunk_iterations": args.chunk},
        "served": served, "plain_fp64": plain, "capacity": capacity,
        "table_fluid_cases": int(sum("table-fluid" in sp.name for sp in specs)),
        "converged_rel_diff_vs_plain": {"mass_flow_max": max(dm) if dm else None,
                                        "mass_flow_median": float(np.median(dm)) if dm else None,
                                        "prtt_max": max(dpr) if dpr else None,
                                        "prtt_median": float(np.median(dpr)) if dpr else None,
                                        "cases": len(dm)},
        "equal_iterations_rel_diff_vs_fp64": {"mass_flow_max": max(em) if em else None,
                                              "mass_flow_median": float(np.median(em)) if em else None,
                                              "prtt_max": max(epr) if epr else None,
                                              "prtt_median": float(np.median(epr)) if epr else None,
                                              "cases": len(em)},
        "limit_cycle_mean_rel_diff_vs_fp64": lc,
        "cases_per_hour_ratio": served["cases_per_hour"] / plain["cases_per_hour"],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--mode", choices=["stage", "batch", "serve"], default="stage",
                    help="stage: the default single-case line; batch: throttle sweep; serve: served stream")
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--batch-cases", type=int, default=64)
    ap.add_argument("--serve-cases", type=int, default=40)
    ap.add_argument("--serve-leg", type=int, default=12,
                    help="stage: cases of the bounded served stream whose time-to-converge the line reports (0: none)")
    ap.add_argument("--rate", type=float, default=2.0)
    ap.add_argument("--chunk", type=int, default=100, help="serve: iterations per batch advance between admissions")
    ap.add_argument("--drop", type=float, default=3.0)
    ap.add_argument("--max-iter", type=int, default=20000)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", default="193x49x129")
    ap.add_argument("--scheme", default="JST")
    ap.add_argument("--limiter", default="",
                    help="stage, Roe: MUSCL reconstruction with this limiter (e.g. 'van albada'), CFL 1.2")
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--viscous", action="store_true",
                    help="stage: the same rotor with VISCOSITY_ACTIVE (Sutherland air, walls unchanged)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only run the plain copy kernel of this many bytes (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (requested --gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    from cturbobfm_amd import Solver, cases, copy_calibration

    if args.mode == "batch":
        print(json.dumps(bench_batch(args)), flush=True)
        return
    if args.mode == "serve":
        print(json.dumps(bench_serve(args)), flush=True)
        return

    if args.copy_calibration:
        ms = copy_calibration(args.copy_calibration, 20)
        print(json.dumps({"copy_bytes": args.copy_calibration, "ms_per_copy": ms,
                          "GBps": 2 * args.copy_calibration / ms / 1e6}))
        return

    ni, nj, nk = (int(v) for v in args.size.split("x"))
    muscl = bool(args.limiter) and args.scheme.upper() == "ROE"
    cfl = 1.2 if muscl else 2.0
    case = cases.rotor3d(ni, nj, nk, n_iterations=args.warmup + 3 * args.steps + 10, scheme=args.scheme,
                         muscl=muscl, limiter=args.limiter if muscl else "None", cfl=cfl)
    if args.viscous:
        case = cases.viscous(case, walls={})
    overrides = {"GPU_PRECISION": args.precision}
    s = Solver.from_case(case, overrides)
    nodes = ni * nj * nk
    stages = 4
    s.advance(args.warmup)
    s.sync()
    s.profile(True, True)
    t0 = time.perf_counter()
    s.advance(args.steps)
    s.sync()
    t1 = time.perf_counter()
    stage_ms, launches, alg_bytes = s.profile_get()
    pre_ms, pre_bytes = s.profile_get_prestage()
    s.profile(False, True)
    # the same number of steps through the production path (captured graphs of one and
    # eight iterations, no per-kernel events): the whole-job rate.  The steps above, with
    # HIP events around every stage and pre-stage launch, give the kernel times.
    g0 = time.perf_counter()
    s.advance(args.steps)
    s.sync()
    g1 = time.perf_counter()
    # the timed steps must have run: a stopped (converged or diverged) case skips its
    # kernels and would report an inflated rate
    its, conv, stopped, last = s.status()
    if stopped or conv or not np.all(np.isfinite(last)):
        raise SystemExit("bench case stopped or diverged inside the timed region (iterations %d, converged %d, "
                         "last log residuals %s)" % (its, conv, last))
    wall = g1 - g0
    value = nodes * stages * args.steps / wall
    nbf = body_force_nodes(case)
    avg_ms = stage_ms / max(launches, 1)
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    avg_pre_ms = pre_ms / max(launches, 1)
    # the committed counter summary covers the default workload (FP64 JST rotor) only
    default_case = (args.precision == "fp64" and args.size == "193x49x129" and args.scheme == "JST"
                    and not args.viscous and not muscl)
    traffic = traffic_from_profiles("k_stage") if default_case else None
    out = {
        "metric": "cell-stage updates/s + HBM roofline %; p50/p99 time-to-converge vs plain FP64",
        "value": value,
        "unit": "cell-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64-state/f32-flux",
        "data": "synthetic annular rotor grid and blade fields (cturbobfm_amd/cases.py), free-stream start",
        "config": {"workload": "annular rotor %dx%dx%d nodes, Hall body force, %s%s, RK41A, CFL %g, local dt, %s%s"
                               % (ni, nj, nk, args.scheme, (" + MUSCL " + args.limiter) if muscl else "", cfl,
                                  args.precision,
                                  ", laminar viscous terms (Sutherland)" if args.viscous else ""),
                   "nodes": nodes, "stages_per_step": stages, "body_force_nodes": nbf,
                   "stage_tiling": s.tuning().split("->")[-1].strip(),
                   "autotune_ms": s.tuning().split("->")[0].strip(),
                   "ms_per_step_with_kernel_events": (t1 - t0) / args.steps * 1e3},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_stage (fused residual + RK stage, alone)", "avg_launch_ms": avg_ms,
                     "algorithmic_bytes_per_launch": alg_bytes, "launches": launches},
        # the kernels launched before it in each stage (boundary fluxes and blade-row
        # sources; viscous gradients and faces), timed by the same events
        "prestage": {"kernels": ("k_prestage + k_vgrad + k_vface" if args.viscous else "k_prestage"),
                     "avg_launch_ms": avg_pre_ms, "body_force_bytes_per_launch": pre_bytes,
                     "stage_pair_achieved_GBps": (alg_bytes + pre_bytes) / ((avg_ms + avg_pre_ms) * 1e-3) / 1e9},
    }
    s.close()
    if args.serve_leg > 0:
        # the metric's second half: time-to-converge of a bounded served stream against the
        # plain FP64 path serving the same stream one case at a time
        specs, served, plain, dm, dpr, _, capacity = serve_leg(args, args.serve_leg)
        out["time_to_converge"] = {
            "workload": "%d served cases (bump 65x33 / 129x65, rotor 97x25x65 / 129x33x97 / 193x49x129, "
                        "5 %% table fluid), Poisson %.1f/s (seed %d), retired at a %.1f-order density drop"
                        % (len(specs), args.rate, args.seed, args.drop),
            "served_p50_s": served["p50_s"], "served_p99_s": served["p99_s"],
            "served_cases_per_hour": served["cases_per_hour"], "served_converged": served["converged"],
            "plain_fp64_p50_s": plain["p50_s"], "plain_fp64_p99_s": plain["p99_s"],
            "plain_fp64_cases_per_hour": plain["cases_per_hour"],
            "cases_per_hour_ratio": served["cases_per_hour"] / plain["cases_per_hour"],
            "mass_flow_rel_diff_max": max(dm) if dm else None, "prtt_rel_diff_max": max(dpr) if dpr else None,
            "capacity": capacity}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(case, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Solve-serving loop: steady cases arrive as a stream, are admitted to one device batch
as memory frees up, advance together (one launch per kernel over all resident cases,
replayed as a captured HIP graph), and retire when their density residual has dropped
the requested number of orders below the baseline of iterations 1-5.

The reference runs one case per process (apps/turbobfm.cpp; Solver::solve,
src/solver.cpp:1298-1424) and has no serving layer; what is reused from it is the
convergence rule (Solver::checkConvergence, solver.cpp:2985-3015) restricted to the
density equation, evaluated on the device for every member.

    specs = mixed_stream(40, seed=7)
    arrivals = poisson_arrivals(len(specs), rate=2.0, seed=7)
    report = Server(drop=3.0).run(specs, arrivals)

Every case computes exactly the same numbers as it would alone (each keeps its own
state, counters and reductions; tests/test_gpu_batch.py and tests/test_gpu_serve.py).
"""
from __future__ import annotations

import dataclasses
import time
from collections import deque
from typing import Callable, Dict, List, Optional

import numpy as np

from . import Batch, Solver, SolverError, cases, device_memory


@dataclasses.dataclass
class CaseSpec:
    """One request: a synthetic case builder and its input.ini overrides."""
    name: str
    build: Callable[[], Dict]
    overrides: Dict[str, str]
    nodes: int
    three_d: bool


@dataclasses.dataclass
class CaseResult:
    name: str
    nodes: int
    arrival: float
    admitted: float = -1.0
    finished: float = -1.0
    iterations: int = 0
    converged: bool = False
    diverged: bool = False
    state: Optional[np.ndarray] = None
    turbo: Optional[np.ndarray] = None     # Solver.turbo() at retirement

    @property
    def time_to_converge(self) -> float:
        return self.finished - self.arrival


def _fluid(case: Dict, lut_path: Optional[str]) -> Dict:
    return cases.with_real_gas(case, lut_path) if lut_path else case


def rotor_spec(ni: int, nj: int, nk: int, outlet_pressure: float, precision: str, scheme: str = "JST",
               max_iter: int = 20000, lut_path: Optional[str] = None) -> CaseSpec:
    def build():
        return _fluid(cases.rotor3d(ni, nj, nk, n_iterations=max_iter, scheme=scheme,
                                    outlet_pressure=outlet_pressure), lut_path)
    # the radial-equilibrium outlet takes its hub pressure from INIT_PRESSURE, which
    # cases.rotor3d sets to the outlet pressure
    return CaseSpec("rotor %dx%dx%d p_out=%.0f %s%s" % (ni, nj, nk, outlet_pressure, precision,
                                                        " table-fluid" if lut_path else ""), build,
                    {"GPU_PRECISION": precision}, ni * nj * nk, True)


def bump_spec(ni: int, nj: int, mach_out: float, precision: str, max_iter: int = 20000,
              lut_path: Optional[str] = None) -> CaseSpec:
    def build():
        return _fluid(cases.bump2d(ni, nj, n_iterations=max_iter, mach_out=mach_out), lut_path)
    return CaseSpec("bump %dx%d M_out=%.3f %s%s" % (ni, nj, mach_out, precision,
                                                    " table-fluid" if lut_path else ""), build,
                    {"GPU_PRECISION": precision}, ni * nj, False)


def stream_table(directory: str, seed: int = 5) -> str:
    """Writes the stream's look-up table (a 60 x 60 rho-e covolume CO2-like gas from a
    seeded smooth analytic equation of state, cases.covolume_lut) and returns its path."""
    import os
    return cases.write_lut(cases.covolume_lut(seed=seed), os.path.join(directory, "covolume_seed%d.lut" % seed))


def mixed_stream(n: int, seed: int = 7, precision: str = "mixed", real_fraction: float = 0.05,
                 lut_path: Optional[str] = None) -> List[CaseSpec]:
    """A seeded mix of 2-D bump and 3-D rotor requests: half bumps (65x33 or 129x65,
    outlet Mach 0.45-0.55), the rest rotors weighted toward the small sizes (97x25x65,
    129x33x97, rarely 193x49x129) with outlet pressure drawn from 92-98 kPa.  A fraction
    real_fraction of the requests (drawn from the same seed) run on the look-up-table
    fluid of lut_path; these compute in FP64, the look-up-table path having no FP32 form."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        u = rng.random()
        table = lut_path if (lut_path and rng.random() < real_fraction) else None
        prec = "fp64" if table else precision
        if u < 0.5:
            ni, nj = (65, 33) if rng.random() < 0.6 else (129, 65)
            out.append(bump_spec(ni, nj, float(rng.uniform(0.45, 0.55)), prec, lut_path=table))
        else:
            v = rng.random()
            size = (97, 25, 65) if v < 0.65 else ((129, 33, 97) if v < 0.95 else (193, 49, 129))
            out.append(rotor_spec(*size, outlet_pressure=float(rng.uniform(92000.0, 98000.0)), precision=prec,
                                  lut_path=table))
    return out


def poisson_arrivals(n: int, rate: float, seed: int = 7) -> np.ndarray:
    """Arrival times (s) of a Poisson stream of the given mean rate."""
    rng = np.random.default_rng(seed + 1)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


class Server:
    """Continuous batching of steady solves on one GPU.

    drop        orders of density-residual drop that retire a case
    max_iter    iteration cap: a case that has not converged by then retires unconverged
    chunk       iterations per batch advance between admission / retirement decisions
    budget      device bytes the resident cases may hold (default: 80 % of free memory)
    lanes       node-count upper bounds of the batches that run side by side: a request
                joins the first lane its grid fits.  Each lane is one batch advanced by its
                own host thread on its own stream, so cheap small grids iterate at their own
                rate instead of at the pace of the largest resident grid.
    """

    def __init__(self, drop: float = 8.0, max_iter: int = 20000, chunk: int = 100,
                 budget: Optional[int] = None, keep_states: bool = False, keep_turbo: bool = False,
                 log: Optional[Callable[[str], None]] = None, lanes=(50000, None)):
        self.drop = drop
        self.max_iter = max_iter
        self.chunk = chunk
        self.budget = budget if budget is not None else int(0.8 * device_memory()[0])
        self.keep_states = keep_states
        self.keep_turbo = keep_turbo
        self.log = log
        self.lanes = list(lanes)
        self.bytes_per_node = {True: 1500.0, False: 1500.0}   # refined from the first cases

    def _estimate(self, spec: CaseSpec) -> int:
        return int(self.bytes_per_node[spec.three_d] * spec.nodes) + (1 << 20)

    def _lane_of(self, spec: CaseSpec) -> int:
        for q, cap in enumerate(self.lanes):
            if cap is None or spec.nodes <= cap:
                return q
        return len(self.lanes) - 1

    def _build(self, spec: CaseSpec) -> Solver:
        s = Solver.from_case(spec.build(), spec.overrides)
        s.set_convergence(True, self.drop)
        return s

    def run(self, specs: List[CaseSpec], arrivals: np.ndarray) -> List[CaseResult]:
        """Serves the stream.  The main thread takes arrivals and sets requests up (grid,
        geometry, device buffers) on a worker pool as soon as memory allows; each lane
        thread admits its set-up requests in arrival order at its chunk boundaries,
        advances its batch, and retires converged, diverged or capped cases."""
        import threading
        from concurrent.futures import ThreadPoolExecutor
        results = [CaseResult(s.name, s.nodes, float(a)) for s, a in zip(specs, arrivals)]
        order = list(np.argsort(arrivals))
        nl = len(self.lanes)
        lock = threading.Lock()
        queues = [deque() for _ in range(nl)]       # arrived requests of each lane, in order
        building: Dict[int, object] = {}            # request -> future of its Solver
        reserved: Dict[int, int] = {}               # estimated bytes of requests in set-up or not yet admitted
        state = {"used": 0, "done": 0, "error": None}
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        pool = ThreadPoolExecutor(max_workers=2)

        def lane_main(li: int):
            batch = Batch()
            resident: Dict[int, Solver] = {}
            footprint: Dict[int, int] = {}
            try:
                while True:
                    with lock:
                        if state["done"] == len(specs) or state["error"]:
                            return
                        ready = []
                        while queues[li] and queues[li][0] in building and building[queues[li][0]].done():
                            q = queues[li].popleft()
                            ready.append((q, building.pop(q)))
                    for q, fut in ready:
                        s = fut.result()
                        sp = specs[q]
                        b = s.device_bytes()
                        batch.admit(s)
                        with lock:
                            self.bytes_per_node[sp.three_d] = b / sp.nodes
                            resident[q] = s
                            footprint[q] = b
                            # the estimate stays reserved until the measured footprint
                            # is counted, so the budget never sees the case as absent
                            state["used"] += b
                            reserved.pop(q)
                            results[q].admitted = now()
                    if not resident:
                        time.sleep(0.002)
                        continue
                    batch.advance(self.chunk)
                    for q in list(resident):
                        s = resident[q]
                        r = results[q]
                        try:
                            it, conv, stopped, _ = s.status()
                        except SolverError:
                            it, conv, stopped = s.iterations, 0, True
                            r.diverged = True
                        if stopped or it >= self.max_iter:
                            r.finished = now()
                            r.iterations = it
                            r.converged = conv > 0
                            if self.keep_states:
                                r.state = s.state()
                            if self.keep_turbo and not r.diverged:
                                r.turbo = s.turbo()
                            batch.retire(s)
                            s.close()
                            del resident[q]
                            with lock:
                                state["used"] -= footprint.pop(q)
                                state["done"] += 1
                                done = state["done"]
                            if self.log:
                                self.log("served %d/%d: %s, %d iterations, %.2f s after arrival"
                                         % (done, len(specs), r.name, r.iterations, r.time_to_converge))
            except BaseException as e:   # surfaced by the main thread
                with lock:
                    state["error"] = e
            finally:
                batch.close()

        threads = [threading.Thread(target=lane_main, args=(li,), daemon=True) for li in range(nl)]
        for t in threads:
            t.start()
        nxt = 0
        try:
            while True:
                with lock:
                    if state["error"]:
                        raise state["error"]
                    if state["done"] == len(specs):
                        break
                    while nxt < len(order) and arrivals[order[nxt]] <= now():
                        q = order[nxt]
                        queues[self._lane_of(specs[q])].append(q)
                        nxt += 1
                    # set-up ahead of admission, first come first served over all lanes,
                    # while the estimated footprint of resident and in-set-up cases fits
                    pending = sorted((q for qu in queues for q in qu if q not in building),
                                     key=lambda q: arrivals[q])
                    for q in pending:
                        est = self._estimate(specs[q])
                        busy = state["used"] > 0 or bool(reserved)
                        if busy and state["used"] + sum(reserved.values()) + est > self.budget:
                            break
                        building[q] = pool.submit(self._build, specs[q])
                        reserved[q] = est
                time.sleep(0.001)
        finally:
            with lock:
                if state["done"] != len(specs) and not state["error"]:
                    state["error"] = RuntimeError("serving stopped")
            for t in threads:
                t.join()
            pool.shutdown(wait=True)
        return results


def sequential(specs: List[CaseSpec], drop: float, max_iter: int = 20000, chunk: int = 500,
               precision: Optional[str] = None, keep_states: bool = False,
               keep_turbo: bool = False, log: Optional[Callable[[str], None]] = None) -> List[CaseResult]:
    """The plain path: each case alone, one after another, from set-up to convergence
    (precision overrides the spec's GPU_PRECISION when given)."""
    out = []
    t0 = time.perf_counter()
    for sp in specs:
        r = CaseResult(sp.name, sp.nodes, time.perf_counter() - t0)
        ov = dict(sp.overrides)
        if precision is not None:
            ov["GPU_PRECISION"] = precision
        s = Solver.from_case(sp.build(), ov)
        s.set_convergence(True, drop)
        r.admitted = time.perf_counter() - t0
        try:
            while s.iterations < max_iter and s.converged_at == 0:
                s.advance(min(chunk, max_iter - s.iterations))
        except SolverError:
            r.diverged = True
        r.iterations = s.iterations
        r.converged = s.converged_at > 0
        if keep_states:
            r.state = s.state()
        if keep_turbo and not r.diverged:
            r.turbo = s.turbo()
        s.close()
        r.finished = time.perf_counter() - t0
        out.append(r)
        if log:
            log("alone %d/%d: %s, %d iterations, %.2f s" % (len(out), len(specs), r.name, r.iterations,
                                                            r.finished - r.admitted))
    return out


def summary(results: List[CaseResult], wall: float) -> Dict:
    ttc = np.array([r.time_to_converge for r in results])
    return {"cases": len(results), "converged": int(sum(r.converged for r in results)),
            "diverged": int(sum(r.diverged for r in results)),
            "p50_s": float(np.percentile(ttc, 50)), "p99_s": float(np.percentile(ttc, 99)),
            "cases_per_hour": len(results) / wall * 3600.0, "wall_s": wall,
            "iterations": int(sum(r.iterations for r in results)),
            "cell_iterations": float(sum(r.iterations * r.nodes for r in results))}

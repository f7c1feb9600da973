"""Serving form: a request loop that runs implicit Navier-Stokes time steps for many meshes at once.

Each request is a unit-cube tetrahedral channel case (cells per side log-uniform in [12, 64], Re
uniform in [10, 1000], inflow perturbation amplitude uniform in [0, 0.1], all drawn from one SplitMix64
stream; inflow on x = 1, outlet p = 0 on x = 0, no-slip on the side faces) that runs `steps` implicit
steps (dt = 1e-2) in the precision form (FP32 element matrices, BiCGStab on the block-scaled FP32
operator with replacements of its residual by the matrix-free FP64 one, to 1e-8; a step whose solve
fails is redone in FP64 and counted).

Continuous batching (`Server.run(..., batched=True)`):
  * admission is event driven: a request is admitted as soon as it has arrived, a slot is free and its
    device-memory estimate fits the HBM budget (the same estimate is charged on admission and returned
    when the request leaves); a request whose estimate exceeds the whole budget is rejected, not queued;
    its mesh, CSR pattern, scatter maps and state are created on the device and stay there until its
    last step;
  * every admitted request advances at its own pace on its own host thread and stream: LinearizeU,
    then it waits for the next batched assembly, then BoundaryConditions / Krylov solve / update.  An
    assembler thread launches ONE batched assembly (a persistent kernel replayed from a hipGraph) for
    whichever requests are waiting when it is free, so a request whose solve ends early starts its
    next step at once instead of waiting for the slowest request of a tick;
  * requests leave after their last step, freeing slot and memory for the queue.
The sequential reference (`batched=False`) serves one request at a time with the plain per-request
calls.  Each request's results are independent of the mix: the batched assembly runs the same
per-workgroup code on the same maps, every solve has its own stream and scalars.
"""
import collections
import math
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import cabi, cases, meshgen


@dataclass
class Request:
    rid: int
    arrival: float     # seconds after the start of the run
    cells: int
    re: float
    amp: float
    steps: int = 10
    dt: float = 1e-2


def make_requests(count=512, rate=200.0, seed=cases.DEFAULT_SEED, cells=(12, 64), re=(10.0, 1000.0),
                  amp=(0.0, 0.1), steps=10):
    """Poisson arrivals at `rate` per second; 4 draws per request in request order."""
    u = meshgen.uniform(seed, 4 * count, 0.0, 1.0).reshape(count, 4)
    t = 0.0
    out = []
    for i in range(count):
        t += -math.log(1.0 - u[i, 0]) / rate
        n = int(round(math.exp(math.log(cells[0]) + (math.log(cells[1]) - math.log(cells[0])) * u[i, 1])))
        out.append(Request(i, t, n, re[0] + (re[1] - re[0]) * u[i, 2], amp[0] + (amp[1] - amp[0]) * u[i, 3], steps))
    return out


def device_bytes_estimate(cells):
    """Device memory of one request, an upper estimate: per node block (at most 15 per row on these
    meshes) the FP64, FP32 and block-scaled FP32 values, its column and staging bytes; per element the
    connectivity, incidences, staging positions and element-to-slot map; per node the BiCGStab
    workspace (10 vectors), the FGMRES workspace of the fallback (35 vectors), D^-1, the fields, the
    packed node records and the rhs / solution / residual vectors."""
    N = (cells + 1) ** 3
    E = 6 * cells ** 3
    nb = 15 * N
    return nb * (128 + 64 + 64 + 4 + 6) + E * (16 + 16 + 16 + 64) + N * (320 + 1120 + 128 + 160 + 96 + 96 + 16)


class _Case:
    def __init__(self, req, seed, tol):
        self.req = req
        self.ctx = cabi.Context(0)
        self.ctx.create_stream()
        case = cases.inflow_cube_tets(req.cells, seed=seed + req.rid, nu=1.0 / req.re, dt=req.dt, amp=req.amp,
                                      outlet=True)
        case.tol = tol
        self.g = cases.GpuCase(self.ctx, case, precision=cabi.FP32)
        self.steps_done = 0
        self.closed = False
        self.iterations = 0
        self.converged = True
        self.prm = cabi.NSParams.from_buffer_copy(self.g.prm)
        self.prm.precision = cabi.FP32

    def solve_step(self):
        st = self.g.post_assembly_step(cabi.solver_params(tol=self.g.case.tol, precision=cabi.FP32),
                                       raise_on_failure=False)
        self.steps_done += 1
        self.iterations += st.iterations
        self.converged = self.converged and bool(st.converged)

    def plain_step(self):
        st = self.g.step(cabi.solver_params(tol=self.g.case.tol, precision=cabi.FP32), raise_on_failure=False)
        self.steps_done += 1
        self.iterations += st.iterations
        self.converged = self.converged and bool(st.converged)

    def close(self):
        if not self.closed:
            self.closed = True
            self.g.close()
            self.ctx.close()


@dataclass
class Record:
    rid: int
    cells: int
    arrival: float
    admitted: float
    done: float
    iterations: int
    converged: bool
    fallbacks: int = 0     # steps whose precision-form solve stagnated and were redone in FP64
    rejected: bool = False  # its device-memory estimate exceeds the whole HBM budget
    device_bytes: int = 0   # device memory of the request's context after its last step
    bytes_estimate: int = 0
    u: np.ndarray = field(default=None, repr=False)
    p: np.ndarray = field(default=None, repr=False)

    @property
    def latency(self):
        return self.done - self.arrival


class Server:
    def __init__(self, max_active=16, hbm_budget=200e9, tol=1e-8, seed=cases.DEFAULT_SEED, keep_fields=False,
                 log=None):
        self.max_active = max_active
        self.hbm_budget = hbm_budget
        self.tol = tol
        self.seed = seed
        self.keep_fields = keep_fields   # True: keep every request's final fields; a set: those of these ids
        self.log = log   # optional callable(str): one line per finished request

    def run(self, requests, batched=True, time_scale=1.0):
        """Serve `requests` (arrival times scaled by time_scale); returns (records, summary)."""
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        records = []
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        if batched:
            self._run_batched(pending, records, now, time_scale)
        else:
            self._run_sequential(pending, records, now, time_scale)
        return records, summarize(records, batched, self.max_active)

    def _finish(self, c, records, now, lock=None):
        keep = self.keep_fields is True or (self.keep_fields not in (None, False) and c.req.rid in self.keep_fields)
        u = c.g.u.copy() if keep else None
        p = c.g.p.copy() if keep else None
        rec = Record(c.req.rid, c.req.cells, c.req.arrival * self._ts, c.admitted, now(), c.iterations,
                     c.converged, c.g.fallbacks, False, c.ctx.memory(), device_bytes_estimate(c.req.cells), u, p)
        if lock is not None:
            with lock:
                records.append(rec)
        else:
            records.append(rec)
        if self.log:
            self.log("request %d (%d^3 cells, Re %.0f): %d Krylov iterations in %d steps, latency %.3f s%s%s"
                     % (rec.rid, rec.cells, c.req.re, rec.iterations, c.steps_done, rec.latency,
                        ", %d FP64 fallback steps" % rec.fallbacks if rec.fallbacks else "",
                        "" if rec.converged else ", NOT converged"))
        c.close()

    def _run_sequential(self, pending, records, now, ts):
        self._ts = ts
        while pending:
            r = pending.popleft()
            wait = r.arrival * ts - now()
            if wait > 0:
                time.sleep(wait)
            c = _Case(r, self.seed, self.tol)
            c.admitted = now()
            while c.steps_done < r.steps:
                c.plain_step()
            self._finish(c, records, now)

    def _run_batched(self, pending, records, now, ts):
        self._ts = ts
        ctx = cabi.Context(0)
        batch = cabi.Batch(ctx, self.max_active)
        cond = threading.Condition()
        ready = []                                      # requests waiting for the batched assembly
        st = {"active": 0, "in_use": 0, "stop": False, "error": None}

        def fail(e):
            with cond:
                if st["error"] is None:
                    st["error"] = e
                cond.notify_all()

        def assembler():
            try:
                while True:
                    with cond:
                        while not ready and not st["stop"] and st["error"] is None:
                            cond.wait()
                        if st["error"] is not None or (st["stop"] and not ready):
                            return
                        group = list(ready)
                        ready.clear()
                    batch.assemble([c.g.system for c in group], [c.g.fields for c in group],
                                   [c.prm for c in group])
                    with cond:
                        for c in group:
                            c.go = True
                        cond.notify_all()
            except BaseException as e:   # noqa: BLE001 -- handed to the main thread
                fail(e)

        def worker(c):
            try:
                while c.steps_done < c.req.steps:
                    c.g.fields.linearize()                  # LinearizeU on the request's stream
                    with cond:
                        c.go = False
                        ready.append(c)
                        cond.notify_all()
                        while not c.go and st["error"] is None:
                            cond.wait()
                        if st["error"] is not None:
                            return
                    c.solve_step()                          # BCs, solve, update: its own stream
                self._finish(c, records, now, cond)
                with cond:
                    st["active"] -= 1
                    st["in_use"] -= c.estimate
                    cond.notify_all()
            except BaseException as e:   # noqa: BLE001 -- handed to the main thread
                fail(e)
            finally:
                c.close()

        asm = threading.Thread(target=assembler, daemon=True)
        asm.start()
        pool = ThreadPoolExecutor(self.max_active)
        try:
            while True:
                admit = []
                with cond:
                    if st["error"] is not None:
                        raise st["error"]
                    while pending and pending[0].arrival * ts <= now() and st["active"] < self.max_active:
                        est = device_bytes_estimate(pending[0].cells)
                        if est > self.hbm_budget:           # can never fit: rejected, not queued
                            r = pending.popleft()
                            records.append(Record(r.rid, r.cells, r.arrival * ts, now(), now(), 0, False,
                                                  rejected=True, bytes_estimate=est))
                            continue
                        if st["in_use"] + est > self.hbm_budget:
                            break
                        r = pending.popleft()
                        st["active"] += 1
                        st["in_use"] += est
                        admit.append((r, est))
                    if not admit:
                        if not pending and st["active"] == 0:
                            break
                        # wait for a departure, or for the next arrival if a slot and memory are free
                        timeout = None
                        if pending and st["active"] < self.max_active:
                            timeout = max(pending[0].arrival * ts - now(), 0.0)
                        if timeout is None or timeout > 0:
                            cond.wait(timeout)
                        continue
                for r, est in admit:                        # device set-up outside the lock
                    c = _Case(r, self.seed, self.tol)
                    c.admitted = now()
                    c.estimate = est
                    c.go = False
                    pool.submit(worker, c)
        finally:
            with cond:
                st["stop"] = True
                if st["error"] is None and st["active"]:
                    st["error"] = RuntimeError("serving loop left with active requests")
                cond.notify_all()
            pool.shutdown(wait=True)
            asm.join()
            batch.close()
            ctx.close()


def summarize(records, batched, max_active):
    if not records:
        return {}
    lat = np.array([r.latency for r in records])
    first = min(r.arrival for r in records)
    last = max(r.done for r in records)
    return {"mode": "continuous batching" if batched else "sequential", "requests": len(records),
            "max_active": max_active if batched else 1,
            "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(lat.mean()), "throughput_req_s": len(records) / (last - first),
            "makespan_s": last - first, "all_converged": bool(all(r.converged for r in records if not r.rejected)),
            "rejected": int(sum(r.rejected for r in records)),
            "krylov_iterations": int(sum(r.iterations for r in records)),
            "fp64_fallback_steps": int(sum(r.fallbacks for r in records))}

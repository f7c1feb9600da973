"""Serving form: a request loop that runs implicit Navier-Stokes time steps for many meshes at once.

Each request is a unit-cube tetrahedral channel case (cells per side log-uniform in [12, 64], Re
uniform in [10, 1000], inflow perturbation amplitude uniform in [0, 0.1], all drawn from one SplitMix64
stream; inflow on x = 1, outlet p = 0 on x = 0, no-slip on the side faces) that runs `steps` implicit
steps (dt = 1e-2) in the precision form (FP32 element matrices, FP32 BiCGStab inner solves under FP64
flexible GMRES on the matrix-free FP64 operator, to 1e-8; a step whose solve fails is redone in FP64
and counted).

Continuous batching (`Server.run(..., batched=True)`):
  * a request is admitted when it has arrived, a slot is free and its device memory fits the HBM
    budget; its mesh, CSR pattern, scatter maps and state are created on the device and stay there
    until its last step;
  * every tick advances all admitted requests by one step: LinearizeU per request, then ONE batched
    assembly launch for all of them (a persistent kernel replayed from a hipGraph), then the
    BoundaryConditions / Krylov solve / update of each request on its own stream from a pool of host
    threads, so their kernels overlap on the GPU;
  * requests leave after their last step, freeing slot and memory for the queue.
The sequential reference (`batched=False`) serves one request at a time with the plain per-request
calls.  Each request's results are independent of the mix: the batched assembly runs the same
per-workgroup code on the same maps, every solve has its own stream and scalars.
"""
import collections
import math
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
    """Device memory of one request (FP64 + FP32 values, scatter maps, fields, Krylov workspace)."""
    N = (cells + 1) ** 3
    E = 6 * cells ** 3
    nb = 15 * N
    return nb * (128 + 64 + 4 + 6) + E * (64 + 16 + 16 + 4 * 4) + N * (4 * 8 * 10 + 8 * 4 * 10 + 64)


class _Case:
    def __init__(self, req, seed, tol):
        self.req = req
        self.ctx = cabi.Context(0)
        self.ctx.create_stream()
        case = cases.inflow_cube_tets(req.cells, seed=seed + req.rid, nu=1.0 / req.re, dt=req.dt, amp=req.amp,
                                      outlet=True)
        case.tol = tol
        self.g = cases.GpuCase(self.ctx, case, precision=cabi.FP32)
        self.bytes = self.ctx.memory()
        self.steps_done = 0
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
        self.keep_fields = keep_fields
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

    def _finish(self, c, records, now):
        u = c.g.u.copy() if self.keep_fields else None
        p = c.g.p.copy() if self.keep_fields else None
        rec = Record(c.req.rid, c.req.cells, c.req.arrival * self._ts, c.admitted, now(), c.iterations,
                     c.converged, c.g.fallbacks, u, p)
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
        pool = ThreadPoolExecutor(self.max_active)
        active = []
        in_use = 0
        try:
            while pending or active:
                # admission: arrived, a free slot, and the HBM budget
                while (pending and pending[0].arrival * ts <= now() and len(active) < self.max_active and
                       in_use + device_bytes_estimate(pending[0].cells) <= self.hbm_budget):
                    r = pending.popleft()
                    c = _Case(r, self.seed, self.tol)
                    c.admitted = now()
                    in_use += c.bytes
                    active.append(c)
                if not active:
                    wait = pending[0].arrival * ts - now()
                    if wait > 0:
                        time.sleep(wait)
                    continue
                # one tick: LinearizeU, one batched assembly, per-request BCs + solve + update
                for c in active:
                    c.g.fields.linearize()
                batch.assemble([c.g.system for c in active], [c.g.fields for c in active], [c.prm for c in active])
                list(pool.map(lambda c: c.solve_step(), active))
                still = []
                for c in active:
                    if c.steps_done >= c.req.steps:
                        in_use -= c.bytes
                        self._finish(c, records, now)
                    else:
                        still.append(c)
                active = still
        finally:
            for c in active:
                c.close()
            pool.shutdown()
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
            "makespan_s": last - first, "all_converged": bool(all(r.converged for r in records)),
            "krylov_iterations": int(sum(r.iterations for r in records)),
            "fp64_fallback_steps": int(sum(r.fallbacks for r in records))}

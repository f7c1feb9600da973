"""Serving form: request generation (CPU), batched assembly and continuous batching (GPU)."""
import numpy as np
import pytest

from coolfluid3_amd import cabi, cases, serving


def test_request_stream():
    reqs = serving.make_requests(512)
    arr = np.array([r.arrival for r in reqs])
    cells = np.array([r.cells for r in reqs])
    assert np.all(np.diff(arr) > 0)
    assert 150 < 512 / arr[-1] < 260                       # Poisson arrivals at 200 req/s
    assert cells.min() >= 12 and cells.max() <= 64
    assert 20 < np.exp(np.log(cells).mean()) < 36          # log-uniform: geometric mean ~ sqrt(12 * 64)
    re = np.array([r.re for r in reqs])
    assert re.min() >= 10 and re.max() <= 1000
    again = serving.make_requests(512)
    assert [(r.arrival, r.cells, r.re, r.amp) for r in again] == [(r.arrival, r.cells, r.re, r.amp) for r in reqs]


@pytest.mark.gpu
def test_batched_assembly_equals_single_assembly():
    """One persistent batched launch gives every system the bits of its own single-system assembly."""
    ctx = cabi.Context(0)
    mk = lambda: [cases.GpuCase(ctx, cases.inflow_cube_tets(n, seed=100 + n, nu=1e-3 * n), precision=cabi.FP32)
                  for n in (4, 6, 9)]
    ref, gs = mk(), mk()
    single = []
    for g in ref:
        g.fields.linearize()
        g.assemble(cabi.FP32)
        single.append(g.system.values_fp32().copy())
    batch = cabi.Batch(ctx, 8)
    prms = []
    for g in gs:
        g.fields.linearize()
        p = cabi.NSParams.from_buffer_copy(g.prm)
        p.precision = cabi.FP32
        prms.append(p)
    batch.assemble([g.system for g in gs], [g.fields for g in gs], prms)
    for g, v in zip(gs, single):
        assert np.abs(v).max() > 0 and np.array_equal(g.system.values_fp32(), v)
    # a second, smaller batch replays the captured graph with a new table
    batch.assemble([ref[1].system], [ref[1].fields], prms[1:2])
    assert np.array_equal(ref[1].system.values_fp32(), single[1])
    batch.close()
    for g in ref + gs:
        g.close()
    ctx.close()


@pytest.mark.gpu
def test_result_alone_equals_result_in_mix():
    """Each request gives the same fields served alone and in a mix (continuous batching)."""
    reqs = [serving.Request(i, 0.0, n, re, amp, steps=3) for i, (n, re, amp) in
            enumerate([(6, 100.0, 0.05), (5, 500.0, 0.0), (8, 20.0, 0.1), (7, 900.0, 0.02)])]
    srv = serving.Server(max_active=4, keep_fields=True)
    alone, _ = srv.run(reqs[:1])
    mix, summary = srv.run(reqs)
    a = alone[0]
    m = [r for r in mix if r.rid == 0][0]
    du = np.abs(a.u - m.u).max() / np.abs(a.u).max()
    print(f"alone vs mix: max relative velocity difference {du:.2e}; {summary}")
    assert du <= 1e-12                                    # stated tolerance (the paths are deterministic)
    assert np.abs(a.p - m.p).max() <= 1e-12 * max(np.abs(a.p).max(), 1.0)
    assert summary["all_converged"]
    # the sequential path serves the same request to the same fields
    seq, _ = srv.run(reqs[:1], batched=False)
    assert np.abs(seq[0].u - a.u).max() <= 1e-12 * np.abs(a.u).max()


@pytest.mark.gpu
def test_mix_of_requests_is_deterministic():
    """Eight mid-sized requests solved concurrently (continuous batching) give every request the
    iteration counts and the fields of one-at-a-time serving: the concurrent solves share nothing and
    every reduction has a fixed order (a stop flag raised while other workgroups of the same launch
    start must not split a workgroup)."""
    reqs = [serving.Request(i, 0.0, n, re, amp, steps=3) for i, (n, re, amp) in
            enumerate([(24, 120.0, 0.05), (30, 260.0, 0.0), (27, 80.0, 0.1), (22, 700.0, 0.02),
                       (28, 300.0, 0.07), (20, 950.0, 0.01), (26, 55.0, 0.04), (25, 410.0, 0.09)])]
    srv = serving.Server(max_active=8, keep_fields=True)
    mix, summary = srv.run(reqs)
    seq, _ = srv.run(reqs, batched=False)
    by_id = {r.rid: r for r in seq}
    for m in mix:
        s = by_id[m.rid]
        assert m.iterations == s.iterations, (m.rid, m.iterations, s.iterations)
        assert np.abs(m.u - s.u).max() <= 1e-12 * np.abs(s.u).max()
        assert np.abs(m.p - s.p).max() <= 1e-12 * max(np.abs(s.p).max(), 1.0)
    assert summary["all_converged"]


@pytest.mark.gpu
def test_continuous_progress_and_admission():
    """No tick barrier: a small request admitted together with a large one finishes long before it; a
    request whose memory estimate exceeds the whole HBM budget is rejected instead of blocking the queue;
    the estimate charged on admission covers the device memory each request ends with."""
    big, small = serving.Request(0, 0.0, 48, 200.0, 0.05, steps=3), serving.Request(1, 0.0, 5, 200.0, 0.05, steps=3)
    huge = serving.Request(2, 0.0, 300, 200.0, 0.05, steps=3)
    budget = 4 * serving.device_bytes_estimate(48)
    assert serving.device_bytes_estimate(300) > budget
    srv = serving.Server(max_active=4, hbm_budget=budget)
    recs, summary = srv.run([big, huge, small])
    by = {r.rid: r for r in recs}
    print(f"big done {by[0].done:.3f} s, small done {by[1].done:.3f} s; {summary}")
    assert by[2].rejected and not by[0].rejected and not by[1].rejected
    assert summary["rejected"] == 1 and summary["all_converged"]
    assert by[1].done < 0.5 * by[0].done
    for r in (by[0], by[1]):
        assert 0 < r.device_bytes <= r.bytes_estimate, (r.rid, r.device_bytes, r.bytes_estimate)


@pytest.mark.gpu
@pytest.mark.timeout(900)
def test_serving_stream_512():
    """The 512-request stream (Poisson arrivals at 200 req/s, cells log-uniform in [12, 64], Re in
    [10, 1000], 10 steps each) through the continuous server: every request converges at every step, and
    a sample of requests gives the same fields served alone as in the mix."""
    reqs = serving.make_requests(512)
    sample = [r.rid for r in sorted(reqs, key=lambda r: r.cells)[::128]]
    srv = serving.Server(max_active=16, keep_fields=set(sample), log=lambda m: print("serving: " + m, flush=True))
    recs, summary = srv.run(reqs)
    print("stream:", summary, flush=True)
    assert summary["requests"] == 512 and summary["rejected"] == 0 and summary["all_converged"]
    by = {r.rid: r for r in recs}
    alone_srv = serving.Server(max_active=1, keep_fields=True)
    for rid in sample:
        r = [q for q in reqs if q.rid == rid][0]
        a = alone_srv.run([serving.Request(r.rid, 0.0, r.cells, r.re, r.amp, r.steps)])[0][0]
        m = by[rid]
        du = np.abs(a.u - m.u).max() / np.abs(a.u).max()
        print(f"request {rid} ({r.cells}^3): alone vs mix velocity difference {du:.1e}", flush=True)
        assert a.iterations == m.iterations
        assert du <= 1e-12 and np.abs(a.p - m.p).max() <= 1e-12 * max(np.abs(a.p).max(), 1.0)

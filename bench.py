"""Benchmark of the UFEM Navier-Stokes hot path on one MI355X.

A step is one Assembly action of the reference (ZeroLSS + element assembly + scatter into the block-CSR
system, the quantity timed by plugins/UFEM/test/ptest-navier-stokes-assembly.py) over the 128^3-hexahedra
unit cube split in 6 P1 tetrahedra with 1% skewed elements (12,582,912 elements, 8,586,756 dofs), FP64
values, all inputs resident in HBM.  value = elements assembled per second (whole job).

Also reported: the dominant kernel's roofline (HIP-event duration of the assembly kernel against its
algorithmic bytes, and the HBM traffic per launch from the committed rocprofv3 counter summary), the FP32
assembly of the same mesh, the SpMV bandwidth, the time to a 1e-8 relative residual per implicit step of
the FP64 path and of the precision form (FP32 element matrices, FP64 refinement), the CPU baseline (the
oracle restatement on the host cores: the same assembly, and the implicit steps of the 32x32 triangle
cavity and of the 64^3 tetrahedral cube), and the other workloads: the FP64 drop-in on the 64^3 cube
(assembly and BiCGStab/Jacobi step to 1e-10), the 160^3 Q1 hexahedral lid-driven cavity (assembly and
Picard iterations in both forms), the 32x32 triangle cavity step (GMRES + ILU0), the reference's
Taylor-Green acceptance run, the semi-implicit solver, and the serving loop (p50/p99 latency).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Melem/s assembly; time-to-1e-8 residual (FP32+refine vs FP64); p99 latency"
HBM_PEAK_GBS = 8000.0


def algorithmic_assembly_bytes(n_nodes, n_elems, k, dim, n_blocks, value_bytes):
    neq = dim + 1
    nnz = n_blocks * neq * neq
    return n_nodes * (3 * dim * 8 + 16) + n_elems * k * 4 + nnz * value_bytes + n_nodes * neq * 8


def algorithmic_spmv_bytes(n_nodes, dim, n_blocks, value_bytes):
    neq = dim + 1
    return n_blocks * neq * neq * value_bytes + n_blocks * 4 + (n_nodes + 1) * 4 + 2 * n_nodes * neq * value_bytes


def host_cpu():
    """Threads for the host baselines: the CPU share of this process (its affinity mask, capped by
    OMP_NUM_THREADS when that is set, as on a shared GPU node), the processor model and the machine's
    logical CPU count."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return n, model, os.cpu_count()


def cpu_baseline(case, threads, model, budget_s=20.0):
    """Oracle restatement of the same FP64 assembly on the host cores (element colouring, OpenMP), plus
    one thread (what one reference MPI rank does) on the first eighth of the elements."""
    from oracle import pyoracle
    o = pyoracle.get()
    start, cols, slots = o.build_sparsity(case.conn, case.n_nodes)
    prm = pyoracle.ns_params(dt=case.dt)
    nu = np.full(case.n_nodes, case.nu)
    reps, t_used = 0, 0.0
    while t_used < budget_s / 2 and reps < 5:
        t0 = time.perf_counter()
        o.ns_assemble(pyoracle.TETRA3D, prm, case.coords, case.conn, case.u0, case.p0, case.u0, nu, start, cols,
                      slots, nthreads=threads)
        t_used += time.perf_counter() - t0
        reps += 1
    e1 = case.n_elems // 8
    t1 = time.perf_counter()
    o.ns_assemble(pyoracle.TETRA3D, prm, case.coords, case.conn[:e1], case.u0, case.p0, case.u0, nu, start, cols,
                  slots[:e1 * 16], nthreads=1)
    single = e1 / (time.perf_counter() - t1) / 1e6
    return {"value": case.n_elems * reps / t_used / 1e6, "unit": "Melem/s", "cores": threads, "kind": "port",
            "cpu_model": model,
            "sample": "%d full FP64 assemblies of the same %d-element mesh (oracle restatement, element colouring, "
                      "OpenMP on %d threads); one thread: the first %d elements" % (reps, case.n_elems, threads, e1),
            "single_thread_value": single}


def cpu_steps(threads, gpu_iterations=None):
    """The oracle's implicit step (assembly, Dirichlet conditions, Krylov solve, all FP64) of the 32x32
    triangle cavity (GMRES(30) + ILU0 to 1e-10, one thread: the ILU0 solves are sequential) and of the
    64^3 tetrahedral cube (BiCGStab + Jacobi to 1e-10, assembly and Krylov kernels on `threads`)."""
    from coolfluid3_amd import cases
    from oracle import pyoracle
    o = pyoracle.get()
    out = {}
    for key, case, nthr in (("cavity_2d_32", cases.lid_cavity_2d(32), 1), ("cube_tets_64", cases.inflow_cube_tets(64),
                                                                              threads)):
        et = pyoracle.TRIAG2D if case.dim == 2 else pyoracle.TETRA3D
        neq = case.dim + 1
        start, cols, slots = o.build_sparsity(case.conn, case.n_nodes)
        prm = pyoracle.ns_params(dt=case.dt)
        nu = np.full(case.n_nodes, case.nu)
        o.set_krylov_threads(nthr)
        t0 = time.perf_counter()
        vals, rhs = o.ns_assemble(et, prm, case.coords, case.conn, case.u0, case.p0, case.u0, nu, start, cols, slots,
                                  nthreads=nthr)
        t1 = time.perf_counter()
        n, e, v = cases.bc_lists(case, case.u0, case.p0)
        sol = np.zeros_like(rhs)
        o.dirichlet(neq, start, cols, vals, rhs, sol, n, e, v)
        t2 = time.perf_counter()
        if case.dim == 2:
            x, st = o.gmres_ilu0(neq, start, cols, vals, rhs, sol, restart=30, tol=1e-10)
            method = "GMRES(30) + ILU0"
        else:
            x, st = o.bicgstab(neq, start, cols, vals, rhs, sol, tol=1e-10, maxit=3000)
            method = "BiCGStab + Jacobi"
        t3 = time.perf_counter()
        o.set_krylov_threads(1)
        out[key] = {"step_ms": (t3 - t0) * 1e3, "assembly_ms": (t1 - t0) * 1e3, "dirichlet_ms": (t2 - t1) * 1e3,
                    "solve_ms": (t3 - t2) * 1e3, "method": method, "iterations": int(st.iterations),
                    "rel_residual": float(st.rel_residual), "converged": bool(st.converged), "threads": nthr,
                    "elements": case.n_elems, "dofs": neq * case.n_nodes}
        print("bench: cpu %s step: %.1f ms (%s %d it, %d threads)" % (key, out[key]["step_ms"], method,
                                                                     st.iterations, nthr), file=sys.stderr, flush=True)
    return out


TRAFFIC_RE = r"traffic_per_launch_bytes\s+(\S+)\s*=\s*([0-9]+)"


def profiled_traffic(kernel_key, profile="assembly_tet128.md"):
    """HBM bytes per launch of a kernel from the committed rocprofv3 counter summary (FETCH_SIZE and
    WRITE_SIZE scaled by a copy kernel of known size), or None."""
    import re
    path = os.path.join(ROOT, "profiles", profile)
    try:
        with open(path) as f:
            for m in re.finditer(TRAFFIC_RE, f.read()):
                if m.group(1) == kernel_key:
                    return int(m.group(2))
    except OSError:
        pass
    return None


def timed_step(ctx, gc, sp, picard=False, label=""):
    ctx.synchronize()
    t0 = time.perf_counter()
    st = gc.picard_step(sp, raise_on_failure=False) if picard else gc.step(sp, raise_on_failure=False)
    ctx.synchronize()
    ms = (time.perf_counter() - t0) * 1e3
    if label:
        print("bench: %s: %.1f ms, %d iterations, residual %.2e%s" % (label, ms, st.iterations, st.rel_residual,
                                                                    "" if st.converged else ", NOT converged"),
              file=sys.stderr, flush=True)
    return {"ms": ms, "iterations": st.iterations, "outer_steps": st.outer, "spmvs": st.spmvs,
            "rel_residual": st.rel_residual, "converged": bool(st.converged)}


def assembly_rates(ctx, gc, reps, k, dim):
    """HIP-event kernel times of the FP64 and FP32 assembly on a resident case."""
    from coolfluid3_amd import cabi
    s = gc.system
    out = {}
    for label, prec, vb in (("fp64", cabi.FP64, 8), ("fp32", cabi.FP32, 4)):
        gc.assemble(prec)
        ctx.synchronize()
        ctx.enable_timing(True)
        for _ in range(reps):
            gc.assemble(prec)
        ctx.synchronize()
        ms, n = ctx.kernel_time(cabi.KERNEL_ASSEMBLY)
        kms = ms / max(n, 1)
        alg = algorithmic_assembly_bytes(gc.case.n_nodes, gc.case.n_elems, k, dim, s.n_blocks, vb)
        out[label] = {"kernel_ms": kms, "Melem_s": gc.case.n_elems / kms / 1e3,
                      "GBps": alg / (kms * 1e-3) / 1e9, "hbm_frac": alg / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    ctx.enable_timing(False)
    return out


def precision_workload(ctx, cells, steps=4):
    """Precision form on the tetrahedral cube channel (inflow x = 1, outlet p = 0 on x = 0) with ~1%
    elements of aspect ratio ~20 (nu = 5e-4), `steps` implicit steps: FP32 element matrices, BiCGStab on
    the block-column scaled FP32 operator fl32(A32 D^-1) (D the 4x4 diagonal blocks, x = D^-1 y) with FP64
    matrix-free residual replacement, against the FP64 assembly and BiCGStab + point Jacobi (the drop-in
    solve), both to a 1e-8 relative residual of the FP64 system.  Step 1 includes one-time allocations
    of both paths and is left out of the summed speedup."""
    from coolfluid3_amd import cabi, cases
    case = cases.inflow_cube_tets(cells, nu=5e-4, dt=5e-3, skew=True, outlet=True)
    res = {"workload": "unit cube %d^3 hexes x 6 P1 tets (%d elements, %d dofs), 1%% skewed (aspect ratio ~20), "
                       "channel inflow x=1 / outlet p=0 x=0, nu=5e-4 dt=5e-3, %d steps" % (cells, case.n_elems,
                                                                                         4 * case.n_nodes, steps)}
    runs = {}
    for label, prec in (("fp64", cabi.FP64), ("fp32_refined", cabi.FP32)):
        gc = cases.GpuCase(ctx, case, precision=prec)
        sp = cabi.solver_params(tol=1e-8, precision=prec, max_iter=5000)
        runs[label] = [timed_step(ctx, gc, sp, label="precision %s step %d" % (label, k)) for k in range(steps)]
        runs[label + "_u"] = gc.u.copy()
        if label == "fp32_refined":
            res["fp64_fallback_steps"] = gc.fallbacks
            res["assembly"] = assembly_rates(ctx, gc, 5, 4, 3)
        gc.close()
    f64, f32 = runs["fp64"], runs["fp32_refined"]
    res["fp64_steps"] = f64
    res["fp32_refined_steps"] = f32
    res["speedup_per_step"] = [a["ms"] / b["ms"] for a, b in zip(f64, f32)]
    res["speedup_steps_2_to_%d" % steps] = sum(a["ms"] for a in f64[1:]) / sum(b["ms"] for b in f32[1:])
    res["ms_per_iteration"] = {"fp64": sum(a["ms"] for a in f64[1:]) / sum(a["iterations"] for a in f64[1:]),
                               "fp32_refined": sum(b["ms"] for b in f32[1:]) / sum(b["iterations"] for b in f32[1:])}
    u64, u32 = runs["fp64_u"], runs["fp32_refined_u"]
    res["velocity_rel_error_vs_fp64"] = {"inf": float(np.abs(u32 - u64).max() / np.abs(u64).max()),
                                         "l2": float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64))}
    return res


HEX_KFLOP = 34.0   # algorithmic FP64 kflop per Q1 element (8 Gauss points, element-level counting)


def hexa_workload(ctx, cells, picard):
    """Q1 hexahedral lid-driven cavity (nu = 2e-3, 5% jitter): assembly rates, then Picard iterations
    (u_adv := u, reassembly, solve to 1e-8) in FP64 and in the precision form."""
    from coolfluid3_amd import cabi, cases
    case = cases.lid_cavity_cube_hexas(cells)
    res = {"workload": "unit cube %d^3 Q1 hexahedra (%d elements, %d dofs), lid-driven cavity, nu=2e-3, "
                       "Picard step dt=5e-2" % (cells, case.n_elems, 4 * case.n_nodes)}
    runs = {}
    for label, prec in (("fp64", cabi.FP64), ("fp32_refined", cabi.FP32)):
        gc = cases.GpuCase(ctx, case, precision=prec, use_specializations=False)
        if label == "fp32_refined":
            res["assembly"] = assembly_rates(ctx, gc, 3, 8, 3)
            k64 = res["assembly"]["fp64"]["kernel_ms"]
            res["assembly"]["fp64"]["fp64_alu_frac"] = case.n_elems * HEX_KFLOP * 1e3 / (k64 * 1e-3) / 78.6e12
        sp = cabi.solver_params(tol=1e-8, precision=prec, max_iter=5000)
        runs[label] = [timed_step(ctx, gc, sp, picard=True, label="hexa %s Picard %d" % (label, k))
                       for k in range(picard)]
        runs[label + "_u"] = gc.u.copy()
        if label == "fp32_refined":
            res["fp64_fallback_iterations"] = gc.fallbacks
        gc.close()
    if picard == 0:
        return res
    f64, f32 = runs["fp64"], runs["fp32_refined"]
    res["fp64_picard"] = f64
    res["fp32_refined_picard"] = f32
    res["speedup_per_iteration"] = [a["ms"] / b["ms"] for a, b in zip(f64, f32)]
    res["speedup_total"] = sum(a["ms"] for a in f64) / sum(b["ms"] for b in f32)
    u64, u32 = runs["fp64_u"], runs["fp32_refined_u"]
    res["velocity_rel_error_vs_fp64"] = {"inf": float(np.abs(u32 - u64).max() / max(np.abs(u64).max(), 1e-300)),
                                         "l2": float(np.linalg.norm(u32 - u64) / max(np.linalg.norm(u64), 1e-300))}
    return res


def cavity2d_workload(ctx):
    """32x32 triangle cavity: one implicit FP64 step, GMRES(30) + ILU0 to 1e-10."""
    from coolfluid3_amd import cabi, cases
    case = cases.lid_cavity_2d(32)
    gc = cases.GpuCase(ctx, case)
    steps = [timed_step(ctx, gc, None) for _ in range(2)]
    gc.close()
    return {"workload": "unit square 32x32 x 2 P1 triangles (2048 elements, 3267 dofs), nu=0.01 dt=0.01",
            "steps": steps}


def taylor_green_workload(ctx, element):
    """The reference's Taylor-Green acceptance run (plugins/UFEM/test/atest-ufem-navier-stokes-taylor-green.py
    with --dt=0.004 --segs=32 --theta=0.5 --tsteps=5): periodic square, generic expression, GMRES(30) +
    ILU0 to 1e-12 per step; the maximum errors against the exact vortex after each step and the
    reference's bounds after the fifth (u, v < 1.4e-3; p < 3.3e-3)."""
    from coolfluid3_amd import cabi, cases
    case = cases.taylor_green_case(32, element=element)
    gc = cases.GpuCase(ctx, case, use_specializations=False)
    sp = cabi.solver_params(cabi.GMRES_ILU0, tol=1e-12, restart=30)
    steps = []
    for _ in range(5):
        st = timed_step(ctx, gc, sp)
        ue, pe = cases.taylor_green_exact(case.coords, case.t)
        st["max_error"] = {"u": float(np.abs(gc.u[:, 0] - ue[:, 0]).max()),
                           "v": float(np.abs(gc.u[:, 1] - ue[:, 1]).max()), "p": float(np.abs(gc.p - pe).max())}
        steps.append(st)
    gc.close()
    last = steps[-1]["max_error"]
    return {"workload": "Taylor-Green vortex, [-0.5,0.5]^2 periodic, 32 segments (%s), nu=1e-3 dt=0.004 theta=0.5, "
                        "5 steps" % ("1024 Q1 quadrilaterals" if element == "quad" else "2048 P1 triangles"),
            "steps": steps, "within_reference_bounds": bool(max(last["u"], last["v"]) < 1.4e-3 and last["p"] < 3.3e-3)}


def semi_implicit_workload(ctx, scale_segments=512, scale_steps=2, with_cpu=True):
    """The semi-implicit half of the Taylor-Green script (NavierStokesSemiImplicit, two inner passes,
    BiCGStab + Jacobi to 1e-13 for the velocity and pressure systems): errors after each step at 32
    segments, then the same step on scale_segments^2 quadrilaterals with the time of the velocity-matrix
    assembly (wall time of the call, and the kernel time of the fused row-owner kernel) and of one step."""
    from coolfluid3_amd import cabi, cases
    sp = cabi.solver_params(cabi.BICGSTAB_JACOBI, tol=1e-13, max_iter=20000)

    def setup(segments, element):
        case = cases.taylor_green_case(segments, element=element)
        mesh = cabi.Mesh(ctx, case.etype, case.coords, case.conn)
        us = cabi.System(ctx, mesh, neq=2, periodic=case.periodic)
        ps = cabi.System(ctx, mesh, neq=1, periodic=case.periodic)
        f = cabi.Fields(mesh)
        f.set(cabi.FIELD_U, case.u0)
        f.set(cabi.FIELD_P, case.p0)
        f.initialize(case.nu)
        prm = cabi.ns_params(dt=case.dt, theta=case.theta)
        si = cabi.SemiImplicit(us, ps)
        si.initial(f, prm)
        return case, (mesh, us, ps, f, si), prm

    def close(objs):
        mesh, us, ps, f, si = objs
        si.close()
        for x in (f, ps, us, mesh):
            x.close()

    def step(case, objs, prm):
        f, si = objs[3], objs[4]
        center = case.bc_regions[0][1]
        pv = cases.taylor_green_exact(case.coords[center], case.t + case.dt)[1]
        ctx.synchronize()
        t0 = time.perf_counter()
        stats = si.step(f, prm, sp, sp, pbc=(center, pv))
        ms = (time.perf_counter() - t0) * 1e3
        case.t += case.dt
        return ms, [int(s.iterations) for s in stats]

    out = {}
    for element in ("quad", "triag"):
        case, objs, prm = setup(32, element)
        steps = []
        for _ in range(5):
            ms, its = step(case, objs, prm)
            ue, pe = cases.taylor_green_exact(case.coords, case.t)
            u, p = objs[3].get(cabi.FIELD_U), objs[3].get(cabi.FIELD_P)
            steps.append({"ms": ms, "iterations_u_p_per_pass": its,
                          "max_error": {"u": float(np.abs(u[:, 0] - ue[:, 0]).max()),
                                        "v": float(np.abs(u[:, 1] - ue[:, 1]).max()),
                                        "p": float(np.abs(p - pe).max())}})
        close(objs)
        last = steps[-1]["max_error"]
        out[element] = {"steps": steps, "within_reference_bounds":
                        bool(max(last["u"], last["v"]) < 1.4e-3 and last["p"] < 3.3e-3)}
    case, objs, prm = setup(scale_segments, "quad")
    mesh, us, ps, f, si = objs
    steps = [step(case, objs, prm) for _ in range(scale_steps)]
    reps = 10
    si.velocity_assemble(f, prm)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        si.velocity_assemble(f, prm)
    asm_ms = (time.perf_counter() - t0) * 1e3 / reps
    # kernel time of the fused row-owner velocity assembly (HIP events inside the library)
    ctx.enable_timing(True)
    for _ in range(reps):
        si.velocity_assemble(f, prm)
    ctx.synchronize()
    kms, kn = ctx.kernel_time(cabi.KERNEL_ASSEMBLY)
    ctx.enable_timing(False)
    kms = kms / kn if kn else None
    E, N = case.conn.shape[0], case.n_nodes
    # compulsory bytes: coordinates, advection velocity, viscosity, connectivity, 2x2 blocks written
    alg = N * (2 * 8 + 2 * 8 + 8) + E * 4 * 4 + us.n_blocks * 4 * 8
    p2m = us.p2m()
    close(objs)
    cpu = None
    if with_cpu:
        # the oracle's velocity-matrix assembly on the same mesh (one thread, element order)
        from oracle import pyoracle
        from oracle.semi_implicit import SemiImplicit as OracleSI
        osi = OracleSI(pyoracle.get(), case.etype, case.coords, case.conn, p2m, case.nu, case.dt, case.theta, 2,
                       case.u0, case.p0)
        t0 = time.perf_counter()
        osi.velocity_matrix()
        cpu_ms = (time.perf_counter() - t0) * 1e3
        cpu = {"value": E / cpu_ms / 1e3, "unit": "Melem/s", "cores": 1, "kind": "port",
               "sample": "one velocity-matrix assembly of the same mesh (oracle restatement, one thread)"}
    out["scale"] = {"workload": "Taylor-Green semi-implicit step, %d^2 periodic Q1 quadrilaterals (%d elements, "
                                "%d velocity rows)" % (scale_segments, E, us.n_nodes),
                    "step_ms": [s[0] for s in steps], "iterations_u_p_per_pass": [s[1] for s in steps],
                    "velocity_assembly_ms": asm_ms, "velocity_assembly_melem_s": E / asm_ms / 1e3,
                    "velocity_assembly_algorithmic_gb_s": alg / asm_ms / 1e6,
                    "velocity_assembly_hbm_frac": alg / asm_ms / 1e6 / 8000.0,
                    "velocity_assembly_kernel_ms": kms,
                    "velocity_assembly_kernel_hbm_frac": alg / kms / 1e6 / 8000.0 if kms else None,
                    "velocity_assembly_cpu_baseline": cpu}
    return out


def serving_workload(count, max_active, mode="both"):
    """Continuous batching against one-request-at-a-time serving on the same request stream
    (mode: both, sequential or batched)."""
    from coolfluid3_amd import serving
    reqs = serving.make_requests(count)
    srv = serving.Server(max_active=max_active, log=lambda m: print("serving: " + m, file=sys.stderr, flush=True))
    out = {"workload": "%d requests, Poisson arrivals 200 req/s, cube tet channels with cells/side log-uniform in "
                       "[12,64], Re U[10,1000], inflow amplitude U[0,0.1], 10 steps dt=1e-2 each, FP32 element "
                       "matrices with FP64 residual to 1e-8, HBM budget 200 GB" % count}
    if mode in ("both", "sequential"):
        out["sequential"] = srv.run(reqs, batched=False)[1]
    if mode in ("both", "batched"):
        out["continuous_batching"] = srv.run(reqs, batched=True)[1]
    if mode == "both":
        out["throughput_ratio"] = out["continuous_batching"]["throughput_req_s"] / out["sequential"]["throughput_req_s"]
    return out


def dropin_workload(ctx, cells=64, reps=10, solve=True):
    """The FP64 drop-in on the inflow cube (64^3 x 6 tetrahedra): assembly kernel rate, SpMV bandwidth
    and the implicit step (FP64 assembly, conditions, BiCGStab/Jacobi to 1e-10)."""
    from coolfluid3_amd import cabi, cases
    case = cases.inflow_cube_tets(cells)
    g = cases.GpuCase(ctx, case)
    g.fields.linearize()
    s = g.system
    res = {"workload": "unit cube %d^3 hexes x 6 P1 tets (%d elements, %d dofs), FP64 drop-in, nu=1e-3 dt=5e-3, "
                       "no-slip walls and parabolic inflow" % (cells, case.n_elems, s.dofs)}
    g.assemble(cabi.FP64)
    ctx.synchronize()
    ctx.enable_timing(True)
    for _ in range(reps):
        g.assemble(cabi.FP64)
    ctx.synchronize()
    ms, n = ctx.kernel_time(cabi.KERNEL_ASSEMBLY)
    kms = ms / max(n, 1)
    alg = algorithmic_assembly_bytes(case.n_nodes, case.n_elems, 4, 3, s.n_blocks, 8)
    res["assembly_fp64"] = {"kernel_ms": kms, "Melem_s": case.n_elems / kms / 1e3, "GBps": alg / kms / 1e6,
                            "hbm_frac": alg / kms / 1e6 / HBM_PEAK_GBS, "algorithmic_bytes": alg}
    ctx.enable_timing(True)
    for _ in range(reps):
        s.apply_device(1.0, s.vector_ptr(cabi.SOLUTION), 0.0, s.vector_ptr(cabi.RHS))
    ctx.synchronize()
    ms, n = ctx.kernel_time(cabi.KERNEL_SPMV)
    sms = ms / max(n, 1)
    sb = algorithmic_spmv_bytes(case.n_nodes, 3, s.n_blocks, 8)
    res["spmv_fp64"] = {"kernel_ms": sms, "GBps": sb / sms / 1e6, "hbm_frac": sb / sms / 1e6 / HBM_PEAK_GBS}
    ctx.enable_timing(False)
    g.close()
    if solve:
        gc = cases.GpuCase(ctx, case)
        res["step_fp64"] = timed_step(ctx, gc, cabi.solver_params(tol=1e-10), label="drop-in 64^3 FP64 step")
        gc.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cells", type=int, default=128, help="cube cells per side of the headline workload "
                                                           "(128 = the precision-form tetrahedral cube)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-workloads", action="store_true", help="skip the 64^3 / 128^3-steps / 160^3 / 2D / serving "
                                                               "workloads")
    ap.add_argument("--workloads", default="dropin_64,cavity_2d,taylor_green,semi_implicit,precision_tets,hexa_cavity,"
                                           "serving", help="comma-separated subset of the extra workloads")
    ap.add_argument("--precision-cells", type=int, default=128)
    ap.add_argument("--hexa-cells", type=int, default=160)
    ap.add_argument("--picard", type=int, default=2)
    ap.add_argument("--serving-requests", type=int, default=512)
    ap.add_argument("--serving-slots", type=int, default=16)
    ap.add_argument("--serving-mode", default="both", choices=("both", "sequential", "batched"))
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this benchmark measures one GPU only; multi-GPU runs are not part of it (use --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    from coolfluid3_amd import cabi, cases

    # the precision-form tetrahedral cube: 1% skewed elements, channel inflow x = 1 / outlet x = 0
    case = cases.inflow_cube_tets(args.cells, nu=5e-4, dt=5e-3, skew=True, outlet=True)
    ctx = cabi.Context(0)
    g = cases.GpuCase(ctx, case, precision=cabi.FP32)
    g.fields.linearize()
    s = g.system
    E, N, nb = case.n_elems, case.n_nodes, s.n_blocks
    print("bench: headline workload %d elements, %d dofs, %d node blocks" % (E, s.dofs, nb), file=sys.stderr,
          flush=True)
    ctx.set_async(True)
    for _ in range(args.warmup):
        g.assemble(cabi.FP64)
    ctx.synchronize()
    ctx.enable_timing(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.assemble(cabi.FP64)
    ctx.synchronize()
    t1 = time.perf_counter()
    asm_ms, asm_n = ctx.kernel_time(cabi.KERNEL_ASSEMBLY)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    kern_ms = asm_ms / max(asm_n, 1)
    alg = algorithmic_assembly_bytes(N, E, 4, 3, nb, 8)
    achieved = alg / (kern_ms * 1e-3) / 1e9

    # FP32 assembly of the precision form
    ctx.enable_timing(True)
    for _ in range(args.steps):
        g.assemble(cabi.FP32)
    ctx.synchronize()
    a32_ms, a32_n = ctx.kernel_time(cabi.KERNEL_ASSEMBLY)
    a32_kern = a32_ms / max(a32_n, 1)
    alg32 = algorithmic_assembly_bytes(N, E, 4, 3, nb, 4)

    # SpMV bandwidth (FP64 values) on the same system
    ctx.enable_timing(True)
    x = s.vector_ptr(cabi.SOLUTION)
    y = s.vector_ptr(cabi.RHS)
    for _ in range(args.steps):
        s.apply_device(1.0, x, 0.0, y)
    ctx.synchronize()
    sp_ms, sp_n = ctx.kernel_time(cabi.KERNEL_SPMV)
    sp_kern = sp_ms / max(sp_n, 1)
    sp_bytes = algorithmic_spmv_bytes(N, 3, nb, 8)
    # plain copies of a known size (dofs doubles read + written): the k_copy kernel calibrates the
    # FETCH_SIZE / WRITE_SIZE counters of the profiled runs
    for _ in range(3):
        s.assign(cabi.RHS, cabi.SOLUTION)
    ctx.synchronize()
    ctx.set_async(False)
    ctx.enable_timing(False)
    g.close()

    extra = {
        "elements": E, "nodes": N, "dofs": s.dofs, "node_blocks": nb,
        "assembly_kernel_ms": kern_ms, "assembly_fp32_kernel_ms": a32_kern,
        "assembly_fp32_GBps": alg32 / (a32_kern * 1e-3) / 1e9,
        "assembly_fp32_hbm_frac": alg32 / (a32_kern * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "spmv_fp64_ms": sp_kern, "spmv_fp64_GBps": sp_bytes / (sp_kern * 1e-3) / 1e9,
        "dofs_per_s_assembly": s.dofs / (ms_per_step * 1e-3),
    }
    traffic = profiled_traffic("k_assemble_fp64_tet%d" % args.cells)
    out = {
        "metric": METRIC, "value": E / (ms_per_step * 1e-3) / 1e6, "unit": "Melem/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "UFEM NS assembly+scatter, unit cube %d^3 hexes x 6 P1 tets, 1%% skewed (aspect "
                               "ratio ~20) (%d elements, %d dofs), FP64 values, channel inflow x=1 / outlet x=0, "
                               "nu=5e-4 dt=5e-3, seed 20261202" % (args.cells, E, s.dofs),
                   "elements": E, "element_type": "P1 tetrahedron (specialised kernel)"},
        "fp32_assembly_Melem_s": E / a32_kern / 1e3,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_assemble<SpecP1<3,double>>", "algorithmic_bytes": alg,
                     "kernel_ms": kern_ms},
        "extra": extra,
    }
    threads, model, ncpu = host_cpu()
    if not args.no_workloads:
        wl = args.workloads.split(",")
        for key, fn in (("dropin_64", lambda: dropin_workload(ctx, solve=not args.no_solve)),
                        ("precision_tets", lambda: precision_workload(ctx, args.precision_cells)),
                        ("cavity_2d", lambda: cavity2d_workload(ctx)),
                        ("taylor_green", lambda: {e: taylor_green_workload(ctx, e) for e in ("quad", "triag")}),
                        ("semi_implicit", lambda: semi_implicit_workload(ctx, with_cpu=not args.no_cpu_baseline)),
                        ("hexa_cavity", lambda: hexa_workload(ctx, args.hexa_cells, args.picard)),
                        ("serving", lambda: serving_workload(args.serving_requests, args.serving_slots,
                                                             args.serving_mode))):
            if key not in wl:
                continue
            t0 = time.perf_counter()
            extra[key] = fn()
            extra[key]["wall_s"] = time.perf_counter() - t0
            print("bench: %s done in %.1f s: %s" % (key, extra[key]["wall_s"], json.dumps(extra[key])),
                  file=sys.stderr, flush=True)
        pt = extra.get("precision_tets")
        if pt:
            f64, f32 = pt["fp64_steps"], pt["fp32_refined_steps"]
            out["time_to_1e-8"] = {
                "workload": pt["workload"],
                "fp64_ms_per_step": [st["ms"] for st in f64], "fp32_refined_ms_per_step": [st["ms"] for st in f32],
                "fp64_iterations": [st["iterations"] for st in f64],
                "fp32_refined_iterations": [st["iterations"] for st in f32],
                "speedup_steps_2_on": pt["speedup_steps_2_to_%d" % len(f64)],
                "fp32_vs_fp64_velocity_rel_error": pt["velocity_rel_error_vs_fp64"]}
        sv = extra.get("serving")
        if sv and "continuous_batching" in sv:
            out["serving"] = {k: sv["continuous_batching"][k] for k in ("p50_latency_s", "p99_latency_s",
                                                                          "throughput_req_s")}
            if "throughput_ratio" in sv:
                out["serving"]["throughput_ratio_vs_sequential"] = sv["throughput_ratio"]
    ctx.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(case, threads, model)
        out["cpu_baseline"]["logical_cpus_of_the_machine"] = ncpu
        extra["cpu_steps"] = cpu_steps(threads)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

"""Parity at the full sizes of the benchmark configurations, against the oracle on the same seeded inputs.

  * config 2: 64^3 hexahedra x 6 P1 tetrahedra (1,572,864 elements, 1,098,500 dofs), FP64 drop-in:
    sparsity and element-to-slot map bit-exact, matrix and rhs to 1e-12 (row-relative), the
    Dirichlet-constrained system to 1e-12, and the BiCGStab/Jacobi step to a 1e-10 relative residual
    evaluated with the oracle's FP64 matrix;
  * config 3: 128^3 x 6 tetrahedra (12,582,912 elements, 8,586,756 dofs), 1% skewed (aspect ratio ~20):
    sparsity and slots bit-exact, FP64 matrix and rhs to 1e-12, FP32 element matrices to 1e-6
    (single precision, error printed), and the precision-form step (FP32 element matrices, FP64
    matrix-free residual) reaching 1e-8 of the oracle's FP64 system;
  * config 5: 160^3 Q1 hexahedra (4,096,000 elements, 16,693,124 dofs): sparsity and slots bit-exact,
    FP64 matrix and rhs to 1e-12 and FP32 to 1e-6 (error printed) over every row, then two Picard
    iterations of the lid-driven cavity in the precision form, each reaching 1e-8 of the FP64 system
    that the oracle assembles from the device state of that iteration.

The oracle runs on up to 16 host threads (element colouring), so these tests take tens of seconds each.
"""
import os

import numpy as np
import pytest

from coolfluid3_amd import cabi, cases
from oracle import pyoracle

pytestmark = pytest.mark.gpu

NT = max(1, min(16, os.cpu_count() or 1))
ETYPE = {cabi.TRIAG2D: pyoracle.TRIAG2D, cabi.TETRA3D: pyoracle.TETRA3D, cabi.HEXA3D: pyoracle.HEXA3D,
         cabi.QUAD2D: pyoracle.QUAD2D}


@pytest.fixture(scope="module")
def ctx():
    c = cabi.Context(0)
    yield c
    c.close()


def row_rel_err(gpu, ref, neq, start, rows_per_chunk=1 << 18):
    """max over block rows and equations of max|gpu - ref| / max|ref| along the scalar row, in chunks of
    block rows (the blocks of node row i are start[i] .. start[i+1], every row holds its diagonal)."""
    nn = neq * neq
    n = len(start) - 1
    worst = 0.0
    for i0 in range(0, n, rows_per_chunk):
        i1 = min(n, i0 + rows_per_chunk)
        b0, b1 = int(start[i0]), int(start[i1])
        g = gpu[b0 * nn:b1 * nn].reshape(-1, neq, neq)
        r = ref[b0 * nn:b1 * nn].reshape(-1, neq, neq)
        if g.dtype != np.float64:
            g = g.astype(np.float64)
        d = np.abs(g - r).max(axis=2)
        a = np.abs(r).max(axis=2)
        seg = (start[i0:i1] - b0).astype(np.int64)
        err = np.maximum.reduceat(d, seg, axis=0)
        rowmax = np.maximum.reduceat(a, seg, axis=0)
        worst = max(worst, float((err / np.maximum(rowmax, 1e-300)).max()))
    return worst


def oracle_sparsity_check(oracle, g, case):
    start, cols, slots = oracle.build_sparsity(case.conn, case.n_nodes)
    s, c = g.system.graph()
    assert np.array_equal(s, start), "starting_indices differ from the oracle"
    assert np.array_equal(c, cols), "node_connectivity differs from the oracle"
    assert np.array_equal(g.system.slots(), slots), "element-to-slot map differs from the oracle"
    return start, cols, slots


def oracle_assemble(oracle, case, start, cols, slots, u, p, ua, spec=True):
    prm = pyoracle.ns_params(rho=case.rho, dt=case.dt, theta=case.theta, use_specializations=spec)
    return oracle.ns_assemble(ETYPE[case.etype], prm, case.coords, case.conn, u, p, ua,
                              np.full(case.n_nodes, case.nu), start, cols, slots, nthreads=NT)


def rel_residual(oracle, neq, start, cols, vals, b, x):
    r = b - oracle.spmv(neq, start, cols, vals, x)
    return float(np.linalg.norm(r) / np.linalg.norm(b))


def test_config2_fp64_step(ctx, oracle):
    case = cases.inflow_cube_tets(64)
    print("config 2:", case.n_elems, "elements", 4 * case.n_nodes, "dofs", flush=True)
    assert case.n_elems == 1572864 and 4 * case.n_nodes == 1098500
    g = cases.GpuCase(ctx, case)
    start, cols, slots = oracle_sparsity_check(oracle, g, case)
    assert int(start[-1]) == 4018753
    vref, rref = oracle_assemble(oracle, case, start, cols, slots, case.u0, case.p0, case.u0)
    g.fields.linearize()
    g.assemble(cabi.FP64)
    assert row_rel_err(g.system.values(), vref, 4, start) <= 1e-12
    r = g.system.vector(cabi.RHS)
    assert np.abs(r - rref).max() <= 1e-12 * np.abs(rref).max()
    # Dirichlet conditions of the inflow cube, in increment form, on both systems
    n, e, v = cases.bc_lists(case, case.u0, case.p0)
    sol = np.zeros_like(rref)
    oracle.dirichlet(4, start, cols, vref, rref, sol, n, e, v)
    g.system.dirichlet(n, e, v)
    assert row_rel_err(g.system.values(), vref, 4, start) <= 1e-12
    assert np.abs(g.system.vector(cabi.RHS) - rref).max() <= 1e-12 * np.abs(rref).max()
    assert np.array_equal(g.system.vector(cabi.SOLUTION), sol)
    st = g.system.solve(cabi.solver_params(tol=1e-10, max_iter=5000))
    x = g.system.vector(cabi.SOLUTION)
    res = rel_residual(oracle, 4, start, cols, vref, rref, x)
    print(f"config 2: BiCGStab/Jacobi {st.iterations} iterations, residual on the oracle's matrix {res:.3e}")
    assert st.converged and res <= 1e-10
    g.close()


def test_config3_assembly_and_precision_step(ctx, oracle):
    case = cases.inflow_cube_tets(128, nu=5e-4, dt=5e-3, skew=True, outlet=True)
    print("config 3:", case.n_elems, "elements", 4 * case.n_nodes, "dofs", flush=True)
    assert case.n_elems == 12582912 and 4 * case.n_nodes == 8586756
    g = cases.GpuCase(ctx, case, precision=cabi.FP32)
    start, cols, slots = oracle_sparsity_check(oracle, g, case)
    assert int(start[-1]) == 31802497
    vref, rref = oracle_assemble(oracle, case, start, cols, slots, case.u0, case.p0, case.u0)
    del slots
    g.fields.linearize()
    g.assemble(cabi.FP64)
    e64 = row_rel_err(g.system.values(), vref, 4, start)
    r = g.system.vector(cabi.RHS)
    assert e64 <= 1e-12
    assert np.abs(r - rref).max() <= 1e-12 * np.abs(rref).max()
    g.assemble(cabi.FP32)
    e32 = row_rel_err(g.system.values_fp32(), vref, 4, start)
    print(f"config 3: FP64 assembly row-relative error {e64:.2e}, FP32 element matrices {e32:.2e}", flush=True)
    assert e32 <= 1e-6
    g.close()
    # the precision-form step against the oracle's FP64 system (same state: first step, u_adv = u)
    n, e, v = cases.bc_lists(case, case.u0, case.p0)
    sol = np.zeros_like(rref)
    oracle.dirichlet(4, start, cols, vref, rref, sol, n, e, v)
    g32 = cases.GpuCase(ctx, case, precision=cabi.FP32)
    st = g32.step(cabi.solver_params(tol=1e-8, precision=cabi.FP32, max_iter=5000))
    dx = g32.system.vector(cabi.SOLUTION)
    g32.close()
    res = rel_residual(oracle, 4, start, cols, vref, rref, dx)
    g64 = cases.GpuCase(ctx, case)
    st64 = g64.step(cabi.solver_params(tol=1e-8, max_iter=5000))
    dx64 = g64.system.vector(cabi.SOLUTION)
    g64.close()
    res64 = rel_residual(oracle, 4, start, cols, vref, rref, dx64)
    err = np.abs(dx - dx64).max() / np.abs(dx64).max()
    print(f"config 3: precision form {st.iterations} it ({st.outer} replacements / inner solves), residual on the "
          f"oracle's FP64 system {res:.3e}; FP64 {st64.iterations} it, residual {res64:.3e}; increment difference "
          f"FP32 form vs FP64 {err:.2e} (inf-norm, relative)")
    assert g32.fallbacks == 0
    assert st.converged and res <= 1e-8 * 1.001
    assert st64.converged and res64 <= 1e-8 * 1.001


@pytest.fixture(scope="module")
def hexa160(oracle):
    case = cases.lid_cavity_cube_hexas(160)
    start, cols, slots = oracle.build_sparsity(case.conn, case.n_nodes)
    return case, start, cols, slots


def test_config5_sparsity_and_assembly(ctx, oracle, hexa160):
    case0, start, cols, slots = hexa160
    # the same mesh with a random velocity field, so every term of the generic expression is exercised
    case = cases.lid_cavity_cube_hexas(160, random_velocity=True)
    assert np.array_equal(case.coords, case0.coords)
    print("config 5:", case.n_elems, "elements", 4 * case.n_nodes, "dofs", flush=True)
    assert case.n_elems == 4096000 and 4 * case.n_nodes == 16693124
    g = cases.GpuCase(ctx, case, precision=cabi.FP32, use_specializations=False)
    s, c = g.system.graph()
    assert np.array_equal(s, start) and np.array_equal(c, cols)
    del s, c
    assert np.array_equal(g.system.slots(), slots)
    assert int(start[-1]) == 111284641
    vref, rref = oracle_assemble(oracle, case, start, cols, slots, case.u0, case.p0, case.u0, spec=False)
    g.fields.linearize()
    g.assemble(cabi.FP64)
    e64 = row_rel_err(g.system.values(), vref, 4, start)
    r = g.system.vector(cabi.RHS)
    assert np.abs(r - rref).max() <= 1e-12 * np.abs(rref).max()
    g.assemble(cabi.FP32)
    e32 = row_rel_err(g.system.values_fp32(), vref, 4, start)
    print(f"config 5: FP64 assembly row-relative error {e64:.2e}, FP32 element matrices {e32:.2e}", flush=True)
    assert e64 <= 1e-12 and e32 <= 1e-6
    g.close()


def test_config5_picard_precision_form(ctx, oracle, hexa160):
    """Two Picard iterations (u_adv := u, reassembly, solve to 1e-8) of the lid-driven cavity in the
    precision form; the solution increment of each is checked on the oracle's FP64 system assembled
    from the device state of that iteration."""
    case, start, cols, slots = hexa160
    g = cases.GpuCase(ctx, case, precision=cabi.FP32, use_specializations=False)
    s = g.system
    sp = cabi.solver_params(tol=1e-8, precision=cabi.FP32, max_iter=5000)
    prm = cabi.NSParams.from_buffer_copy(g.prm)
    prm.precision = cabi.FP32
    for it in range(2):
        g.fields.copy(cabi.FIELD_UADV, cabi.FIELD_U)
        g.assemble(cabi.FP32)
        n, e, v = cases.bc_lists(case, g.u, g.p)
        s.dirichlet(n, e, v)
        st = s.ns_solve(g.fields, prm, sp)
        dx = s.vector(cabi.SOLUTION)
        ua = g.fields.get(cabi.FIELD_UADV)
        vref, rref = oracle_assemble(oracle, case, start, cols, slots, g.u, g.p, ua, spec=False)
        sol = np.zeros_like(rref)
        oracle.dirichlet(4, start, cols, vref, rref, sol, n, e, v)
        res = rel_residual(oracle, 4, start, cols, vref, rref, dx)
        del vref
        print(f"config 5 Picard {it}: precision form {st.iterations} it ({st.outer} replacements / inner solves), "
              f"residual on the oracle's FP64 system {res:.3e}", flush=True)
        assert st.converged and res <= 1e-8 * 1.001
        s.update(g.fields)
        g.u = g.fields.get(cabi.FIELD_U)
        g.p = g.fields.get(cabi.FIELD_P)
    assert np.abs(g.u).max() > 0.5   # the lid has set the fluid moving
    g.close()

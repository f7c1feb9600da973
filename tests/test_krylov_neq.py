"""BiCGStab/Jacobi on systems with 1, 2 and 3 equations per node (the fused small-neq products), in FP64
and in the precision form on stored FP32 values, and GMRES(30)+ILU0 on 1 to 4 equations per node (the
single-workgroup solver), against the oracle.

Bars: the FP64 solution of the device solve matches the oracle's BiCGStab solution to 1e-8 (relative,
inf-norm) and its residual on the oracle's matrix is <= 1e-10; iteration counts agree within 10%
(the reductions sum in another order); the precision-form solve reaches 1e-8 of the FP64 system.
"""
import numpy as np
import pytest

from coolfluid3_amd import cabi, cases, meshgen
from oracle import pyoracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = cabi.Context(0)
    yield c
    c.close()


def p1_element_blocks(coords, conn, neq, seed):
    """Diagonally weighted, non-symmetric element matrices of P1 tetrahedra: a stiffness term, a mass
    term, a convection-like term and a coupling of the equations, node-major (k neq) x (k neq)."""
    rng = np.random.default_rng(seed)
    k = conn.shape[1]
    mats = np.zeros((conn.shape[0], k * neq, k * neq))
    c = rng.uniform(0.5, 1.5, (neq, neq)) * 0.1 + np.eye(neq)
    for e, nodes in enumerate(conn):
        X = coords[nodes]
        J = (X[1:] - X[0]).T
        G = np.linalg.inv(J).T @ np.hstack([-np.ones((3, 1)), np.eye(3)])
        V = abs(np.linalg.det(J)) / 6
        K = V * (G.T @ G) + V / 20 * (np.ones((k, k)) + np.eye(k)) * 50
        w = rng.uniform(-1, 1, 3)
        K += V * np.outer(np.full(k, 0.25), w @ G) * 2.0
        mats[e] = np.kron(K, c)
    return mats


def build_system(ctx, oracle, neq, n=8, seed=3):
    coords, conn = meshgen.box_tetras(n, n, n)
    N = coords.shape[0]
    mesh = cabi.Mesh(ctx, cabi.TETRA3D, coords, conn)
    s = cabi.System(ctx, mesh, neq=neq)
    start, cols, slots = oracle.build_sparsity(conn, N)
    mats = p1_element_blocks(coords, conn, neq, seed)
    vals = np.zeros(int(start[-1]) * neq * neq)
    k = conn.shape[1]
    nn = neq * neq
    s.reset(0.0)
    for e in range(conn.shape[0]):
        s.add_values(conn[e], mats[e])
        for a in range(k):
            for b in range(k):
                blk = mats[e][a * neq:(a + 1) * neq, b * neq:(b + 1) * neq]
                sl = int(slots[(e * k + a) * k + b])
                vals[sl * nn:(sl + 1) * nn] += blk.ravel()
    b = np.random.default_rng(seed + 1).uniform(-1, 1, N * neq)
    return mesh, s, start, cols, vals, b


@pytest.mark.parametrize("neq", [1, 2, 3])
def test_bicgstab_fp64_small_neq(ctx, oracle, neq):
    mesh, s, start, cols, vals, b = build_system(ctx, oracle, neq)
    print("neq", neq, "dofs", s.dofs, flush=True)
    assert np.abs(s.values() - vals).max() <= 1e-14 * np.abs(vals).max()
    s.set_vector(cabi.RHS, b)
    s.reset_vector(cabi.SOLUTION, 0.0)
    st = s.solve(cabi.solver_params(tol=1e-10, max_iter=2000))
    x = s.vector(cabi.SOLUTION)
    xo, sto = oracle.bicgstab(neq, start, cols, vals, b, np.zeros_like(b), tol=1e-10)
    res = np.linalg.norm(b - oracle.spmv(neq, start, cols, vals, x)) / np.linalg.norm(b)
    diff = np.abs(x - xo).max() / np.abs(xo).max()
    print(f"neq {neq}: gpu {st.iterations} it, oracle {sto.iterations} it, residual {res:.2e}, solution diff {diff:.2e}")
    assert st.converged and res <= 1e-10
    assert diff <= 1e-8
    assert abs(st.iterations - sto.iterations) <= max(2, sto.iterations // 10)
    s.close()
    mesh.close()


@pytest.mark.parametrize("neq", [1, 2, 3, 4])
def test_gmres_ilu0_small_neq(ctx, oracle, neq):
    """GMRES(30)+ILU0 on 1 to 4 equations per node (the single-workgroup solver, one kernel per block
    size): the solution matches the oracle's GMRES+ILU0 solution to 1e-8, its residual on the oracle's
    matrix is <= 1e-10, and the iteration counts agree within 10%."""
    mesh, s, start, cols, vals, b = build_system(ctx, oracle, neq)
    print("neq", neq, "dofs", s.dofs, flush=True)
    s.set_vector(cabi.RHS, b)
    s.reset_vector(cabi.SOLUTION, 0.0)
    st = s.solve(cabi.solver_params(cabi.GMRES_ILU0, tol=1e-10, restart=30, max_iter=2000))
    x = s.vector(cabi.SOLUTION)
    xo, sto = oracle.gmres_ilu0(neq, start, cols, vals, b, np.zeros_like(b), restart=30, tol=1e-10)
    res = np.linalg.norm(b - oracle.spmv(neq, start, cols, vals, x)) / np.linalg.norm(b)
    diff = np.abs(x - xo).max() / np.abs(xo).max()
    print(f"neq {neq}: gpu {st.iterations} it, oracle {sto.iterations} it, residual {res:.2e}, solution diff {diff:.2e}")
    assert st.converged and res <= 1e-10
    assert diff <= 1e-8
    assert abs(st.iterations - sto.iterations) <= max(2, sto.iterations // 10)
    s.close()
    mesh.close()


def oracle_system(o, case):
    start, cols, slots = o.build_sparsity(case.conn, case.n_nodes)
    prm = pyoracle.ns_params(rho=case.rho, dt=case.dt, theta=case.theta,
                             use_specializations=case.etype != cabi.QUAD2D)
    et = {cabi.TRIAG2D: pyoracle.TRIAG2D, cabi.QUAD2D: pyoracle.QUAD2D}[case.etype]
    vals, rhs = o.ns_assemble(et, prm, case.coords, case.conn, case.u0, case.p0, case.u0,
                              np.full(case.n_nodes, case.nu), start, cols, slots)
    return start, cols, vals, rhs


def quads_case(n=9, seed=11):
    rng = np.random.default_rng(seed)
    coords, conn = meshgen.rect_quads(n, n)
    inner = np.all((coords > 1e-12) & (coords < 1 - 1e-12), axis=1)
    coords[inner] += rng.uniform(-0.2, 0.2, (int(inner.sum()), 2)) / n
    N = coords.shape[0]
    dims = (n, n)
    zero = {0: 0.0, 1: 0.0}
    regions = [("left", cases._face_nodes(dims, 0, 0), zero), ("right", cases._face_nodes(dims, 0, 1), zero),
               ("bottom", cases._face_nodes(dims, 1, 0), zero),
               ("top", cases._face_nodes(dims, 1, 1), {0: 1.0, 1: 0.0}), ("pressure", np.array([0], np.uint32), {2: 0.0})]
    return cases.Case("quads_%d" % n, cabi.QUAD2D, coords, conn, dims, 1e-2, 1e-2,
                      u0=rng.uniform(-1.0, 1.0, (N, 2)), p0=np.zeros(N), bc_regions=regions)


@pytest.mark.parametrize("name", ["tri8", "quad9"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_bicgstab_neq3_navier_stokes(ctx, oracle, name, precision):
    """The 2D Navier-Stokes systems (3 equations per node): FP64 BiCGStab, and the precision form on the
    stored FP32 values with FP64 residual replacement, both checked on the oracle's FP64 system."""
    case = cases.lid_cavity_2d(8) if name == "tri8" else quads_case()
    start, cols, vref, rref = oracle_system(oracle, case)
    fp32 = precision == "fp32"
    g = cases.GpuCase(ctx, case, precision=cabi.FP32 if fp32 else cabi.FP64,
                      use_specializations=case.etype != cabi.QUAD2D)
    g.fields.linearize()
    if fp32:
        g.assemble(cabi.FP32)
    g.assemble(cabi.FP64)
    n, e, v = cases.bc_lists(case, case.u0, case.p0)
    g.system.dirichlet(n, e, v)
    sol = np.zeros_like(rref)
    oracle.dirichlet(3, start, cols, vref, rref, sol, n, e, v)
    tol = 1e-8 if fp32 else 1e-10
    st = g.system.solve(cabi.solver_params(tol=tol, max_iter=5000, precision=cabi.FP32 if fp32 else cabi.FP64))
    x = g.system.vector(cabi.SOLUTION)
    res = np.linalg.norm(rref - oracle.spmv(3, start, cols, vref, x)) / np.linalg.norm(rref)
    print(f"{name} {precision}: {st.iterations} it ({st.outer} outer), residual on the oracle's system {res:.2e}")
    assert st.converged and res <= tol * 1.001
    if not fp32:
        xo, sto = oracle.bicgstab(3, start, cols, vref, rref, sol, tol=1e-10)
        assert abs(st.iterations - sto.iterations) <= max(3, sto.iterations // 10)
    g.close()

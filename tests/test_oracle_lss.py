"""Pins the oracle's sparsity, scatter, symmetric Dirichlet, SpMV and Krylov solvers.

Reference data: test/math/utest-lss-test-matrix.hpp (per-rank element lists and node-graph CSR),
utest-lss-symmetric-dirichlet.cpp (exact entries), utest-lss-vector.cpp (apply on an all-ones matrix).
Mathematical pins: brute-force node-pair enumeration, a dense global assembly built here from the
oracle's element matrices, the exact solution of the Dirichlet-constrained dense system, dense direct
solves for the Krylov methods, and the closed-form node-block counts of the Kuhn meshes.
"""
import numpy as np
import pytest

from oracle.pyoracle import HEXA3D, TETRA3D, TRIAG2D, ns_params


def brute_graph(conn, n_nodes):
    k = conn.shape[1]
    a = np.repeat(conn.astype(np.int64), k, axis=1).ravel()
    b = np.tile(conn.astype(np.int64), (1, k)).ravel()
    keys = np.unique(a * n_nodes + b)
    rows, cols = keys // n_nodes, keys % n_nodes
    start = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=n_nodes))])
    return start, cols


def kuhn_blocks(n):
    N = (n + 1) ** 3
    edges = 3 * n * (n + 1) ** 2 + 3 * n * n * (n + 1) + n ** 3
    return N + 2 * edges


@pytest.mark.parametrize("mesh", ["tri", "tet", "hex"])
def test_sparsity_bruteforce_and_slots(oracle, mesh):
    if mesh == "tri":
        c, e = oracle.rect_tris(5, 4)
    elif mesh == "tet":
        c, e = oracle.box_tetras(3, 4, 2)
    else:
        c, e = oracle.box_hexas(3, 2, 4)
    start, cols, slots = oracle.build_sparsity(e, c.shape[0])
    s2, c2 = brute_graph(e, c.shape[0])
    assert np.array_equal(start, s2) and np.array_equal(cols, c2)
    k = e.shape[1]
    sl = slots.reshape(-1, k, k)
    for a in range(k):
        for b in range(k):
            assert np.array_equal(cols[sl[:, a, b]], e[:, b])
            assert np.all(sl[:, a, b] >= start[e[:, a]]) and np.all(sl[:, a, b] < start[e[:, a] + 1])


def test_sparsity_counts_match_config_table(oracle):
    c, t = oracle.rect_tris(32, 32)
    start, cols, _ = oracle.build_sparsity(t, c.shape[0], with_slots=False)
    assert (c.shape[0], t.shape[0], len(cols)) == (1089, 2048, 7361)
    for n in (2, 3, 5):
        c, t = oracle.box_tetras(n, n, n)
        _, cols, _ = oracle.build_sparsity(t, c.shape[0], with_slots=False)
        assert len(cols) == kuhn_blocks(n)
    assert kuhn_blocks(64) == 4018753 and kuhn_blocks(128) == 31802497
    c, h = oracle.box_hexas(3, 3, 3)
    _, cols, _ = oracle.build_sparsity(h, c.shape[0], with_slots=False)
    assert len(cols) == sum(min(i, 1) + 1 + min(3 - i, 1) for i in range(4)) ** 3


def test_sparsity_reference_fixture(oracle, fixtures):
    """Element lists of utest-lss-test-matrix.hpp reproduce the fixture's node-graph rows."""
    f = fixtures["lss_test_matrix"]
    for rk in f["ranks"]:
        en = np.array(rk["elem_nodes"], dtype=np.uint32).reshape(-1, 3)
        rs = np.array(rk["rowstart_positions"]); ci = np.array(rk["column_indices"])
        n = len(rs) - 1
        start, cols, _ = oracle.build_sparsity(en, n, with_slots=False)
        periodic = set(rk["periodic_pairs"])
        checked = 0
        for i in range(n):
            if i in periodic:       # periodic rows are merged with their partner (next row of the plan)
                continue
            assert sorted(ci[rs[i]:rs[i + 1]]) == cols[start[i]:start[i + 1]].tolist()
            checked += 1
        assert checked == n - len(periodic)


def dense_from_blocks(neq, start, cols, values):
    n = (len(start) - 1) * neq
    D = np.zeros((n, n))
    nn = neq * neq
    for i in range(len(start) - 1):
        for q in range(start[i], start[i + 1]):
            D[i * neq:(i + 1) * neq, cols[q] * neq:(cols[q] + 1) * neq] = values[q * nn:(q + 1) * nn].reshape(neq, neq)
    return D


def small_case(oracle, kind, seed=3):
    rng = np.random.default_rng(seed)
    if kind == "tri":
        c, e = oracle.rect_tris(4, 3); et, dim = TRIAG2D, 2
    elif kind == "tet":
        c, e = oracle.box_tetras(2, 2, 2); c = oracle.jitter(c, 3, 2, 2, 2, 0.1, seed); et, dim = TETRA3D, 3
    else:
        c, e = oracle.box_hexas(2, 2, 2); c = oracle.jitter(c, 3, 2, 2, 2, 0.05, seed); et, dim = HEXA3D, 3
    N = c.shape[0]
    u = rng.uniform(-1, 1, (N, dim)); ua = rng.uniform(-1, 1, (N, dim))
    p = rng.uniform(-1, 1, N); nu = np.full(N, 0.01)
    return et, dim, c, e, u, ua, p, nu


@pytest.mark.parametrize("kind", ["tri", "tet", "hex"])
@pytest.mark.parametrize("spec", [True, False])
def test_assembly_vs_dense_global(oracle, kind, spec):
    et, dim, c, e, u, ua, p, nu = small_case(oracle, kind)
    prm = ns_params(dt=0.01, theta=0.7, use_specializations=spec)
    start, cols, slots = oracle.build_sparsity(e, c.shape[0])
    values, rhs = oracle.ns_assemble(et, prm, c, e, u, p, ua, nu, start, cols, slots)
    neq = dim + 1
    n = c.shape[0] * neq
    D = np.zeros((n, n)); r = np.zeros(n)
    for el in e:
        M, b = oracle.ns_element_system(et, prm, c[el], u[el], p[el], ua[el], nu[el])
        idx = (el[:, None] * neq + np.arange(neq)).ravel()
        D[np.ix_(idx, idx)] += M
        r[idx] += b
    assert np.abs(dense_from_blocks(neq, start, cols, values) - D).max() <= 1e-13 * np.abs(D).max()
    assert np.abs(rhs - r).max() <= 1e-13 * np.abs(r).max()
    # the threaded (coloured) baseline changes only the summation order
    v2, r2 = oracle.ns_assemble(et, prm, c, e, u, p, ua, nu, start, cols, slots, nthreads=4)
    assert np.abs(v2 - values).max() <= 1e-13 * np.abs(values).max()


def test_symmetric_dirichlet_reference_fixture(oracle, fixtures):
    f = fixtures["symmetric_dirichlet"]
    start = np.array(f["starting_indices"], np.uint32); cols = np.array(f["node_connectivity"], np.uint32)
    values = np.zeros(len(cols)); rhs = np.zeros(4); sol = np.zeros(4)
    for i in range(4):
        oracle.set_row(1, start, cols, values, i, 0, f["set_row"]["diag"], f["set_row"]["offdiag"])
    bc = f["bc"]
    oracle.dirichlet(1, start, cols, values, rhs, sol, [bc["node"]], [bc["eq"]], [bc["value"]])
    assert np.array_equal(dense_from_blocks(1, start, cols, values), np.array(f["expected_matrix"], float))
    assert np.array_equal(rhs, np.array(f["expected_rhs"]))
    assert sol[1] == 10.0


def test_vector_apply_reference_fixture(oracle):
    c, t = oracle.rect_tris(3, 3)
    start, cols, _ = oracle.build_sparsity(t, c.shape[0], with_slots=False)
    neq = 3
    values = np.ones(len(cols) * neq * neq)
    x = np.random.default_rng(0).uniform(-1, 1, c.shape[0] * neq)
    y = oracle.spmv(neq, start, cols, values, x, alpha=2.0, beta=1.0, y=np.ones(len(x)))
    D = dense_from_blocks(neq, start, cols, values)
    assert np.allclose(y, 1 + 2 * D @ x, rtol=1e-14, atol=1e-14)
    rp, ci, v = oracle.point_csr(neq, start, cols, values)
    import scipy.sparse as sp
    assert np.array_equal(sp.csr_matrix((v, ci, rp), shape=D.shape).toarray(), D)


@pytest.mark.parametrize("kind", ["tri", "tet"])
def test_symmetric_dirichlet_exact_elimination(oracle, kind):
    et, dim, c, e, u, ua, p, nu = small_case(oracle, kind)
    neq = dim + 1
    start, cols, slots = oracle.build_sparsity(e, c.shape[0])
    values, rhs = oracle.ns_assemble(et, ns_params(), c, e, u, p, ua, nu, start, cols, slots)
    D0 = dense_from_blocks(neq, start, cols, values); b0 = rhs.copy()
    rng = np.random.default_rng(9)
    bn = rng.choice(c.shape[0], 5, replace=False)
    be = rng.integers(0, neq, 5)
    bv = rng.uniform(-1, 1, 5)
    sol = np.zeros_like(rhs)
    oracle.dirichlet(neq, start, cols, values, rhs, sol, bn, be, bv)
    D = dense_from_blocks(neq, start, cols, values)
    bcd = bn * neq + be
    free = np.setdiff1d(np.arange(len(rhs)), bcd)
    x = np.linalg.solve(D, rhs)
    xr = np.zeros_like(x); xr[bcd] = bv
    xr[free] = np.linalg.solve(D0[np.ix_(free, free)], b0[free] - D0[np.ix_(free, bcd)] @ bv)
    assert np.allclose(x, xr, rtol=1e-10, atol=1e-12)
    assert np.allclose(D[bcd][:, bcd], np.eye(5)) and np.all(D[free][:, bcd] == 0)
    assert np.array_equal(sol[bcd], bv)


def test_dirichlet_repeated_dof_keeps_first_elimination(oracle):
    """A dof that receives two BCs keeps the first value in the other rows' rhs and the last in its own."""
    c, t = oracle.rect_tris(2, 2)
    start, cols, _ = oracle.build_sparsity(t, c.shape[0], with_slots=False)
    vals = np.random.default_rng(1).uniform(1, 2, len(cols) * 9)
    rhs = np.zeros(c.shape[0] * 3); sol = np.zeros_like(rhs)
    D0 = dense_from_blocks(3, start, cols, vals)
    oracle.dirichlet(3, start, cols, vals, rhs, sol, [4, 4], [0, 0], [1.0, 5.0])
    other = np.array([i for i in range(len(rhs)) if i != 12])
    assert np.allclose(rhs[other], -D0[other, 12] * 1.0)
    assert rhs[12] == 5.0 and sol[12] == 5.0


@pytest.mark.parametrize("kind", ["tri", "tet", "hex"])
def test_krylov_vs_dense(oracle, kind):
    et, dim, c, e, u, ua, p, nu = small_case(oracle, kind)
    neq = dim + 1
    start, cols, slots = oracle.build_sparsity(e, c.shape[0])
    values, rhs = oracle.ns_assemble(et, ns_params(dt=0.01), c, e, u, p, ua, nu, start, cols, slots)
    sol = np.zeros_like(rhs)
    # no-slip on the x=0 wall (without velocity BCs a uniform velocity with a balancing linear
    # pressure is a null vector of the unsteady incompressible operator) and p pinned at node 0
    wall = np.where(np.isclose(c[:, 0], 0.0))[0]
    bn = np.concatenate([np.repeat(wall, dim), [0]])
    be = np.concatenate([np.tile(np.arange(dim), len(wall)), [dim]])
    oracle.dirichlet(neq, start, cols, values, rhs, sol, bn, be, np.zeros(len(bn)))
    D = dense_from_blocks(neq, start, cols, values)
    xd = np.linalg.solve(D, rhs)
    x0 = np.random.default_rng(2).uniform(-0.1, 0.1, len(rhs))
    x1, st = oracle.bicgstab(neq, start, cols, values, rhs, x0, tol=1e-10)
    assert st.converged and st.rel_residual <= 1e-10
    assert np.linalg.norm(rhs - D @ x1) <= 1e-10 * np.linalg.norm(rhs) * 1.01
    x2, st2 = oracle.gmres_ilu0(neq, start, cols, values, rhs, x0, restart=30, tol=1e-10)
    assert st2.rel_residual <= 1e-10
    assert np.abs(x2 - xd).max() <= 1e-6 * np.abs(xd).max()


def test_ilu0_exact_on_tridiagonal(oracle):
    """ILU(0) of a tridiagonal matrix is its exact LU factorisation: GMRES converges in one iteration."""
    n = 20
    cols = []; start = [0]
    for i in range(n):
        cols += [j for j in (i - 1, i, i + 1) if 0 <= j < n]; start.append(len(cols))
    start = np.array(start, np.uint32); cols = np.array(cols, np.uint32)
    rng = np.random.default_rng(4)
    values = rng.uniform(-1, 1, len(cols))
    for i in range(n):
        values[start[i] + (1 if i > 0 else 0)] = 4.0
    b = rng.uniform(-1, 1, n)
    x, st = oracle.gmres_ilu0(1, start, cols, values, b, np.zeros(n), restart=30, tol=1e-12)
    assert st.iterations == 1 and st.rel_residual < 1e-13


def test_ilu0_exact_on_block_tridiagonal(oracle):
    """With 3 equations per node and full blocks coupling each node to its neighbours only, ILU(0) drops
    no fill: it is the exact LU factorisation and GMRES converges in one iteration (the node-block order
    of the triangular sweeps included)."""
    N, neq = 12, 3
    cols = []; start = [0]
    for i in range(N):
        cols += [j for j in (i - 1, i, i + 1) if 0 <= j < N]; start.append(len(cols))
    start = np.array(start, np.uint32); cols = np.array(cols, np.uint32)
    rng = np.random.default_rng(5)
    values = rng.uniform(-1, 1, len(cols) * neq * neq)
    for i in range(N):
        s = start[i] + (1 if i > 0 else 0)
        values[s * neq * neq:(s + 1) * neq * neq] += 6.0 * np.eye(neq).ravel()
    b = rng.uniform(-1, 1, N * neq)
    x, st = oracle.gmres_ilu0(neq, start, cols, values, b, np.zeros(N * neq), restart=30, tol=1e-12)
    assert st.iterations == 1 and st.rel_residual < 1e-13


def test_node_updates(oracle):
    rng = np.random.default_rng(8)
    u, u1, u2, u3 = (rng.uniform(-1, 1, (7, 3)) for _ in range(4))
    # LinearizeU (NavierStokes.cpp:105) extrapolates the last four states, a unit time apart, to half a
    # step ahead: exact for every cubic in time and for no quartic.  Four weights at the abscissae
    # 0, -1, -2, -3 are fixed by exactness on 1, t, t^2, t^3, so this pins them and the argument order.
    t = np.array([0.0, -1.0, -2.0, -3.0])
    c = rng.uniform(-1, 1, (4, 7, 3))            # per entry: f(t) = c0 + c1 t + c2 t^2 + c3 t^3
    f = [c[0] + c[1] * tk + c[2] * tk ** 2 + c[3] * tk ** 3 for tk in t]
    ua = oracle.linearize_u(*f)
    assert np.allclose(ua, c[0] + c[1] * 0.5 + c[2] * 0.25 + c[3] * 0.125, rtol=1e-13, atol=1e-13)
    # t^4: the extrapolation error is -(1/2)(3/2)(5/2)(7/2) = -105/16
    q = [np.full((7, 3), tk ** 4) for tk in t]
    assert np.allclose(oracle.linearize_u(*q), 0.5 ** 4 - 105.0 / 16.0, rtol=1e-14)
    # first step: u1 = u2 = u3 = u gives u_adv = u
    assert np.allclose(oracle.linearize_u(u, u, u, u), u, rtol=1e-15, atol=1e-15)
    p = rng.uniform(-1, 1, 7); sol = rng.uniform(-1, 1, 28)
    uu, pp, a1, a2, a3 = u.copy(), p.copy(), u1.copy(), u2.copy(), u3.copy()
    oracle.update(uu, pp, a1, a2, a3, sol)
    assert np.array_equal(a3, u2) and np.array_equal(a2, u1) and np.array_equal(a1, u)
    assert np.allclose(uu, u + sol.reshape(7, 4)[:, :3]) and np.allclose(pp, p + sol.reshape(7, 4)[:, 3])

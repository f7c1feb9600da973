"""The reference's 4-rank Stokes system with Dirichlet and periodic conditions
(test/math/utest-lss-test-matrix.hpp, test/math/utest-lss-distributed-matrix.cpp:102-166), replayed on
one process: each rank's owned rows are filled with the values before the conditions, its periodic ties
(System::periodicity) and Dirichlet conditions are applied as that rank applies them, the owned rows are
compared with the fixture's values after the conditions, then merged into one global system whose
solution must match the reference result within 1e-3 % (the reference test's BOOST_CHECK_CLOSE).

CPU tests run the oracle's restatement of tie_blockrow_pairs / periodicity; the GPU test runs the same
replay through the C ABI (cf3hip_lss_periodicity, cf3hip_matrix_set_row, GMRES(30) + ILU0).
"""
import numpy as np
import pytest

import lss_fixture as lf
from coolfluid3_amd import cabi

NEQ = lf.NEQ


def oracle_replay(oracle, fx):
    out = []
    for r, rk in enumerate(fx["lss_test_matrix"]["ranks"]):
        start, cols, vals, rhs, sol = lf.local_system(rk, r)
        pr = lf.owned_pairs(rk, r)
        assert oracle.periodicity(NEQ, start, cols, vals, rhs, sol, [t for t, _ in pr], [f for _, f in pr]) == 0
        for n, e, v in lf.owned_conditions(rk, r):   # System::dirichlet without symmetry preservation
            oracle.set_row(NEQ, start, cols, vals, n, e, 1.0, 0.0)
            rhs[n * NEQ + e] = v
            sol[n * NEQ + e] = v
        out.append((start, cols, vals, rhs, sol))
    return out


def test_fixture_values_after_conditions_oracle(oracle, fixtures):
    reps = oracle_replay(oracle, fixtures)
    for r, (rk, (start, cols, vals, rhs, sol)) in enumerate(zip(fixtures["lss_test_matrix"]["ranks"], reps)):
        err = lf.owned_block_error(rk, r, start, vals)
        own = np.array([i * NEQ + j for i in lf.owned(rk, r) for j in range(NEQ)])
        print(f"rank {r}: {len(lf.owned_pairs(rk, r))} ties, {len(lf.owned_conditions(rk, r))} conditions, "
              f"owned rows vs mat_presolve {err:.1e}")
        assert err <= 1e-11   # the fixture prints 13 significant digits
        assert np.array_equal(rhs[own], np.array(rk["rhs_presolve"])[own])
        assert np.array_equal(sol[own], np.array(rk["sol_presolve"])[own])


def test_fixture_solution_oracle(oracle, fixtures):
    start, cols, vals, rhs, sol, res = lf.merge(fixtures, oracle_replay(oracle, fixtures))
    x, st = oracle.gmres_ilu0(NEQ, start, cols, vals, rhs, sol, restart=30, tol=1e-12, maxit=5000)
    ok, worst = lf.close_to_result(x, res)
    print(f"oracle GMRES(30)+ILU0: {st.iterations} it, largest relative difference to the result {worst:.2e}")
    assert st.converged and ok


def test_tie_oracle_errors_and_chain(oracle):
    """Rows with different block counts or columns are refused; a chain of ties (a row tied as 'to' and
    then as 'from') follows the sequential definition, checked against dense arithmetic."""
    # 1D chain 0-1-2-3-4 with periodic-style extra columns so rows 0, 2 and 4 share columns {0, 2, 4}
    graph = [[0, 2, 4], [1], [0, 2, 4], [3], [0, 2, 4]]
    start = np.concatenate([[0], np.cumsum([len(g) for g in graph])]).astype(np.uint32)
    cols = np.array([c for g in graph for c in g], np.uint32)
    rng = np.random.default_rng(5)
    neq = 2
    vals = rng.uniform(-1, 1, len(cols) * neq * neq)
    rhs = rng.uniform(-1, 1, 5 * neq)
    sol = rng.uniform(-1, 1, 5 * neq)
    dense = np.zeros((5 * neq, 5 * neq))
    for i in range(5):
        for q in range(start[i], start[i + 1]):
            dense[i * neq:(i + 1) * neq, cols[q] * neq:(cols[q] + 1) * neq] = vals[q * 4:(q + 1) * 4].reshape(2, 2)
    d_rhs, d_sol = rhs.copy(), sol.copy()
    for t, f in ((2, 4), (0, 2)):   # tie 4 -> 2, then 2 -> 0
        T, F = slice(t * neq, (t + 1) * neq), slice(f * neq, (f + 1) * neq)
        dense[T] += dense[F]
        dense[F] = 0
        dense[F, F] = np.eye(neq)
        dense[F, T] = -np.eye(neq)
        dense[T, T] += dense[T, F]
        dense[T, F] = 0
        d_rhs[T] += d_rhs[F]
        d_rhs[F] = 0
        m = 0.5 * (d_sol[T] + d_sol[F])
        d_sol[T] = m
        d_sol[F] = m
    assert oracle.periodicity(neq, start, cols, vals, rhs, sol, [2, 0], [4, 2]) == 0
    got = np.zeros_like(dense)
    for i in range(5):
        for q in range(start[i], start[i + 1]):
            got[i * neq:(i + 1) * neq, cols[q] * neq:(cols[q] + 1) * neq] = vals[q * 4:(q + 1) * 4].reshape(2, 2)
    assert np.allclose(got, dense, rtol=0, atol=1e-15)
    assert np.allclose(rhs, d_rhs, rtol=0, atol=1e-15) and np.allclose(sol, d_sol, rtol=0, atol=1e-15)
    assert oracle.periodicity(neq, start, cols, vals, None, None, [0], [1]) == -1   # 3 blocks vs 1
    graph2 = [[0, 1], [1, 2], [1, 2]]
    s2 = np.array([0, 2, 4, 6], np.uint32)
    c2 = np.array([c for g in graph2 for c in g], np.uint32)
    assert oracle.periodicity(1, s2, c2, np.ones(6), None, None, [0], [1]) == -2   # columns differ


@pytest.mark.gpu
def test_fixture_replay_gpu(oracle, fixtures):
    ctx = cabi.Context(0)
    reps = oracle_replay(oracle, fixtures)
    dev = []
    for r, rk in enumerate(fixtures["lss_test_matrix"]["ranks"]):
        start, cols, vals, rhs, sol = lf.local_system(rk, r)
        s = cabi.System(ctx, neq=NEQ, graph=(start, cols, len(start) - 1))
        s.reset(0.0)
        # the fill loop of utest-lss-distributed-matrix.cpp:120-125: set_value(column, row, value)
        nn = NEQ * NEQ
        for i in lf.owned(rk, r):
            for q in range(int(start[i]), int(start[i + 1])):
                for j in range(NEQ):
                    for l in range(NEQ):
                        s.set_value(int(cols[q]) * NEQ + l, i * NEQ + j, vals[q * nn + j * NEQ + l])
        s.set_vector(cabi.RHS, rhs)
        s.set_vector(cabi.SOLUTION, sol)
        pr = lf.owned_pairs(rk, r)
        s.periodicity([t for t, _ in pr], [f for _, f in pr])
        for n, e, v in lf.owned_conditions(rk, r):
            s.set_row(n, e, 1.0, 0.0)
            s.set_vector_value(cabi.RHS, n, e, v)
            s.set_vector_value(cabi.SOLUTION, n, e, v)
        gv, grhs, gsol = s.values(), s.vector(cabi.RHS), s.vector(cabi.SOLUTION)
        ostart, ocols, ovals, orhs, osol = reps[r]
        own = np.array([i * NEQ + j for i in lf.owned(rk, r) for j in range(NEQ)])
        err = lf.owned_block_error(rk, r, start, gv)
        assert np.array_equal(gv, ovals), "device ties differ from the oracle's"
        assert err <= 1e-11
        assert np.array_equal(grhs[own], orhs[own]) and np.array_equal(gsol[own], osol[own])
        dev.append((start, cols, gv, grhs, gsol))
        s.close()
    start, cols, vals, rhs, sol, res = lf.merge(fixtures, dev)
    g = cabi.System(ctx, neq=NEQ, graph=(start, cols, len(start) - 1))
    g.reset(0.0)
    nn = NEQ * NEQ
    for i in range(len(start) - 1):
        for q in range(int(start[i]), int(start[i + 1])):
            for j in range(NEQ):
                for l in range(NEQ):
                    v = vals[q * nn + j * NEQ + l]
                    if v != 0.0:
                        g.set_value(int(cols[q]) * NEQ + l, i * NEQ + j, v)
    g.set_vector(cabi.RHS, rhs)
    g.set_vector(cabi.SOLUTION, sol)
    st = g.solve(cabi.solver_params(cabi.GMRES_ILU0, tol=1e-12, restart=30, max_iter=5000))
    x = g.vector(cabi.SOLUTION)
    ok, worst = lf.close_to_result(x, res)
    print(f"device GMRES(30)+ILU0: {st.iterations} it, residual {st.rel_residual:.1e}, largest relative "
          f"difference to the reference result {worst:.2e}")
    assert st.converged and ok
    g.close()
    ctx.close()

"""The CRS matrix builder (TrilinosCrsMatrix semantics on the node-block storage): symmetric Dirichlet
with cached columns (cf3/math/LSS/Trilinos/TrilinosCrsMatrix.cpp:325-402), clone_to (:518-538),
read_native (:540-546) and get_column_and_replace_to_zero refused (:318-321).

CPU tests pin the oracle's restatement (oracle.pyoracle.CrsConditions): a first application equals the
block builder's symmetric elimination (oracle orc_dirichlet), and repeated applications subtract the
ORIGINAL column once per application -- checked with dense linear algebra on the matrix before any
condition.  GPU tests compare the device with that restatement and exercise clone / read / refusals."""
import os

import numpy as np
import pytest

from coolfluid3_amd import cabi
from oracle import pyoracle


@pytest.fixture(scope="module")
def ctx():
    c = cabi.Context(0)
    yield c
    c.close()


def random_graph_system(neq, n=12, seed=7):
    rng = np.random.default_rng(seed)
    graph = []
    for i in range(n):
        nb = {i, (i + 1) % n, (i - 1) % n, (i + 5) % n, (i - 5) % n}   # structurally symmetric
        graph.append(sorted(nb))
    start = np.concatenate([[0], np.cumsum([len(g) for g in graph])]).astype(np.uint32)
    cols = np.array([c for g in graph for c in g], np.uint32)
    vals = rng.uniform(-1, 1, len(cols) * neq * neq)
    rhs = rng.uniform(-1, 1, n * neq)
    sol = rng.uniform(-1, 1, n * neq)
    return start, cols, vals, rhs, sol


def dense(neq, start, cols, vals):
    n = len(start) - 1
    A = np.zeros((n * neq, n * neq))
    nn = neq * neq
    for i in range(n):
        for q in range(int(start[i]), int(start[i + 1])):
            A[i * neq:(i + 1) * neq, cols[q] * neq:(cols[q] + 1) * neq] = vals[q * nn:(q + 1) * nn].reshape(neq, neq)
    return A


CONDS = [(3, 0, 0.7), (5, 1, -0.4), (3, 0, 1.3), (4, 2, 0.25), (5, 1, 2.0), (3, 0, -0.5)]


@pytest.mark.parametrize("neq", [1, 3])
def test_crs_oracle_first_application_is_the_block_elimination(oracle, neq):
    start, cols, vals, rhs, sol = random_graph_system(neq)
    v1, r1, s1 = vals.copy(), rhs.copy(), sol.copy()
    v2, r2, s2 = vals.copy(), rhs.copy(), sol.copy()
    crs = pyoracle.CrsConditions()
    seen = set()
    nodes, eqs, vs = [], [], []
    for node, eq, v in CONDS:
        e = min(eq, neq - 1)
        if (node, e) in seen:
            continue
        seen.add((node, e))
        crs.apply(neq, start, cols, v1, r1, s1, node, e, v)
        nodes.append(node); eqs.append(e); vs.append(v)
    oracle.dirichlet(neq, start, cols, v2, r2, s2, nodes, eqs, vs)
    assert np.array_equal(v1, v2) and np.allclose(r1, r2, rtol=0, atol=1e-15) and np.array_equal(s1, s2)


@pytest.mark.parametrize("neq", [1, 3])
def test_crs_oracle_repeats_subtract_the_original_column(oracle, neq):
    """Dense model of the same semantics: the column cached at a dof's first application is A's column
    at that moment (A0's, with the rows of dofs conditioned before it already identity rows, and the
    dof's own row excluded); every application subtracts that cached column times its value from the rhs
    (also from rows of other conditioned dofs, as the reference does) and then sets rhs = value at the
    dof.  The block builder subtracts only at the first application."""
    start, cols, vals, rhs, sol = random_graph_system(neq)
    A = dense(neq, start, cols, vals)
    crs = pyoracle.CrsConditions()
    v1, r1, s1 = vals.copy(), rhs.copy(), sol.copy()
    conds = [(node, min(eq, neq - 1), v) for node, eq, v in CONDS]
    for node, e, v in conds:
        crs.apply(neq, start, cols, v1, r1, s1, node, e, v)
    expect = rhs.copy()
    cache = {}
    for node, e, v in conds:
        d = node * neq + e
        if d not in cache:
            col = A[:, d].copy()
            col[d] = 0.0
            cache[d] = col
            A[np.arange(A.shape[0]) != d, d] = 0.0
            A[d, :] = 0.0
            A[d, d] = 1.0
        expect -= cache[d] * v
        expect[d] = v
    assert np.allclose(r1, expect, rtol=0, atol=1e-14)
    assert np.array_equal(dense(neq, start, cols, v1), A)
    # the block builder: one subtraction per dof, the same matrix, a different rhs once conditions repeat
    v2, r2, s2 = vals.copy(), rhs.copy(), sol.copy()
    oracle.dirichlet(neq, start, cols, v2, r2, s2, [c[0] for c in conds], [c[1] for c in conds], [c[2] for c in conds])
    assert np.array_equal(v1, v2)
    assert not np.allclose(r1, r2)


def device_system(ctx, neq, start, cols, vals, rhs, sol, builder):
    s = cabi.System(ctx, neq=neq, graph=(start, cols, len(start) - 1))
    s.set_builder(builder)
    s.reset(0.0)
    nn = neq * neq
    for i in range(len(start) - 1):
        for q in range(int(start[i]), int(start[i + 1])):
            for j in range(neq):
                for l in range(neq):
                    s.set_value(int(cols[q]) * neq + l, i * neq + j, vals[q * nn + j * neq + l])
    s.set_vector(cabi.RHS, rhs)
    s.set_vector(cabi.SOLUTION, sol)
    return s


@pytest.mark.gpu
@pytest.mark.parametrize("neq", [1, 3, 4])
def test_crs_dirichlet_device_vs_oracle(ctx, neq):
    start, cols, vals, rhs, sol = random_graph_system(neq, seed=neq)
    crs = pyoracle.CrsConditions()
    v1, r1, s1 = vals.copy(), rhs.copy(), sol.copy()
    s = device_system(ctx, neq, start, cols, vals, rhs, sol, cabi.BUILDER_CRS)
    assert s.builder() == cabi.BUILDER_CRS
    # two calls: the second repeats conditions of the first and adds one
    calls = [CONDS[:3], CONDS[3:] + [(7, 1, 0.9)]]
    for call in calls:
        c = [(node, min(eq, neq - 1), v) for node, eq, v in call]
        for node, e, v in c:
            crs.apply(neq, start, cols, v1, r1, s1, node, e, v)
        s.dirichlet([x[0] for x in c], [x[1] for x in c], [x[2] for x in c])
    assert np.array_equal(s.values(), v1)
    assert np.allclose(s.vector(cabi.RHS), r1, rtol=0, atol=1e-14 * np.abs(r1).max())
    assert np.array_equal(s.vector(cabi.SOLUTION), s1)
    # a reset clears the cache: the next condition is a first application again
    s.reset(0.0)
    crs.reset()
    s.close()


@pytest.mark.gpu
def test_crs_clone_read_native_and_refusals(ctx, tmp_path):
    neq = 3
    start, cols, vals, rhs, sol = random_graph_system(neq, seed=11)
    a = device_system(ctx, neq, start, cols, vals, rhs, sol, cabi.BUILDER_CRS)
    a.dirichlet([3], [1], [0.5])
    b = cabi.System(ctx, neq=neq, graph=(start, cols, len(start) - 1))
    b.set_builder(cabi.BUILDER_CRS)
    b.reset(0.0)
    a.clone_to(b)
    assert np.array_equal(b.values(), a.values())
    # the clone carries the cached column: a repeat on both gives the same rhs change
    b.set_vector(cabi.RHS, a.vector(cabi.RHS))
    a.dirichlet([3], [1], [1.5])
    b.dirichlet([3], [1], [1.5])
    assert np.array_equal(a.vector(cabi.RHS), b.vector(cabi.RHS))
    # the block builder: clone_to and read_native are not implemented (TrilinosFEVbrMatrix.cpp:798, 805),
    # and a CRS source needs a CRS destination (SetupError)
    v = device_system(ctx, neq, start, cols, vals, rhs, sol, cabi.BUILDER_VBR)
    with pytest.raises(cabi.Cf3HipError) as e1:
        v.clone_to(b)
    assert e1.value.code == 7
    with pytest.raises(cabi.Cf3HipError) as e2:
        a.clone_to(v)
    assert e2.value.code == 2
    with pytest.raises(cabi.Cf3HipError) as e3:
        v.read_native(str(tmp_path / "m.mtx"))
    assert e3.value.code == 7
    with pytest.raises(cabi.Cf3HipError) as e4:
        a.column_zero(2, 0)
    assert e4.value.code == 7
    # read_native: a Matrix Market file of the matrix values (general, 1-based), into a fresh CRS system
    rp, ci, pv = a.debug_data()
    path = tmp_path / "a.mtx"
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n% written by the test\n")
        f.write(f"{len(rp) - 1} {len(rp) - 1} {len(pv)}\n")
        for r in range(len(rp) - 1):
            for q in range(int(rp[r]), int(rp[r + 1])):
                f.write(f"{r + 1} {int(ci[q]) + 1} {float(pv[q])!r}\n")
    c = cabi.System(ctx, neq=neq, graph=(start, cols, len(start) - 1))
    c.set_builder(cabi.BUILDER_CRS)
    c.read_native(str(path))
    assert np.array_equal(c.values(), a.values())
    bad = tmp_path / "bad.mtx"
    with open(bad, "w") as f:   # (row 1, col 7) is outside the pattern: node 0 is not connected to node 2
        f.write("%%MatrixMarket matrix coordinate real general\n")
        f.write(f"{len(rp) - 1} {len(rp) - 1} 1\n1 7 1.0\n")
    with pytest.raises(cabi.Cf3HipError) as e5:
        c.read_native(str(bad))
    assert e5.value.code == 6
    for x in (a, b, v, c):
        x.close()

"""Recycling conjugate gradients (RCGStrategy, cf3/math/LSS/Trilinos/RCGStrategy.cpp:59-171, over
Belos::RCGSolMgr): the oracle's restatement (oracle/rcg.py) pinned on the CPU, then the device solver
(CF3HIP_RCG_JACOBI) against it.

The system is the pressure-type operator of the semi-implicit solver: a P1 Laplacian on a randomly
perturbed triangulated square (no symmetry, so no repeated eigenvalues at the recycle cut) with the
left edge eliminated symmetrically (identity rows and columns), which is symmetric positive definite."""
import numpy as np
import pytest
import scipy.linalg
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from coolfluid3_amd import cabi
from oracle import rcg


def laplacian_system(nx=28, seed=3):
    rng = np.random.default_rng(seed)
    xs = np.linspace(0.0, 1.0, nx + 1)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    h = 1.0 / nx
    X = X + rng.uniform(-0.2, 0.2, X.shape) * h * ((X > 0) & (X < 1))
    Y = Y + rng.uniform(-0.2, 0.2, Y.shape) * h * ((Y > 0) & (Y < 1))
    pts = np.column_stack([X.ravel(), Y.ravel()])
    idx = lambda i, j: i * (nx + 1) + j
    tris = []
    for i in range(nx):
        for j in range(nx):
            tris.append((idx(i, j), idx(i + 1, j), idx(i + 1, j + 1)))
            tris.append((idx(i, j), idx(i + 1, j + 1), idx(i, j + 1)))
    n = len(pts)
    rows, cols, vals = [], [], []
    for t in tris:
        P = pts[list(t)]
        J = np.array([P[1] - P[0], P[2] - P[0]]).T
        area = 0.5 * abs(np.linalg.det(J))
        G = np.linalg.solve(J.T, np.array([[-1.0, 1.0, 0.0], [-1.0, 0.0, 1.0]]))
        K = area * (G.T @ G)
        for a in range(3):
            for b in range(3):
                rows.append(t[a]); cols.append(t[b]); vals.append(K[a, b])
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n)).tolil()
    fixed = [idx(0, j) for j in range(nx + 1)]
    for f in fixed:   # symmetric elimination: identity row and column
        A[f, :] = 0.0
        A[:, f] = 0.0
        A[f, f] = 1.0
    A = A.tocsr()
    A.sort_indices()
    # the node graph of the mesh (self included), which holds every nonzero of A
    nb = [set([i]) for i in range(n)]
    for t in tris:
        for a in t:
            nb[a].update(t)
    start = np.concatenate([[0], np.cumsum([len(s) for s in nb])]).astype(np.uint32)
    gcols = np.array([c for s in nb for c in sorted(s)], np.uint32)
    return A, start, gcols


def rhs_sequence(n, count, seed=9):
    rng = np.random.default_rng(seed)
    return [rng.uniform(-1, 1, n) for _ in range(count)]


# ------------------------------------------------------------------------------------------------
# the oracle
# ------------------------------------------------------------------------------------------------

def test_oracle_without_recycling_is_jacobi_cg():
    A, _, _ = laplacian_system()
    b = rhs_sequence(A.shape[0], 1)[0]
    d = A.diagonal()
    its = []
    xs, info = spla.cg(A, b, rtol=1e-10, atol=0.0, maxiter=5000, M=sp.diags(1.0 / d),
                       callback=lambda xk: its.append(1))
    assert info == 0
    x, it, conv, rel = rcg.rcg_solve(A, b, np.zeros_like(b), rcg.Space(), 1e-10, 5000, cycle=10 ** 6, recycle=0)
    assert conv and rel <= 1e-10
    assert it == len(its)
    assert np.allclose(x, xs, rtol=0, atol=1e-12 * np.abs(xs).max())


def test_oracle_recycling_converges_and_saves_iterations():
    A, _, _ = laplacian_system()
    n = A.shape[0]
    space = rcg.Space()
    plain, recycled = [], []
    for b in rhs_sequence(n, 6):
        xd = spla.spsolve(A.tocsc(), b)
        x, it, conv, rel = rcg.rcg_solve(A, b, np.zeros(n), space, 1e-10, 5000, cycle=40, recycle=8)
        assert conv and rel <= 1e-10
        assert np.allclose(x, xd, rtol=0, atol=1e-7 * np.abs(xd).max())
        recycled.append(it)
        plain.append(rcg.rcg_solve(A, b, np.zeros(n), rcg.Space(), 1e-10, 5000, cycle=40, recycle=0)[1])
    assert space.k == 8
    # from the second solve on, the recycled space removes the slowest modes
    assert all(r < 0.85 * p for r, p in zip(recycled[1:], plain[1:])), (recycled, plain)


def test_oracle_space_invariants_and_deflation():
    A, _, _ = laplacian_system(nx=20)
    n = A.shape[0]
    space = rcg.Space()
    for b in rhs_sequence(n, 4, seed=5):
        rcg.rcg_solve(A, b, np.zeros(n), space, 1e-10, 5000, cycle=400, recycle=6)
    W, AW, E = space.W.copy(), space.AW.copy(), space.E.copy()
    assert np.linalg.norm(A @ W - AW) <= 1e-10 * np.linalg.norm(AW)
    assert np.linalg.norm(W.T @ (A @ W) - E) <= 1e-10 * np.linalg.norm(E)
    # every search direction of the next solve is A-orthogonal to the space it started with (one cycle)
    trace = []
    b = rhs_sequence(n, 1, seed=17)[0]
    rcg.rcg_solve(A, b, np.zeros(n), space, 1e-10, 5000, cycle=400, recycle=6, trace=trace)
    assert trace
    for p in trace:
        assert np.abs(AW.T @ p).max() <= 1e-8 * np.linalg.norm(AW, axis=0).max() * np.linalg.norm(p)
    # b = A w with w in span W is solved by the projection alone
    w = space.W @ np.linspace(1.0, 2.0, space.k)
    x, it, conv, rel = rcg.rcg_solve(A, A @ w, np.zeros(n), space, 1e-10, 5000, cycle=400, recycle=6)
    assert conv and it == 0
    assert np.allclose(x, w, rtol=0, atol=1e-9 * np.abs(w).max())
    # the space approaches the lowest eigenvectors: its smallest Ritz value is the smallest eigenvalue
    lam = scipy.linalg.eigvalsh(A.toarray())
    ritz = scipy.linalg.eigvalsh(space.W.T @ (A @ space.W), space.W.T @ space.W)
    assert abs(ritz[0] - lam[0]) <= 1e-3 * lam[0], (ritz[:3], lam[:3])


# ------------------------------------------------------------------------------------------------
# the device
# ------------------------------------------------------------------------------------------------

def device_system(ctx, A, start, gcols):
    n = A.shape[0]
    s = cabi.System(ctx, neq=1, graph=(start, gcols, n))
    s.reset(0.0)
    Ac = A.tocoo()
    for i, j, v in zip(Ac.row, Ac.col, Ac.data):
        s.set_value(int(j), int(i), float(v))
    return s


@pytest.fixture(scope="module")
def ctx():
    c = cabi.Context(0)
    yield c
    c.close()


@pytest.mark.gpu
def test_rcg_device_vs_oracle(ctx):
    A, start, gcols = laplacian_system()
    n = A.shape[0]
    s = device_system(ctx, A, start, gcols)
    space = rcg.Space()
    prm = cabi.solver_params(method=cabi.RCG_JACOBI, tol=1e-10, max_iter=5000, restart=40, recycle=8, rcg_space=0)
    dev_its, orc_its = [], []
    for b in rhs_sequence(n, 5):
        s.set_vector(cabi.RHS, b)
        s.set_vector(cabi.SOLUTION, np.zeros(n))
        st = s.solve(prm)
        xo, ito, convo, _ = rcg.rcg_solve(A, b, np.zeros(n), space, 1e-10, 5000, cycle=40, recycle=8)
        assert st.converged and convo
        assert st.rel_residual <= 1e-10
        xd = s.vector(cabi.SOLUTION)
        assert np.allclose(xd, xo, rtol=0, atol=1e-7 * np.abs(xo).max())
        dev_its.append(st.iterations)
        orc_its.append(ito)
    print("RCG iterations device", dev_its, "oracle", orc_its, flush=True)
    assert all(abs(d - o) <= 2 for d, o in zip(dev_its, orc_its)), (dev_its, orc_its)
    assert dev_its[-1] < 0.85 * dev_its[0]
    # the second space of the system starts empty: its first solve is a first solve
    b = rhs_sequence(n, 1, seed=21)[0]
    s.set_vector(cabi.RHS, b)
    s.set_vector(cabi.SOLUTION, np.zeros(n))
    st1 = s.solve(cabi.solver_params(method=cabi.RCG_JACOBI, tol=1e-10, max_iter=5000, restart=40, recycle=8,
                                     rcg_space=1))
    _, it1, _, _ = rcg.rcg_solve(A, b, np.zeros(n), rcg.Space(), 1e-10, 5000, cycle=40, recycle=8)
    assert st1.converged and abs(st1.iterations - it1) <= 2
    # without recycling: Jacobi CG (in space 1, which it empties: a request for fewer recycled vectors
    # than a space holds starts that space afresh)
    s.set_vector(cabi.SOLUTION, np.zeros(n))
    st0 = s.solve(cabi.solver_params(method=cabi.RCG_JACOBI, tol=1e-10, max_iter=5000, restart=96, recycle=0,
                                     rcg_space=1))
    _, it0, _, _ = rcg.rcg_solve(A, b, np.zeros(n), rcg.Space(), 1e-10, 5000, cycle=96, recycle=0)
    assert st0.converged and abs(st0.iterations - it0) <= 1
    # new values in the same system: the space of slot 0 is carried over, AW and E recomputed from W
    A2 = A.copy()
    A2.data = A2.data * np.linspace(1.0, 1.5, A2.nnz)
    A2 = (0.5 * (A2 + A2.T)).tocsr()
    A2c = A2.tocoo()
    for i, j, v in zip(A2c.row, A2c.col, A2c.data):
        s.set_value(int(j), int(i), float(v))
    b = rhs_sequence(n, 1, seed=23)[0]
    s.set_vector(cabi.RHS, b)
    s.set_vector(cabi.SOLUTION, np.zeros(n))
    st2 = s.solve(prm)
    _, it2, _, _ = rcg.rcg_solve(A2, b, np.zeros(n), space, 1e-10, 5000, cycle=40, recycle=8)
    assert st2.converged and st2.rel_residual <= 1e-10
    assert abs(st2.iterations - it2) <= 2, (st2.iterations, it2)
    s.close()


@pytest.mark.gpu
def test_rcg_refusals(ctx):
    A, start, gcols = laplacian_system(nx=6)
    s = device_system(ctx, A, start, gcols)
    s.set_vector(cabi.RHS, np.ones(A.shape[0]))
    for kw, code in [(dict(recycle=33), 8), (dict(restart=97), 8), (dict(rcg_space=2), 1),
                     (dict(precision=cabi.FP32), 7)]:
        p = dict(method=cabi.RCG_JACOBI, tol=1e-10, max_iter=100, restart=40, recycle=4)
        p.update(kw)
        with pytest.raises(cabi.Cf3HipError) as e:
            s.solve(cabi.solver_params(**p))
        assert e.value.code == code
    s.close()

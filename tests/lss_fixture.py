"""The reference's 4-rank Stokes test system (test/math/utest-lss-test-matrix.hpp, driven by
test/math/utest-lss-distributed-matrix.cpp:102-166) read from the committed fixture, with helpers that
replay it on one process: the owned rows of each rank merged into one global block system, and the
per-rank local node graphs for replaying the conditions the way each rank applies them.

Owned rows: a rank's local node i is updatable when irank_updatable[i] equals the rank.  The values
arrays list, for each owned local row i, equation j, block column k of the row and equation l, the value
of row (i, j) and column (cols[k], l) -- the loop of utest-lss-distributed-matrix.cpp:120-125 -- where
the block columns of a row are those of column_indices plus, for a node of a periodic pair, its partner,
in ascending local order: the sizes of the values arrays (3312, 3951, 3510, 3429) count exactly these
blocks, and the tied rows of mat_presolve hold +1 at their own column and -1 at their partner's there.

The fixture's values after the conditions (mat_presolve, rhs_presolve) are those of the periodic ties
applied first (System::periodicity, System.cpp:259-282 with TrilinosFEVbrMatrix::tie_blockrow_pairs,
TrilinosFEVbrMatrix.cpp:558-590) and the Dirichlet conditions second, without symmetry preservation
(System::dirichlet with preserve_symmetry false, System.cpp:240-255: Matrix::set_row(row, eq, 1, 0),
rhs = sol = value); the other orders and the symmetric elimination reproduce neither array.
"""
import numpy as np

NEQ = 3


def owned(rank, r):
    return [i for i, o in enumerate(rank["irank_updatable"]) if o == r]


def pairs(rank):
    pp = rank["periodic_pairs"]
    return [(pp[a], pp[a + 1]) for a in range(0, len(pp), 2)]


def rank_graph(rank):
    """Block columns of every local row: column_indices plus the periodic partners, ascending."""
    rs, ci = rank["rowstart_positions"], rank["column_indices"]
    partner = {}
    for t, f in pairs(rank):
        partner.setdefault(t, set()).add(f)
        partner.setdefault(f, set()).add(t)
    return [sorted(set(ci[rs[i]:rs[i + 1]]) | partner.get(i, set())) for i in range(len(rs) - 1)]


def local_system(rank, r, key="mat_prebc", rhs_key="rhs_prebc", sol_key="sol_prebc"):
    """(start, cols, block values, rhs, sol) of one rank's local graph (all local rows; the rows of nodes
    the rank does not own stay zero), values of the owned rows from array `key`."""
    graph = rank_graph(rank)
    start = np.concatenate([[0], np.cumsum([len(g) for g in graph])]).astype(np.uint32)
    cols = np.array([c for g in graph for c in g], np.uint32)
    vals = np.zeros(int(start[-1]) * NEQ * NEQ)
    it = iter(rank[key])
    for i in owned(rank, r):
        nb = len(graph[i])
        m = np.array([[next(it) for _ in range(NEQ * nb)] for _ in range(NEQ)])
        blk = m.reshape(NEQ, nb, NEQ).transpose(1, 0, 2)          # [block k][row eq j][col eq l]
        vals[int(start[i]) * NEQ * NEQ:int(start[i + 1]) * NEQ * NEQ] = blk.ravel()
    assert next(it, None) is None, "values left over: the owned rows do not describe the array"
    return start, cols, vals, np.array(rank[rhs_key], float), np.array(rank[sol_key], float)


def owned_pairs(rank, r):
    """Periodic pairs whose two rows the rank owns (TrilinosFEVbrMatrix.cpp:569-570 ties only those)."""
    upd = rank["irank_updatable"]
    return [(t, f) for t, f in pairs(rank) if upd[t] == r and upd[f] == r]


def owned_conditions(rank, r):
    """The Dirichlet conditions of the rank's owned nodes, in order (utest-lss-distributed-matrix.cpp:136-138)."""
    upd = rank["irank_updatable"]
    return [(n, e, v) for n, e, v in zip(rank["bc_node"], rank["bc_eqn"], rank["bc_value"]) if upd[n] == r]


def merge(fx, per_rank):
    """Global block system from per-rank local systems [(start, cols, vals, rhs, sol)]: the owned rows of
    every rank by global node number, columns sorted; plus the reference result."""
    ranks = fx["lss_test_matrix"]["ranks"]
    n = 1 + max(max(rk["global_numbering"]) for rk in ranks)
    rows = [None] * n
    rhs, sol, res = np.zeros(n * NEQ), np.zeros(n * NEQ), np.zeros(n * NEQ)
    nn = NEQ * NEQ
    for r, (rk, (start, cols, vals, lrhs, lsol)) in enumerate(zip(ranks, per_rank)):
        gid = rk["global_numbering"]
        for i in owned(rk, r):
            g = gid[i]
            assert rows[g] is None, "global node owned twice"
            rows[g] = {gid[int(cols[q])]: vals[q * nn:(q + 1) * nn] for q in range(int(start[i]), int(start[i + 1]))}
            rhs[g * NEQ:(g + 1) * NEQ] = lrhs[i * NEQ:(i + 1) * NEQ]
            sol[g * NEQ:(g + 1) * NEQ] = lsol[i * NEQ:(i + 1) * NEQ]
            res[g * NEQ:(g + 1) * NEQ] = rk["result"][i * NEQ:(i + 1) * NEQ]
    assert all(rw is not None for rw in rows), "a global node has no owner"
    gstart, gcols, gvals = [0], [], []
    for g in range(n):
        for c in sorted(rows[g]):
            gcols.append(c)
            gvals.append(rows[g][c])
        gstart.append(len(gcols))
    return (np.array(gstart, np.uint32), np.array(gcols, np.uint32), np.concatenate(gvals), rhs, sol, res)


def owned_block_error(rank, r, start, vals, key="mat_presolve"):
    """max |vals - fixture| over the owned rows, relative to each scalar row's largest fixture value."""
    ref = local_system(rank, r, key)[2]
    nn = NEQ * NEQ
    worst = 0.0
    for i in owned(rank, r):
        a, b = int(start[i]) * nn, int(start[i + 1]) * nn
        g = vals[a:b].reshape(-1, NEQ, NEQ)
        f = ref[a:b].reshape(-1, NEQ, NEQ)
        for j in range(NEQ):
            worst = max(worst, float(np.abs(g[:, j] - f[:, j]).max() / max(np.abs(f[:, j]).max(), 1e-300)))
    return worst


def close_to_result(x, res, rel=1e-5):
    """BOOST_CHECK_CLOSE(x, result, 1e-3) (percent, both-sided) of utest-lss-distributed-matrix.cpp:162-165;
    entries whose reference value is exactly zero (fixed dofs) must be zero within 1e-12."""
    nz = res != 0.0
    d = np.abs(x - res)
    ok_nz = np.all(d[nz] <= rel * np.abs(res[nz])) and np.all(d[nz] <= rel * np.abs(x[nz]))
    ok_z = np.all(np.abs(x[~nz]) <= 1e-12)
    worst = float((d[nz] / np.abs(res[nz])).max())
    return bool(ok_nz and ok_z), worst

"""Options of the precision form (FP32 Krylov phase under FP64 residual replacement): block-column
scaling of the FP32 operator (B = fl32(A32 D^-1), x = D^-1 y, D the 4x4 / 3x3 / ... diagonal blocks) and
FP32 inner vectors.  Each combination must reach the FP64 tolerance of the FP64 system and give the
FP64 solution within what that tolerance allows.

Stored systems: random block systems with a dominant block diagonal (coupled within each block, so the
block scaling is not a point scaling) are set through the entry interface in FP64, rounded into the FP32
mirror (cf3hip_lss_round_fp32) and solved in the precision form; the solution is checked against a
dense FP64 solve (numpy) of the same matrix.  Element systems: one implicit step of the skewed
tetrahedral channel in each combination against the FP64 step."""
import numpy as np
import pytest

from coolfluid3_amd import cabi, cases

from test_gpu_parity import CASES, gpu_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = cabi.Context(0)
    yield c
    c.close()


def random_block_system(neq, n=40, seed=3):
    """Block tridiagonal-plus-random graph, values with a dominant, internally coupled diagonal block."""
    rng = np.random.default_rng(seed)
    graph = []
    for i in range(n):
        nb = {i, (i + 1) % n, (i - 1) % n}
        nb |= set(rng.integers(0, n, 2).tolist())
        graph.append(sorted(nb))
    start = np.concatenate([[0], np.cumsum([len(g) for g in graph])]).astype(np.uint32)
    cols = np.array([c for g in graph for c in g], np.uint32)
    dense = np.zeros((n * neq, n * neq))
    for i in range(n):
        for c in graph[i]:
            blk = rng.uniform(-1, 1, (neq, neq))
            if c == i:
                blk = blk + np.diag(rng.uniform(4 * neq, 8 * neq, neq)) * 10.0 ** rng.uniform(-3, 3)
            dense[i * neq:(i + 1) * neq, c * neq:(c + 1) * neq] = blk
    return start, cols, dense


@pytest.mark.gpu
@pytest.mark.parametrize("neq", [1, 2, 3, 4])
@pytest.mark.parametrize("block_scale,fp32_vectors", [(0, 0), (1, 0), (1, 1)])
def test_stored_system_options(ctx, neq, block_scale, fp32_vectors):
    print("case neq", neq, "block_scale", block_scale, "fp32_vectors", fp32_vectors, flush=True)
    start, cols, dense = random_block_system(neq)
    n = len(start) - 1
    s = cabi.System(ctx, neq=neq, graph=(start, cols, n))
    s.reset(0.0)
    for i in range(n):
        for q in range(int(start[i]), int(start[i + 1])):
            c = int(cols[q])
            for a in range(neq):
                for b in range(neq):
                    s.set_value(c * neq + b, i * neq + a, dense[i * neq + a, c * neq + b])
    s.enable_fp32()
    s.round_fp32()
    v32 = s.values_fp32()
    assert np.array_equal(v32, s.values().astype(np.float32))   # round to nearest, value by value
    rng = np.random.default_rng(11)
    xs = rng.uniform(-1, 1, n * neq)
    b = dense @ xs
    s.set_vector(cabi.RHS, b)
    s.set_vector(cabi.SOLUTION, np.zeros(n * neq))
    st = s.solve(cabi.solver_params(tol=1e-10, precision=cabi.FP32, block_scale=block_scale,
                                    fp32_vectors=fp32_vectors, max_iter=2000))
    x = s.vector(cabi.SOLUTION)
    res = np.linalg.norm(b - dense @ x) / np.linalg.norm(b)
    xd = np.linalg.solve(dense, b)
    err = np.linalg.norm(x - xd) / np.linalg.norm(xd)
    kappa = np.linalg.cond(dense)
    print(f"neq {neq}: {st.iterations} it, {st.outer} replacements, residual {res:.2e}, error {err:.2e}, "
          f"cond {kappa:.1e}")
    assert st.converged and res <= 1e-10 * 1.001
    assert err <= 2 * kappa * res   # the forward error a residual of that size allows
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("block_scale,fp32_vectors", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_element_step_options(ctx, block_scale, fp32_vectors):
    """tet12s: ~1% elements of aspect ratio ~20; the FP32 element matrices (matrix-free FP64 residual)."""
    print("case tet12s block_scale", block_scale, "fp32_vectors", fp32_vectors, flush=True)
    case = CASES["tet12s"]()
    g = gpu_case(ctx, case, fp32=True)
    st = g.step(cabi.solver_params(tol=1e-8, precision=cabi.FP32, block_scale=block_scale,
                                   fp32_vectors=fp32_vectors))
    dx = g.system.vector(cabi.SOLUTION)
    g64 = gpu_case(ctx, CASES["tet12s"]())
    st64 = g64.step(cabi.solver_params(tol=1e-8))
    dx64 = g64.system.vector(cabi.SOLUTION)
    err = np.abs(dx - dx64).max() / np.abs(dx64).max()
    print(f"precision form: {st.iterations} it ({st.outer} replacements) vs FP64 {st64.iterations} it; "
          f"residual {st.rel_residual:.2e}; increment difference to FP64 {err:.2e}")
    assert st.converged and st.rel_residual <= 1e-8 * 1.001
    assert err <= 1e-5
    assert st.iterations <= 1.5 * st64.iterations + 5
    g.close()
    g64.close()

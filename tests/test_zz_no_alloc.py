"""The allocation-free residual entry point (cf3hip_lss_residual_noalloc) against the allocating one
(cf3hip_lss_residual) and the oracle, through the C ABI: two calls in a row, the second on a smaller
system, each within 1e-12 (relative) of the allocating entry point's value."""
import numpy as np
import pytest

from coolfluid3_amd import cabi, cases

pytestmark = pytest.mark.gpu


def constrained_system(ctx, n, seed):
    case = cases.inflow_cube_tets(n)
    g = cases.GpuCase(ctx, case)
    g.fields.linearize()
    g.assemble(cabi.FP64)
    nd, e, v = cases.bc_lists(case, case.u0, case.p0)
    g.system.dirichlet(nd, e, v)
    g.system.set_vector(cabi.SOLUTION, np.random.default_rng(seed).uniform(-1, 1, g.system.dofs))
    return g


def test_residual_noalloc_matches(oracle):
    ctx = cabi.Context(0)
    for n, seed in ((12, 1), (5, 2)):   # the second call on a smaller system
        print("case", n, seed, flush=True)
        g = constrained_system(ctx, n, seed)
        s = g.system
        r_na = s.residual_noalloc()
        r_al = s.residual()
        start, cols = s.graph()
        b = s.vector(cabi.RHS)
        x = s.vector(cabi.SOLUTION)
        r_or = np.linalg.norm(b - oracle.spmv(4, start, cols, s.values(), x)) / np.linalg.norm(b)
        print(f"n={n}: noalloc {r_na:.15e}, allocating {r_al:.15e}, oracle {r_or:.15e}")
        assert abs(r_na - r_al) <= 1e-12 * r_al
        assert abs(r_na - r_or) <= 1e-12 * r_or
        g.close()
    ctx.close()

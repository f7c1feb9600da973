"""CPU-side checks of the product: the C ABI library loads and exports every declared symbol, argument
errors are reported without touching a device, and the host-side mesh generators reproduce the
oracle's generators bit for bit."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from coolfluid3_amd import cabi, cases, meshgen


def declared_functions():
    text = open(cabi.HEADER).read()
    return sorted(set(re.findall(r"\b(cf3hip_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = cabi.lib()
    names = declared_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.cf3hip_version().decode().startswith("cf3hip")


def test_errors_without_device():
    lib = cabi.lib()
    assert lib.cf3hip_matrix_reset(None, 0.0) == 2                     # SetupError: not created
    assert lib.cf3hip_lss_solve(None, None, None) == 1
    assert "null" in lib.cf3hip_last_error().decode() or lib.cf3hip_last_error()
    assert lib.cf3hip_context_create(0, None) == 1
    assert lib.cf3hip_ns_assemble(None, None, None) == 1
    assert lib.cf3hip_si_create(None, None, None) == 1
    assert lib.cf3hip_si_step(None, None, None, 2, None, None, 0, None, None, 0, None, None, None) == 1


def test_meshgen_matches_oracle(oracle):
    c1, t1 = meshgen.rect_tris(5, 4)
    c2, t2 = oracle.rect_tris(5, 4)
    assert np.array_equal(c1, c2) and np.array_equal(t1, t2)
    c1, h1 = meshgen.box_hexas(3, 2, 4)
    c2, h2 = oracle.box_hexas(3, 2, 4)
    assert np.array_equal(c1, c2) and np.array_equal(h1, h2)
    c1, t1 = meshgen.box_tetras(3, 2, 4)
    c2, t2 = oracle.box_tetras(3, 2, 4)
    assert np.array_equal(c1, c2) and np.array_equal(t1, t2)
    j1, _ = meshgen.jitter(c1, (3, 2, 4), 0.1, 77)
    j2 = oracle.jitter(c2, 3, 3, 2, 4, 0.1, 77)
    assert np.array_equal(j1, j2)
    assert np.array_equal(meshgen.uniform(5, 10, -1, 1), oracle.uniform_stream(5, 10, -1, 1))


def test_case_definitions():
    c = cases.lid_cavity_2d(32)
    assert (c.n_nodes, c.n_elems) == (1089, 2048)
    n, e, v = cases.bc_lists(c, c.u0, c.p0)
    # lid applied last: the top corners end with u = 1, and p is pinned at node 0
    top_left = 32 * 33
    last_u = {int(a): float(b) for a, b, q in zip(n, v, e) if q == 0}
    assert last_u[top_left] == pytest.approx(1.0 - c.u0[top_left, 0])
    assert (n[-1], e[-1]) == (0, 2)
    t = cases.inflow_cube_tets(6)
    assert (t.n_nodes, t.n_elems) == (343, 1296)
    assert meshgen.tet_volumes(t.coords, t.conn).min() > 0


def test_skewed_tetra_mesh():
    """About 1% of the elements flattened to aspect ratio ~20, all valid, deterministic; the skew
    draws come after the jitter and velocity draws, so the velocity field is unchanged."""
    t = cases.inflow_cube_tets(24, skew=True)
    t0 = cases.inflow_cube_tets(24)
    ar = meshgen.tet_quality(t.coords, t.conn)
    assert 0.008 <= (ar > 15).mean() <= 0.012
    assert 18 <= ar.max() <= 40
    assert meshgen.tet_quality(t0.coords, t0.conn).max() < 6
    assert meshgen.tet_volumes(t.coords, t.conn).min() > 0
    assert np.array_equal(t.u0, t0.u0)
    t2 = cases.inflow_cube_tets(24, skew=True)
    assert np.array_equal(t.coords, t2.coords)
    # a moved node sits 93% of the way to its Kuhn-diagonal neighbour
    moved = np.nonzero((t.coords != t0.coords).any(1))[0]
    s = moved[0]
    d = s + 1 + 25 + 25 * 25
    assert np.allclose(t.coords[s], t0.coords[s] + 0.93 * (t0.coords[d] - t0.coords[s]), rtol=0, atol=1e-15)


def test_hexa_cavity_case():
    c = cases.lid_cavity_cube_hexas(4)
    assert (c.n_nodes, c.n_elems) == (125, 64)
    x = c.coords[c.conn]
    # trilinear hexahedra stay valid under 5% jitter: Jacobian determinant at the centre > 0
    j = np.stack([(x[:, [1, 2, 5, 6]].sum(1) - x[:, [0, 3, 4, 7]].sum(1)),
                  (x[:, [2, 3, 6, 7]].sum(1) - x[:, [0, 1, 4, 5]].sum(1)),
                  (x[:, [4, 5, 6, 7]].sum(1) - x[:, [0, 1, 2, 3]].sum(1))], axis=1) / 8
    assert np.linalg.det(j).min() > 0
    n, e, v = cases.bc_lists(c, c.u0, c.p0)
    lid = {int(a) for a, q in zip(n, e) if q == 0}
    assert (n[-1], e[-1]) == (0, 3)
    top = set(range(4 * 25, 125))
    assert top <= lid


def test_outlet_channel_boundary_conditions():
    """Outlet variant: p = 0 on every node of the x = 0 face and no velocity condition there except on
    its edges with the no-slip side faces; the inflow face is applied last; the closed box instead
    walls x = 0 and pins p at one node."""
    n = 6
    t = cases.inflow_cube_tets(n, outlet=True, amp=0.05)
    nodes, eqs, vals = cases.bc_lists(t, t.u0, t.p0)
    x = t.coords[:, 0]
    on_outlet = np.isclose(x, 0.0)
    side = np.zeros(t.n_nodes, bool)
    for ax in (1, 2):
        side |= np.isclose(t.coords[:, ax], 0.0) | np.isclose(t.coords[:, ax], 1.0)
    p_nodes = set(nodes[eqs == 3].tolist())
    assert p_nodes == set(np.nonzero(on_outlet)[0].tolist())
    vel_nodes = set(nodes[eqs == 0].tolist())
    assert not (vel_nodes & set(np.nonzero(on_outlet & ~side)[0].tolist()))
    inflow = np.nonzero(np.isclose(x, 1.0))[0]
    # the inflow region comes last: its nodes' last x-velocity condition is the inflow profile
    last = {}
    for a, q, v in zip(nodes, eqs, vals):
        if q == 0:
            last[int(a)] = v
    y, z = t.coords[inflow, 1], t.coords[inflow, 2]
    prof = -16 * y * (1 - y) * z * (1 - z) * (1 + 0.05 * np.sin(np.pi * y) * np.sin(np.pi * z))
    assert np.allclose([last[int(i)] for i in inflow], prof - t.u0[inflow, 0], rtol=0, atol=1e-15)
    c = cases.inflow_cube_tets(n)
    cn, ce, _ = cases.bc_lists(c, c.u0, c.p0)
    assert (ce == 3).sum() == 1 and np.isclose(c.coords[cn[ce == 3][0], 0], 0.0)
    assert set(cn[ce == 0].tolist()) >= set(np.nonzero(np.isclose(c.coords[:, 0], 0.0))[0].tolist())

"""The options of NavierStokesSemiImplicit (ns_semi_implicit/implementation/NavierStokesSemiImplicit.cpp:
309-317): enable_body_force (the body_force field in the velocity rhs, MatrixAssembly.hpp:339-363,
:460-477) and pressure_rcg_solve (two RCG strategies for the first and the later pressure solves,
:158-166, :532-537).

CPU: the oracle's body-force term is the consistent mass matrix applied to the nodal force, with the
mass matrix from textbook closed forms (simplices) or a 3-point-per-direction Gauss rule written here
(quadrilaterals, hexahedra on affine meshes, where the second-order rule of the element is exact).
GPU: the device URHS action with the force against the oracle within 1e-12 on all four element types,
and Taylor-Green steps with a force and the RCG pressure solves against the oracle's direct solves."""
import numpy as np
import pytest

from coolfluid3_amd import cabi, cases, meshgen
from oracle import pyoracle
from oracle.semi_implicit import SemiImplicit


def _mesh(name):
    if name == "triag":
        c, e = meshgen.rect_tris(4, 3, 1.0, 0.8)
        return pyoracle.TRIAG2D, np.asarray(c, float), np.asarray(e, np.uint32)
    if name == "quad":
        c, e = meshgen.rect_quads(4, 3, 1.0, 0.8)
        return pyoracle.QUAD2D, np.asarray(c, float), np.asarray(e, np.uint32)
    if name == "tetra":
        case = cases.inflow_cube_tets(2, nu=1e-2, dt=1e-2)
        return case.etype, case.coords, case.conn
    case = cases.lid_cavity_cube_hexas(2, nu=1e-2, dt=1e-2, jitter=0.0)
    return case.etype, case.coords, case.conn


def _linear_force(coords, dim):
    B = np.array([[0.3, -1.2, 0.7], [2.0, 0.4, -0.5], [-0.8, 1.1, 0.9]])[:dim, :dim]
    c = np.array([1.5, -0.25, 0.6])[:dim]
    return c + coords[:, :dim] @ B.T


def _mass_matrix(etype, coords, conn):
    """Consistent mass matrix int N_a N_b (dense, nodes x nodes)."""
    n = coords.shape[0]
    M = np.zeros((n, n))
    g3 = np.array([-np.sqrt(0.6), 0.0, np.sqrt(0.6)])
    w3 = np.array([5.0, 8.0, 5.0]) / 9.0
    for el in conn:
        X = coords[el]
        if etype == pyoracle.TRIAG2D:
            vol = 0.5 * abs(np.linalg.det(np.array([X[1] - X[0], X[2] - X[0]])))
            Me = vol / 12.0 * (np.ones((3, 3)) + np.eye(3))
        elif etype == pyoracle.TETRA3D:
            vol = abs(np.linalg.det(np.array([X[1] - X[0], X[2] - X[0], X[3] - X[0]]))) / 6.0
            Me = vol / 20.0 * (np.ones((4, 4)) + np.eye(4))
        elif etype == pyoracle.QUAD2D:
            sx, sy = np.array([-1, 1, 1, -1]), np.array([-1, -1, 1, 1])
            Me = np.zeros((4, 4))
            for i, s in enumerate(g3):
                for j, t in enumerate(g3):
                    N = 0.25 * (1 + sx * s) * (1 + sy * t)
                    dN = np.array([0.25 * sx * (1 + sy * t), 0.25 * sy * (1 + sx * s)])
                    Me += w3[i] * w3[j] * abs(np.linalg.det(dN @ X)) * np.outer(N, N)
        else:
            sx = np.array([-1, 1, 1, -1, -1, 1, 1, -1])
            sy = np.array([-1, -1, 1, 1, -1, -1, 1, 1])
            sz = np.array([-1, -1, -1, -1, 1, 1, 1, 1])
            Me = np.zeros((8, 8))
            for i, s in enumerate(g3):
                for j, t in enumerate(g3):
                    for k, u in enumerate(g3):
                        a, b, c = 1 + sx * s, 1 + sy * t, 1 + sz * u
                        N = 0.125 * a * b * c
                        dN = 0.125 * np.array([sx * b * c, sy * a * c, sz * a * b])
                        Me += w3[i] * w3[j] * w3[k] * abs(np.linalg.det(dN @ X)) * np.outer(N, N)
        M[np.ix_(el, el)] += Me
    return M


@pytest.mark.parametrize("name", ["triag", "quad", "tetra", "hexa"])
def test_oracle_body_force_is_the_consistent_mass_times_the_force(name):
    etype, coords, conn = _mesh(name)
    n, dim = coords.shape[0], pyoracle.DIM[etype]
    g = _linear_force(coords, dim)
    zeros_u = np.zeros((n, dim))
    si = SemiImplicit(pyoracle.get(), etype, coords, conn, np.arange(n, dtype=np.uint32), 1e-2, 1e-2, 1.0, 2,
                      zeros_u, np.zeros(n), body_force=g)
    z = np.zeros(n * dim)
    rhs = si.velocity_rhs(z, z, np.zeros(n), np.zeros(n)).reshape(n, dim)
    M = _mass_matrix(etype, coords, conn)
    expect = M @ g
    assert np.abs(rhs - expect).max() <= 1e-13 * np.abs(expect).max()
    # without the force the same inputs give zero
    si.g = None
    assert not si.velocity_rhs(z, z, np.zeros(n), np.zeros(n)).any()


# ------------------------------------------------------------------------------------------------
# device
# ------------------------------------------------------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("name", ["triag", "quad", "tetra", "hexa"])
def test_device_body_force_urhs_vs_oracle(name):
    etype, coords, conn = _mesh(name)
    n, dim = coords.shape[0], pyoracle.DIM[etype]
    rng = np.random.default_rng(5)
    u0 = rng.normal(size=(n, dim)) * 0.3
    p0 = rng.normal(size=n)
    g = _linear_force(coords, dim) + rng.normal(size=(n, dim)) * 0.1
    si = SemiImplicit(pyoracle.get(), etype, coords, conn, np.arange(n, dtype=np.uint32), 1e-2, 1e-2, 1.0, 2, u0, p0,
                      body_force=g)
    ctx = cabi.Context(0)
    mesh = cabi.Mesh(ctx, etype, coords, conn)
    us, ps = cabi.System(ctx, mesh, neq=dim), cabi.System(ctx, mesh, neq=1)
    f = cabi.Fields(mesh)
    f.set(cabi.FIELD_U, u0)
    f.set(cabi.FIELD_P, p0)
    f.set(cabi.FIELD_G, g)
    f.initialize(1e-2)
    prm = cabi.ns_params(dt=1e-2, theta=1.0)
    dsi = cabi.SemiImplicit(us, ps)
    try:
        assert dsi.options() == {"enable_body_force": False, "pressure_rcg_solve": False}
        dsi.set_options(enable_body_force=True)
        assert dsi.options() == {"enable_body_force": True, "pressure_rcg_solve": False}
        dsi.initial(f, prm)
        si.uadv = si.u.copy()
        a, U, P, dps = rng.normal(size=dim * n), rng.normal(size=dim * n), rng.normal(size=n), rng.normal(size=n)
        for w, v in ((cabi.SI_VEC_A, a), (cabi.SI_VEC_U, U), (cabi.SI_VEC_P, P), (cabi.SI_VEC_DPS, dps)):
            dsi.set(w, v)
        us.reset_vector(cabi.RHS)
        dsi.action(f, prm, cabi.SI_URHS)
        expect = si.velocity_rhs(a, U, dps, P)
        assert np.abs(us.vector(cabi.RHS) - expect).max() <= 1e-12 * np.abs(expect).max()
        # the option off: the field is ignored
        dsi.set_options(enable_body_force=False)
        us.reset_vector(cabi.RHS)
        dsi.action(f, prm, cabi.SI_URHS)
        si.g = None
        expect = si.velocity_rhs(a, U, dps, P)
        assert np.abs(us.vector(cabi.RHS) - expect).max() <= 1e-12 * np.abs(expect).max()
    finally:
        dsi.close()
        for x in (f, ps, us, mesh, ctx):
            x.close()


@pytest.mark.gpu
def test_semi_implicit_laminar_channel_acceptance():
    """atest-ufem-navier-stokes-semi-implicit-laminar-channel-2d.py: both options on (a body force drives
    the periodic channel, the pressure solves by the two RCG strategies), theta 0.5, two inner passes,
    alpha_su 0, 50 steps of 0.5; the script's check: |u - y (2 - y)| <= 1e-2 and |v| <= 1e-3 at every
    node.  Velocity solves to 1e-6 (the script's Block CG tolerance), pressure solves to 1e-8."""
    case = cases.si_laminar_channel_case()
    ctx = cabi.Context(0)
    mesh = cabi.Mesh(ctx, case.etype, case.coords, case.conn)
    us = cabi.System(ctx, mesh, neq=2, periodic=case.periodic)
    ps = cabi.System(ctx, mesh, neq=1, periodic=case.periodic)
    f = cabi.Fields(mesh)
    f.set(cabi.FIELD_U, case.u0)
    f.set(cabi.FIELD_P, case.p0)
    f.set(cabi.FIELD_G, case.body_force)
    f.initialize(case.nu)
    prm = cabi.ns_params(dt=case.dt, theta=case.theta, alpha_su=0.0)
    dsi = cabi.SemiImplicit(us, ps)
    try:
        dsi.set_options(enable_body_force=True, pressure_rcg_solve=True)
        dsi.initial(f, prm)
        walls, center = case.bc_regions[0][1], case.bc_regions[1][1]
        usp = cabi.solver_params(cabi.BICGSTAB_JACOBI, tol=1e-6, max_iter=3000)
        psp = cabi.solver_params(cabi.RCG_JACOBI, tol=1e-8, max_iter=3000, restart=40, recycle=16)
        its = []
        for step in range(50):
            stats = dsi.step(f, prm, usp, psp, ubc=(walls, np.zeros((len(walls), 2))), pbc=(center, [10.0]))
            its.append([s.iterations for s in stats])
        u = f.get(cabi.FIELD_U)
        y = case.coords[:, 1]
        eu = np.abs(u[:, 0] - y * (2.0 - y)).max()
        ev = np.abs(u[:, 1]).max()
        print("pressure iterations (first, second pass): steps 1-3", its[:3], "last", its[-1], flush=True)
        print(f"max |u - y(2-y)| {eu:.3e}, max |v| {ev:.3e}", flush=True)
        assert eu <= 1e-2 and ev <= 1e-3
    finally:
        dsi.close()
        for x in (f, ps, us, mesh, ctx):
            x.close()


@pytest.mark.gpu
def test_device_steps_with_body_force_and_rcg_pressure_follow_oracle():
    """Taylor-Green quadrilaterals with a body force, both options on: four device steps (velocity by
    BiCGStab to 1e-13, pressure by the two RCG strategies to 1e-13) follow the oracle's direct solves
    within 1e-9; the recycle spaces carry over, so later steps' pressure solves take fewer iterations."""
    case = cases.taylor_green_case(24, element="quad")
    p2m = cases.periodic_p2m(*case.periodic)
    center = case.bc_regions[0][1]
    g = np.column_stack([np.sin(2 * np.pi * case.coords[:, 1]), 0.5 * np.cos(2 * np.pi * case.coords[:, 0])])
    si = SemiImplicit(pyoracle.get(), case.etype, case.coords, case.conn, p2m, case.nu, case.dt, case.theta, 2,
                      case.u0, case.p0, pbc_nodes=center, body_force=g)
    ctx = cabi.Context(0)
    mesh = cabi.Mesh(ctx, case.etype, case.coords, case.conn)
    us = cabi.System(ctx, mesh, neq=2, periodic=case.periodic)
    ps = cabi.System(ctx, mesh, neq=1, periodic=case.periodic)
    f = cabi.Fields(mesh)
    f.set(cabi.FIELD_U, case.u0)
    f.set(cabi.FIELD_P, case.p0)
    f.set(cabi.FIELD_G, g)
    f.initialize(case.nu)
    prm = cabi.ns_params(dt=case.dt, theta=case.theta)
    dsi = cabi.SemiImplicit(us, ps)
    try:
        dsi.set_options(enable_body_force=True, pressure_rcg_solve=True)
        dsi.initial(f, prm)
        usp = cabi.solver_params(cabi.BICGSTAB_JACOBI, tol=1e-13, max_iter=5000)
        psp = cabi.solver_params(cabi.RCG_JACOBI, tol=1e-13, max_iter=5000, restart=40, recycle=16)
        t = 0.0
        first_p = []
        for step in range(4):
            pv = cases.taylor_green_exact(case.coords[center], t + case.dt)[1]
            stats = dsi.step(f, prm, usp, psp, pbc=(center, pv))
            si.step(pv)
            t += case.dt
            u, p = f.get(cabi.FIELD_U), f.get(cabi.FIELD_P)
            du = np.abs(u - si.u).max() / np.abs(si.u).max()
            dpp = np.abs(p - si.p).max() / np.abs(si.p).max()
            its = [s.iterations for s in stats]
            print("step", step + 1, f"du {du:.2e} dp {dpp:.2e}", its, flush=True)
            assert du < 1e-9 and dpp < 1e-9
            assert all(s.converged for s in stats)
            first_p.append(stats[1].iterations)
        assert first_p[-1] < first_p[0]
    finally:
        dsi.close()
        for x in (f, ps, us, mesh, ctx):
            x.close()

"""Pins the oracle's NS element kernels (specialised P1 tri/tet, generic quadrature, tau, combine).

The reference's only comparison test of these kernels is commented out
(plugins/UFEM/test/utest-proto-ns-assembly.cpp:162-202), so the pins are mathematical:
  * closed forms of the P1 integrals (mass V/20(1+d), grad-grad V gradNi.gradNj) with gradients
    computed here by numpy from the nodes, independent of the oracle's normals;
  * METRIC tau on the unit reference tetrahedron, where gij is the identity;
  * specialised vs generic agreement for uniform fields, and the exact documented differences
    (stress-form diffusion, the vk/wk substitution of NavierStokesSpecializations.hpp:224-226,236-238);
  * generic hexahedron terms against an independent 4x4x4 tensor Gauss integration on affine hexes,
    and generic quadrilateral (Quad2D) terms against a 4x4 rule on parallelograms, with its shape
    functions, Gauss rule, volume and METRIC tau checked against closed forms;
  * generic prism (Prism3D) terms against a collapsed-Gauss triangle rule times a 4-point rule in zeta
    on affine prisms, and METRIC tau at the prism centre (1/3, 1/3, 0) against its closed form;
  * the combine step (p rows / theta, rhs formed before the division, node-major reorder).
"""
import numpy as np
import pytest

from oracle.pyoracle import (HEXA3D, PRISM3D, QUAD2D, SUPG_CF2, SUPG_METRIC, SUPG_TEZDUYAR, TETRA3D, TRIAG2D,
                             ns_params)


def p1_grads(nodes):
    d = nodes.shape[1]
    J = (nodes[1:] - nodes[0]).T            # x = x0 + J xi
    dN = np.hstack([-np.ones((d, 1)), np.eye(d)])   # d x k reference gradients
    G = np.linalg.inv(J).T @ dN                     # physical gradients, d x k
    vol = abs(np.linalg.det(J)) / (2 if d == 2 else 6)
    return G, vol


def rand_tet(rng, scale=1.0):
    base = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    while True:
        X = (base + rng.uniform(-0.2, 0.2, (4, 3))) * scale
        if np.linalg.det((X[1:] - X[0]).T) > 0.05 * scale ** 3:
            return X


def var_major_block(M, k, R, C):
    return M[R * k:(R + 1) * k, C * k:(C + 1) * k]


def test_tau_metric_unit_tet(oracle):
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    u = np.tile([0.3, -0.2, 0.5], (4, 1))
    nu, dt = 0.02, 0.01
    prm = ns_params(dt=dt)
    _, _, _, tau = oracle.ns_element(TETRA3D, prm, X, u, u, np.full(4, nu))
    adv = (u[0] ** 2).sum()
    tau_su = 1 / np.sqrt(4 / dt ** 2 + adv + 16 * 3 * nu ** 2)
    tau_bulk = np.sqrt(4 / dt ** 2 + adv + 16 * 3 * nu ** 2) / 3
    assert tau[1] == pytest.approx(tau_su, rel=1e-14)
    assert tau[0] == pytest.approx(tau_su, rel=1e-14)
    assert tau[2] == pytest.approx(tau_bulk, rel=1e-14)


def test_tet_specialised_closed_forms(oracle):
    rng = np.random.default_rng(11)
    X = rand_tet(rng)
    G, V = p1_grads(X)
    nu, dt, rho = 0.05, 0.02, 1.3
    prm = ns_params(rho=rho, dt=dt)
    zero = np.zeros((4, 3))
    A, T, a, tau = oracle.ns_element(TETRA3D, prm, X, zero, zero, np.full(4, nu))
    k = 4
    mass = V / 20 * (np.ones((4, 4)) + np.eye(4))
    for R in range(3):
        assert np.allclose(var_major_block(T, k, R, R), mass, rtol=1e-13, atol=0)
    # diffusion in stress form + bulk viscosity (u = 0, so no convection terms)
    tb = tau[2]
    for R in range(3):
        for C in range(3):
            exp = nu * V * (np.outer(G[R], G[C]) / 3 + (R == C) * G.T @ G) + tb * V * np.outer(G[R], G[C])
            assert np.allclose(var_major_block(A, k, R, C), exp, rtol=1e-12, atol=1e-14 * np.abs(exp).max())
        # pressure gradient: A(u_R, p) = int N_i dN_j/dx_R / rho = V/4 G[R, j] / rho
        exp = np.tile(V / 4 * G[R] / rho, (4, 1))
        assert np.allclose(var_major_block(A, k, R, 3), exp, rtol=1e-12)
        # continuity: A(p, u_C) = V/4 G[C, j]
        assert np.allclose(var_major_block(A, k, 3, R), np.tile(V / 4 * G[R], (4, 1)), rtol=1e-12)
    # PSPG Laplacian
    assert np.allclose(var_major_block(A, k, 3, 3), tau[0] / rho * V * G.T @ G, rtol=1e-12)
    assert np.all(a == 0)


@pytest.mark.parametrize("nu", [0.0, 0.03])
@pytest.mark.parametrize("vw_equal", [True, False])
def test_tet_specialised_vs_generic(oracle, nu, vw_equal):
    rng = np.random.default_rng(5)
    X = rand_tet(rng, 0.3)
    uu = np.array([0.7, -0.4, -0.4 if vw_equal else 0.9])
    u = np.tile(uu, (4, 1))
    nus = np.full(4, nu)
    A_s, T_s, _, tau_s = oracle.ns_element(TETRA3D, ns_params(dt=0.01), X, u, u, nus)
    A_g, T_g, _, tau_g = oracle.ns_element(TETRA3D, ns_params(dt=0.01, use_specializations=False), X, u, u, nus)
    assert np.allclose(tau_s, tau_g, rtol=1e-14)
    G, V = p1_grads(X)
    k = 4
    scale = np.abs(A_g).max()
    D = A_s - A_g
    expected = np.zeros_like(D)
    n = 3 * V * G.T                       # area vectors n_i = 3 V grad N_i
    un = n @ uu
    for R in range(3):
        for C in range(3):
            expected[R * k:(R + 1) * k, C * k:(C + 1) * k] += nu * V / 3 * np.outer(G[R], G[C])
    dv = uu[1] - uu[2]
    for C in range(3):
        # W rows use vk for the skew-symmetric convection (NavierStokesSpecializations.hpp:224-226)
        expected[2 * k:3 * k, C * k:(C + 1) * k] += np.outer(0.5 * (1 / 12 + tau_s[1] * un / (9 * V)), dv * n[:, C])
        # ZZ PSPG skew term uses vk (NavierStokesSpecializations.hpp:236-238)
        expected[3 * k:4 * k, C * k:(C + 1) * k] += 0.5 * tau_s[0] / (9 * V) * np.outer(n[:, 2], dv * n[:, C])
    assert np.abs(D - expected).max() <= 1e-12 * scale
    assert np.abs(T_s - T_g).max() <= 1e-12 * np.abs(T_g).max()


def test_tri_specialised_vs_generic(oracle):
    X = np.array([[0.1, 0.05], [0.9, 0.2], [0.3, 0.8]])
    uu = np.array([0.6, -0.3])
    u = np.tile(uu, (3, 1))
    nu = 0.02
    A_s, T_s, _, _ = oracle.ns_element(TRIAG2D, ns_params(), X, u, u, np.full(3, nu))
    A_g, T_g, _, _ = oracle.ns_element(TRIAG2D, ns_params(use_specializations=False), X, u, u, np.full(3, nu))
    G, V = p1_grads(X)
    k = 3
    expected = np.zeros_like(A_s)
    for R in range(2):
        for C in range(2):
            expected[R * k:(R + 1) * k, C * k:(C + 1) * k] = nu * V / 3 * np.outer(G[R], G[C])
    assert np.abs(A_s - A_g - expected).max() <= 1e-12 * np.abs(A_g).max()
    assert np.abs(T_s - T_g).max() <= 1e-12 * np.abs(T_g).max()
    # tri mass matrix V/12 (1 + delta)
    A0, T0, _, _ = oracle.ns_element(TRIAG2D, ns_params(), X, 0 * u, 0 * u, np.full(3, nu))
    assert np.allclose(T0[:3, :3], V / 12 * (np.ones((3, 3)) + np.eye(3)), rtol=1e-13)


# ---- generic hexahedron vs an independent 4x4x4 tensor Gauss integration on affine hexes ----
_LOC = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1.]])


def q1(xi):
    N = np.prod(1 + _LOC * xi, axis=1) / 8
    dN = np.zeros((3, 8))
    for d in range(3):
        o = [e for e in range(3) if e != d]
        dN[d] = _LOC[:, d] * (1 + _LOC[:, o[0]] * xi[o[0]]) * (1 + _LOC[:, o[1]] * xi[o[1]]) / 8
    return N, dN


def hex_reference_terms(X, ua, nu, tau_ps, tau_su, tau_bulk, rho):
    x, w = np.polynomial.legendre.leggauss(4)
    k, n = 8, 32
    A = np.zeros((n, n)); T = np.zeros((n, n))
    for i, a in enumerate(x):
        for j, b in enumerate(x):
            for l, c in enumerate(x):
                xi = np.array([a, b, c])
                N, dN = q1(xi)
                J = dN @ X                       # J(ksi, x)
                G = np.linalg.solve(J, dN)       # physical gradients, 3 x 8
                wq = w[i] * w[j] * w[l] * np.linalg.det(J)
                L = ua @ G
                SU = N + tau_su * L
                P = 3 * k
                for R in range(3):
                    A[P:, R * k:(R + 1) * k] += wq * (np.outer(N + 0.5 * tau_ps * L, G[R]) + tau_ps * np.outer(G[R], L))
                    A[R * k:(R + 1) * k, R * k:(R + 1) * k] += wq * (nu * G.T @ G + np.outer(SU, L))
                    A[R * k:(R + 1) * k, P:] += wq * np.outer(SU, G[R]) / rho
                    for C in range(3):
                        A[R * k:(R + 1) * k, C * k:(C + 1) * k] += wq * np.outer(tau_bulk * G[R] + 0.5 * ua[R] * SU, G[C])
                    T[P:, R * k:(R + 1) * k] += wq * tau_ps * np.outer(G[R], N)
                    T[R * k:(R + 1) * k, R * k:(R + 1) * k] += wq * np.outer(SU, N)
                A[P:, P:] += wq * tau_ps * G.T @ G / rho
    return A, T


@pytest.mark.parametrize("case", [0, 1])
def test_generic_hex_vs_tensor_gauss(oracle, case):
    rng = np.random.default_rng(40 + case)
    B = np.eye(3) * 0.2 + rng.uniform(-0.03, 0.03, (3, 3))   # affine map: Gauss 2x2x2 is exact
    X = (_LOC + 1) / 2 @ B.T + rng.uniform(0, 1, 3)
    ua = np.array([0.4, -0.7, 0.2]) * case
    nu, rho = 0.01, 1.1
    u = np.tile(ua, (8, 1))
    A, T, a, tau = oracle.ns_element(HEXA3D, ns_params(rho=rho, dt=0.02), X, u, u, np.full(8, nu))
    Ar, Tr = hex_reference_terms(X, ua, nu, tau[0], tau[1], tau[2], rho)
    assert np.abs(A - Ar).max() <= 1e-12 * np.abs(Ar).max()
    assert np.abs(T - Tr).max() <= 1e-12 * np.abs(Tr).max()
    # the mass matrix of every velocity block sums to the volume (partition of unity)
    if case == 0:
        assert T[:8, :8].sum() == pytest.approx(abs(np.linalg.det(B)), rel=1e-13)


# ---- generic prism (Prism3D) vs a collapsed-Gauss triangle rule x Gauss in zeta on affine prisms ----
def prism_sf(xi):
    """Wedge shape functions written from their definition: the P1 triangle (0,0), (1,0), (0,1) times
    the linear interval (1 - zeta)/2, (1 + zeta)/2, bottom nodes first."""
    t = np.array([1 - xi[0] - xi[1], xi[0], xi[1]])
    dt = np.array([[-1.0, 1.0, 0.0], [-1.0, 0.0, 1.0]])
    z = np.array([(1 - xi[2]) / 2, (1 + xi[2]) / 2])
    dz = np.array([-0.5, 0.5])
    N = np.concatenate([t * z[0], t * z[1]])
    dN = np.zeros((3, 6))
    dN[0] = np.concatenate([dt[0] * z[0], dt[0] * z[1]])
    dN[1] = np.concatenate([dt[1] * z[0], dt[1] * z[1]])
    dN[2] = np.concatenate([t * dz[0], t * dz[1]])
    return N, dN


def prism_reference_terms(X, ua, nu, tau_ps, tau_su, tau_bulk, rho):
    x, w = np.polynomial.legendre.leggauss(4)
    u01, w01 = (x + 1) / 2, w / 2                # Gauss-Legendre on [0, 1]
    k, n = 6, 24
    A = np.zeros((n, n)); T = np.zeros((n, n))
    for i, a in enumerate(u01):
        for j, b in enumerate(u01):
            for l, c in enumerate(x):
                xi = np.array([a, (1 - a) * b, c])   # collapsed map of the unit square onto the triangle
                N, dN = prism_sf(xi)
                J = dN @ X
                G = np.linalg.solve(J, dN)
                wq = w01[i] * w01[j] * (1 - a) * w[l] * np.linalg.det(J)
                L = ua @ G
                SU = N + tau_su * L
                P = 3 * k
                for R in range(3):
                    A[P:, R * k:(R + 1) * k] += wq * (np.outer(N + 0.5 * tau_ps * L, G[R]) + tau_ps * np.outer(G[R], L))
                    A[R * k:(R + 1) * k, R * k:(R + 1) * k] += wq * (nu * G.T @ G + np.outer(SU, L))
                    A[R * k:(R + 1) * k, P:] += wq * np.outer(SU, G[R]) / rho
                    for C in range(3):
                        A[R * k:(R + 1) * k, C * k:(C + 1) * k] += wq * np.outer(tau_bulk * G[R] + 0.5 * ua[R] * SU, G[C])
                    T[P:, R * k:(R + 1) * k] += wq * tau_ps * np.outer(G[R], N)
                    T[R * k:(R + 1) * k, R * k:(R + 1) * k] += wq * np.outer(SU, N)
                A[P:, P:] += wq * tau_ps * G.T @ G / rho
    return A, T


_PLOC = np.array([[0, 0, -1], [1, 0, -1], [0, 1, -1], [0, 0, 1], [1, 0, 1], [0, 1, 1.]])


@pytest.mark.parametrize("case", [0, 1])
def test_generic_prism_vs_collapsed_gauss(oracle, case):
    rng = np.random.default_rng(60 + case)
    B = np.eye(3) * 0.2 + rng.uniform(-0.03, 0.03, (3, 3))   # affine map: the 6-point rule is exact
    X = _PLOC @ B.T + rng.uniform(0, 1, 3)
    ua = np.array([0.4, -0.7, 0.2]) * case
    nu, rho = 0.01, 1.1
    u = np.tile(ua, (6, 1))
    A, T, a, tau = oracle.ns_element(PRISM3D, ns_params(rho=rho, dt=0.02, use_specializations=False), X, u, u,
                                     np.full(6, nu))
    Ar, Tr = prism_reference_terms(X, ua, nu, tau[0], tau[1], tau[2], rho)
    assert np.abs(A - Ar).max() <= 1e-12 * np.abs(Ar).max()
    assert np.abs(T - Tr).max() <= 1e-12 * np.abs(Tr).max()
    if case == 0:   # the mass matrix of a velocity block sums to the volume
        assert T[:6, :6].sum() == pytest.approx(abs(np.linalg.det(B)), rel=1e-13)


def test_tau_metric_sheared_prism(oracle):
    """Affine prism x = c + B xi: J = B^T everywhere, so gij = J^-T J^-1 = B^-1 B^-T at the centre."""
    h, s = 0.1, 0.04
    B = np.array([[h, s, 0], [0, h, 0], [0, 0, h / 2]])
    X = _PLOC @ B.T + np.array([0.3, 0.2, 0.1])
    Binv = np.array([[1 / h, -s / h ** 2, 0], [0, 1 / h, 0], [0, 0, 2 / h]])
    g = Binv @ Binv.T
    uc = np.array([0.4, 0.9, -0.2])
    nu, dt = 0.005, 0.01
    u = np.tile(uc, (6, 1))
    _, _, _, tau = oracle.ns_element(PRISM3D, ns_params(dt=dt, use_specializations=False), X, u, u, np.full(6, nu))
    ts, tb = _metric_tau(g, uc, nu, dt)
    assert tau[1] == pytest.approx(ts, rel=1e-13) and tau[2] == pytest.approx(tb, rel=1e-13)
    assert abs(_metric_tau(Binv.T @ Binv, uc, nu, dt)[0] - ts) > 1e-6 * ts


# ---- generic quadrilateral (Quad2D) ----
_QLOC = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1.]])


def q1_2d(xi):
    N = np.prod(1 + _QLOC * xi, axis=1) / 4
    dN = np.stack([_QLOC[:, 0] * (1 + _QLOC[:, 1] * xi[1]), _QLOC[:, 1] * (1 + _QLOC[:, 0] * xi[0])]) / 4
    return N, dN


def test_quad_shapes_gauss_volume(oracle):
    """Quad.cpp shape functions: Kronecker at the nodes, partition of unity, gradients equal to finite
    differences; the 2 x 2 rule (points (+-x, +-x) in the order of GaussImplementation.hpp:566-600,
    w = 1) integrates the Jacobian determinant of a non-affine quadrilateral to its shoelace area, which
    Quad2D::volume gives too."""
    for j, node in enumerate(_QLOC):
        assert np.allclose(oracle.shape_values(QUAD2D, node), np.eye(4)[j], atol=0)
    rng = np.random.default_rng(5)
    for _ in range(5):
        xi = rng.uniform(-1, 1, 2)
        N = oracle.shape_values(QUAD2D, xi)
        assert abs(N.sum() - 1) < 1e-15 and np.allclose(N, q1_2d(xi)[0], rtol=1e-15)
        h = 1e-6
        fd = np.stack([(oracle.shape_values(QUAD2D, xi + h * e) - oracle.shape_values(QUAD2D, xi - h * e)) / (2 * h)
                       for e in np.eye(2)])
        assert np.allclose(oracle.shape_gradients(QUAD2D, xi), fd, atol=1e-9)
    c, w = oracle.gauss_points(QUAD2D, 2)
    x = 1 / np.sqrt(3)
    assert np.allclose(c, [[x, x], [-x, x], [x, -x], [-x, -x]], rtol=1e-15) and np.all(w == 1)
    X = np.array([[0.0, 0.0], [1.2, 0.1], [1.0, 0.9], [-0.1, 0.8]])
    area = 0.5 * abs(sum(X[i, 0] * X[(i + 1) % 4, 1] - X[(i + 1) % 4, 0] * X[i, 1] for i in range(4)))
    assert oracle.element_volume(QUAD2D, X) == pytest.approx(area, rel=1e-15)
    integ = sum(wq * np.linalg.det(oracle.element_jacobian(QUAD2D, cq, X)) for cq, wq in zip(c, w))
    assert integ == pytest.approx(area, rel=1e-14)


def test_quad_tau_metric_square(oracle):
    """METRIC tau on the square of side h: J = h/2 I at the centre, so gij = 4/h^2 I."""
    h, dt, nu = 0.05, 0.01, 2e-3
    X = (_QLOC + 1) / 2 * h
    uc = np.array([0.7, -0.4])
    u = np.tile(uc, (4, 1))
    _, _, _, tau = oracle.ns_element(QUAD2D, ns_params(dt=dt, use_specializations=False), X, u, u, np.full(4, nu))
    g = 4 / h ** 2
    adv = g * (uc @ uc)
    diff = nu ** 2 * 2 * g ** 2
    assert tau[1] == pytest.approx(1 / np.sqrt(4 / dt ** 2 + adv + 16 * diff), rel=1e-14)
    assert tau[0] == tau[1]
    assert tau[2] == pytest.approx(np.sqrt(4 / dt ** 2 + adv + 16 * diff) / (2 * g), rel=1e-14)


def quad_reference_terms(X, ua, nu, tau_ps, tau_su, tau_bulk, rho):
    x, w = np.polynomial.legendre.leggauss(4)
    k, n, P = 4, 12, 8
    A = np.zeros((n, n)); T = np.zeros((n, n))
    for i, a in enumerate(x):
        for j, b in enumerate(x):
            N, dN = q1_2d(np.array([a, b]))
            J = dN @ X
            G = np.linalg.solve(J, dN)
            wq = w[i] * w[j] * np.linalg.det(J)
            L = ua @ G
            SU = N + tau_su * L
            for R in range(2):
                A[P:, R * k:(R + 1) * k] += wq * (np.outer(N + 0.5 * tau_ps * L, G[R]) + tau_ps * np.outer(G[R], L))
                A[R * k:(R + 1) * k, R * k:(R + 1) * k] += wq * (nu * G.T @ G + np.outer(SU, L))
                A[R * k:(R + 1) * k, P:] += wq * np.outer(SU, G[R]) / rho
                for C in range(2):
                    A[R * k:(R + 1) * k, C * k:(C + 1) * k] += wq * np.outer(tau_bulk * G[R] + 0.5 * ua[R] * SU, G[C])
                T[P:, R * k:(R + 1) * k] += wq * tau_ps * np.outer(G[R], N)
                T[R * k:(R + 1) * k, R * k:(R + 1) * k] += wq * np.outer(SU, N)
            A[P:, P:] += wq * tau_ps * G.T @ G / rho
    return A, T


@pytest.mark.parametrize("case", [0, 1])
def test_generic_quad_vs_tensor_gauss(oracle, case):
    rng = np.random.default_rng(60 + case)
    B = np.eye(2) * 0.2 + rng.uniform(-0.04, 0.04, (2, 2))   # parallelogram: the 2 x 2 rule is exact
    X = (_QLOC + 1) / 2 @ B.T + rng.uniform(0, 1, 2)
    ua = np.array([0.4, -0.7]) * case
    nu, rho = 0.01, 1.1
    u = np.tile(ua, (4, 1))
    A, T, a, tau = oracle.ns_element(QUAD2D, ns_params(rho=rho, dt=0.02, use_specializations=False), X, u, u,
                                     np.full(4, nu))
    Ar, Tr = quad_reference_terms(X, ua, nu, tau[0], tau[1], tau[2], rho)
    assert np.abs(A - Ar).max() <= 1e-12 * np.abs(Ar).max()
    assert np.abs(T - Tr).max() <= 1e-12 * np.abs(Tr).max()
    if case == 0:   # mass matrix of a parallelogram: |det B| / 36 [4 2 1 2; ...]
        m = np.array([[4, 2, 1, 2], [2, 4, 2, 1], [1, 2, 4, 2], [2, 1, 2, 4]]) / 36 * abs(np.linalg.det(B))
        assert np.allclose(T[:4, :4], m, rtol=1e-13)


def test_hex_body_force_and_tezduyar(oracle):
    X = (_LOC + 1) / 2 * 0.1
    g = np.tile([0.0, 0.0, -9.81], (8, 1))
    zero = np.zeros((8, 3))
    _, _, a, _ = oracle.ns_element(HEXA3D, ns_params(), X, zero, zero, np.full(8, 1e-3), g)
    # int N_r g dV = g V / 8 per node for a cube when u_adv = 0
    assert np.allclose(a[16:24], -9.81 * 1e-3 / 8, rtol=1e-13)
    assert np.all(a[:16] == 0)
    # TEZDUYAR tau (SUPG.hpp:149-175): h = min edge, tau_su = 1/(2|u|/h + 2/dt + 4 nu/h^2)
    u = np.tile([1.0, 0.0, 0.0], (8, 1))
    prm = ns_params(dt=0.01, use_specializations=False, supg_type=SUPG_TEZDUYAR)
    _, _, _, tau = oracle.ns_element(HEXA3D, prm, X, u, u, np.full(8, 1e-3))
    h = 0.1
    ts = 1 / (2 / h + 2 / 0.01 + 4e-3 / h ** 2)
    assert tau[1] == pytest.approx(ts, rel=1e-14) and tau[2] == pytest.approx(ts, rel=1e-14)


@pytest.mark.parametrize("etype", [TRIAG2D, TETRA3D, HEXA3D])
def test_combine_theta_invariants(oracle, etype):
    """Invariants of the combine step (NavierStokesAssembly.hpp:69-71) over theta, with no restated formula:
      * the rhs -A x + a is formed before the p rows are divided by theta: it does not depend on theta;
      * the p rows invdt T_p + theta (A_p / theta) do not depend on theta;
      * the velocity rows invdt T_u + theta A_u are affine in theta, and their slope S = A_u gives the
        velocity rhs exactly: rhs_u = -S x (no body force), with x the element (u, p) in node-major order;
      * the intercept invdt T_u is block diagonal in the velocity components (T[u_i, u_j] = 0 for i != j,
        T[u_i, p] = 0), the same for every component, and sums to V / dt over one component's block
        (sum_a N_a = 1, so the SUPG part of sum_ab (N_a + tau L_a) N_b integrates to 0);
      * the slope is the var-major A of the element kernel moved to node-major order by the block
        assignment row -> (row % k) neq + row / k (BlockAccumulator.hpp:134-184)."""
    rng = np.random.default_rng(etype + 100)
    k, dim = {TRIAG2D: (3, 2), TETRA3D: (4, 3), HEXA3D: (8, 3)}[etype]
    if etype == TRIAG2D:
        X = np.array([[0, 0], [1, 0.1], [0.2, 0.9]]) * 0.1
    elif etype == TETRA3D:
        X = rand_tet(rng, 0.1)
    else:
        X = (_LOC + 1) / 2 * 0.1 + rng.uniform(-0.005, 0.005, (8, 3))
    u = rng.uniform(-1, 1, (k, dim)); ua = rng.uniform(-1, 1, (k, dim))
    p = rng.uniform(-1, 1, k); nu = np.full(k, 0.01)
    dt = 0.01
    neq = dim + 1
    thetas = (0.3, 0.6, 1.0)
    sys_ = [oracle.ns_element_system(etype, ns_params(dt=dt, theta=t, use_specializations=etype != HEXA3D),
                                     X, u, p, ua, nu) for t in thetas]
    Ms = [m for m, _ in sys_]
    rs = [r for _, r in sys_]
    scale = np.abs(Ms[2]).max()
    assert np.array_equal(rs[0], rs[1]) and np.array_equal(rs[0], rs[2])
    prow = np.array([a * neq + dim for a in range(k)])
    urow = np.array([a * neq + i for a in range(k) for i in range(dim)])
    for m in Ms[:2]:
        assert np.abs(m[prow] - Ms[2][prow]).max() <= 1e-14 * scale
    S = (Ms[2][urow] - Ms[0][urow]) / (thetas[2] - thetas[0])
    assert np.abs((Ms[1][urow] - Ms[0][urow]) - (thetas[1] - thetas[0]) * S).max() <= 1e-13 * scale
    x = np.concatenate([np.concatenate([u[a], [p[a]]]) for a in range(k)])
    assert np.abs(rs[0][urow] + S @ x).max() <= 1e-11 * (np.abs(S) @ np.abs(x)).max()
    T = (Ms[0][urow] - thetas[0] * S) * dt          # T_u, rows urow, all columns (node-major)
    if etype == HEXA3D:
        # the Gauss-rule volume sum_q det J_q (exact for a trilinear map), from this file's own Q1 shapes
        g1 = 1 / np.sqrt(3)
        V = sum(np.linalg.det(q1(np.array([x, y, z]))[1] @ X) for x in (-g1, g1) for y in (-g1, g1) for z in (-g1, g1))
    else:
        V = oracle.element_volume(etype, X)
    for i in range(dim):
        ri = np.array([a * dim + i for a in range(k)])
        for j in range(neq):
            cj = np.array([b * neq + j for b in range(k)])
            blk = T[np.ix_(ri, cj)]
            if j == i:
                assert blk.sum() == pytest.approx(V, rel=1e-10)
                assert np.abs(blk - T[np.ix_(np.array([a * dim for a in range(k)]),
                                              np.array([b * neq for b in range(k)]))]).max() <= 1e-10 * V
            else:
                assert np.abs(blk).max() <= 1e-10 * V
    A, _, _, _ = oracle.ns_element(etype, ns_params(dt=dt, use_specializations=etype != HEXA3D), X, u, ua, nu)
    perm = np.array([(r % k) * neq + r // k for r in range(k * neq)])   # var-major -> node-major
    A_nm = np.zeros_like(A); A_nm[np.ix_(perm, perm)] = A
    assert np.abs(S - A_nm[urow]).max() <= 1e-11 * np.abs(A).max()


def _metric_tau(g, uc, nu, dt):
    """tau_su and tau_bulk from an explicit metric g (SUPG.hpp:176-190, c1 = 1, c2 = 16)."""
    adv = abs(uc @ g @ uc)
    diff = nu ** 2 * (g ** 2).sum()
    s = np.sqrt(4 / dt ** 2 + adv + 16 * diff)
    return 1 / s, s / np.trace(g)


@pytest.mark.parametrize("spec", [True, False])
def test_tau_metric_sheared_tet(oracle, spec):
    """METRIC tau on a sheared tetrahedron, nodes 0, (a,0,0), (s,b,0), (0,0,c): J (rows d x / d xi_i) is
    [[a,0,0],[s,b,0],[0,0,c]], whose inverse is written out by hand below, and gij = J^-T J^-1
    (SUPG.hpp:178).  The transposed product J^-1 J^-T has the same trace and Frobenius norm but gives
    another u g u^T for a velocity in the sheared plane, so an operand swap fails here."""
    a, b, c, s = 0.2, 0.1, 0.15, 0.12
    X = np.array([[0, 0, 0], [a, 0, 0], [s, b, 0], [0, 0, c]])
    Jinv = np.array([[1 / a, 0, 0], [-s / (a * b), 1 / b, 0], [0, 0, 1 / c]])
    g = np.array([[1 / a ** 2 + s ** 2 / (a * b) ** 2, -s / (a * b ** 2), 0],
                  [-s / (a * b ** 2), 1 / b ** 2, 0],
                  [0, 0, 1 / c ** 2]])
    assert np.allclose(Jinv.T @ Jinv, g, rtol=1e-14)   # the hand-written forms agree with each other
    uc = np.array([0.8, -0.5, 0.3])
    nu, dt = 0.01, 0.02
    u = np.tile(uc, (4, 1))
    _, _, _, tau = oracle.ns_element(TETRA3D, ns_params(dt=dt, use_specializations=spec), X, u, u, np.full(4, nu))
    ts, tb = _metric_tau(g, uc, nu, dt)
    assert tau[1] == pytest.approx(ts, rel=1e-13) and tau[0] == tau[1]
    assert tau[2] == pytest.approx(tb, rel=1e-13)
    gt = Jinv @ Jinv.T
    assert abs(_metric_tau(gt, uc, nu, dt)[0] - ts) > 1e-6 * ts   # the transposed metric is told apart


def test_tau_metric_sheared_tri(oracle):
    """Sheared triangle 0, (a,0), (s,b): J = [[a,0],[s,b]], gij = J^-T J^-1 written out by hand."""
    a, b, s = 0.2, 0.1, 0.13
    X = np.array([[0, 0], [a, 0], [s, b]])
    g = np.array([[1 / a ** 2 + s ** 2 / (a * b) ** 2, -s / (a * b ** 2)], [-s / (a * b ** 2), 1 / b ** 2]])
    uc = np.array([0.7, -0.6])
    nu, dt = 0.02, 0.01
    u = np.tile(uc, (3, 1))
    for spec in (True, False):
        _, _, _, tau = oracle.ns_element(TRIAG2D, ns_params(dt=dt, use_specializations=spec), X, u, u, np.full(3, nu))
        ts, tb = _metric_tau(g, uc, nu, dt)
        assert tau[1] == pytest.approx(ts, rel=1e-13) and tau[2] == pytest.approx(tb, rel=1e-13)


def test_tau_metric_sheared_hex(oracle):
    """Parallelepiped x = c + B xi / 2 (xi in [-1,1]^3) with B = [[h,s,0],[0,h,0],[0,0,h]]: J = B^T / 2 at
    the centre, so gij = J^-T J^-1 = 4 B^-1 B^-T, with B^-1 = [[1/h, -s/h^2, 0],[0,1/h,0],[0,0,1/h]]."""
    h, s = 0.1, 0.04
    B = np.array([[h, s, 0], [0, h, 0], [0, 0, h]])
    X = 0.5 * _LOC @ B.T + np.array([0.3, 0.2, 0.1])
    Binv = np.array([[1 / h, -s / h ** 2, 0], [0, 1 / h, 0], [0, 0, 1 / h]])
    g = 4 * Binv @ Binv.T
    uc = np.array([0.4, 0.9, -0.2])
    nu, dt = 0.005, 0.01
    u = np.tile(uc, (8, 1))
    _, _, _, tau = oracle.ns_element(HEXA3D, ns_params(dt=dt, use_specializations=False), X, u, u, np.full(8, nu))
    ts, tb = _metric_tau(g, uc, nu, dt)
    assert tau[1] == pytest.approx(ts, rel=1e-13) and tau[2] == pytest.approx(tb, rel=1e-13)
    assert abs(_metric_tau(4 * Binv.T @ Binv, uc, nu, dt)[0] - ts) > 1e-6 * ts


@pytest.mark.parametrize("etype", [TETRA3D, TRIAG2D, HEXA3D, QUAD2D])
@pytest.mark.parametrize("nu", [0.02, 0.1])
def test_tau_cf2_closed_forms(oracle, etype, nu):
    """CF2 tau (SUPG.hpp:191-210) on a reference-shaped element: he from the volume (pi written as
    3.141592654 there), Re_e = u_ref he / (2 nu), xi = min(Re_e / 3, 1), tau_ps = he xi / (2 u_ref),
    tau_bulk = he u_ref / xi; tau_su = 1 / (1/tau_adv + 2/dt + 1/tau_diff) with h = 2|u| / sum_n u.grad N_n,
    where the gradients sum to zero, so tau_su is dt / 2 up to rounding; and 0 for |u| <= 1e-10."""
    pi = 3.141592654
    k, dim = {TETRA3D: (4, 3), TRIAG2D: (3, 2), HEXA3D: (8, 3), QUAD2D: (4, 2)}[etype]
    h = 0.1
    if etype == TETRA3D:
        X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]]) * h
        V = h ** 3 / 6
    elif etype == TRIAG2D:
        X = np.array([[0, 0], [1, 0], [0, 1.0]]) * h
        V = h ** 2 / 2
    elif etype == HEXA3D:
        X = (_LOC + 1) / 2 * h
        V = h ** 3
    else:
        X = (_QLOC + 1) / 2 * h
        V = h ** 2
    he = np.sqrt(4 / pi * V) if dim == 2 else (3 / 4 / pi * V) ** (1 / 3)
    ree = he / (2 * nu)
    xi = min(ree / 3, 1.0)
    dt = 0.01
    prm = ns_params(dt=dt, use_specializations=False, supg_type=SUPG_CF2)
    u = np.tile([0.7, -0.4, 0.2][:dim], (k, 1))
    _, _, _, tau = oracle.ns_element(etype, prm, X, u, u, np.full(k, nu))
    assert tau[0] == pytest.approx(he * xi / 2, rel=1e-14)
    assert tau[2] == pytest.approx(he / xi, rel=1e-14)
    assert tau[1] == pytest.approx(dt / 2, rel=1e-13)
    _, _, _, tau0 = oracle.ns_element(etype, prm, X, 0 * u, 0 * u, np.full(k, nu))
    assert tau0[1] == 0.0 and tau0[0] == tau[0] and tau0[2] == tau[2]

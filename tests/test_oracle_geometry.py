"""Pins the oracle's element geometry, shape functions and Gauss rules.

Known answers come from the reference's own unit tests (tests/golden/reference_fixtures.json):
test/mesh/utest-mesh-lagrangep1-tetra3d.cpp, -hexa3d.cpp, -triag2d.cpp, -prism3d.cpp.  Gauss rules are pinned by
exactness (a plausible wrong abscissa or weight fails the polynomial integrals).
"""
import numpy as np
import pytest

from oracle.pyoracle import HEXA3D, PRISM3D, TETRA3D, TRIAG2D


def ulps(a, b):
    a = np.asarray(a, float).ravel(); b = np.asarray(b, float).ravel()
    return np.max(np.abs(a - b) / np.spacing(np.maximum(np.abs(b), 1e-300)))


def test_tetra_fixture(oracle, fixtures):
    f = fixtures["tetra3d"]
    nodes = np.array(f["nodes"])
    # utest: BOOST_CHECK_CLOSE(volume, 1.4518e-4, 1e-5 %)
    assert abs(oracle.tetra_volume(nodes) - f["volume"]) <= 1e-7 * f["volume"]
    assert ulps(oracle.tetra_jacobian(nodes), f["jacobian"]) < 2
    assert ulps(oracle.tetra_jacobian_adjoint(nodes), f["jacobian_adjoint"]) < 5
    assert ulps(oracle.shape_values(TETRA3D, f["mapped_coords"]), f["shape_values"]) < 10


def test_hexa_fixture(oracle, fixtures):
    f = fixtures["hexa3d"]
    m = f["mapped_coords"]
    unit, skew = np.array(f["unit_nodes"], float), np.array(f["skewed_nodes"], float)
    assert ulps(oracle.shape_values(HEXA3D, m), f["shape_values"]) < 10
    assert ulps(oracle.shape_gradients(HEXA3D, m), f["shape_gradients"]) < 2
    assert oracle.hexa_jacobian_determinant(m, unit) == pytest.approx(f["det_unit"], rel=1e-7)
    assert oracle.hexa_jacobian_determinant(m, skew) == pytest.approx(f["det_skewed"], rel=1e-7)
    assert ulps(oracle.hexa_jacobian(m, skew), f["jacobian_skewed"]) < 2
    assert ulps(oracle.hexa_jacobian_adjoint(m, skew), f["adjoint_skewed"]) < 2
    assert oracle.hexa_volume(unit) == f["volume_unit"]
    assert oracle.hexa_volume(skew) == pytest.approx(f["volume_skewed"], rel=1e-6)
    # determinant formula against the Jacobian matrix at random mapped points
    rng = np.random.default_rng(3)
    for _ in range(5):
        mm = rng.uniform(-1, 1, 3)
        assert oracle.hexa_jacobian_determinant(mm, skew) == pytest.approx(
            np.linalg.det(oracle.hexa_jacobian(mm, skew)), rel=1e-12)


def test_prism_fixture(oracle, fixtures):
    f = fixtures["prism3d"]
    m = f["mapped_coords"]
    unit, skew = np.array(f["unit_nodes"], float), np.array(f["skewed_nodes"], float)
    assert ulps(oracle.shape_values(PRISM3D, m), f["shape_values"]) < 10
    assert ulps(oracle.shape_gradients(PRISM3D, m), f["shape_gradients"]) < 3
    J = oracle.element_jacobian(PRISM3D, m, skew)
    assert ulps(J, f["jacobian_skewed"]) < 2
    # compute_jacobian_adjoint = inverse * det (Prism3D.cpp:131-140)
    assert ulps(np.linalg.inv(J) * np.linalg.det(J), f["adjoint_skewed"]) < 2
    assert np.linalg.det(oracle.element_jacobian(PRISM3D, m, unit)) == pytest.approx(f["det_unit"], rel=1e-7)
    assert np.linalg.det(J) == pytest.approx(f["det_skewed"], rel=1e-7)
    assert oracle.element_volume(PRISM3D, unit) == pytest.approx(f["volume_unit"], rel=1e-6)
    assert oracle.element_volume(PRISM3D, skew) == pytest.approx(f["volume_skewed"], rel=1e-6)
    # gauss_integrate of det J at the fixture's mapped point (a constant) with both rules: 0.5 times the
    # sum of the weights, which is the reference prism's volume 1
    for order, key in ((1, "integral_det_order1"), (2, "integral_det_order2")):
        _, w = oracle.gauss_points(PRISM3D, order)
        assert w.sum() * np.linalg.det(J) == pytest.approx(f[key], rel=1e-8)
    # the Jacobian is the shape-function gradients applied to the nodes, at random points
    rng = np.random.default_rng(5)
    for _ in range(4):
        mm = np.array([rng.uniform(0, 0.5), rng.uniform(0, 0.5), rng.uniform(-1, 1)])
        assert np.allclose(oracle.element_jacobian(PRISM3D, mm, skew), oracle.shape_gradients(PRISM3D, mm) @ skew,
                           rtol=0, atol=1e-15)


def test_triag_fixture(oracle, fixtures):
    f = fixtures["triag2d"]
    nodes = np.array(f["nodes"])
    assert ulps(oracle.element_jacobian(TRIAG2D, [0.1, 0.8], nodes), f["jacobian"]) < 2
    for nd, vol in f["volume_cases"]:
        assert oracle.element_volume(TRIAG2D, np.array(nd, float)) == vol


@pytest.mark.parametrize("etype", [TRIAG2D, TETRA3D, HEXA3D, PRISM3D])
def test_shape_partition_of_unity(oracle, etype):
    rng = np.random.default_rng(etype)
    dim = 2 if etype == TRIAG2D else 3
    for _ in range(4):
        m = rng.uniform(0, 0.3, dim)
        assert abs(oracle.shape_values(etype, m).sum() - 1) < 1e-15
        assert np.abs(oracle.shape_gradients(etype, m).sum(axis=1)).max() < 1e-15


def test_gauss_exactness(oracle):
    # triangle order 2 (3 edge midpoints, w=1/6): exact for degree 2 on the reference triangle
    c, w = oracle.gauss_points(TRIAG2D, 2)
    assert w.sum() == pytest.approx(0.5, abs=1e-15)
    assert (w * c[:, 0] ** 2).sum() == pytest.approx(1 / 12, abs=1e-15)
    assert (w * c[:, 0] * c[:, 1]).sum() == pytest.approx(1 / 24, abs=1e-15)
    # tetrahedron order 2 (4 points, w=1/24): integral of xi^2 = 1/60, xi*eta = 1/120
    c, w = oracle.gauss_points(TETRA3D, 2)
    assert w.sum() == pytest.approx(1 / 6, abs=1e-15)
    assert (w * c[:, 0] ** 2).sum() == pytest.approx(1 / 60, abs=1e-15)
    assert (w * c[:, 0] * c[:, 1]).sum() == pytest.approx(1 / 120, abs=1e-15)
    # hexahedron 2x2x2: exact for degree 3 per axis over [-1,1]^3
    c, w = oracle.gauss_points(HEXA3D, 2)
    assert w.sum() == 8.0
    assert (w * c[:, 0] ** 2 * c[:, 1] ** 2).sum() == pytest.approx(8 / 9, abs=1e-14)
    assert abs((w * c[:, 0] ** 3 * c[:, 2]).sum()) < 1e-15
    # prism 6 points (triangle edge midpoints x 2-point Gauss in zeta, w 1/6): exact for degree 2 in
    # (xi, eta) times degree 3 in zeta over the reference prism (volume 1)
    c, w = oracle.gauss_points(PRISM3D, 2)
    assert w.sum() == pytest.approx(1.0, abs=1e-15)
    assert (w * c[:, 0] ** 2 * c[:, 2] ** 2).sum() == pytest.approx(1 / 12 * 2 / 3, abs=1e-15)
    assert (w * c[:, 0] * c[:, 1]).sum() == pytest.approx(2 / 24, abs=1e-15)
    assert abs((w * c[:, 1] * c[:, 2] ** 3).sum()) < 1e-15
    assert np.allclose(oracle.gauss_points(PRISM3D, 1)[0], [[1 / 3, 1 / 3, 0.0]])
    # one-point rules are the centroids (tau evaluation point, SUPG.hpp:145)
    assert np.allclose(oracle.gauss_points(TETRA3D, 1)[0], 0.25)
    assert np.allclose(oracle.gauss_points(HEXA3D, 1)[0], 0.0)
    assert np.allclose(oracle.gauss_points(TRIAG2D, 1)[0], 1 / 3)


def test_meshes_positive_and_conforming(oracle):
    c, t = oracle.box_tetras(3, 2, 2)
    c = oracle.jitter(c, 3, 3, 2, 2, 0.1, 20261202)
    vols = np.array([oracle.tetra_volume(c[e]) for e in t])
    assert vols.min() > 0 and abs(vols.sum() - 1.0) < 1e-13
    # conforming: every interior face is shared by exactly two tets, boundary faces tile the cube surface
    faces = {}
    for e in t:
        for f in ([0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]):
            key = tuple(sorted(e[f]))
            faces[key] = faces.get(key, 0) + 1
    assert max(faces.values()) == 2
    n_boundary = sum(1 for v in faces.values() if v == 1)
    assert n_boundary == 2 * 2 * (3 * 2 + 3 * 2 + 2 * 2)
    # hexahedra of create_box have positive volume and fill the unit cube
    c, h = oracle.box_hexas(2, 3, 2)
    assert sum(oracle.hexa_volume(c[e]) for e in h) == pytest.approx(1.0, abs=1e-14)
    # rectangle triangles: create_rectangle_tris numbering (MeshGeneration.cpp:240-252)
    c, t = oracle.rect_tris(2, 2)
    assert t[0].tolist() == [0, 4, 3] and t[1].tolist() == [0, 1, 4]
    assert all(oracle.element_volume(TRIAG2D, c[e]) > 0 for e in t)


def test_jitter_keeps_boundary_planes(oracle):
    n = 4
    c, _ = oracle.box_tetras(n, n, n)
    cj = oracle.jitter(c, 3, n, n, n, 0.1, 7)
    on_bnd = np.isclose(c, 0) | np.isclose(c, 1)
    assert np.array_equal(cj[on_bnd], c[on_bnd])
    d = np.abs(cj - c)
    assert d.max() <= 0.1 / n + 1e-15 and d[~on_bnd].min() > 0

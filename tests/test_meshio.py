"""Mesh file reading (CPU): the reference's Gambit neutral and Gmsh files, read into device-mesh arrays."""
import os

import numpy as np
import pytest

from coolfluid3_amd import cases, meshio
from oracle import pyoracle

MESHES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "meshes")


def test_ring2d_triangles():
    """plugins/UFEM/test/meshes/ring2d-tris.neu: the annulus 1 <= r <= 2 in 1204 triangles."""
    m = meshio.read_neu(os.path.join(MESHES, "ring2d-tris.neu"))
    assert (m.name, m.dim, m.n_nodes) == ("ring2d", 2, 669)
    t = m.blocks["Triag2D"]
    assert t.shape == (1204, 3) and list(m.blocks) == ["Triag2D"]
    assert m.groups == {"fluid": list(range(1, 1205))}
    c = m.coords
    area = 0.5 * ((c[t[:, 1], 0] - c[t[:, 0], 0]) * (c[t[:, 2], 1] - c[t[:, 0], 1]) -
                  (c[t[:, 2], 0] - c[t[:, 0], 0]) * (c[t[:, 1], 1] - c[t[:, 0], 1]))
    assert area.min() > 0                                    # counter-clockwise cells
    # inscribed polygons: 70 outer and 64 inner sides
    exact = 0.5 * 70 * 4 * np.sin(2 * np.pi / 70) - 0.5 * 64 * np.sin(2 * np.pi / 64)
    assert abs(area.sum() - exact) < 1e-6 * exact           # sides not exactly equal in the file
    for name, faces, r in (("inner", 64, 1.0), ("outer", 70, 2.0)):
        f = m.boundaries[name]["Line2D"]
        assert f.shape == (faces, 2)
        nodes = m.boundary_nodes(name)
        assert len(nodes) == faces and np.all(np.diff(nodes.astype(np.int64)) > 0)
        assert np.allclose(np.hypot(c[nodes, 0], c[nodes, 1]), r, atol=1e-9)


def test_ring2d_quadrilaterals():
    """plugins/UFEM/test/meshes/ring2d-quads.neu (atest-ufem-heat2d-disk): the annulus 1 <= r <= 2 in 490
    quadrilaterals between two 70-gons."""
    m = meshio.read_neu(os.path.join(MESHES, "ring2d-quads.neu"))
    assert (m.name, m.dim, m.n_nodes) == ("ring2d-quads", 2, 560)
    q = m.blocks["Quad2D"]
    assert q.shape == (490, 4) and list(m.blocks) == ["Quad2D"]
    c = m.coords
    area = 0.5 * ((c[q[:, 2], 0] - c[q[:, 0], 0]) * (c[q[:, 3], 1] - c[q[:, 1], 1]) -
                  (c[q[:, 2], 1] - c[q[:, 0], 1]) * (c[q[:, 3], 0] - c[q[:, 1], 0]))
    assert area.min() > 0
    exact = 0.5 * 70 * (4 - 1) * np.sin(2 * np.pi / 70)
    assert abs(area.sum() - exact) < 1e-6 * exact             # the file's coordinates carry ~8 digits
    for name, r in (("inner", 1.0), ("outer", 2.0)):
        assert m.boundaries[name]["Line2D"].shape == (70, 2)
        nodes = m.boundary_nodes(name)
        assert np.allclose(np.hypot(c[nodes, 0], c[nodes, 1]), r, atol=1e-9)


def test_kvs15_cylinder_channel():
    """plugins/UFEM/test/meshes/kvs15.neu (atest-ufem-navier-stokes-cylinder2d): the box [-25.5, 34.5] x
    [-30, 30] around a cylinder of radius 0.5 (a 200-gon) in 33538 triangles."""
    m = meshio.read_neu(os.path.join(MESHES, "kvs15.neu"))
    assert (m.dim, m.n_nodes) == (2, 16953)
    t = m.blocks["Triag2D"]
    assert t.shape == (33538, 3)
    c = m.coords
    area = 0.5 * ((c[t[:, 1], 0] - c[t[:, 0], 0]) * (c[t[:, 2], 1] - c[t[:, 0], 1]) -
                  (c[t[:, 2], 0] - c[t[:, 0], 0]) * (c[t[:, 1], 1] - c[t[:, 0], 1]))
    assert area.min() > 0
    exact = 60.0 * 60.0 - 0.5 * 200 * 0.25 * np.sin(2 * np.pi / 200)
    assert abs(area.sum() - exact) < 1e-9 * exact
    assert {k: m.boundaries[k]["Line2D"].shape[0] for k in m.boundaries} == {"in": 30, "out": 74, "symm": 64,
                                                                             "wall": 200}
    assert np.all(c[m.boundary_nodes("in"), 0] == -25.5) and np.all(c[m.boundary_nodes("out"), 0] == 34.5)
    assert np.all(np.abs(c[m.boundary_nodes("symm"), 1]) == 30.0)
    w = m.boundary_nodes("wall")
    assert np.allclose(np.hypot(c[w, 0], c[w, 1]), 0.5, atol=1e-9)


def test_hextet_mixed_cells():
    """resources/hextet.neu: hexahedra (node order mapped to cf3's) and tetrahedra, three boundary sets
    with quadrilateral and triangular sides."""
    m = meshio.read_neu(os.path.join(MESHES, "hextet.neu"))
    assert (m.dim, m.n_nodes) == (3, 35)
    assert m.blocks["Hexa3D"].shape == (8, 8) and m.blocks["Tetra3D"].shape == (6, 4)
    o = pyoracle.get()
    vh = [o.element_volume(pyoracle.HEXA3D, m.coords[e]) for e in m.blocks["Hexa3D"]]
    vt = [o.element_volume(pyoracle.TETRA3D, m.coords[e]) for e in m.blocks["Tetra3D"]]
    assert np.allclose(vh, 125.0) and np.allclose(vt, 1000.0 / 6.0)   # positive: cf3 orientation
    assert {k: {t: a.shape for t, a in v.items()} for k, v in m.boundaries.items()} == {
        "wall": {"Quad3D": (16, 4), "Triag3D": (8, 3)}, "outlet": {"Triag3D": (2, 3)}, "inlet": {"Quad3D": (4, 4)}}
    # a quadrilateral side of a hexahedron lies in one coordinate plane of the box
    q = m.coords[m.boundaries["inlet"]["Quad3D"]]
    assert all(np.any(np.ptp(face, axis=0) == 0.0) for face in q)
    with pytest.raises(ValueError):
        cases.mesh_case(os.path.join(MESHES, "hextet.neu"))   # two element types: one must be named


def test_unsupported_cells_rejected(tmp_path):
    text = ("        CONTROL INFO 2.3.16\n** GAMBIT NEUTRAL FILE\nprism\nPROGRAM:  Gambit\n date\n"
            "     NUMNP     NELEM     NGRPS    NBSETS     NDFCD     NDFVL\n"
            "         6         1         0         0         3         3\nENDOFSECTION\n"
            "   NODAL COORDINATES 2.3.16\n"
            "         1 0 0 0\n         2 1 0 0\n         3 0 1 0\n         4 0 0 1\n         5 1 0 1\n"
            "         6 0 1 1\nENDOFSECTION\n      ELEMENTS/CELLS 2.3.16\n"
            "       1  5  6        1       2       3       4       5       6\nENDOFSECTION\n")
    p = tmp_path / "prism.neu"
    p.write_text(text)
    with pytest.raises(ValueError):
        meshio.read_neu(str(p))


def test_ring_case_oracle_system():
    """The ring case assembles in the oracle (CPU) with the boundary conditions of its sets."""
    case = cases.mesh_case(os.path.join(MESHES, "ring2d-tris.neu"),
                          bcs=[("inner", {0: 0.0, 1: 0.0}), ("outer", {0: 1.0, 1: 0.0})], pin=("outer", 2))
    n, e, v = cases.bc_lists(case, case.u0, case.p0)
    assert len(n) == 2 * 64 + 2 * 70 + 1
    o = pyoracle.get()
    start, cols, slots = o.build_sparsity(case.conn, case.n_nodes)
    prm = pyoracle.ns_params(dt=case.dt)
    vals, rhs = o.ns_assemble(pyoracle.TRIAG2D, prm, case.coords, case.conn, case.u0, case.p0, case.u0,
                              np.full(case.n_nodes, case.nu), start, cols, slots)
    assert np.all(np.isfinite(vals)) and np.abs(vals).max() > 0


def test_gmsh_rectangle_triangles():
    """resources/rectangle-tg-p1.msh: the rectangle [-1,1] x [0,1] in 929 triangles, three named line
    regions (Inlet, Outlet on y = 0, Wall on the other sides) and one cell region."""
    m = meshio.read_mesh(os.path.join(MESHES, "rectangle-tg-p1.msh"))
    assert (m.name, m.dim, m.n_nodes) == ("rectangle-tg-p1", 2, 507)
    t = m.blocks["Triag2D"]
    assert t.shape == (929, 3) and list(m.blocks) == ["Triag2D"]
    assert list(m.groups) == ["InnerCells"] and len(m.groups["InnerCells"]) == 929
    o = pyoracle.get()
    area = np.array([o.element_volume(pyoracle.TRIAG2D, m.coords[e]) for e in t])
    assert area.min() > 0 and abs(area.sum() - 2.0) < 1e-12
    c = m.coords
    assert {k: {f: a.shape for f, a in v.items()} for k, v in m.boundaries.items()} == {
        "Inlet": {"Line2D": (15, 2)}, "Outlet": {"Line2D": (15, 2)}, "Wall": {"Line2D": (53, 2)}}
    inlet, outlet, wall = (m.boundary_nodes(b) for b in ("Inlet", "Outlet", "Wall"))
    assert np.all(c[inlet, 1] == 0) and np.all(c[inlet, 0] <= 0) and np.all(c[outlet, 0] >= 0)
    on_wall = (np.abs(c[wall, 0]) == 1) | (c[wall, 1] == 1)
    assert on_wall.all() and len(wall) == 54
    # every node of the rectangle's perimeter is on one of the three regions (two corners on two)
    perim = np.flatnonzero((np.abs(c[:, 0]) == 1) | (c[:, 1] == 0) | (c[:, 1] == 1))
    assert np.array_equal(np.union1d(np.union1d(inlet, outlet), wall), perim)


def test_gmsh_rectangle_quadrilaterals():
    """resources/rectangle-qd-p1.msh: the rectangle [-1,1] x [0,1] in 42 quadrilaterals (gmsh node order
    is cf3's: positive Quad2D volumes)."""
    m = meshio.read_mesh(os.path.join(MESHES, "rectangle-qd-p1.msh"))
    assert (m.dim, m.n_nodes) == (2, 56) and m.blocks["Quad2D"].shape == (42, 4)
    o = pyoracle.get()
    vol = np.array([o.element_volume(pyoracle.QUAD2D, m.coords[e]) for e in m.blocks["Quad2D"]])
    assert vol.min() > 0 and abs(vol.sum() - 2.0) < 1e-12
    assert sorted(m.boundaries) == ["Outlet", "Wall", "inlet"]
    assert np.all(m.coords[m.boundary_nodes("inlet"), 0] == -1.0)


def test_gmsh_square_duct_hexahedra():
    """resources/sqduct_3000e.msh: the duct [-0.5,0.5]^2 x [0,3] in 3000 hexahedra (gmsh node order
    is cf3's), the inlet region's quadrilaterals on z = 0."""
    m = meshio.read_mesh(os.path.join(MESHES, "sqduct_3000e.msh"))
    assert (m.dim, m.n_nodes) == (3, 3751) and m.blocks["Hexa3D"].shape == (3000, 8)
    o = pyoracle.get()
    vol = np.array([o.element_volume(pyoracle.HEXA3D, m.coords[e]) for e in m.blocks["Hexa3D"]])
    assert vol.min() > 0 and abs(vol.sum() - 3.0) < 1e-12            # positive: cf3 orientation
    q = m.boundaries["inlet"]["Quad3D"]
    assert q.shape == (100, 4) and np.all(m.coords[q][..., 2] == 0.0)
    assert len(m.boundary_nodes("inlet")) == 121


def test_gmsh_files_rejected(tmp_path):
    head = "$MeshFormat\n2.2 0 8\n$EndMeshFormat\n"
    nodes = "$Nodes\n6\n1 0 0 0\n2 1 0 0\n3 0 1 0\n4 0 0 1\n5 1 0 1\n6 0 1 1\n$EndNodes\n"
    prism = "$Elements\n1\n1 6 2 1 1 1 2 3 4 5 6\n$EndElements\n"
    names = '$PhysicalNames\n1\n3 1 "fluid"\n$EndPhysicalNames\n'
    for i, text in enumerate([head + nodes + prism,                     # no physical names
                              head + names + nodes + prism,             # prisms: no kernel
                              head + names + nodes + prism.replace(" 6 2 1 1", " 4 2 2 1")]):   # unnamed tag
        p = tmp_path / ("m%d.msh" % i)
        p.write_text(text)
        with pytest.raises(ValueError):
            meshio.read_msh(str(p))
    p = tmp_path / "ok.msh"
    p.write_text(head + names + nodes + "$Elements\n1\n7 4 2 1 1 1 2 3 4\n$EndElements\n")
    m = meshio.read_mesh(str(p))
    assert m.blocks["Tetra3D"].tolist() == [[0, 1, 2, 3]] and m.groups == {"fluid": [7]}
    with pytest.raises(ValueError):
        meshio.read_mesh(str(tmp_path / "ok.vtk"))

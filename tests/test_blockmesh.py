"""Block-structured meshes (coolfluid3_amd/blockmesh.py, BlockArrays of cf3/mesh/BlockMesh/BlockData.cpp):
node sharing between blocks, geometric grading, patch faces and their orientation (CPU)."""
import numpy as np
import pytest

from coolfluid3_amd import blockmesh, cases, meshgen
from oracle import pyoracle


def square_blocks(segments):
    """The two blocks and four patches of plugins/UFEM/test/atest-ufem-navier-stokes-taylor-green.py:39-76."""
    b = blockmesh.BlockArrays()
    p = b.create_points(2, 6)
    p[:] = [[-0.5, -0.5], [0.5, -0.5], [-0.5, 0.0], [0.5, 0.0], [-0.5, 0.5], [0.5, 0.5]]
    bl = b.create_blocks(2)
    bl[0], bl[1] = [0, 1, 3, 2], [2, 3, 5, 4]
    sd = b.create_block_subdivisions()
    sd[0], sd[1] = [segments, segments // 2], [segments, segments // 2]
    b.create_block_gradings()
    lp = b.create_patch_nb_faces("left", 2)
    lp[0], lp[1] = [2, 0], [4, 2]
    b.create_patch_nb_faces("bottom", 1)[0] = [0, 1]
    b.create_patch_nb_faces("top", 1)[0] = [5, 4]
    rp = b.create_patch_nb_faces("right", 2)
    rp[0], rp[1] = [1, 3], [3, 5]
    return b


def test_two_blocks_share_their_face():
    """The lower block does not own the nodes of the face it shares with the upper one: 33 x 33 nodes,
    numbered x fastest over both blocks (the numbering of one 32 x 32 grid), cells counter-clockwise."""
    m = square_blocks(32).create_mesh()
    assert m.n_nodes == 33 * 33 and m.blocks["Quad2D"].shape == (1024, 4)
    ref, conn = meshgen.rect_quads(32, 32)
    assert np.allclose(m.coords, ref - 0.5, atol=1e-15, rtol=0)
    assert np.array_equal(m.blocks["Quad2D"], conn)
    o = pyoracle.get()
    area = np.array([o.element_volume(pyoracle.QUAD2D, m.coords[e]) for e in m.blocks["Quad2D"]])
    assert area.min() > 0 and abs(area.sum() - 1.0) < 1e-13
    for name, (axis, value) in {"left": (0, -0.5), "right": (0, 0.5), "bottom": (1, -0.5), "top": (1, 0.5)}.items():
        f = m.boundaries[name]["Line2D"]
        assert f.shape == (32, 2) and np.all(m.coords[f][..., axis] == value)
        # patch segments run as the patch points were given (domain boundary counter-clockwise)
        d = m.coords[f[:, 1]] - m.coords[f[:, 0]]
        t = {"left": (0, -1), "right": (0, 1), "bottom": (1, 0), "top": (-1, 0)}[name]
        assert np.all(d @ np.array(t) > 0)


def test_grading_and_hexahedra():
    """Grading g: the last cell of an edge is g times the first (r = g^(1/(n-1))); a hexahedral block
    with six patches: positive volumes summing to the block's, every patch face pointing outward."""
    b = blockmesh.BlockArrays()
    b.create_points(2, 4)[:] = [[0, 0], [1, 0], [1, 1], [0, 1]]
    b.create_blocks(1)[0] = [0, 1, 2, 3]
    b.create_block_subdivisions()[0] = [10, 3]
    b.create_block_gradings()[0] = [4, 4, 1, 1]
    for n, f in {"s": (0, 1), "e": (1, 2), "n": (2, 3), "w": (3, 0)}.items():
        b.create_patch_nb_faces(n, 1)[0] = f
    m = b.create_mesh()
    dx = np.diff(np.unique(np.round(m.coords[:, 0], 13)))   # one x per column, to rounding of the blending
    assert len(dx) == 10 and dx[-1] / dx[0] == pytest.approx(4.0, rel=1e-11) and np.all(np.diff(dx) > 0)
    assert np.allclose(np.unique(np.round(m.coords[:, 1], 13)), [0, 1 / 3, 2 / 3, 1], rtol=0, atol=1e-13)

    h = blockmesh.BlockArrays()
    h.create_points(3, 8)[:] = [[0, 0, 0], [2, 0, 0], [2, 1, 0], [0, 1, 0], [0, 0, 1], [2, 0, 1], [2, 1, 1], [0, 1, 1]]
    h.create_blocks(1)[0] = range(8)
    h.create_block_subdivisions()[0] = [3, 4, 5]
    h.create_block_gradings()
    faces = {"zneg": (0, 3, 2, 1), "zpos": (4, 5, 6, 7), "yneg": (0, 1, 5, 4), "xpos": (1, 2, 6, 5),
             "ypos": (3, 7, 6, 2), "xneg": (0, 4, 7, 3)}
    for n, f in faces.items():
        h.create_patch_nb_faces(n, 1)[0] = f
    m = h.create_mesh()
    o = pyoracle.get()
    vol = np.array([o.element_volume(pyoracle.HEXA3D, m.coords[e]) for e in m.blocks["Hexa3D"]])
    assert m.n_nodes == 4 * 5 * 6 and vol.min() > 0 and abs(vol.sum() - 2.0) < 1e-13
    centre = np.array([1.0, 0.5, 0.5])
    for n, v in m.boundaries.items():
        X = m.coords[v["Quad3D"]]
        nrm = np.cross(X[:, 1] - X[:, 0], X[:, 3] - X[:, 0])
        assert np.all((nrm * (X.mean(axis=1) - centre)).sum(axis=1) > 0), n
    assert sum(len(v["Quad3D"]) for v in m.boundaries.values()) == 2 * (12 + 15 + 20)


def test_open_face_rejected():
    b = blockmesh.BlockArrays()
    b.create_points(2, 4)[:] = [[0, 0], [1, 0], [1, 1], [0, 1]]
    b.create_blocks(1)[0] = [0, 1, 2, 3]
    b.create_block_subdivisions()[0] = [2, 2]
    b.create_block_gradings()
    b.create_patch_nb_faces("s", 1)[0] = [0, 1]
    with pytest.raises(ValueError):
        b.create_mesh()


def test_taylor_green_case_on_block_mesh():
    """The Taylor-Green case builds its quadrilaterals with the script's blocks and links the patches."""
    case = cases.taylor_green_case(32, element="quad")
    m = square_blocks(32).create_mesh()
    assert np.array_equal(case.conn, m.blocks["Quad2D"]) and np.array_equal(case.coords, m.coords)
    links, active = case.periodic
    assert set(np.flatnonzero(active)) == set(m.boundary_nodes("right")) | set(m.boundary_nodes("top"))

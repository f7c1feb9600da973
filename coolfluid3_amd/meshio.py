"""Mesh files into the host arrays of the device mesh: Gambit neutral (.neu) and Gmsh 2.2 ASCII (.msh).

Gambit neutral files: serial restatement of the reading in cf3/mesh/neu/Reader.cpp: the header (NUMNP, NELEM, NGRPS, NBSETS,
NDFCD; :205-255), the nodal coordinates (node k of the file is node k-1 here; :300-358), the element
connectivity in the cf3 node order (cf_element[neu_to_cf[j]] = node j; :362-430), the element groups
(ELMMAT names; :437-533) and the boundary sets (ITYPE 1 / IBCODE1 6: element sides, each side turned
into the nodes of the cf3 face; :535-605).  Node and face orders: cf3/mesh/neu/Shared.cpp:28-230; face
nodes of the cf3 element types: cf3/mesh/LagrangeP1/Triag2D.cpp:57-70, Quad2D.cpp:34-46,
Tetra3D.cpp:30-43, Hexa3D.cpp:36-52.

Gmsh files: serial restatement of cf3/mesh/gmsh/Reader.cpp: the physical groups ($PhysicalNames: the
mesh dimension is the largest group dimension, group i names region i; :180-205), the nodes in file order
(gmsh node number -> row; x, y and, in 3D, z; :398-450) and the elements, each in the region of its first
tag, with the gmsh-to-cf3 node order of cf3/mesh/gmsh/Shared.cpp:85-152 (the identity for the P1 line,
triangle, quadrilateral, tetrahedron and hexahedron; Reader.cpp:560-650).  Elements of the mesh
dimension are cells, those one dimension lower are boundary faces.
"""
import os
from dataclasses import dataclass, field

import numpy as np

# neu element codes (Shared.hpp:42) -> (cf3 element type, nodes, neu-to-cf node order, neu-to-cf face number)
_TYPES = {
    2: ("Quad2D", 4, (0, 1, 2, 3), {1: 0, 2: 1, 3: 2, 4: 3}),
    3: ("Triag2D", 3, (0, 1, 2), {1: 0, 2: 1, 3: 2}),
    4: ("Hexa3D", 8, (3, 2, 7, 6, 0, 1, 4, 5), {1: 0, 2: 3, 3: 1, 4: 5, 5: 4, 6: 2}),
    6: ("Tetra3D", 4, (0, 1, 2, 3), {1: 0, 2: 1, 3: 2, 4: 3}),
}

# cf3 face type of each element type (LagrangeP1 face_type)
_FACE_TYPE = {"Triag2D": "Line2D", "Quad2D": "Line2D", "Tetra3D": "Triag3D", "Hexa3D": "Quad3D"}

# cf3 face -> element-local nodes
_FACES = {
    "Triag2D": ((0, 1), (1, 2), (2, 0)),
    "Quad2D": ((0, 1), (1, 2), (2, 3), (3, 0)),
    "Tetra3D": ((0, 2, 1), (0, 1, 3), (1, 2, 3), (0, 3, 2)),
    "Hexa3D": ((0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (3, 7, 6, 2), (0, 4, 7, 3)),
}


@dataclass
class FileMesh:
    name: str
    dim: int
    coords: np.ndarray                                   # (N, dim) float64
    blocks: dict = field(default_factory=dict)           # cf3 type -> (E, k) uint32, file order
    elem_index: dict = field(default_factory=dict)       # file element number -> (cf3 type, row)
    groups: dict = field(default_factory=dict)           # element group (region) name -> file element numbers
    boundaries: dict = field(default_factory=dict)       # set name -> {cf3 face type: (faces, nodes) uint32}

    @property
    def n_nodes(self):
        return self.coords.shape[0]

    def boundary_nodes(self, name):
        """Nodes of a boundary set, ascending (the order in which cf3 applies a region's conditions)."""
        return np.unique(np.concatenate([f.ravel() for f in self.boundaries[name].values()])).astype(np.uint32)


def _sections(lines):
    """Start index of every section (the line after its title) by title keyword."""
    out = []
    for i, l in enumerate(lines):
        s = l.strip()
        for key in ("NODAL COORDINATES", "ELEMENTS/CELLS", "ELEMENT GROUP", "BOUNDARY CONDITIONS"):
            if s.startswith(key):
                out.append((key, i + 1))
    return out


def read_neu(path):
    """Read a Gambit neutral file.  Cells are triangles, quadrilaterals, tetrahedra or hexahedra (the
    reference also reads lines and prisms, for which this path has no kernel: ValueError); boundary
    sets must be element sides (ValueError otherwise, as in the reference)."""
    with open(path) as f:
        lines = f.read().splitlines()
    name = lines[2].strip()
    hdr = lines[6].split()
    numnp, nelem, ngrps, nbsets, ndfcd = (int(x) for x in hdr[:5])
    secs = _sections(lines)
    starts = {}
    for key, i in secs:
        starts.setdefault(key, []).append(i)

    # nodal coordinates: one node per line, "id x y [z]"
    i0 = starts["NODAL COORDINATES"][0]
    coords = np.empty((numnp, ndfcd))
    for k in range(numnp):
        tok = lines[i0 + k].split()
        if int(tok[0]) != k + 1:
            raise ValueError("nodes must be numbered 1..NUMNP in order")
        coords[k] = [float(t) for t in tok[1:1 + ndfcd]]

    # elements: a token stream "id type nnodes n1 .. nk" (long rows wrap onto the next line)
    i0 = starts["ELEMENTS/CELLS"][0]
    toks = []
    k = i0
    while not lines[k].strip().startswith("ENDOFSECTION"):
        toks.extend(lines[k].split())
        k += 1
    rows = {}
    mesh = FileMesh(name, ndfcd, coords)
    p = 0
    for _ in range(nelem):
        num, etype, nn = int(toks[p]), int(toks[p + 1]), int(toks[p + 2])
        if etype not in _TYPES or _TYPES[etype][1] != nn:
            raise ValueError("unsupported element type %d with %d nodes" % (etype, nn))
        tname, _, order, _ = _TYPES[etype]
        neu_nodes = [int(t) - 1 for t in toks[p + 3:p + 3 + nn]]
        cf = [0] * nn
        for j in range(nn):
            cf[order[j]] = neu_nodes[j]
        rows.setdefault(tname, []).append(cf)
        mesh.elem_index[num] = (tname, len(rows[tname]) - 1)
        p += 3 + nn
    mesh.blocks = {t: np.asarray(r, dtype=np.uint32) for t, r in rows.items()}

    # element groups: "GROUP: NGP ELEMENTS: NELGP MATERIAL: MTYP NFLAGS: NFLAGS", name, flags, element ids
    for i0 in starts.get("ELEMENT GROUP", []):
        head = lines[i0].replace(":", " ").split()
        nelgp = int(head[head.index("ELEMENTS") + 1])
        nflags = int(head[head.index("NFLAGS") + 1])
        gname = lines[i0 + 1].strip()
        toks = []
        k = i0 + 2
        while not lines[k].strip().startswith("ENDOFSECTION"):
            toks.extend(lines[k].split())
            k += 1
        mesh.groups[gname] = [int(t) for t in toks[nflags:nflags + nelgp]]

    # boundary sets: "NAME ITYPE NENTRY NVALUES IBCODE1 ..", then NENTRY rows "element type face"
    for i0 in starts.get("BOUNDARY CONDITIONS", []):
        head = lines[i0].split()
        bname, itype, nentry = head[0], int(head[1]), int(head[2])
        if itype != 1 or int(head[4]) != 6:
            raise ValueError("boundary set %s: only element sides (ITYPE 1, IBCODE1 6) are read" % bname)
        faces = {}
        for k in range(nentry):
            el, et, fc = (int(t) for t in lines[i0 + 1 + k].split()[:3])
            tname, row = mesh.elem_index[el]
            cf_face = _TYPES[et][3][fc]
            conn = mesh.blocks[tname][row]
            faces.setdefault(_FACE_TYPE[tname], []).append([conn[n] for n in _FACES[tname][cf_face]])
        mesh.boundaries[bname] = {t: np.asarray(v, dtype=np.uint32) for t, v in faces.items()}
    return mesh


# gmsh element type -> (geometry, nodes, dimensionality) for the P1 cells and faces the path reads
# (Shared.cpp:17-36: node counts and geometry names of the gmsh types)
_GMSH = {1: ("Line", 2, 1), 2: ("Triag", 3, 2), 3: ("Quad", 4, 2), 4: ("Tetra", 4, 3), 5: ("Hexa", 8, 3),
         15: ("Point", 1, 0)}


def _block(lines, title):
    """Lines between "$title" and "$Endtitle", or None when the file has no such section."""
    try:
        i = lines.index("$" + title)
    except ValueError:
        return None
    j = lines.index("$End" + title, i)
    return lines[i + 1:j]


def read_msh(path):
    """Read a Gmsh 2.2 ASCII file.  Every element's first tag names a physical group of $PhysicalNames
    (required, as in the reference); elements of the mesh dimension must be triangles, quadrilaterals,
    tetrahedra or hexahedra (ValueError otherwise).  The reference also reads pyramids and higher orders,
    which have no kernel here, and prisms, through the node order 0, 2, 1, 3, 5, 4
    (gmsh/Shared.cpp:375-390), which turns a prism of positive gmsh orientation into one of negative
    Jacobian; prisms for this path come from meshgen.box_prisms."""
    with open(path) as f:
        lines = [l.strip() for l in f.read().splitlines()]
    fmt = _block(lines, "MeshFormat")
    if fmt is None or not fmt[0].split()[0].startswith("2") or fmt[0].split()[1] != "0":
        raise ValueError("only Gmsh 2.x ASCII files are read")
    names = _block(lines, "PhysicalNames")
    if names is None:
        raise ValueError("the file has no $PhysicalNames (every element must belong to a named group)")
    regions = {}
    for l in names[1:1 + int(names[0])]:
        gdim, gidx, gname = l.split(None, 2)
        regions[int(gidx)] = (gname.strip().strip('"'), int(gdim))
    dim = max(d for _, d in regions.values())

    nodes = _block(lines, "Nodes")
    nn = int(nodes[0])
    coords = np.empty((nn, dim))
    row_of = {}
    for k in range(nn):
        tok = nodes[1 + k].split()
        row_of[int(tok[0])] = k
        coords[k] = [float(t) for t in tok[1:1 + dim]]

    els = _block(lines, "Elements")
    mesh = FileMesh(os.path.splitext(os.path.basename(path))[0], dim, coords)
    rows, faces = {}, {}
    suffix = "%dD" % dim
    for l in els[1:1 + int(els[0])]:
        tok = [int(t) for t in l.split()]
        num, gtype, ntags = tok[0], tok[1], tok[2]
        if ntags < 1 or tok[3] not in regions:
            raise ValueError("element %d: its first tag names no physical group" % num)
        gname, _ = regions[tok[3]]
        if gtype not in _GMSH:
            raise ValueError("element %d: unsupported gmsh element type %d" % (num, gtype))
        geo, nk, edim = _GMSH[gtype]
        conn = [row_of[g] for g in tok[3 + ntags:3 + ntags + nk]]
        if len(conn) != nk:
            raise ValueError("element %d: %d nodes expected" % (num, nk))
        tname = geo + suffix
        if edim == dim:
            if tname not in _FACES:
                raise ValueError("element %d: %s cells have no kernel" % (num, tname))
            rows.setdefault(tname, []).append(conn)
            mesh.elem_index[num] = (tname, len(rows[tname]) - 1)
            mesh.groups.setdefault(gname, []).append(num)
        else:
            faces.setdefault(gname, {}).setdefault(tname, []).append(conn)
    mesh.blocks = {t: np.asarray(r, dtype=np.uint32) for t, r in rows.items()}
    mesh.boundaries = {g: {t: np.asarray(v, dtype=np.uint32) for t, v in d.items()} for g, d in faces.items()}
    return mesh


def read_mesh(path):
    """Read a .neu or .msh file by its extension."""
    ext = os.path.splitext(path)[1].lower()
    if ext == ".neu":
        return read_neu(path)
    if ext == ".msh":
        return read_msh(path)
    raise ValueError("unknown mesh file extension %r (.neu or .msh)" % ext)

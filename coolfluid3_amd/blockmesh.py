"""Block-structured meshes of quadrilaterals (2D) or hexahedra (3D): serial restatement of BlockArrays,
cf3/mesh/BlockMesh/BlockData.cpp.

A user gives points, blocks (4 or 8 point indices in the cf3 Quad2D / Hexa3D order), subdivisions per
block and direction, gradings per block edge, and named patches (boundary faces as point lists, in the
cf3 face orientation).  create_mesh follows Implementation::create_blocks (:313-422), add_block
(:472-512), fill_block_coordinates_2d / _3d (:514-622) and add_patch (:624-690), on one rank with the
blocks in one region ("interior"):
  * a block owns its nodes except those on a positive-side face shared with another block, which the
    neighbour owns (nb_points = segments, or segments + 1 when a patch bounds the positive side); nodes
    are numbered block after block, x fastest (strides), and an index equal to nb_points in a direction
    continues in the neighbour of that direction (Block::operator[], :176-193);
  * mapped coordinates along each block edge follow the grading (detail::create_mapped_coords, :64-96:
    r = grading^(1 / (segments - 1)), xi_i = 2 (1 - r^i) / (1 - grading r) - 1), blended over the
    edges with the BlockMesh weights and mapped by the Q1 shape functions of the block corners;
  * cells are numbered block after block, x fastest; patch faces keep the reference's node order.
Partitioning (partition_blocks), extrusion and region boundaries between blocks are not restated.
"""
import numpy as np

from .meshio import FileMesh

# faces of the cf3 element types (LagrangeP1 Quad2D / Hexa3D face connectivity)
_QUAD_FACES = ((0, 1), (1, 2), (2, 3), (3, 0))
_HEXA_FACES = ((0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (3, 7, 6, 2), (0, 4, 7, 3))
# positive / negative face per direction: Quad2D faces 1, 2 / 3, 0; Hexa KSI_POS, ETA_POS, ZTA_POS /
# KSI_NEG, ETA_NEG, ZTA_NEG with FaceNumbering { ZTA_NEG, ZTA_POS, ETA_NEG, KSI_POS, ETA_POS, KSI_NEG }
_POS = {2: (1, 2), 3: (3, 4, 1)}
_NEG = {2: (3, 0), 3: (5, 2, 0)}
# node offsets of a 3D patch face per orientation (add_patch, :633-640)
_IDX_OFFSETS = (((0, 0), (0, 1), (1, 1), (1, 0)), ((0, 0), (0, 1), (1, 1), (1, 0)), ((0, 0), (0, 1), (1, 1), (1, 0)),
                ((0, 0), (1, 0), (1, 1), (0, 1)), ((0, 0), (0, 1), (1, 1), (1, 0)), ((0, 0), (1, 0), (1, 1), (0, 1)))


def mapped_coords(segments, grading):
    """Mapped coordinates -1 .. 1 of the segments + 1 points of one edge (create_mapped_coords)."""
    if abs(grading - 1.0) > 1e-6:
        r = grading ** (1.0 / (segments - 1))
        out = np.array([2.0 * (1.0 - r ** i) / (1.0 - grading * r) - 1.0 for i in range(segments + 1)])
    else:
        step = 2.0 / segments
        out = np.array([i * step - 1.0 for i in range(segments + 1)])
    out[0], out[-1] = -1.0, 1.0
    return out


class BlockArrays:
    """The BlockArrays component's data (create_points, create_blocks, create_block_subdivisions,
    create_block_gradings, create_patch_nb_faces) and create_mesh."""

    def __init__(self):
        self.points = None
        self.blocks = None
        self.block_subdivisions = None
        self.block_gradings = None
        self.patches = {}

    def create_points(self, dimensions, nb_points):
        self.points = np.zeros((nb_points, dimensions))
        return self.points

    def create_blocks(self, nb_blocks):
        self.blocks = np.zeros((nb_blocks, 4 if self.points.shape[1] == 2 else 8), dtype=np.int64)
        return self.blocks

    def create_block_subdivisions(self):
        self.block_subdivisions = np.zeros((len(self.blocks), self.points.shape[1]), dtype=np.int64)
        return self.block_subdivisions

    def create_block_gradings(self):
        self.block_gradings = np.ones((len(self.blocks), 4 if self.points.shape[1] == 2 else 12))
        return self.block_gradings

    def create_patch_nb_faces(self, name, nb_faces):
        self.patches[name] = np.zeros((nb_faces, 2 if self.points.shape[1] == 2 else 4), dtype=np.int64)
        return self.patches[name]

    # ---- create_mesh ----
    def _adjacency(self):
        """Per block and face: ("block", index) or ("patch", name) (the face connectivity of the block
        mesh: faces match by their point set)."""
        dim = self.points.shape[1]
        faces = _QUAD_FACES if dim == 2 else _HEXA_FACES
        owners = {}
        for b, row in enumerate(self.blocks):
            for f, fn in enumerate(faces):
                owners.setdefault(frozenset(int(row[m]) for m in fn), []).append(("block", b, f))
        for name, pf in self.patches.items():
            for face in pf:
                owners.setdefault(frozenset(int(p) for p in face), []).append(("patch", name, -1))
        adj = []
        for b, row in enumerate(self.blocks):
            a = []
            for f, fn in enumerate(faces):
                others = [o for o in owners[frozenset(int(row[m]) for m in fn)] if not (o[0] == "block" and o[1] == b)]
                if len(others) != 1:
                    raise ValueError("block %d has no single adjacent element for face %s. Did you flip the ordering "
                                     "of patch nodes?" % (b, [int(row[m]) for m in fn]))
                a.append(others[0])
            adj.append(a)
        return adj

    def create_mesh(self, name="blocks"):
        dim = self.points.shape[1]
        nblk = len(self.blocks)
        segs = self.block_subdivisions
        adj = self._adjacency()
        # Implementation::create_blocks: points, strides, neighbours, start indices
        nbp, strides, nbr, start = [], [], [], []
        total = 0
        for b in range(nblk):
            s, stride, cnt, pts, ns = [], 1, 1, [], []
            for d in range(dim):
                kind, other, f = adj[b][_POS[dim][d]]
                bounded = kind == "patch"
                if not bounded and f != _NEG[dim][d]:
                    raise ValueError("block %d and block %d are not aligned in direction %d" % (b, other, d))
                pts.append(int(segs[b][d]) + (1 if bounded else 0))
                s.append(stride)
                stride *= pts[-1]
                cnt *= pts[-1]
                ns.append(None if bounded else other)
            nbp.append(pts); strides.append(s); nbr.append(ns); start.append(total)
            total += cnt

        def step(state, i):   # Block::operator[]
            b, ind = state
            d = len(ind)
            if i == nbp[b][d]:
                nstate = (nbr[b][d], [])
                for j in range(d):
                    nstate = step(nstate, ind[j])
                return (nstate[0], nstate[1] + [0])
            return (b, ind + [i])

        def gidx(b, idx):
            state = (b, [])
            for i in idx:
                state = step(state, i)
            bb, ind = state
            return start[bb] + sum(strides[bb][d] * ind[d] for d in range(dim))

        # cells (add_block)
        conn = []
        for b in range(nblk):
            sx = segs[b]
            if dim == 2:
                for j in range(sx[1]):
                    for i in range(sx[0]):
                        conn.append([gidx(b, (i, j)), gidx(b, (i + 1, j)), gidx(b, (i + 1, j + 1)), gidx(b, (i, j + 1))])
            else:
                for k in range(sx[2]):
                    for j in range(sx[1]):
                        for i in range(sx[0]):
                            conn.append([gidx(b, (i, j, k)), gidx(b, (i + 1, j, k)), gidx(b, (i + 1, j + 1, k)),
                                         gidx(b, (i, j + 1, k)), gidx(b, (i, j, k + 1)), gidx(b, (i + 1, j, k + 1)),
                                         gidx(b, (i + 1, j + 1, k + 1)), gidx(b, (i, j + 1, k + 1))])
        # coordinates (fill_block_coordinates_2d / _3d)
        coords = np.zeros((total, dim))
        for b in range(nblk):
            X = self.points[self.blocks[b]]
            g = self.block_gradings[b]
            if dim == 2:
                ksi = np.stack([mapped_coords(segs[b][0], g[e]) for e in (0, 1)], axis=1)
                eta = np.stack([mapped_coords(segs[b][1], g[e]) for e in (2, 3)], axis=1)
                for j in range(segs[b][1] + 1):
                    for i in range(segs[b][0] + 1):
                        k_, e_ = ksi[i], eta[j]
                        w0k = (1. - k_[0]) * (1. - e_[0]) + (1. + k_[0]) * (1. - e_[1])
                        w1k = (1. - k_[1]) * (1. + e_[0]) + (1. + k_[1]) * (1. + e_[1])
                        w0e = (1. - e_[0]) * (1. - k_[0]) + (1. + e_[0]) * (1. - k_[1])
                        w1e = (1. - e_[1]) * (1. + k_[0]) + (1. + e_[1]) * (1. + k_[1])
                        xi = (w0k * k_[0] + w1k * k_[1]) / (w0k + w1k)
                        et = (w0e * e_[0] + w1e * e_[1]) / (w0e + w1e)
                        N = np.array([0.25 * (1.0 - xi) * (1.0 - et), 0.25 * (1.0 + xi) * (1.0 - et),
                                      0.25 * (1.0 + xi) * (1.0 + et), 0.25 * (1.0 - xi) * (1.0 + et)])
                        coords[gidx(b, (i, j))] = N @ X
            else:
                ksi = np.stack([mapped_coords(segs[b][0], g[e]) for e in range(0, 4)], axis=1)
                eta = np.stack([mapped_coords(segs[b][1], g[e]) for e in range(4, 8)], axis=1)
                zta = np.stack([mapped_coords(segs[b][2], g[e]) for e in range(8, 12)], axis=1)
                for k in range(segs[b][2] + 1):
                    for j in range(segs[b][1] + 1):
                        for i in range(segs[b][0] + 1):
                            K, E, Z = ksi[i], eta[j], zta[k]
                            wk = [(1. - K[0]) * (1. - E[0]) * (1. - Z[0]) + (1. + K[0]) * (1. - E[1]) * (1. - Z[1]),
                                  (1. - K[1]) * (1. + E[0]) * (1. - Z[3]) + (1. + K[1]) * (1. + E[1]) * (1. - Z[2]),
                                  (1. - K[2]) * (1. + E[3]) * (1. + Z[3]) + (1. + K[2]) * (1. + E[2]) * (1. + Z[2]),
                                  (1. - K[3]) * (1. - E[3]) * (1. + Z[0]) + (1. + K[3]) * (1. - E[2]) * (1. + Z[1])]
                            we = [(1. - E[0]) * (1. - K[0]) * (1. - Z[0]) + (1. + E[0]) * (1. - K[1]) * (1. - Z[3]),
                                  (1. - E[1]) * (1. + K[0]) * (1. - Z[1]) + (1. + E[1]) * (1. + K[1]) * (1. - Z[2]),
                                  (1. - E[2]) * (1. + K[3]) * (1. + Z[1]) + (1. + E[2]) * (1. + K[2]) * (1. + Z[2]),
                                  (1. - E[3]) * (1. - K[3]) * (1. + Z[0]) + (1. + E[3]) * (1. - K[2]) * (1. + Z[3])]
                            wz = [(1. - Z[0]) * (1. - K[0]) * (1. - E[0]) + (1. + Z[0]) * (1. - K[3]) * (1. - E[3]),
                                  (1. - Z[1]) * (1. + K[0]) * (1. - E[1]) + (1. + Z[1]) * (1. + K[3]) * (1. - E[2]),
                                  (1. - Z[2]) * (1. + K[1]) * (1. + E[1]) + (1. + Z[2]) * (1. + K[2]) * (1. + E[2]),
                                  (1. - Z[3]) * (1. - K[1]) * (1. + E[0]) + (1. + Z[3]) * (1. - K[2]) * (1. + E[3])]
                            xi = (wk[0] * K[0] + wk[1] * K[1] + wk[2] * K[2] + wk[3] * K[3]) / (wk[0] + wk[1] + wk[2] + wk[3])
                            et = (we[0] * E[0] + we[1] * E[1] + we[2] * E[2] + we[3] * E[3]) / (we[0] + we[1] + we[2] + we[3])
                            ze = (wz[0] * Z[0] + wz[1] * Z[1] + wz[2] * Z[2] + wz[3] * Z[3]) / (wz[0] + wz[1] + wz[2] + wz[3])
                            a1, a2, b1, b2, c1, c2 = 1 + xi, 1 - xi, 1 + et, 1 - et, 1 + ze, 1 - ze
                            N = np.array([a2 * b2 * c2, a1 * b2 * c2, a1 * b1 * c2, a2 * b1 * c2,
                                          a2 * b2 * c1, a1 * b2 * c1, a1 * b1 * c1, a2 * b1 * c1]) * 0.125
                            coords[gidx(b, (i, j, k))] = N @ X
        # patches (create_blocks patch list, add_patch)
        plist = {}
        for b in range(nblk):
            for d in range(dim):
                for side, face in ((0, _NEG[dim][d]), (1, _POS[dim][d])):
                    kind, pname, _ = adj[b][face]
                    if kind == "patch":
                        plist.setdefault(pname, []).append((b, d, side * (nbp[b][d] - 1), face))
        boundaries = {}
        for pname, items in plist.items():
            faces = []
            for b, fd, fidx, orient in items:
                ps = [int(segs[b][d]) for d in range(dim) if d != fd]
                if dim == 2:
                    def pg(i):
                        i = i if fidx else ps[0] - i
                        return gidx(b, (fidx if fd == 0 else i, fidx if fd == 1 else i))
                    first, second = (1, 0) if fd == 0 else (0, 1)
                    for i in range(ps[0]):
                        faces.append([pg(i + second), pg(i + first)])
                else:
                    def pg(i, j):
                        if fd != 2:
                            i = i if fidx else ps[0] - i
                        else:
                            i = ps[0] - i if fidx else i
                        return gidx(b, (fidx, i, j) if fd == 0 else ((i, fidx, j) if fd == 1 else (i, j, fidx)))
                    off = _IDX_OFFSETS[orient]
                    for i in range(ps[0]):
                        for j in range(ps[1]):
                            faces.append([pg(i + off[m][0], j + off[m][1]) for m in range(4)])
            boundaries[pname] = {("Line2D" if dim == 2 else "Quad3D"): np.asarray(faces, dtype=np.uint32)}
        etype = "Quad2D" if dim == 2 else "Hexa3D"
        mesh = FileMesh(name, dim, coords)
        mesh.blocks = {etype: np.asarray(conn, dtype=np.uint32)}
        mesh.groups = {"interior": list(range(len(conn)))}
        mesh.elem_index = {e: (etype, e) for e in range(len(conn))}
        mesh.boundaries = boundaries
        return mesh

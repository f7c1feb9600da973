"""Synthetic meshes and seeded fields for the UFEM Navier-Stokes workloads (numpy, host side).

Numbering follows the reference generators:
  * rectangle of triangles: cf3/Tools/MeshGeneration/MeshGeneration.cpp:209-254 (node j*(nx+1)+i,
    cell (i,j) -> (n0,n2,n3), (n0,n1,n2));
  * box of hexahedra: cf3/mesh/SimpleMeshGenerator.cpp:680-805 (node (k*(ny+1)+j)*(nx+1)+i, element
    (k*ny+j)*nx+i, local nodes (i,j,k),(i+1,j,k),(i+1,j+1,k),(i,j+1,k), then k+1);
  * tetrahedra: each hexahedron split into the 6 Kuhn tetrahedra along its local 0 -> 6 diagonal
    (conforming between neighbours, positive volume).
Random numbers: one SplitMix64 stream per seed, uniform doubles (x >> 11) * 2^-53 mapped to [a, b).
"""
import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)

KUHN = np.array([[0, 1, 2, 6], [0, 5, 1, 6], [0, 2, 3, 6], [0, 3, 7, 6], [0, 4, 5, 6], [0, 7, 4, 6]])


def splitmix64(seed, n, offset=0):
    """Outputs offset+1 .. offset+n of the SplitMix64 stream started at `seed`."""
    k = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + k * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
    return z ^ (z >> np.uint64(31))


def uniform(seed, n, a, b, offset=0):
    u = (splitmix64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return a + (b - a) * u


def rect_quads(nx, ny, lx=1.0, ly=1.0):
    """nx x ny quadrilaterals on [0, lx] x [0, ly], nodes numbered x fastest, cells counter-clockwise from
    their lower-left node (the cf3 Quad2D node order)."""
    j, i = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    coords = np.stack([i.ravel() * (lx / nx), j.ravel() * (ly / ny)], axis=1).astype(np.float64)
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    n0 = (jj * (nx + 1) + ii).ravel()
    n3 = ((jj + 1) * (nx + 1) + ii).ravel()
    conn = np.stack([n0, n0 + 1, n3 + 1, n3], axis=1).astype(np.uint32)
    return coords, conn


def rect_tris(nx, ny, lx=1.0, ly=1.0):
    j, i = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    coords = np.stack([i.ravel() * (lx / nx), j.ravel() * (ly / ny)], axis=1).astype(np.float64)
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    n0 = (jj * (nx + 1) + ii).ravel()
    n1 = n0 + 1
    n3 = ((jj + 1) * (nx + 1) + ii).ravel()
    n2 = n3 + 1
    conn = np.empty((2 * nx * ny, 3), dtype=np.uint32)
    conn[0::2] = np.stack([n0, n2, n3], axis=1)
    conn[1::2] = np.stack([n0, n1, n2], axis=1)
    return coords, conn


def _box_nodes(nx, ny, nz, lx, ly, lz):
    k, j, i = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    return np.stack([i.ravel() * (lx / nx), j.ravel() * (ly / ny), k.ravel() * (lz / nz)], axis=1).astype(np.float64)


def _hex_conn(nx, ny, nz):
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    i, j, k = i.ravel().astype(np.int64), j.ravel().astype(np.int64), k.ravel().astype(np.int64)

    def nid(a, b, c):
        return (c * (ny + 1) + b) * (nx + 1) + a
    return np.stack([nid(i, j, k), nid(i + 1, j, k), nid(i + 1, j + 1, k), nid(i, j + 1, k),
                     nid(i, j, k + 1), nid(i + 1, j, k + 1), nid(i + 1, j + 1, k + 1), nid(i, j + 1, k + 1)],
                    axis=1).astype(np.uint32)


def box_hexas(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0):
    return _box_nodes(nx, ny, nz, lx, ly, lz), _hex_conn(nx, ny, nz)


def box_prisms(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0):
    """Each hexahedron of the box split into two prisms along its xy diagonal from (i, j) to (i+1, j+1):
    bottom triangles (i,j), (i+1,j), (i+1,j+1) and (i,j), (i+1,j+1), (i,j+1), counter-clockwise seen from
    +z, each extruded to the layer above (the Prism3D node order: bottom triangle, then top)."""
    h = _hex_conn(nx, ny, nz)
    a = h[:, [0, 1, 2, 4, 5, 6]]
    b = h[:, [0, 2, 3, 4, 6, 7]]
    return _box_nodes(nx, ny, nz, lx, ly, lz), np.ascontiguousarray(np.stack([a, b], axis=1).reshape(-1, 6))


def box_tetras(nx, ny, nz, lx=1.0, ly=1.0, lz=1.0):
    h = _hex_conn(nx, ny, nz)
    conn = h[:, KUHN].reshape(-1, 4)
    return _box_nodes(nx, ny, nz, lx, ly, lz), np.ascontiguousarray(conn)


def jitter(coords, dims, frac, seed, lengths=(1.0, 1.0, 1.0)):
    """Interior nodes move by U[-frac h, frac h] per axis; boundary nodes move only tangentially.
    One draw per node and axis in node order (the stream is consumed for fixed axes too)."""
    dim = coords.shape[1]
    n = coords.shape[0]
    r = uniform(seed, n * dim, 0.0, 1.0).reshape(n, dim)
    out = coords.copy()
    shape = [d + 1 for d in dims]
    idx = np.unravel_index(np.arange(n), shape[::-1])[::-1]     # (i, j[, k]) of each node
    for d in range(dim):
        h = lengths[d] / dims[d]
        delta = -frac * h + (2 * frac * h) * r[:, d]
        free = (idx[d] != 0) & (idx[d] != dims[d])
        out[free, d] += delta[free]
    return out, n * dim


def tet_volumes(coords, conn):
    x = coords[conn]
    return np.einsum("ij,ij->i", np.cross(x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]), x[:, 3] - x[:, 0]) / 6.0


def skew_kuhn(coords, dims, site_prob, seed, offset, frac=0.93):
    """Deterministic skewed elements in a Kuhn-split box: one draw per node in node order (stream
    positions offset+1 ..); an interior node whose draw is below site_prob becomes a site unless a
    site was already taken in its 3x3x3 neighbourhood (node order).  A site moves the fraction `frac`
    of the way to its Kuhn-diagonal neighbour (i+1, j+1, k+1), which flattens the 6 tetrahedra of the
    hexahedron between them to an aspect ratio (longest edge / smallest height) of about 20 for
    frac = 0.93.  Returns (coords, sites, draws used)."""
    nx, ny, nz = dims
    n = coords.shape[0]
    r = uniform(seed, n, 0.0, 1.0, offset=offset)
    idx = np.arange(n)
    i = idx % (nx + 1)
    j = (idx // (nx + 1)) % (ny + 1)
    k = idx // ((nx + 1) * (ny + 1))
    interior = (i > 0) & (i < nx) & (j > 0) & (j < ny) & (k > 0) & (k < nz)
    cand = np.nonzero(interior & (r < site_prob))[0]
    taken = np.zeros(n, bool)
    sites = []
    ring = np.array([(dk * (ny + 1) + dj) * (nx + 1) + di
                     for dk in (-1, 0, 1) for dj in (-1, 0, 1) for di in (-1, 0, 1)])
    for s in cand:
        if taken[s + ring].any():
            continue
        taken[s] = True
        sites.append(s)
    sites = np.array(sites, dtype=np.int64)
    out = coords.copy()
    diag = sites + 1 + (nx + 1) + (nx + 1) * (ny + 1)
    out[sites] += frac * (coords[diag] - coords[sites])
    return out, sites, n


def tet_quality(coords, conn):
    """Aspect ratio (longest edge / smallest height) of each tetrahedron."""
    x = coords[conn]
    v = np.abs(tet_volumes(coords, conn))
    lmax = np.max([np.linalg.norm(x[:, a] - x[:, b], axis=1) for a in range(4) for b in range(a + 1, 4)], axis=0)
    faces = ((1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2))
    amax = np.max([0.5 * np.linalg.norm(np.cross(x[:, f[1]] - x[:, f[0]], x[:, f[2]] - x[:, f[0]]), axis=1)
                   for f in faces], axis=0)
    return lmax * amax / (3.0 * v)

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

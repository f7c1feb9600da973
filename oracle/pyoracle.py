"""ctypes front-end of the oracle library (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: the oracle is the CPU FP64 restatement of the reference hot path used to
check results by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.  The product
package (coolfluid3_amd) never imports this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")

TRIAG2D, TETRA3D, HEXA3D, QUAD2D, PRISM3D = 0, 1, 2, 3, 4
SUPG_TEZDUYAR, SUPG_METRIC, SUPG_CF2 = 0, 1, 2
NODES = {TRIAG2D: 3, TETRA3D: 4, HEXA3D: 8, QUAD2D: 4, PRISM3D: 6}
DIM = {TRIAG2D: 2, TETRA3D: 3, HEXA3D: 3, QUAD2D: 2, PRISM3D: 3}

_d = C.POINTER(C.c_double)
_u32 = C.POINTER(C.c_uint32)
_u64 = C.POINTER(C.c_uint64)


class NSParams(C.Structure):
    _fields_ = [("rho", C.c_double), ("dt", C.c_double), ("theta", C.c_double),
                ("use_specializations", C.c_int), ("supg_type", C.c_int),
                ("c1", C.c_double), ("c2", C.c_double),
                ("alpha_ps", C.c_double), ("alpha_su", C.c_double), ("alpha_bu", C.c_double),
                ("u_ref", C.c_double)]


def ns_params(rho=1.0, dt=0.01, theta=1.0, use_specializations=True, supg_type=SUPG_METRIC, u_ref=1.0):
    return NSParams(rho, dt, theta, int(use_specializations), supg_type, 1.0, 16.0, 1.0, 1.0, 1.0, u_ref)


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("rel_residual", C.c_double),
                ("spmvs", C.c_int)]


def _p(a, t=_d):
    return a.ctypes.data_as(t) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _u32a(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


class Oracle:
    def __init__(self, path=LIB_PATH):
        if not os.path.exists(path):
            raise FileNotFoundError(f"oracle library not built: {path} (run make -C oracle)")
        self.lib = L = C.CDLL(path)
        L.orc_splitmix64_next.restype = C.c_uint64
        L.orc_splitmix64_next.argtypes = [_u64]
        L.orc_uniform.restype = C.c_double
        L.orc_uniform.argtypes = [_u64, C.c_double, C.c_double]
        L.orc_tetra_volume.restype = C.c_double
        L.orc_hexa_volume.restype = C.c_double
        L.orc_hexa_jacobian_determinant.restype = C.c_double
        L.orc_build_sparsity.restype = C.c_int64
        L.orc_build_sparsity.argtypes = [C.c_int, C.c_uint32, C.c_uint32, _u32, _u32, _u32, C.c_int64, _u32]
        for name in ("orc_mesh_rect_tris",):
            getattr(L, name).argtypes = [C.c_uint32, C.c_uint32, C.c_double, C.c_double, _d, _u32]
        for name in ("orc_mesh_box_hexas", "orc_mesh_box_tetras"):
            getattr(L, name).argtypes = [C.c_uint32] * 3 + [C.c_double] * 3 + [_d, _u32]
        L.orc_jitter_structured.argtypes = [C.c_int] + [C.c_uint32] * 3 + [C.c_double] * 4 + [C.c_uint64, _d]
        L.orc_ns_element.argtypes = [C.c_int, C.POINTER(NSParams)] + [_d] * 9
        L.orc_ns_element_system.argtypes = [C.c_int, C.POINTER(NSParams)] + [_d] * 8
        L.orc_ns_assemble.argtypes = ([C.c_int, C.POINTER(NSParams), C.c_uint32, C.c_uint32, _d, _u32]
                                      + [_d] * 5 + [_u32] * 3 + [_d, _d, C.c_int])
        L.orc_ns_assemble_abs.argtypes = L.orc_ns_assemble.argtypes
        L.orc_ns_assemble_p2m.argtypes = ([C.c_int, C.POINTER(NSParams), C.c_uint32, C.c_uint32, _d, _u32, _u32]
                                          + [_d] * 5 + [_u32] * 2 + [_d, _d])
        L.orc_dirichlet.argtypes = [C.c_int, C.c_uint32, _u32, _u32, _d, _d, _d, C.c_uint32, _u32, _u32, _d]
        L.orc_set_row.argtypes = [C.c_int, _u32, _u32, _d, C.c_uint32, C.c_uint32, C.c_double, C.c_double]
        L.orc_tie_blockrow_pairs.restype = C.c_int
        L.orc_tie_blockrow_pairs.argtypes = [C.c_int, _u32, _u32, _d, _d, _d, C.c_uint32, _u32, _u32]
        L.orc_spmv.argtypes = [C.c_int, C.c_uint32, _u32, _u32, _d, C.c_double, _d, C.c_double, _d]
        L.orc_point_csr.argtypes = [C.c_int, C.c_uint32, _u32, _u32, _d, _u32, _u32, _d]
        L.orc_bicgstab_jacobi.argtypes = [C.c_int, C.c_uint32, _u32, _u32, _d, _d, _d, C.c_double, C.c_int,
                                          C.POINTER(SolveStats)]
        L.orc_gmres_ilu0.argtypes = [C.c_int, C.c_uint32, _u32, _u32, _d, _d, _d, C.c_int, C.c_double, C.c_int,
                                     C.POINTER(SolveStats)]
        L.orc_linearize_u.argtypes = [C.c_uint32, C.c_int, _d, _d, _d, _d, _d]
        L.orc_update.argtypes = [C.c_uint32, C.c_int, _d, _d, _d, _d, _d, _d]
        L.orc_shape_values.argtypes = [C.c_int, _d, _d]
        L.orc_shape_gradients.argtypes = [C.c_int, _d, _d]
        L.orc_gauss_points.argtypes = [C.c_int, C.c_int, _d, _d]
        L.orc_tetra_volume.argtypes = [_d]
        L.orc_tetra_jacobian.argtypes = [_d, _d]
        L.orc_tetra_jacobian_adjoint.argtypes = [_d, _d]
        L.orc_hexa_jacobian_determinant.argtypes = [_d, _d]
        L.orc_hexa_jacobian.argtypes = [_d, _d, _d]
        L.orc_hexa_jacobian_adjoint.argtypes = [_d, _d, _d]
        L.orc_hexa_volume.argtypes = [_d]
        L.orc_element_volume.restype = C.c_double
        L.orc_element_volume.argtypes = [C.c_int, _d]
        L.orc_element_jacobian.argtypes = [C.c_int, _d, _d, _d]
        L.orc_si_lumped_mass.argtypes = [C.c_int, C.c_uint32, _d, _u32, _u32, _d]
        for name in ("orc_si_velocity_matrix", "orc_si_pressure_matrix"):
            getattr(L, name).argtypes = [C.c_int, C.POINTER(NSParams), C.c_uint32, _d, _u32, _d, _d, _u32, _d]
        L.orc_si_velocity_rhs.argtypes = [C.c_int, C.POINTER(NSParams), C.c_uint32, _d, _u32, _u32] + [_d] * 8
        L.orc_si_pressure_rhs.argtypes = [C.c_int, C.POINTER(NSParams), C.c_uint32, _d, _u32, _u32] + [_d] * 8
        L.orc_si_apply_aup.argtypes = [C.c_int, C.POINTER(NSParams), C.c_uint32, _d, _u32, _u32] + [_d] * 4

    # ---------------- PRNG / meshes ----------------
    def uniform_stream(self, seed, n, a=0.0, b=1.0):
        s = C.c_uint64(seed)
        return np.array([self.lib.orc_uniform(C.byref(s), a, b) for _ in range(n)])

    def rect_tris(self, nx, ny, lx=1.0, ly=1.0):
        coords = np.zeros(((nx + 1) * (ny + 1), 2))
        conn = np.zeros((2 * nx * ny, 3), dtype=np.uint32)
        self.lib.orc_mesh_rect_tris(nx, ny, lx, ly, _p(coords), _p(conn, _u32))
        return coords, conn

    def box_hexas(self, nx, ny, nz, lx=1.0, ly=1.0, lz=1.0):
        coords = np.zeros(((nx + 1) * (ny + 1) * (nz + 1), 3))
        conn = np.zeros((nx * ny * nz, 8), dtype=np.uint32)
        self.lib.orc_mesh_box_hexas(nx, ny, nz, lx, ly, lz, _p(coords), _p(conn, _u32))
        return coords, conn

    def box_tetras(self, nx, ny, nz, lx=1.0, ly=1.0, lz=1.0):
        coords = np.zeros(((nx + 1) * (ny + 1) * (nz + 1), 3))
        conn = np.zeros((6 * nx * ny * nz, 4), dtype=np.uint32)
        self.lib.orc_mesh_box_tetras(nx, ny, nz, lx, ly, lz, _p(coords), _p(conn, _u32))
        return coords, conn

    def jitter(self, coords, dim, nx, ny, nz, frac, seed, lx=1.0, ly=1.0, lz=1.0):
        c = _f64(coords).copy()
        self.lib.orc_jitter_structured(dim, nx, ny, nz, lx, ly, lz, frac, seed, _p(c))
        return c

    # ---------------- sparsity ----------------
    def build_sparsity(self, conn, n_nodes, with_slots=True):
        conn = _u32a(conn)
        ne, k = conn.shape
        start = np.zeros(n_nodes + 1, dtype=np.uint32)
        cap = ne * k * k
        cols = np.zeros(cap, dtype=np.uint32)
        slots = np.zeros(ne * k * k, dtype=np.uint32) if with_slots else None
        nb = self.lib.orc_build_sparsity(k, n_nodes, ne, _p(conn, _u32), _p(start, _u32), _p(cols, _u32), cap,
                                         _p(slots, _u32) if with_slots else None)
        assert nb >= 0
        return start, cols[:nb].copy(), slots

    # ---------------- element kernels ----------------
    def ns_element(self, etype, prm, nodes, u, u_adv, nu, g=None):
        k, dim = NODES[etype], DIM[etype]
        n = k * (dim + 1)
        A = np.zeros((n, n)); T = np.zeros((n, n)); a = np.zeros(n); tau = np.zeros(3)
        g = np.zeros((k, dim)) if g is None else g
        err = self.lib.orc_ns_element(etype, C.byref(prm), _p(_f64(nodes)), _p(_f64(u)), _p(_f64(u_adv)),
                                      _p(_f64(nu)), _p(_f64(g)), _p(A), _p(T), _p(a), _p(tau))
        if err:
            raise RuntimeError(f"orc_ns_element error {err}")
        return A, T, a, tau

    def ns_element_system(self, etype, prm, nodes, u, p, u_adv, nu, g=None):
        k, dim = NODES[etype], DIM[etype]
        n = k * (dim + 1)
        M = np.zeros((n, n)); rhs = np.zeros(n)
        g = np.zeros((k, dim)) if g is None else g
        err = self.lib.orc_ns_element_system(etype, C.byref(prm), _p(_f64(nodes)), _p(_f64(u)), _p(_f64(p)),
                                             _p(_f64(u_adv)), _p(_f64(nu)), _p(_f64(g)), _p(M), _p(rhs))
        if err:
            raise RuntimeError(f"orc_ns_element_system error {err}")
        return M, rhs

    # ---------------- assembly / LSS ----------------
    def ns_assemble(self, etype, prm, coords, conn, u, p, u_adv, nu, start, cols, slots, g=None, nthreads=1,
                    magnitudes=False):
        """Values and rhs of the assembled system; with magnitudes, per entry the sum of the magnitudes of its
        element contributions instead (the scale of the entry's rounding error)."""
        dim = DIM[etype]
        neq = dim + 1
        n_nodes = coords.shape[0]
        conn = _u32a(conn)
        values = np.zeros(int(start[-1]) * neq * neq)
        rhs = np.zeros(n_nodes * neq)
        gg = _f64(g) if g is not None else np.zeros((n_nodes, dim))
        fn = self.lib.orc_ns_assemble_abs if magnitudes else self.lib.orc_ns_assemble
        err = fn(etype, C.byref(prm), n_nodes, conn.shape[0], _p(_f64(coords)),
                                       _p(conn, _u32), _p(_f64(u)), _p(_f64(p)), _p(_f64(u_adv)), _p(_f64(nu)),
                                       _p(gg), _p(_u32a(start), _u32), _p(_u32a(cols), _u32),
                                       _p(_u32a(slots), _u32), _p(values), _p(rhs), nthreads)
        if err:
            raise RuntimeError(f"orc_ns_assemble error {err}")
        return values, rhs

    def ns_assemble_p2m(self, etype, prm, coords, conn, p2m, u, p, u_adv, nu, start, slots, g=None):
        """Assembly with periodic links: p2m maps mesh nodes to block rows (len(start) - 1 of them)."""
        dim = DIM[etype]
        neq = dim + 1
        n_rows = len(start) - 1
        conn = _u32a(conn)
        values = np.zeros(int(start[-1]) * neq * neq)
        rhs = np.zeros(n_rows * neq)
        gg = _f64(g) if g is not None else np.zeros((coords.shape[0], dim))
        err = self.lib.orc_ns_assemble_p2m(etype, C.byref(prm), n_rows, conn.shape[0], _p(_f64(coords)),
                                           _p(conn, _u32), _p(_u32a(p2m), _u32), _p(_f64(u)), _p(_f64(p)),
                                           _p(_f64(u_adv)), _p(_f64(nu)), _p(gg), _p(_u32a(start), _u32),
                                           _p(_u32a(slots), _u32), _p(values), _p(rhs))
        if err:
            raise RuntimeError(f"orc_ns_assemble_p2m error {err}")
        return values, rhs

    def dirichlet(self, neq, start, cols, values, rhs, sol, bc_node, bc_eq, bc_val):
        n_nodes = len(start) - 1
        bn, be, bv = _u32a(bc_node), _u32a(bc_eq), _f64(bc_val)
        self.lib.orc_dirichlet(neq, n_nodes, _p(_u32a(start), _u32), _p(_u32a(cols), _u32), _p(values), _p(rhs),
                               _p(sol), len(bn), _p(bn, _u32), _p(be, _u32), _p(bv))

    def set_row(self, neq, start, cols, values, blockrow, ieq, diag, offdiag):
        self.lib.orc_set_row(neq, _p(_u32a(start), _u32), _p(_u32a(cols), _u32), _p(values), blockrow, ieq,
                             diag, offdiag)

    def periodicity(self, neq, start, cols, values, rhs, sol, to, frm):
        """In place: ties of (to[i], frm[i]) block rows in order (rhs / sol may be None: the matrix only).
        Returns 0, -1 (block counts differ) or -2 (block columns differ)."""
        t, f = _u32a(to), _u32a(frm)
        return self.lib.orc_tie_blockrow_pairs(neq, _p(_u32a(start), _u32), _p(_u32a(cols), _u32), _p(values),
                                               _p(rhs) if rhs is not None else None,
                                               _p(sol) if sol is not None else None, len(t), _p(t, _u32), _p(f, _u32))

    def spmv(self, neq, start, cols, values, x, alpha=1.0, beta=0.0, y=None):
        n_nodes = len(start) - 1
        y = np.zeros(n_nodes * neq) if y is None else _f64(y).copy()
        self.lib.orc_spmv(neq, n_nodes, _p(_u32a(start), _u32), _p(_u32a(cols), _u32), _p(_f64(values)), alpha,
                          _p(_f64(x)), beta, _p(y))
        return y

    def point_csr(self, neq, start, cols, values):
        n_nodes = len(start) - 1
        nnz = int(start[-1]) * neq * neq
        rp = np.zeros(n_nodes * neq + 1, dtype=np.uint32)
        ci = np.zeros(nnz, dtype=np.uint32)
        v = np.zeros(nnz)
        self.lib.orc_point_csr(neq, n_nodes, _p(_u32a(start), _u32), _p(_u32a(cols), _u32), _p(_f64(values)),
                               _p(rp, _u32), _p(ci, _u32), _p(v))
        return rp, ci, v

    def set_krylov_threads(self, n):
        """Threads of the oracle's SpMV and dot products (1 by default; CPU-baseline timing only)."""
        self.lib.orc_set_krylov_threads(int(n))

    def bicgstab(self, neq, start, cols, values, b, x0, tol=1e-10, maxit=2000):
        x = _f64(x0).copy()
        st = SolveStats()
        self.lib.orc_bicgstab_jacobi(neq, len(start) - 1, _p(_u32a(start), _u32), _p(_u32a(cols), _u32),
                                     _p(_f64(values)), _p(_f64(b)), _p(x), tol, maxit, C.byref(st))
        return x, st

    def gmres_ilu0(self, neq, start, cols, values, b, x0, restart=30, tol=1e-10, maxit=2000):
        x = _f64(x0).copy()
        st = SolveStats()
        self.lib.orc_gmres_ilu0(neq, len(start) - 1, _p(_u32a(start), _u32), _p(_u32a(cols), _u32),
                                _p(_f64(values)), _p(_f64(b)), _p(x), restart, tol, maxit, C.byref(st))
        return x, st

    def linearize_u(self, u, u1, u2, u3):
        n, dim = u.shape
        out = np.zeros_like(_f64(u))
        self.lib.orc_linearize_u(n, dim, _p(_f64(u)), _p(_f64(u1)), _p(_f64(u2)), _p(_f64(u3)), _p(out))
        return out

    def update(self, u, p, u1, u2, u3, sol):
        n, dim = u.shape
        self.lib.orc_update(n, dim, _p(u), _p(p), _p(u1), _p(u2), _p(u3), _p(_f64(sol)))

    # ---------------- geometry fixtures ----------------
    def tetra_volume(self, nodes):
        return self.lib.orc_tetra_volume(_p(_f64(nodes)))

    def tetra_jacobian(self, nodes):
        J = np.zeros(9); self.lib.orc_tetra_jacobian(_p(_f64(nodes)), _p(J)); return J.reshape(3, 3)

    def tetra_jacobian_adjoint(self, nodes):
        J = np.zeros(9); self.lib.orc_tetra_jacobian_adjoint(_p(_f64(nodes)), _p(J)); return J.reshape(3, 3)

    def hexa_jacobian_determinant(self, mapped, nodes):
        return self.lib.orc_hexa_jacobian_determinant(_p(_f64(mapped)), _p(_f64(nodes)))

    def hexa_jacobian(self, mapped, nodes):
        J = np.zeros(9); self.lib.orc_hexa_jacobian(_p(_f64(mapped)), _p(_f64(nodes)), _p(J)); return J.reshape(3, 3)

    def hexa_jacobian_adjoint(self, mapped, nodes):
        J = np.zeros(9)
        self.lib.orc_hexa_jacobian_adjoint(_p(_f64(mapped)), _p(_f64(nodes)), _p(J))
        return J.reshape(3, 3)

    def hexa_volume(self, nodes):
        return self.lib.orc_hexa_volume(_p(_f64(nodes)))

    def element_volume(self, etype, nodes):
        return self.lib.orc_element_volume(etype, _p(_f64(nodes)))

    def element_jacobian(self, etype, mapped, nodes):
        d = DIM[etype]
        J = np.zeros(d * d)
        self.lib.orc_element_jacobian(etype, _p(_f64(mapped)), _p(_f64(nodes)), _p(J))
        return J.reshape(d, d)

    def shape_values(self, etype, mapped):
        N = np.zeros(NODES[etype]); self.lib.orc_shape_values(etype, _p(_f64(mapped)), _p(N)); return N

    def shape_gradients(self, etype, mapped):
        k, dim = NODES[etype], DIM[etype]
        g = np.zeros(k * dim); self.lib.orc_shape_gradients(etype, _p(_f64(mapped)), _p(g))
        return g.reshape(dim, k)

    def gauss_points(self, etype, order):
        c = np.zeros(24); w = np.zeros(8)
        n = self.lib.orc_gauss_points(etype, order, _p(c), _p(w))
        return c[:n * DIM[etype]].reshape(n, DIM[etype]), w[:n]


_ORACLE = None


def get():
    global _ORACLE
    if _ORACLE is None:
        _ORACLE = Oracle()
    return _ORACLE


class CrsConditions:
    """TrilinosCrsMatrix::symmetric_dirichlet (cf3/math/LSS/Trilinos/TrilinosCrsMatrix.cpp:325-402) followed by
    the solution value of System::dirichlet (cf3/math/LSS/System.cpp:240-255), on node-block CSR arrays
    (block (c, i) of row c holds A[(c, j), (i, e)] at j * neq + e), in place.  The first application of a dof
    since the reset (`reset`, Matrix::reset clearing m_symmetric_dirichlet_values) visits the point rows of
    the dof's node's neighbours in the order of :345-353: every row but the dof's own has its entry in the
    dof's column cached, eliminated into the rhs and zeroed (:360-376); the dof's own row becomes 1 at its
    column and 0 elsewhere (:377-389).  A later application only subtracts the cached column times the new
    value from those rows' rhs (:392-398).  Both set rhs and solution at the dof to the value.  Pure
    Python loops: for small systems (test oracle only)."""

    def __init__(self):
        self.cache = {}

    def reset(self):
        self.cache.clear()

    @staticmethod
    def _block(start, cols, row, col):
        for q in range(int(start[row]), int(start[row + 1])):
            if int(cols[q]) == col:
                return q
        return -1

    def apply(self, neq, start, cols, values, rhs, sol, node, eq, value):
        nn = neq * neq
        bc = node * neq + eq
        cached = self.cache.get(bc)
        if cached is None:
            cached = {}
            for q in range(int(start[node]), int(start[node + 1])):
                c = int(cols[q])
                for j in range(neq):
                    r = c * neq + j
                    if r != bc:
                        qb = self._block(start, cols, c, node)
                        if qb < 0:
                            continue
                        k = qb * nn + j * neq + eq
                        cached[r] = values[k]
                        rhs[r] -= values[k] * value
                        values[k] = 0.0
                    else:
                        for q2 in range(int(start[node]), int(start[node + 1])):
                            for l in range(neq):
                                values[q2 * nn + eq * neq + l] = 1.0 if (int(cols[q2]) == node and l == eq) else 0.0
            self.cache[bc] = cached
        else:
            for r, v in cached.items():
                rhs[r] -= v * value
        rhs[bc] = value
        sol[bc] = value

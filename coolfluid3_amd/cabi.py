"""ctypes binding of the cf3hip C ABI (include/cf3hip.h), loaded from the in-tree libcf3hip.so.

There is no fallback: if the library is missing, importing the product path raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcf3hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "cf3hip.h")

OK = 0
ERR = {1: "INVALID_ARG", 2: "NOT_CREATED", 3: "HIP", 4: "OUT_OF_MEMORY", 5: "NOT_CONVERGED", 6: "ILLEGAL_ENTRY",
       7: "NOT_IMPLEMENTED", 8: "LIMIT", 9: "SINGULAR"}
TRIAG2D, TETRA3D, HEXA3D, QUAD2D, PRISM3D = 0, 1, 2, 3, 4
SUPG_TEZDUYAR, SUPG_METRIC, SUPG_CF2 = 0, 1, 2
FP64, FP32 = 0, 1
RHS, SOLUTION = 0, 1
BICGSTAB_JACOBI, GMRES_ILU0, RCG_JACOBI = 0, 1, 2
BUILDER_VBR, BUILDER_CRS = 0, 1
FIELD_U, FIELD_P, FIELD_UADV, FIELD_U1, FIELD_U2, FIELD_U3, FIELD_NU, FIELD_G = range(8)
KERNEL_ASSEMBLY, KERNEL_SPMV, KERNEL_SPMV_FP32, KERNEL_RESIDUAL = range(4)


class Cf3HipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cf3hip error {code} ({ERR.get(code, '?')}): {msg}")
        self.code = code


class SolverParams(C.Structure):
    _fields_ = [("method", C.c_int), ("restart", C.c_int), ("max_iter", C.c_int), ("tol", C.c_double),
                ("precision", C.c_int), ("inner_tol", C.c_double), ("max_outer", C.c_int),
                ("block_scale", C.c_int), ("fp32_vectors", C.c_int), ("recycle", C.c_int), ("rcg_space", C.c_int)]


class SiOptions(C.Structure):
    _fields_ = [("enable_body_force", C.c_int), ("pressure_rcg_solve", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("outer", C.c_int), ("spmvs", C.c_int), ("converged", C.c_int),
                ("rel_residual", C.c_double)]


class NSParams(C.Structure):
    _fields_ = [("rho", C.c_double), ("dt", C.c_double), ("theta", C.c_double),
                ("use_specializations", C.c_int), ("supg_type", C.c_int), ("c1", C.c_double), ("c2", C.c_double),
                ("alpha_ps", C.c_double), ("alpha_su", C.c_double), ("alpha_bu", C.c_double), ("precision", C.c_int),
                ("u_ref", C.c_double)]


def ns_params(rho=1.0, dt=0.01, theta=1.0, use_specializations=True, supg_type=SUPG_METRIC, precision=FP64,
              u_ref=1.0, alpha_ps=1.0, alpha_su=1.0, alpha_bu=1.0):
    return NSParams(rho, dt, theta, int(use_specializations), supg_type, 1.0, 16.0, alpha_ps, alpha_su, alpha_bu,
                    precision, u_ref)


def solver_params(method=BICGSTAB_JACOBI, tol=1e-10, max_iter=5000, restart=30, precision=FP64, inner_tol=1e-2,
                  max_outer=0, block_scale=1, fp32_vectors=0, recycle=0, rcg_space=0):
    return SolverParams(method, restart, max_iter, tol, precision, inner_tol, max_outer, block_scale, fp32_vectors,
                        recycle, rcg_space)


_lib = None
_vp = C.c_void_p
_d = C.POINTER(C.c_double)
_f = C.POINTER(C.c_float)
_u32 = C.POINTER(C.c_uint32)


def lib():
    """Load libcf3hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() (no CPU fallback exists)")
    L = C.CDLL(LIB_PATH)
    L.cf3hip_last_error.restype = C.c_char_p
    L.cf3hip_version.restype = C.c_char_p
    sig = {
        "cf3hip_context_create": [C.c_int, C.POINTER(_vp)],
        "cf3hip_context_destroy": [_vp],
        "cf3hip_context_set_stream": [_vp, _vp],
        "cf3hip_context_synchronize": [_vp],
        "cf3hip_context_memory": [_vp, C.POINTER(C.c_uint64)],
        "cf3hip_context_set_async": [_vp, C.c_int],
        "cf3hip_context_enable_timing": [_vp, C.c_int],
        "cf3hip_context_kernel_time": [_vp, C.c_int, _d, C.POINTER(C.c_uint64)],
        "cf3hip_mesh_create": [_vp, C.c_int, C.c_uint32, C.c_uint32, _d, _u32, C.POINTER(_vp)],
        "cf3hip_mesh_destroy": [_vp],
        "cf3hip_mesh_set_coordinates": [_vp, _d],
        "cf3hip_lss_create": [_vp, _vp, C.c_int, C.POINTER(_vp)],
        "cf3hip_lss_create_periodic": [_vp, _vp, C.c_int, _u32, C.POINTER(C.c_uint8), C.POINTER(_vp)],
        "cf3hip_lss_get_p2m": [_vp, _u32],
        "cf3hip_lss_create_from_graph": [_vp, C.c_uint32, C.c_int, _u32, _u32, C.POINTER(_vp)],
        "cf3hip_lss_destroy": [_vp],
        "cf3hip_lss_info": [_vp, _u32, C.POINTER(C.c_int), C.POINTER(C.c_uint64)],
        "cf3hip_lss_get_graph": [_vp, _u32, _u32],
        "cf3hip_lss_get_slots": [_vp, _u32],
        "cf3hip_lss_enable_fp32": [_vp],
        "cf3hip_lss_round_fp32": [_vp],
        "cf3hip_matrix_reset": [_vp, C.c_double],
        "cf3hip_matrix_set_value": [_vp, C.c_uint32, C.c_uint32, C.c_double],
        "cf3hip_matrix_add_value": [_vp, C.c_uint32, C.c_uint32, C.c_double],
        "cf3hip_matrix_get_value": [_vp, C.c_uint32, C.c_uint32, _d],
        "cf3hip_matrix_set_row": [_vp, C.c_uint32, C.c_uint32, C.c_double, C.c_double],
        "cf3hip_matrix_add_values": [_vp, C.c_uint32, _u32, _d],
        "cf3hip_matrix_get_diagonal": [_vp, _d],
        "cf3hip_matrix_apply": [_vp, C.c_double, _vp, C.c_double, _vp],
        "cf3hip_matrix_apply_fp32": [_vp, C.c_float, _vp, C.c_float, _vp],
        "cf3hip_matrix_get_values": [_vp, _d],
        "cf3hip_matrix_get_values_fp32": [_vp, _f],
        "cf3hip_matrix_debug_data": [_vp, _u32, _u32, _d],
        "cf3hip_lss_set_builder": [_vp, C.c_int],
        "cf3hip_matrix_get_column_and_replace_to_zero": [_vp, C.c_uint32, C.c_uint32, _d],
        "cf3hip_lss_get_builder": [_vp, C.POINTER(C.c_int)],
        "cf3hip_matrix_clone_to": [_vp, _vp],
        "cf3hip_matrix_read_native": [_vp, C.c_char_p],
        "cf3hip_vector_reset": [_vp, C.c_int, C.c_double],
        "cf3hip_vector_get": [_vp, C.c_int, _d],
        "cf3hip_vector_set": [_vp, C.c_int, _d],
        "cf3hip_vector_set_value": [_vp, C.c_int, C.c_uint32, C.c_uint32, C.c_double],
        "cf3hip_vector_get_value": [_vp, C.c_int, C.c_uint32, C.c_uint32, _d],
        "cf3hip_vector_device_ptr": [_vp, C.c_int, C.POINTER(_vp)],
        "cf3hip_vector_assign": [_vp, C.c_int, C.c_int],
        "cf3hip_lss_dirichlet": [_vp, C.c_uint32, _u32, _u32, _d],
        "cf3hip_matrix_tie_blockrow_pairs": [_vp, C.c_uint32, _u32, _u32],
        "cf3hip_lss_periodicity": [_vp, C.c_uint32, _u32, _u32],
        "cf3hip_lss_solve": [_vp, C.POINTER(SolverParams), C.POINTER(SolveStats)],
        "cf3hip_lss_residual": [_vp, _d],
        "cf3hip_lss_residual_noalloc": [_vp, _d],
        "cf3hip_fields_create": [_vp, C.POINTER(_vp)],
        "cf3hip_fields_destroy": [_vp],
        "cf3hip_fields_set": [_vp, C.c_int, _d],
        "cf3hip_fields_get": [_vp, C.c_int, _d],
        "cf3hip_fields_initialize": [_vp, C.c_double],
        "cf3hip_fields_copy": [_vp, C.c_int, C.c_int],
        "cf3hip_ns_linearize": [_vp],
        "cf3hip_ns_assemble": [_vp, _vp, C.POINTER(NSParams)],
        "cf3hip_ns_update": [_vp, _vp],
        "cf3hip_ns_solve": [_vp, _vp, C.POINTER(NSParams), C.POINTER(SolverParams), C.POINTER(SolveStats)],
        "cf3hip_ns_residual": [_vp, _vp, C.POINTER(NSParams), _vp, _vp],
        "cf3hip_batch_create": [_vp, C.c_uint32, C.POINTER(_vp)],
        "cf3hip_context_create_stream": [_vp],
        "cf3hip_batch_destroy": [_vp],
        "cf3hip_batch_assemble": [_vp, C.c_uint32, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(NSParams)],
        "cf3hip_si_create": [_vp, _vp, C.POINTER(_vp)],
        "cf3hip_si_destroy": [_vp],
        "cf3hip_si_initial": [_vp, _vp, C.POINTER(NSParams)],
        "cf3hip_si_velocity_assemble": [_vp, _vp, C.POINTER(NSParams)],
        "cf3hip_si_set_solution": [_vp, _vp],
        "cf3hip_si_action": [_vp, _vp, C.POINTER(NSParams), C.c_int],
        "cf3hip_si_set_options": [_vp, C.POINTER(SiOptions)],
        "cf3hip_si_get_options": [_vp, C.POINTER(SiOptions)],
        "cf3hip_si_get": [_vp, C.c_int, _d],
        "cf3hip_si_set": [_vp, C.c_int, _d],
        "cf3hip_si_step": [_vp, _vp, C.POINTER(NSParams), C.c_int, C.POINTER(SolverParams), C.POINTER(SolverParams),
                           C.c_uint32, _u32, _d, C.c_uint32, _u32, _d, C.POINTER(SolveStats)],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    _lib = L
    return L


def check(code):
    if code != OK:
        raise Cf3HipError(code, lib().cf3hip_last_error().decode())


def _p(a, t=_d):
    return a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _u32a(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


class Context:
    def __init__(self, device=0):
        self.h = _vp()
        check(lib().cf3hip_context_create(device, C.byref(self.h)))

    def synchronize(self):
        check(lib().cf3hip_context_synchronize(self.h))

    def set_stream(self, stream_ptr):
        check(lib().cf3hip_context_set_stream(self.h, _vp(stream_ptr)))

    def create_stream(self):
        """Use a private non-blocking stream (work of different contexts then overlaps)."""
        check(lib().cf3hip_context_create_stream(self.h))

    def memory(self):
        b = C.c_uint64()
        check(lib().cf3hip_context_memory(self.h, C.byref(b)))
        return b.value

    def set_async(self, enable):
        check(lib().cf3hip_context_set_async(self.h, int(enable)))

    def enable_timing(self, enable=True):
        check(lib().cf3hip_context_enable_timing(self.h, int(enable)))

    def kernel_time(self, kernel):
        """(total ms, launches) of a hot kernel since timing was enabled (HIP events on the stream)."""
        ms = C.c_double()
        n = C.c_uint64()
        check(lib().cf3hip_context_kernel_time(self.h, kernel, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def close(self):
        if self.h:
            check(lib().cf3hip_context_destroy(self.h))
            self.h = _vp()


class Batch:
    """Batched FP32 assembly of many tetrahedral systems in one persistent launch (hipGraph)."""

    def __init__(self, ctx, max_cases):
        self.h = _vp()
        self.max_cases = max_cases
        check(lib().cf3hip_batch_create(ctx.h, max_cases, C.byref(self.h)))

    def assemble(self, systems, fields, params):
        n = len(systems)
        ls = (_vp * max(n, 1))(*[s.h for s in systems])
        fs = (_vp * max(n, 1))(*[f.h for f in fields])
        ps = (NSParams * max(n, 1))(*params)
        check(lib().cf3hip_batch_assemble(self.h, n, ls, fs, ps))

    def close(self):
        if self.h:
            check(lib().cf3hip_batch_destroy(self.h))
            self.h = _vp()


class Mesh:
    def __init__(self, ctx, etype, coords, conn):
        self.ctx = ctx
        self.etype = etype
        coords = _f64(coords)
        conn = _u32a(conn)
        self.n_nodes, self.dim = coords.shape
        self.n_elems, self.k = conn.shape
        self.h = _vp()
        check(lib().cf3hip_mesh_create(ctx.h, etype, self.n_nodes, self.n_elems, _p(coords), _p(conn, _u32),
                                       C.byref(self.h)))

    def close(self):
        if self.h:
            check(lib().cf3hip_mesh_destroy(self.h))
            self.h = _vp()


class Fields:
    NCOMP = {FIELD_P: 1, FIELD_NU: 1}

    def __init__(self, mesh):
        self.mesh = mesh
        self.h = _vp()
        check(lib().cf3hip_fields_create(mesh.h, C.byref(self.h)))

    def ncomp(self, which):
        return self.NCOMP.get(which, self.mesh.dim)

    def set(self, which, arr):
        a = _f64(arr).reshape(self.mesh.n_nodes, self.ncomp(which))
        check(lib().cf3hip_fields_set(self.h, which, _p(a)))

    def get(self, which):
        a = np.zeros((self.mesh.n_nodes, self.ncomp(which)))
        check(lib().cf3hip_fields_get(self.h, which, _p(a)))
        return a if self.ncomp(which) > 1 else a[:, 0]

    def initialize(self, nu):
        check(lib().cf3hip_fields_initialize(self.h, nu))

    def copy(self, dst, src):
        check(lib().cf3hip_fields_copy(self.h, dst, src))

    def linearize(self):
        check(lib().cf3hip_ns_linearize(self.h))

    def close(self):
        if self.h:
            check(lib().cf3hip_fields_destroy(self.h))
            self.h = _vp()


class System:
    """The LSS system (matrix + rhs + solution) of one mesh."""

    def __init__(self, ctx, mesh=None, neq=None, graph=None, periodic=None):
        """periodic: (links, active) per mesh node (cf3hip_lss_create_periodic); the system then has one
        block row per unlinked node and p2m() maps mesh nodes to block rows."""
        self.ctx = ctx
        self.h = _vp()
        if mesh is not None:
            neq = neq or mesh.dim + 1
            if periodic is None:
                check(lib().cf3hip_lss_create(ctx.h, mesh.h, neq, C.byref(self.h)))
            else:
                links = _u32a(periodic[0])
                active = np.ascontiguousarray(periodic[1], dtype=np.uint8)
                check(lib().cf3hip_lss_create_periodic(ctx.h, mesh.h, neq, _p(links, _u32),
                                                       active.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(self.h)))
        else:
            start, conn, n_nodes = graph
            s, cn = _u32a(start), _u32a(conn)
            check(lib().cf3hip_lss_create_from_graph(ctx.h, n_nodes, neq, _p(s, _u32), _p(cn, _u32), C.byref(self.h)))
        self.mesh = mesh
        n = C.c_uint32()
        q = C.c_int()
        nb = C.c_uint64()
        check(lib().cf3hip_lss_info(self.h, C.byref(n), C.byref(q), C.byref(nb)))
        self.n_nodes, self.neq, self.n_blocks = n.value, q.value, nb.value
        self.dofs = self.n_nodes * self.neq

    def graph(self):
        start = np.zeros(self.n_nodes + 1, dtype=np.uint32)
        cols = np.zeros(self.n_blocks, dtype=np.uint32)
        check(lib().cf3hip_lss_get_graph(self.h, _p(start, _u32), _p(cols, _u32)))
        return start, cols

    def p2m(self):
        """Block row of every mesh node (the identity without periodic links)."""
        m = np.zeros(self.mesh.n_nodes, dtype=np.uint32)
        check(lib().cf3hip_lss_get_p2m(self.h, _p(m, _u32)))
        return m

    def slots(self):
        m = self.mesh
        s = np.zeros(m.n_elems * m.k * m.k, dtype=np.uint32)
        check(lib().cf3hip_lss_get_slots(self.h, _p(s, _u32)))
        return s

    def enable_fp32(self):
        check(lib().cf3hip_lss_enable_fp32(self.h))

    def round_fp32(self):
        """FP32 mirror := the FP64 values rounded to nearest."""
        check(lib().cf3hip_lss_round_fp32(self.h))

    def values(self):
        v = np.zeros(self.n_blocks * self.neq * self.neq)
        check(lib().cf3hip_matrix_get_values(self.h, _p(v)))
        return v

    def values_fp32(self):
        v = np.zeros(self.n_blocks * self.neq * self.neq, dtype=np.float32)
        check(lib().cf3hip_matrix_get_values_fp32(self.h, _p(v, _f)))
        return v

    def debug_data(self):
        nnz = self.n_blocks * self.neq * self.neq
        rp = np.zeros(self.dofs + 1, dtype=np.uint32)
        ci = np.zeros(nnz, dtype=np.uint32)
        v = np.zeros(nnz)
        check(lib().cf3hip_matrix_debug_data(self.h, _p(rp, _u32), _p(ci, _u32), _p(v)))
        return rp, ci, v

    def set_builder(self, builder):
        """Matrix builder semantics: BUILDER_VBR (TrilinosFEVbrMatrix) or BUILDER_CRS (TrilinosCrsMatrix)."""
        check(lib().cf3hip_lss_set_builder(self.h, builder))

    def builder(self):
        b = C.c_int()
        check(lib().cf3hip_lss_get_builder(self.h, C.byref(b)))
        return b.value

    def column_zero(self, iblockcol, ieq):
        """Matrix::get_column_and_replace_to_zero: the column (scalar rows), then zeroed in the matrix."""
        out = np.zeros(self.dofs)
        check(lib().cf3hip_matrix_get_column_and_replace_to_zero(self.h, iblockcol, ieq, _p(out)))
        return out

    def clone_to(self, other):
        check(lib().cf3hip_matrix_clone_to(self.h, other.h))

    def read_native(self, path):
        check(lib().cf3hip_matrix_read_native(self.h, os.fsencode(path)))

    def reset(self, value=0.0):
        check(lib().cf3hip_matrix_reset(self.h, value))

    def set_row(self, blockrow, ieq, diag, offdiag):
        check(lib().cf3hip_matrix_set_row(self.h, blockrow, ieq, diag, offdiag))

    def get_value(self, icol, irow):
        v = C.c_double()
        check(lib().cf3hip_matrix_get_value(self.h, icol, irow, C.byref(v)))
        return v.value

    def set_value(self, icol, irow, v):
        check(lib().cf3hip_matrix_set_value(self.h, icol, irow, v))

    def add_value(self, icol, irow, v):
        check(lib().cf3hip_matrix_add_value(self.h, icol, irow, v))

    def add_values(self, indices, mat):
        idx = _u32a(indices)
        m = _f64(mat)
        check(lib().cf3hip_matrix_add_values(self.h, len(idx), _p(idx, _u32), _p(m)))

    def diagonal(self):
        d = np.zeros(self.dofs)
        check(lib().cf3hip_matrix_get_diagonal(self.h, _p(d)))
        return d

    def apply_device(self, alpha, x_ptr, beta, y_ptr):
        check(lib().cf3hip_matrix_apply(self.h, alpha, _vp(x_ptr), beta, _vp(y_ptr)))

    def apply_fp32_device(self, alpha, x_ptr, beta, y_ptr):
        check(lib().cf3hip_matrix_apply_fp32(self.h, alpha, _vp(x_ptr), beta, _vp(y_ptr)))

    def vector(self, which):
        a = np.zeros(self.dofs)
        check(lib().cf3hip_vector_get(self.h, which, _p(a)))
        return a

    def set_vector(self, which, arr):
        a = _f64(arr)
        assert a.size == self.dofs
        check(lib().cf3hip_vector_set(self.h, which, _p(a)))

    def reset_vector(self, which, value=0.0):
        check(lib().cf3hip_vector_reset(self.h, which, value))

    def vector_ptr(self, which):
        p = _vp()
        check(lib().cf3hip_vector_device_ptr(self.h, which, C.byref(p)))
        return p.value

    def assign(self, to, frm):
        check(lib().cf3hip_vector_assign(self.h, to, frm))

    def dirichlet(self, nodes, eqs, values):
        n, e, v = _u32a(nodes), _u32a(eqs), _f64(values)
        check(lib().cf3hip_lss_dirichlet(self.h, len(n), _p(n, _u32), _p(e, _u32), _p(v)))

    def tie_blockrow_pairs(self, to, frm):
        """Matrix::tie_blockrow_pairs for the (to[i], frm[i]) block-row pairs in order."""
        t, f = _u32a(to), _u32a(frm)
        assert len(t) == len(f)
        check(lib().cf3hip_matrix_tie_blockrow_pairs(self.h, len(t), _p(t, _u32), _p(f, _u32)))

    def periodicity(self, to, frm):
        """System::periodicity for the (to[i], frm[i]) block-row pairs in order (matrix, rhs, solution)."""
        t, f = _u32a(to), _u32a(frm)
        assert len(t) == len(f)
        check(lib().cf3hip_lss_periodicity(self.h, len(t), _p(t, _u32), _p(f, _u32)))

    def set_vector_value(self, which, blockrow, ieq, v):
        check(lib().cf3hip_vector_set_value(self.h, which, blockrow, ieq, v))

    def solve(self, params, raise_on_failure=True):
        st = SolveStats()
        code = lib().cf3hip_lss_solve(self.h, C.byref(params), C.byref(st))
        if code != OK and (raise_on_failure or code != 5):
            check(code)
        return st

    def residual(self):
        r = C.c_double()
        check(lib().cf3hip_lss_residual(self.h, C.byref(r)))
        return r.value

    def residual_noalloc(self):
        """The same residual computed without device allocation (fixed per-GPU scratch)."""
        r = C.c_double()
        check(lib().cf3hip_lss_residual_noalloc(self.h, C.byref(r)))
        return r.value

    def assemble(self, fields, prm):
        check(lib().cf3hip_ns_assemble(self.h, fields.h, C.byref(prm)))

    def update(self, fields):
        check(lib().cf3hip_ns_update(self.h, fields.h))

    def ns_solve(self, fields, ns_prm, params, raise_on_failure=True):
        st = SolveStats()
        code = lib().cf3hip_ns_solve(self.h, fields.h, C.byref(ns_prm), C.byref(params), C.byref(st))
        if code != OK and (raise_on_failure or code != 5):
            check(code)
        return st

    def ns_residual_device(self, fields, ns_prm, x_ptr, r_ptr):
        check(lib().cf3hip_ns_residual(self.h, fields.h, C.byref(ns_prm), _vp(x_ptr), _vp(r_ptr)))

    def close(self):
        if self.h:
            check(lib().cf3hip_lss_destroy(self.h))
            self.h = _vp()


SI_URHS, SI_PRHS, SI_AUP = 0, 1, 2
SI_VEC_A, SI_VEC_U, SI_VEC_P, SI_VEC_DPS, SI_VEC_MINV = 0, 1, 2, 3, 4


class SemiImplicit:
    """NavierStokesSemiImplicit on a velocity system (neq = dim) and a pressure system (neq = 1) of one
    mesh (cf3hip_si_*)."""

    def __init__(self, u_system, p_system):
        self.u_system, self.p_system = u_system, p_system
        self.h = _vp()
        check(lib().cf3hip_si_create(u_system.h, p_system.h, C.byref(self.h)))

    def initial(self, fields, prm):
        check(lib().cf3hip_si_initial(self.h, fields.h, C.byref(prm)))

    def velocity_assemble(self, fields, prm):
        check(lib().cf3hip_si_velocity_assemble(self.h, fields.h, C.byref(prm)))

    def set_solution(self, fields):
        check(lib().cf3hip_si_set_solution(self.h, fields.h))

    def action(self, fields, prm, which):
        check(lib().cf3hip_si_action(self.h, fields.h, C.byref(prm), which))

    def set_options(self, enable_body_force=False, pressure_rcg_solve=False):
        """The enable_body_force and pressure_rcg_solve options of NavierStokesSemiImplicit."""
        check(lib().cf3hip_si_set_options(self.h, C.byref(SiOptions(int(enable_body_force),
                                                                      int(pressure_rcg_solve)))))

    def options(self):
        o = SiOptions()
        check(lib().cf3hip_si_get_options(self.h, C.byref(o)))
        return {"enable_body_force": bool(o.enable_body_force), "pressure_rcg_solve": bool(o.pressure_rcg_solve)}

    def _n(self, which):
        rows = self.u_system.n_nodes
        return rows * self.u_system.neq if which in (SI_VEC_A, SI_VEC_U, SI_VEC_MINV) else rows

    def get(self, which):
        a = np.zeros(self._n(which))
        check(lib().cf3hip_si_get(self.h, which, _p(a)))
        return a

    def set(self, which, arr):
        a = _f64(arr).ravel()
        assert a.size == self._n(which)
        check(lib().cf3hip_si_set(self.h, which, _p(a)))

    def step(self, fields, prm, u_solver, p_solver, nb_iterations=2, ubc=None, pbc=None):
        """ubc / pbc: (mesh nodes, new values) of the velocity (n x dim) and pressure conditions."""
        un, uv = (np.zeros(0, np.uint32), np.zeros(0)) if ubc is None else (_u32a(ubc[0]), _f64(ubc[1]))
        pn, pv = (np.zeros(0, np.uint32), np.zeros(0)) if pbc is None else (_u32a(pbc[0]), _f64(pbc[1]))
        stats = (SolveStats * (2 * nb_iterations))()
        check(lib().cf3hip_si_step(self.h, fields.h, C.byref(prm), nb_iterations, C.byref(u_solver), C.byref(p_solver),
                                   len(un), _p(un, _u32), _p(uv), len(pn), _p(pn, _u32), _p(pv), stats))
        return list(stats)

    def close(self):
        if self.h:
            check(lib().cf3hip_si_destroy(self.h))
            self.h = _vp()

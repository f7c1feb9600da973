/*
 * cf3hip: C ABI of the MI355X (gfx950) implementation of COOLFluiD 3's UFEM incompressible
 * Navier-Stokes hot path: PSPG/SUPG/bulk-viscosity element assembly, its deterministic scatter into
 * the LSS block-CSR matrix, symmetric Dirichlet conditions, SpMV and the Krylov solve.
 *
 * Plain pointers and sizes only.  Host arrays are marked "host"; arrays marked "device" are device
 * pointers (for zero-copy interop with callers that keep vectors on the GPU).  Every function
 * returns CF3HIP_OK (0) or an error code; cf3hip_last_error() gives the message of the last failure
 * in the calling thread.  All calls are synchronous with respect to the host unless stated.
 *
 * Reference interfaces replaced (paths relative to the coolfluid3 source tree) are cited per entry.
 */
#ifndef CF3HIP_H
#define CF3HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif
/* the library is built with hidden visibility: only these C entry points are exported */
#if defined(__GNUC__)
#pragma GCC visibility push(default)
#endif

/* ---- status codes: mapped to the reference exceptions by the C++ host layer ---- */
#define CF3HIP_OK                 0
#define CF3HIP_ERR_INVALID_ARG    1  /* common::BadValue                                   */
#define CF3HIP_ERR_NOT_CREATED    2  /* common::SetupError (LSS not created, bad theta)     */
#define CF3HIP_ERR_HIP            3  /* HIP runtime failure                                 */
#define CF3HIP_ERR_OUT_OF_MEMORY  4  /* device allocation failed                            */
#define CF3HIP_ERR_NOT_CONVERGED  5  /* Krylov solve hit max iterations or stagnated        */
#define CF3HIP_ERR_ILLEGAL_ENTRY  6  /* BadValue "Trying to access an illegal entry"        */
#define CF3HIP_ERR_NOT_IMPLEMENTED 7 /* common::NotImplemented                              */
#define CF3HIP_ERR_LIMIT          8  /* mesh exceeds a documented kernel limit              */
#define CF3HIP_ERR_SINGULAR       9  /* non-invertible element Jacobian (ElementData.hpp:312-313) */

/* element types (cf3/mesh/LagrangeP1: Triag2D, Tetra3D, Hexa3D, Quad2D, Prism3D) */
#define CF3HIP_TRIAG2D 0
#define CF3HIP_TETRA3D 1
#define CF3HIP_HEXA3D  2
#define CF3HIP_QUAD2D  3   /* LagrangeP1 Quad2D: generic expression, 2 x 2 Gauss points */
#define CF3HIP_PRISM3D 4   /* LagrangeP1 Prism3D (NavierStokesPrisms.cpp): generic expression, 6 Gauss points */

/* SUPG tau variants (plugins/UFEM/src/UFEM/SUPG.hpp:62-74) */
#define CF3HIP_SUPG_TEZDUYAR 0
#define CF3HIP_SUPG_METRIC   1
#define CF3HIP_SUPG_CF2      2   /* generic expression only; the specialised tri/tet kernels always use METRIC,
                                     as the reference's specialisations do (NavierStokesSpecializations.hpp:57-58) */

/* value precision of an assembled matrix / Krylov solve */
#define CF3HIP_FP64 0
#define CF3HIP_FP32 1

typedef struct cf3hip_context_s* cf3hip_context;
typedef struct cf3hip_mesh_s*    cf3hip_mesh;
typedef struct cf3hip_lss_s*     cf3hip_lss;
typedef struct cf3hip_fields_s*  cf3hip_fields;

const char* cf3hip_last_error(void);
const char* cf3hip_version(void);

/* ---- context: one device + one HIP stream (no reference counterpart: process/device setup,
 *      the role of common::PE::Comm::init in cf3/common/PE/Comm.cpp:87) ---- */
int cf3hip_context_create(int device, cf3hip_context* ctx);
int cf3hip_context_destroy(cf3hip_context ctx);
/* use a caller-owned hipStream_t (NULL = default stream) for every later launch */
int cf3hip_context_set_stream(cf3hip_context ctx, void* hip_stream);
/* give the context a private non-blocking stream of its own (default priority), used from now on
 * (contexts on separate streams overlap on the device: the serving loop) */
int cf3hip_context_create_stream(cf3hip_context ctx);
int cf3hip_context_synchronize(cf3hip_context ctx);
/* bytes currently held by device allocations of this context */
int cf3hip_context_memory(cf3hip_context ctx, uint64_t* bytes);
/* asynchronous mode: assembly, Dirichlet and apply calls return without waiting for the device;
 * element-validity errors are reported by the next cf3hip_context_synchronize */
int cf3hip_context_set_async(cf3hip_context ctx, int enable);
/* per-launch HIP-event timing of the hot kernels (the role of the per-action timers of
 * cf3/common/TimedComponent.cpp:24-107); enabling resets the accumulated times */
#define CF3HIP_KERNEL_ASSEMBLY   0
#define CF3HIP_KERNEL_SPMV       1
#define CF3HIP_KERNEL_SPMV_FP32  2
#define CF3HIP_KERNEL_RESIDUAL   3
#define CF3HIP_KERNEL_COUNT      4
int cf3hip_context_enable_timing(cf3hip_context ctx, int enable);
/* total milliseconds and number of timed launches of one kernel since timing was enabled */
int cf3hip_context_kernel_time(cf3hip_context ctx, int kernel, double* total_ms, uint64_t* launches);

/* ---- mesh: one Elements component of one LagrangeP1 type, device SoA coordinates
 *      (cf3/mesh Dictionary coordinates + Connectivity, cf3/mesh/Field.hpp:34, Connectivity.hpp:23) ---- */
/* coords: host, node-major n_nodes x dim (cf3 Field table layout); conn: host, n_elems x k */
int cf3hip_mesh_create(cf3hip_context ctx, int etype, uint32_t n_nodes, uint32_t n_elems,
                       const double* coords, const uint32_t* conn, cf3hip_mesh* mesh);
int cf3hip_mesh_destroy(cf3hip_mesh mesh);
/* replace coordinates (same topology), host node-major */
int cf3hip_mesh_set_coordinates(cf3hip_mesh mesh, const double* coords);

/* ---- LSS system: matrix + rhs + solution.
 *      Replaces math::LSS::System::create (cf3/math/LSS/System.cpp:82-110) with the sparsity of
 *      plugins/UFEM/src/UFEM/SparsityBuilder.cpp:31-192 built on the device.  The matrix is block
 *      CSR with neq x neq row-major blocks, node rows contiguous (semantics of
 *      TrilinosFEVbrMatrix, cf3/math/LSS/Trilinos/TrilinosFEVbrMatrix.cpp:71-134). ---- */
int cf3hip_lss_create(cf3hip_context ctx, cf3hip_mesh mesh, int neq, cf3hip_lss* lss);
/* System::create with periodic links (System.cpp:82-99 with periodic_links_nodes / periodic_links_active,
 * the matrix map of cf3/math/LSS/Trilinos/TrilinosDetail.cpp:147-227; the links are those of
 * cf3/mesh/actions/LinkPeriodicNodes.cpp:136-180).  links, active: host, one entry per mesh node; a node
 * with active != 0 is linked to links[node].  The unlinked nodes take the block rows in node order, a
 * linked node shares the row of the end of its chain of links (m_p2m of TrilinosFEVbrMatrix).  Element
 * geometry and fields stay per mesh node; the sparsity, the scatter and cf3hip_ns_update go through the
 * map.  Every index argument of the lss, matrix and vector calls on such a system is a block row, not a
 * mesh node (the reference converts inside, with m_p2m): cf3hip_lss_get_p2m gives the map.  The
 * matrix-free residual (cf3hip_ns_residual, the FP32 form of cf3hip_ns_solve) and batched assembly
 * return CF3HIP_ERR_NOT_IMPLEMENTED on such a system. */
int cf3hip_lss_create_periodic(cf3hip_context ctx, cf3hip_mesh mesh, int neq, const uint32_t* links,
                               const uint8_t* active, cf3hip_lss* lss);
/* block row of every mesh node (host, mesh n_nodes entries; the identity without periodic links) */
int cf3hip_lss_get_p2m(cf3hip_lss lss, uint32_t* p2m);
/* System::create(cp, neq, node_connectivity, starting_indices) for a caller-given node graph
 * (rows must be sorted and contain the diagonal).  No element scatter map: assembly needs lss_create. */
int cf3hip_lss_create_from_graph(cf3hip_context ctx, uint32_t n_nodes, int neq,
                                 const uint32_t* starting_indices, const uint32_t* node_connectivity,
                                 cf3hip_lss* lss);
int cf3hip_lss_destroy(cf3hip_lss lss);                  /* System::destroy (System.cpp:192-199) */
int cf3hip_lss_info(cf3hip_lss lss, uint32_t* n_nodes, int* neq, uint64_t* n_blocks);
/* node graph (starting_indices n_nodes+1, node_connectivity n_blocks), host */
int cf3hip_lss_get_graph(cf3hip_lss lss, uint32_t* starting_indices, uint32_t* node_connectivity);
/* element-to-slot map: block slot of element pair (a, b), host, n_elems*k*k */
int cf3hip_lss_get_slots(cf3hip_lss lss, uint32_t* slots);
/* enable an FP32 mirror of the values (precision form); the FP32 assembly writes it */
int cf3hip_lss_enable_fp32(cf3hip_lss lss);
/* round the FP64 values into the FP32 mirror (each value to nearest): the FP32 operator of a precision
 * solve on a stored system whose values were set or assembled in FP64.  No counterpart in the reference
 * (its Trilinos matrices hold one precision); callers: cf3hip_lss_solve with precision CF3HIP_FP32. */
int cf3hip_lss_round_fp32(cf3hip_lss lss);

/* ---- Matrix interface (cf3/math/LSS/Matrix.hpp:35-180) ---- */
int cf3hip_matrix_reset(cf3hip_lss lss, double value);                          /* Matrix::reset */
int cf3hip_matrix_set_value(cf3hip_lss lss, uint32_t icol, uint32_t irow, double v);  /* column first */
int cf3hip_matrix_add_value(cf3hip_lss lss, uint32_t icol, uint32_t irow, double v);
int cf3hip_matrix_get_value(cf3hip_lss lss, uint32_t icol, uint32_t irow, double* v);
int cf3hip_matrix_set_row(cf3hip_lss lss, uint32_t blockrow, uint32_t ieq, double diag, double offdiag);
/* Matrix::add_values(BlockAccumulator): indices (host, nb node ids), mat (host, row-major
 * (nb*neq)^2, node-major rows) -- BlockAccumulator.hpp:28-74, TrilinosFEVbrMatrix.cpp:315-355 */
int cf3hip_matrix_add_values(cf3hip_lss lss, uint32_t nb, const uint32_t* indices, const double* mat);
/* Matrix::set_values / get_values(BlockAccumulator) (Matrix.hpp:87, 95), same layout as add_values */
int cf3hip_matrix_set_values(cf3hip_lss lss, uint32_t nb, const uint32_t* indices, const double* mat);
int cf3hip_matrix_get_block_values(cf3hip_lss lss, uint32_t nb, const uint32_t* indices, double* mat);
/* Matrix::set_diagonal / add_diagonal (Matrix.hpp:114-117): host, dofs */
int cf3hip_matrix_set_diagonal(cf3hip_lss lss, const double* diag);
int cf3hip_matrix_add_diagonal(cf3hip_lss lss, const double* diag);
/* ---- matrix builder (LSSAction option matrix_builder, LSSAction.cpp:105-109).  The node-block storage
 *      is the same for both; the builder selects the semantics of the reference's two matrices:
 *      CF3HIP_BUILDER_VBR (default, TrilinosFEVbrMatrix) or CF3HIP_BUILDER_CRS (TrilinosCrsMatrix), whose
 *      symmetric Dirichlet (TrilinosCrsMatrix.cpp:325-402) caches each eliminated column until the next
 *      reset and subtracts it from the rhs again when the same condition is applied again, whose
 *      get_column_and_replace_to_zero is not implemented (:318-321), and which implements clone_to
 *      (:518-538) and read_native (:540-546).  Choose before the first condition or tie
 *      (CF3HIP_ERR_INVALID_ARG otherwise).  The matrix-free residual of the precision form refuses a CRS
 *      system. ---- */
#define CF3HIP_BUILDER_VBR 0
#define CF3HIP_BUILDER_CRS 1
int cf3hip_lss_set_builder(cf3hip_lss lss, int builder);
int cf3hip_lss_get_builder(cf3hip_lss lss, int* builder);
/* Matrix::clone_to (TrilinosCrsMatrix.cpp:518-538): values and cached Dirichlet columns into another
 * CRS-builder system of the same graph (CF3HIP_ERR_NOT_CREATED for another builder or graph, as the
 * reference's SetupError); the VBR builder returns CF3HIP_ERR_NOT_IMPLEMENTED (TrilinosFEVbrMatrix.cpp:798) */
int cf3hip_matrix_clone_to(cf3hip_lss src, cf3hip_lss dst);
/* Matrix::read_native (TrilinosCrsMatrix.cpp:540-546): a Matrix Market coordinate file (real, general or
 * symmetric) of dofs x dofs into the values, which are zeroed first; an entry outside the pattern gives
 * CF3HIP_ERR_ILLEGAL_ENTRY; the conditions are cleared.  VBR: CF3HIP_ERR_NOT_IMPLEMENTED
 * (TrilinosFEVbrMatrix.cpp:805) */
int cf3hip_matrix_read_native(cf3hip_lss lss, const char* path);
/* Matrix::get_column_and_replace_to_zero (Matrix.hpp:103, TrilinosFEVbrMatrix.cpp:452-478):
 * values (host, dofs) receive column (iblockcol, ieq), which is then zeroed */
int cf3hip_matrix_get_column_and_replace_to_zero(cf3hip_lss lss, uint32_t iblockcol, uint32_t ieq, double* values);
int cf3hip_matrix_get_diagonal(cf3hip_lss lss, double* diag);                  /* host, dofs */
/* y = alpha*A*x + beta*y (Matrix::apply, TrilinosDetail.cpp:380-411); device pointers, dofs each */
int cf3hip_matrix_apply(cf3hip_lss lss, double alpha, const double* x_dev, double beta, double* y_dev);
/* FP32 mirror apply, float device vectors */
int cf3hip_matrix_apply_fp32(cf3hip_lss lss, float alpha, const float* x_dev, float beta, float* y_dev);
/* raw block values (host, n_blocks*neq*neq) and point-CSR view (Matrix::debug_data) */
int cf3hip_matrix_get_values(cf3hip_lss lss, double* values);
int cf3hip_matrix_get_values_fp32(cf3hip_lss lss, float* values);
int cf3hip_matrix_debug_data(cf3hip_lss lss, uint32_t* row_ptr, uint32_t* col_idx, double* vals);

/* ---- Vector interface (cf3/math/LSS/Vector.hpp:36-173) for the system rhs / solution ---- */
#define CF3HIP_RHS      0
#define CF3HIP_SOLUTION 1
int cf3hip_vector_reset(cf3hip_lss lss, int which, double value);
int cf3hip_vector_get(cf3hip_lss lss, int which, double* host);
int cf3hip_vector_set(cf3hip_lss lss, int which, const double* host);
int cf3hip_vector_set_value(cf3hip_lss lss, int which, uint32_t blockrow, uint32_t ieq, double v);
int cf3hip_vector_get_value(cf3hip_lss lss, int which, uint32_t blockrow, uint32_t ieq, double* v);
/* device pointer of the rhs / solution storage (FP64, dofs entries), owned by the system */
int cf3hip_vector_device_ptr(cf3hip_lss lss, int which, double** dev);
/* Vector::assign (Vector.hpp): copy one system vector into the other */
int cf3hip_vector_assign(cf3hip_lss lss, int to, int from);
/* Vector::add_value(iblockrow, ieq) (Vector.hpp:79) */
int cf3hip_vector_add_value(cf3hip_lss lss, int which, uint32_t blockrow, uint32_t ieq, double v);
/* Vector::set/add/get_rhs_values and _sol_values(BlockAccumulator) (Vector.hpp:90-105): indices
 * (host, nb node ids), vals (host, nb*neq, node-major) */
int cf3hip_vector_set_values(cf3hip_lss lss, int which, uint32_t nb, const uint32_t* indices, const double* vals);
int cf3hip_vector_add_values(cf3hip_lss lss, int which, uint32_t nb, const uint32_t* indices, const double* vals);
int cf3hip_vector_get_values(cf3hip_lss lss, int which, uint32_t nb, const uint32_t* indices, double* vals);
/* Vector::update (this += alpha * source, Vector.hpp:150) and Vector::scale (Vector.hpp:154) */
int cf3hip_vector_update(cf3hip_lss lss, int to, int from, double alpha);
int cf3hip_vector_scale(cf3hip_lss lss, int which, double alpha);

/* ---- System::dirichlet with preserve_symmetry (System.cpp:240-255 ->
 *      TrilinosFEVbrMatrix::symmetric_dirichlet, TrilinosFEVbrMatrix.cpp:482-553), applied to a list
 *      in order: later conditions on a dof set its rhs/solution value, the column elimination
 *      uses the first value (the matrix column is already zero afterwards).  Host arrays. ---- */
int cf3hip_lss_dirichlet(cf3hip_lss lss, uint32_t nbc, const uint32_t* nodes, const uint32_t* eqs,
                         const double* values);

/* ---- Matrix::tie_blockrow_pairs (cf3/math/LSS/Matrix.hpp; TrilinosFEVbrMatrix.cpp:558-590,
 *      TrilinosCrsMatrix.cpp:413-460) for (to[i], from[i]) block-row pairs applied in order: row from is
 *      added into row to and becomes x_from - x_to = 0, and in row to the block of column from moves
 *      onto column to.  Both rows must hold the same block columns (BadValue otherwise).  Also applied
 *      to the FP32 mirror when enabled.  Host arrays. ---- */
int cf3hip_matrix_tie_blockrow_pairs(cf3hip_lss lss, uint32_t npairs, const uint32_t* to, const uint32_t* from);
/* ---- System::periodicity (cf3/math/LSS/System.cpp:259-282) for each pair in order: the tie above,
 *      rhs_to += rhs_from, rhs_from = 0, and both solution entries set to their mean.  A system with
 *      tied rows is solved from its stored values (cf3hip_lss_solve); the matrix-free precision form
 *      of cf3hip_ns_solve refuses it (NotImplemented) until the next assembly or reset. ---- */
int cf3hip_lss_periodicity(cf3hip_lss lss, uint32_t npairs, const uint32_t* to, const uint32_t* from);

/* ---- SolutionStrategy (cf3/math/LSS/SolutionStrategy.hpp:35-59; Trilinos strategy
 *      TrilinosStratimikosStrategy.cpp:135-207) ---- */
#define CF3HIP_BICGSTAB_JACOBI 0
#define CF3HIP_GMRES_ILU0      1
/* recycling conjugate gradients for symmetric positive definite systems (RCGStrategy,
 * cf3/math/LSS/Trilinos/RCGStrategy.cpp:59-171, Belos::RCGSolMgr): Jacobi-preconditioned CG deflated by
 * a recycle space of the system that is kept between solves and refreshed from harmonic Ritz vectors
 * after every cycle of `restart` directions; FP64 only */
#define CF3HIP_RCG_JACOBI      2
typedef struct cf3hip_solver_params {
  int method;        /* CF3HIP_BICGSTAB_JACOBI, CF3HIP_GMRES_ILU0 or CF3HIP_RCG_JACOBI */
  int restart;       /* GMRES restart length; RCG: directions per cycle ("Num Blocks", <= 96, 0: 40) */
  int max_iter;
  double tol;        /* stop on ||b - A x||_2 <= tol ||b||_2 */
  int precision;     /* CF3HIP_FP64, or CF3HIP_FP32 = FP32-matrix BiCGStab with FP64 residual
                        replacement, then FP64 flexible GMRES around FP32 inner solves if that stalls
                        (restart length min(restart, 16)) */
  double inner_tol;  /* relative tolerance of each FP32 inner solve of the GMRES phase (0: 1e-2) */
  int max_outer;     /* cap on the FP32 inner solves of the GMRES phase (0: no cap) */
  int block_scale;   /* precision FP32 only.  0: the FP32 Krylov phase runs on the FP32 values with point
                        Jacobi; 1 (the default of the host layers): on B = fl32(A32 D^-1), D the diagonal
                        blocks of the FP32 values inverted in FP64, with x = D^-1 y (block-Jacobi right
                        preconditioning, applied before the FP32 rounding).  The FP32 mirror itself is
                        left unchanged. */
  int fp32_vectors;  /* precision FP32, 4 equations per node: 1 = the Krylov vectors of the FP32 phase in
                        FP32 and its products accumulated in FP32 (dot products in FP64); 0 = FP64 */
  int recycle;       /* RCG: recycled vectors ("Num Recycled Blocks", 0 to 32; 0 = plain Jacobi CG) */
  int rcg_space;     /* RCG: which of the system's two recycle spaces the solve uses and updates (0, 1) */
} cf3hip_solver_params;
typedef struct cf3hip_solve_stats {
  int iterations;    /* BiCGStab iterations (FP32 form: both phases) */
  int outer;         /* FP32 form: residual replacements + inner solves (1 for FP64) */
  int spmvs;
  int converged;
  double rel_residual;  /* ||b - A x|| / ||b|| in FP64 at exit */
} cf3hip_solve_stats;
/* solve A x = rhs into the solution vector, using it as the initial guess */
int cf3hip_lss_solve(cf3hip_lss lss, const cf3hip_solver_params* prm, cf3hip_solve_stats* stats);
/* ||rhs - A sol||_2 / ||rhs||_2 (TrilinosStratimikosStrategy::compute_residual) */
int cf3hip_lss_residual(cf3hip_lss lss, double* rel);
/* The same residual without any device allocation.  No counterpart in the reference: for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Partial sums live in a fixed per-GPU device array (2 x 2048 + 8 doubles, held until the process
 * ends), whatever the system size; systems of 2^32 dofs or more and neq above 4 are refused
 * (CF3HIP_ERR_LIMIT).  Calls that overlap on one GPU share that array and race. */
int cf3hip_lss_residual_noalloc(cf3hip_lss lss, double* rel);

/* ---- UFEM NavierStokes fields (tags navier_stokes_solution, linearized_velocity,
 *      navier_stokes_viscosity, body_force; ns_implementation/NavierStokes.cpp:49-68) ---- */
#define CF3HIP_FIELD_U      0   /* Velocity (dim)            */
#define CF3HIP_FIELD_P      1   /* Pressure (1)              */
#define CF3HIP_FIELD_UADV   2   /* AdvectionVelocity (dim)   */
#define CF3HIP_FIELD_U1     3   /* AdvectionVelocity1 (dim)  */
#define CF3HIP_FIELD_U2     4
#define CF3HIP_FIELD_U3     5
#define CF3HIP_FIELD_NU     6   /* EffectiveViscosity (1)    */
#define CF3HIP_FIELD_G      7   /* Force (dim), body force   */
int cf3hip_fields_create(cf3hip_mesh mesh, cf3hip_fields* fields);
int cf3hip_fields_destroy(cf3hip_fields fields);
/* host, node-major n_nodes x ncomp */
int cf3hip_fields_set(cf3hip_fields fields, int which, const double* host);
int cf3hip_fields_get(cf3hip_fields fields, int which, double* host);
/* initial conditions of NavierStokes::on_initial_conditions_set (NavierStokes.cpp:194-205):
 * u_adv = u1 = u2 = u3 = u, nu_eff = nu */
int cf3hip_fields_initialize(cf3hip_fields fields, double nu);
/* device copy of one field into another of the same shape (e.g. u_adv := u for a Picard iteration
 * of a steady solve, the linearisation of NavierStokes' steady variants); stream-ordered */
int cf3hip_fields_copy(cf3hip_fields fields, int dst, int src);

/* ---- NavierStokes LSS action (plugins/UFEM/src/UFEM/NavierStokes.hpp:34-118) ---- */
typedef struct cf3hip_ns_params {
  double rho;               /* NavierStokesPhysics "density" */
  double dt;                /* LSSActionUnsteady dt (invdt = 1/dt) */
  double theta;             /* "theta" option, 0 <= theta <= 1 */
  int use_specializations;  /* "use_specializations" option (tri/tet) */
  int supg_type;            /* "supg_type" option for the generic expression */
  double c1, c2;            /* METRIC constants (1, 16) */
  double alpha_ps, alpha_su, alpha_bu;
  int precision;            /* CF3HIP_FP64 (drop-in) or CF3HIP_FP32 (element arithmetic and values) */
  double u_ref;             /* ComputeTau "u_ref" (SUPG.hpp:103), reference velocity of supg_type cf2;
                               must be positive when cf2 is selected */
} cf3hip_ns_params;
/* LinearizeU node loop (NavierStokes.cpp:105) */
int cf3hip_ns_linearize(cf3hip_fields fields);
/* ZeroLSS (matrix and rhs reset, solution kept; ZeroLSS.cpp:56-85) + Assembly
 * (NavierStokesAssembly.hpp:28-75 / NavierStokesSpecializations.hpp) + scatter into the system.
 * precision FP32 writes the FP32 mirror values and the rhs in FP64. */
int cf3hip_ns_assemble(cf3hip_lss lss, cf3hip_fields fields, const cf3hip_ns_params* prm);
/* UpdateActions (NavierStokes.cpp:172-180): u3=u2, u2=u1, u1=u, u += du, p += dp from the solution */
int cf3hip_ns_update(cf3hip_lss lss, cf3hip_fields fields);
/* SolveLSS of the NS system (SolveLSS.cpp:44-55).  precision FP64: the stored FP64 system (as
 * cf3hip_lss_solve).  precision FP32: the FP64 operator is the residual b' - A' x evaluated
 * matrix-free from the FP64 element matrices of the same fields (after the Dirichlet conditions
 * applied with cf3hip_lss_dirichlet); only the FP32 assembly of the step is needed.  BiCGStab +
 * Jacobi on the FP32 element matrices (FP64 vectors) whose residual is replaced by that FP64
 * residual after every 10-fold drop; if a replacement raises the FP64 residual or two in a row fail
 * to halve it, the solve continues with flexible GMRES in FP64 on the FP64 operator, preconditioned
 * by BiCGStab solves on the FP32 values (FP64 vectors) to inner_tol.  Converged when the FP64 residual
 * meets tol.  stats: iterations = BiCGStab iterations of both phases, outer = residual replacements
 * + inner solves. */
int cf3hip_ns_solve(cf3hip_lss lss, cf3hip_fields fields, const cf3hip_ns_params* ns_prm,
                    const cf3hip_solver_params* prm, cf3hip_solve_stats* stats);
/* ---- serving: batched assembly (no reference counterpart; the serving form of cf3hip) ----
 * A batch holds a device table of per-system assembly arguments and a hipGraph with one persistent
 * launch that assembles the FP32 element matrices of every system in the table (tetrahedra,
 * specialised kernel; each system needs cf3hip_lss_enable_fp32).  The table is refreshed on every
 * call, the graph is captured once.  Synchronous: returns when all systems are assembled (ZeroLSS +
 * Assembly of each system; its Dirichlet conditions of the previous step are forgotten). */
typedef struct cf3hip_batch_s* cf3hip_batch;
int cf3hip_batch_create(cf3hip_context ctx, uint32_t max_cases, cf3hip_batch* batch);
int cf3hip_batch_destroy(cf3hip_batch batch);
int cf3hip_batch_assemble(cf3hip_batch batch, uint32_t n, const cf3hip_lss* lss, const cf3hip_fields* fields,
                          const cf3hip_ns_params* params /* n entries */);

/* matrix-free FP64 residual r = b' - A' x of the NS system with its Dirichlet conditions
 * (device pointers, dofs entries) -- no counterpart in the reference */
int cf3hip_ns_residual(cf3hip_lss lss, cf3hip_fields fields, const cf3hip_ns_params* ns_prm, const double* x_dev,
                       double* r_dev);

/* ---- UFEM NavierStokesSemiImplicit (plugins/UFEM/src/UFEM/ns_semi_implicit) on triangles,
 *      quadrilaterals, tetrahedra and hexahedra.  u_lss: the VelocityLSS system (neq = dim), p_lss: the PressureLSS system
 *      (neq = 1), created from the same mesh with the same periodic links.  The "solution" vectors of
 *      the two systems hold delta_a and delta_p inside the inner loop and u, p outside it
 *      (implementation/NavierStokesSemiImplicit.cpp:84-190, :344, :354). ---- */
typedef struct cf3hip_si_s* cf3hip_si;
#define CF3HIP_SI_URHS 0    /* URHSAssembly: velocity rhs += velocity_rhs(...) (MatrixAssembly.hpp:206-330) */
#define CF3HIP_SI_PRHS 1    /* PRHSAssembly: pressure rhs += pressure_rhs(...) (MatrixAssembly.hpp:148-204) */
#define CF3HIP_SI_AUP  2    /* ApplyAup: velocity rhs += apply_aup(delta_p) (MatrixAssembly.hpp:332-376) */
#define CF3HIP_SI_VEC_A    0   /* inner-loop vectors (InnerLoop members a, u, p, delta_p_sum, rows x ncomp) */
#define CF3HIP_SI_VEC_U    1
#define CF3HIP_SI_VEC_P    2
#define CF3HIP_SI_VEC_DPS  3
#define CF3HIP_SI_VEC_MINV 4   /* reciprocal of the lumped velocity mass (SetupInnerLoopData, :236-262) */
int cf3hip_si_create(cf3hip_lss u_lss, cf3hip_lss p_lss, cf3hip_si* si);
int cf3hip_si_destroy(cf3hip_si si);
/* options of NavierStokesSemiImplicit (implementation/NavierStokesSemiImplicit.cpp:309-317), both off
 * by default as there */
typedef struct cf3hip_si_options {
  int enable_body_force;   /* "enable_body_force": URHSAssembly takes the overload of velocity_rhs with the
                              body_force field (CF3HIP_FIELD_G), + int N^T g_i over the second-order rule
                              (MatrixAssembly.hpp:339-363, :460-477) */
  int pressure_rcg_solve;  /* "pressure_rcg_solve": the first pressure solve of each step and the later ones
                              use two RCG strategies, each with its own recycle space kept between steps
                              (:158-166, :532-537): CF3HIP_RCG_JACOBI in recycle spaces 0 and 1 of p_lss,
                              with p_solver's tol and max_iter, and its restart and recycle when p_solver
                              selects CF3HIP_RCG_JACOBI (otherwise 40 and 16) */
} cf3hip_si_options;
int cf3hip_si_set_options(cf3hip_si si, const cf3hip_si_options* options);
int cf3hip_si_get_options(cf3hip_si si, cf3hip_si_options* options);
/* initial conditions: PressureSystemAssembly (PressureSystem.cpp:120-176: p_lss reset, the simplified
 * pressure matrix theta (tau_ps + dt) int G^T G with tau_ps from u), ZeroVelocitySystem +
 * MassMatrixAssembly (lumped mass into the velocity rhs) + SetupInnerLoopData (its reciprocal) */
int cf3hip_si_initial(cf3hip_si si, cf3hip_fields fields, const cf3hip_ns_params* prm);
/* VelocityLSS Assembly: ZeroLSS + velocity matrix T + theta dt M (MatrixAssembly.hpp:49-146, :447-460) */
int cf3hip_si_velocity_assemble(cf3hip_si si, cf3hip_fields fields, const cf3hip_ns_params* prm);
/* SetSolution (:344): solution(u) = u, solution(p) = p (node loop; the last node of a row wins) */
int cf3hip_si_set_solution(cf3hip_si si, cf3hip_fields fields);
/* one element action (CF3HIP_SI_*) accumulated into the rhs of its system from the inner-loop vectors,
 * the solutions (delta_a, delta_p) and the fields */
int cf3hip_si_action(cf3hip_si si, cf3hip_fields fields, const cf3hip_ns_params* prm, int which);
/* inner-loop vectors, host arrays of rows x ncomp */
int cf3hip_si_get(cf3hip_si si, int which, double* host);
int cf3hip_si_set(cf3hip_si si, int which, const double* host);
/* one time step: LinearizeU, velocity assembly, SetSolution, InnerLoop::execute with nb_iterations
 * passes (URHS, velocity BC in acceleration form at the first pass, velocity solve, PRHS, pressure BC
 * at the first pass, pressure solve, ApplyAup, delta_a -= Ml^-1 Aup delta_p, a/u/p/delta_p_sum
 * updates) and Update.  Dirichlet conditions: the new values at mesh nodes (velocity: n x dim);
 * the increments new - old are taken against the fields.  Each degree of freedom may carry one
 * condition (CF3HIP_ERR_INVALID_ARG otherwise): resolve overlapping regions before the call.  stats: 2 * nb_iterations entries
 * (velocity and pressure solve of each pass), may be null. */
int cf3hip_si_step(cf3hip_si si, cf3hip_fields fields, const cf3hip_ns_params* prm, int nb_iterations,
                   const cf3hip_solver_params* u_solver, const cf3hip_solver_params* p_solver, uint32_t n_ubc,
                   const uint32_t* ubc_nodes, const double* ubc_values, uint32_t n_pbc, const uint32_t* pbc_nodes,
                   const double* pbc_values, cf3hip_solve_stats* stats);

#if defined(__GNUC__)
#pragma GCC visibility pop
#endif
#ifdef __cplusplus
}
#endif
#endif

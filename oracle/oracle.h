/*
 * Oracle: CPU FP64 restatement of COOLFluiD 3's UFEM incompressible Navier-Stokes hot path
 * (element assembly -> LSS scatter -> symmetric Dirichlet -> SpMV/Krylov).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the test oracle: it is loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the product path.
 * Every function cites the reference file:line it restates (paths relative to the
 * coolfluid3 source tree).
 *
 * Parity pins (what each function is checked against) are listed in the test files
 * tests/test_oracle_*.py; functions with no pin beyond the restatement say so there.
 */
#ifndef CF3_ORACLE_H
#define CF3_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* element types (cf3/mesh/LagrangeP1) */
enum { ORC_TRIAG2D = 0, ORC_TETRA3D = 1, ORC_HEXA3D = 2, ORC_QUAD2D = 3, ORC_PRISM3D = 4 };

/* SUPG tau variants (plugins/UFEM/src/UFEM/SUPG.hpp:62-74) */
enum { ORC_SUPG_TEZDUYAR = 0, ORC_SUPG_METRIC = 1, ORC_SUPG_CF2 = 2 };

typedef struct orc_ns_params {
  double rho;          /* density (NavierStokesPhysics "density") */
  double dt;           /* LSSActionUnsteady time step */
  double theta;        /* NavierStokes "theta" option, default 1 */
  int use_specializations; /* NavierStokes "use_specializations", default 1 */
  int supg_type;       /* ORC_SUPG_*, default METRIC */
  double c1, c2;       /* METRIC constants (1, 16) */
  double alpha_ps, alpha_su, alpha_bu; /* tau scale factors (1,1,1) */
  double u_ref;        /* CF2 reference velocity (ComputeTau "u_ref", default 1) */
} orc_ns_params;

/* ---------------- PRNG and synthetic meshes ---------------- */
uint64_t orc_splitmix64_next(uint64_t* state);
double orc_uniform(uint64_t* state, double a, double b);

/* cf3/Tools/MeshGeneration/MeshGeneration.cpp:209-254 (create_rectangle_tris) */
void orc_mesh_rect_tris(uint32_t nx, uint32_t ny, double lx, double ly, double* coords, uint32_t* conn);
/* cf3/mesh/SimpleMeshGenerator.cpp:680-805 (create_box hexahedra, node (k*(ny+1)+j)*(nx+1)+i) */
void orc_mesh_box_hexas(uint32_t nx, uint32_t ny, uint32_t nz, double lx, double ly, double lz, double* coords, uint32_t* conn);
/* create_box hexahedra split into 6 Kuhn tetrahedra along the local 0->6 diagonal (conforming, detJ>0) */
void orc_mesh_box_tetras(uint32_t nx, uint32_t ny, uint32_t nz, double lx, double ly, double lz, double* coords, uint32_t* conn);
/* node jitter: interior nodes move by U[-f h, f h] per axis, boundary nodes only tangentially */
void orc_jitter_structured(int dim, uint32_t nx, uint32_t ny, uint32_t nz, double lx, double ly, double lz,
                           double frac, uint64_t seed, double* coords);

/* ---------------- geometry (fixtures) ---------------- */
double orc_tetra_volume(const double* nodes /*4x3*/);
void orc_tetra_jacobian(const double* nodes, double* J /*3x3 row-major, J(ksi,x)*/);
void orc_tetra_jacobian_adjoint(const double* nodes, double* adj);
double orc_hexa_jacobian_determinant(const double* mapped, const double* nodes /*8x3*/);
void orc_hexa_jacobian(const double* mapped, const double* nodes, double* J);
void orc_hexa_jacobian_adjoint(const double* mapped, const double* nodes, double* adj);
double orc_hexa_volume(const double* nodes);
/* any element type: volume (Triag2D.cpp:150-153, Tetra3D.cpp:205-208, Hexa3D.cpp:317-320) and J(mapped) */
double orc_element_volume(int etype, const double* nodes);
void orc_element_jacobian(int etype, const double* mapped, const double* nodes, double* J);
void orc_shape_values(int etype, const double* mapped, double* N);
void orc_shape_gradients(int etype, const double* mapped, double* dN /* dim x k row-major */);
int orc_gauss_points(int etype, int order, double* coords /* npts x dim */, double* weights);

/* ---------------- sparsity ---------------- */
/* plugins/UFEM/src/UFEM/SparsityBuilder.cpp:157-189 (std::set per node, self included, sorted)
 * Returns n_b (number of node blocks) or -1 when cols_cap is too small.
 * slots (optional, E*k*k): block slot of element pair (a,b) = start[conn[a]] + pos(conn[b])  */
int64_t orc_build_sparsity(int k, uint32_t n_nodes, uint32_t n_elems, const uint32_t* conn,
                           uint32_t* start, uint32_t* cols, int64_t cols_cap, uint32_t* slots);

/* ---------------- element kernels ---------------- */
/* Element matrices A, T (var-major, (k*neq)^2, row-major) and element vector a (k*neq).
 * nodes k x dim, u k x dim, u_adv k x dim, nu k, g k x dim (body force).  Returns 0 or error. */
int orc_ns_element(int etype, const orc_ns_params* prm, const double* nodes, const double* u,
                   const double* u_adv, const double* nu, const double* g,
                   double* A, double* T, double* a, double* tau3 /* tau_ps, tau_su, tau_bulk */);
/* Element combine + node-major reorder (NavierStokesAssembly.hpp:69-71, BlockAccumulator.hpp:134-184):
 * M (node-major, row-major) = invdt*T + theta*A' ; rhs (node-major) = -A*x + a */
int orc_ns_element_system(int etype, const orc_ns_params* prm, const double* nodes, const double* u,
                          const double* p, const double* u_adv, const double* nu, const double* g,
                          double* M, double* rhs);

/* ---------------- assembly + LSS ---------------- */
/* Full assembly (reset + element loop + add_values), BSR values with neq x neq row-major blocks.
 * fields are node-major AoS like cf3 Field tables: u (N x dim), p (N), u_adv (N x dim), nu (N), g (N x dim).
 * nthreads>1 uses a colouring of the elements (deterministic).  */
int orc_ns_assemble(int etype, const orc_ns_params* prm, uint32_t n_nodes, uint32_t n_elems,
                    const double* coords, const uint32_t* conn, const double* u, const double* p,
                    const double* u_adv, const double* nu, const double* g,
                    const uint32_t* start, const uint32_t* cols, const uint32_t* slots,
                    double* values, double* rhs, int nthreads);
/* The same with periodic links: the system has n_rows block rows, p2m gives the block row of each mesh
 * node (a linked node shares its master's row; System::create with periodic_links_nodes/active,
 * cf3/math/LSS/Trilinos/TrilinosDetail.cpp:147-227); start/slots come from the sparsity of p2m[conn].
 * Elements and fields are gathered by mesh node.  Serial. */
/* the same scatter of |M_e| and |b_e| for every element contribution: per entry the sum of the magnitudes of
   its contributions (the scale of its rounding error) */
int orc_ns_assemble_abs(int etype, const orc_ns_params* prm, uint32_t n_nodes, uint32_t n_elems,
                        const double* coords, const uint32_t* conn, const double* u, const double* p,
                        const double* u_adv, const double* nu, const double* g,
                        const uint32_t* start, const uint32_t* cols, const uint32_t* slots,
                        double* values, double* rhs, int nthreads);
int orc_ns_assemble_p2m(int etype, const orc_ns_params* prm, uint32_t n_rows, uint32_t n_elems,
                        const double* coords, const uint32_t* conn, const uint32_t* p2m, const double* u,
                        const double* p, const double* u_adv, const double* nu, const double* g,
                        const uint32_t* start, const uint32_t* slots, double* values, double* rhs);

/* symmetric Dirichlet, TrilinosFEVbrMatrix::symmetric_dirichlet (TrilinosFEVbrMatrix.cpp:482-553)
 * + System::dirichlet (System.cpp:240-255), applied sequentially in list order. */
void orc_dirichlet(int neq, uint32_t n_nodes, const uint32_t* start, const uint32_t* cols,
                   double* values, double* rhs, double* sol, uint32_t nbc,
                   const uint32_t* bc_node, const uint32_t* bc_eq, const double* bc_val);
/* Matrix::set_row (Matrix.hpp; TrilinosFEVbrMatrix::set_row): row (blockrow, ieq) -> diag/offdiag */
/* System::periodicity / Matrix::tie_blockrow_pairs for (to[i], from[i]) pairs in order; rhs, sol may be null.
   0 ok, -1 block counts differ, -2 block columns differ */
int orc_tie_blockrow_pairs(int neq, const uint32_t* start, const uint32_t* cols, double* values, double* rhs,
                           double* sol, uint32_t npairs, const uint32_t* to, const uint32_t* from);
void orc_set_row(int neq, const uint32_t* start, const uint32_t* cols, double* values,
                 uint32_t blockrow, uint32_t ieq, double diag, double offdiag);
/* y = alpha*A*x + beta*y  (Matrix::apply, TrilinosDetail.cpp:380-411) */
void orc_spmv(int neq, uint32_t n_nodes, const uint32_t* start, const uint32_t* cols, const double* values,
              double alpha, const double* x, double beta, double* y);
/* point-CSR view of the block matrix (Matrix::debug_data) */
void orc_point_csr(int neq, uint32_t n_nodes, const uint32_t* start, const uint32_t* cols, const double* values,
                   uint32_t* row_ptr, uint32_t* col_idx, double* vals);

/* ---------------- Krylov ---------------- */
typedef struct orc_solve_stats {
  int iterations;
  int converged;
  double rel_residual;   /* true ||b-Ax||/||b|| at exit */
  int spmvs;
} orc_solve_stats;
/* BiCGStab, right Jacobi preconditioning, stop on ||r|| <= tol*||b|| */
/* threads of the Krylov kernels (SpMV, dot products); 1 by default, raised only for CPU-baseline timing */
void orc_set_krylov_threads(int n);
int orc_bicgstab_jacobi(int neq, uint32_t n_nodes, const uint32_t* start, const uint32_t* cols, const double* values,
                        const double* b, double* x, double tol, int maxit, orc_solve_stats* st);
/* restarted GMRES(m), right ILU(0) preconditioning on the point matrix, CGS2 orthogonalisation */
int orc_gmres_ilu0(int neq, uint32_t n_nodes, const uint32_t* start, const uint32_t* cols, const double* values,
                   const double* b, double* x, int restart, double tol, int maxit, orc_solve_stats* st);

/* ---------------- node updates ---------------- */
/* LinearizeU (ns_implementation/NavierStokes.cpp:105) */
void orc_linearize_u(uint32_t n, int dim, const double* u, const double* u1, const double* u2, const double* u3, double* u_adv);
/* UpdateActions (NavierStokes.cpp:172-180): u3=u2; u2=u1; u1=u; u+=du; p+=dp, solution node-major (u..,p) */
void orc_update(uint32_t n, int dim, double* u, double* p, double* u1, double* u2, double* u3, const double* sol);

/* ---------------- NavierStokesSemiImplicit (plugins/UFEM/src/UFEM/ns_semi_implicit) ----------------
 * Element actions over all elements in element order, accumulated (+=) into the caller's arrays.
 * Mesh-node fields: coords, u, uadv (N x dim), nu (N).  LSS vectors by row (p2m[node]): velocity-system
 * vectors rows x dim, pressure-system vectors rows.  Velocity matrix: neq = dim blocks (row-major) at
 * the block slots of the p2m sparsity; pressure matrix: scalar values at those slots.  Uses the
 * METRIC/TEZDUYAR tau of prm (dt, theta, supg_type, c1, c2, alphas; rho is not used). */
/* MassMatrixAssembly: lumped (row-sum) int N^T N into every velocity component (diag: rows x dim) */
int orc_si_lumped_mass(int etype, uint32_t n_elems, const double* coords, const uint32_t* conn,
                       const uint32_t* p2m, double* diag);
/* velocity matrix T + theta dt M (MatrixAssembly.hpp:49-146, :458) */
int orc_si_velocity_matrix(int etype, const orc_ns_params* prm, uint32_t n_elems, const double* coords,
                           const uint32_t* conn, const double* uadv, const double* nu, const uint32_t* slots,
                           double* values);
/* simplified pressure matrix theta (tau_ps + dt) int G^T G, tau_ps from u (PressureSystem.cpp:166-176) */
int orc_si_pressure_matrix(int etype, const orc_ns_params* prm, uint32_t n_elems, const double* coords,
                           const uint32_t* conn, const double* u, const double* nu, const uint32_t* slots,
                           double* values);
/* URHSAssembly: velocity_rhs(u_adv, nu, a, dt (1 - theta) a - u_vec, (1 - theta) delta_p_sum - p_vec)
 * (MatrixAssembly.hpp:206-330, :466-470); body_force (n_nodes x dim, or null) selects the overload of
 * enable_body_force (:339-363, :472-477) */
int orc_si_velocity_rhs(int etype, const orc_ns_params* prm, uint32_t n_elems, const double* coords,
                        const uint32_t* conn, const uint32_t* p2m, const double* uadv, const double* nu,
                        const double* a_lss, const double* u_lss, const double* dps_lss, const double* p_lss,
                        double* rhs, const double* body_force);
/* PRHSAssembly: pressure_rhs(u_adv, u_vec, a, delta_a, p_vec) with tau_ps from u (:148-204, :487-494) */
int orc_si_pressure_rhs(int etype, const orc_ns_params* prm, uint32_t n_elems, const double* coords,
                        const uint32_t* conn, const uint32_t* p2m, const double* u, const double* uadv,
                        const double* nu, const double* u_lss, const double* a_lss, const double* da_lss,
                        const double* p_lss, double* rhs);
/* ApplyAup: apply_aup(u_adv, delta_p) (:332-376, :497-503) */
int orc_si_apply_aup(int etype, const orc_ns_params* prm, uint32_t n_elems, const double* coords,
                     const uint32_t* conn, const uint32_t* p2m, const double* uadv, const double* nu,
                     const double* dp_lss, double* rhs);

#ifdef __cplusplus
}
#endif
#endif

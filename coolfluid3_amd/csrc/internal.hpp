// Internal declarations shared by the cf3hip translation units (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "../../include/cf3hip.h"

namespace cf3hip {

// ---------------------------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------------------------
void set_error(const std::string& msg);

// internal failure carried to the C ABI boundary and turned into a status code there
struct Failure {
  int code;
  std::string msg;
};

#define CF3_THROW(code, msg) throw ::cf3hip::Failure{code, msg}
#define HIP_OK(expr)                                                                             \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if(e_ != hipSuccess)                                                                         \
      CF3_THROW(e_ == hipErrorOutOfMemory ? CF3HIP_ERR_OUT_OF_MEMORY : CF3HIP_ERR_HIP,           \
                std::string(#expr) + ": " + hipGetErrorString(e_));                             \
  } while(0)

// ---------------------------------------------------------------------------------------------
// device buffers
// ---------------------------------------------------------------------------------------------
struct Context;

struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  Context* ctx = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void alloc(Context* c, size_t n);
  void release();
  template<typename T> T* as() const { return static_cast<T*>(ptr); }
};

// HIP-event timing of the hot kernels (per launch, summed at query time)
struct KernelTimer {
  std::vector<hipEvent_t> ev;   // start/stop pairs
  size_t used = 0;              // events recorded since the last reset
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t bytes_in_use = 0;
  bool async_mode = false;      // no host synchronisation inside calls; errors at synchronize
  bool timing = false;
  KernelTimer timers[CF3HIP_KERNEL_COUNT];
  // scratch for reductions: partial sums and device scalars
  DevBuf red_partials;   // double[kMaxRedBlocks * 8]
  DevBuf scalars;        // double[64]; the last slot holds the element-validity flag
  DevBuf fused_partials; // partial sums of the fused Krylov iteration (allocated on first use)
  double* scalars_host = nullptr;   // pinned mirror
  hipStream_t own_stream = nullptr;  // created by cf3hip_context_create_stream, destroyed with the context
  int* status_flag() { return reinterpret_cast<int*>(scalars.as<double>() + 63); }
};

void timer_begin(Context& c, int kernel);
void timer_end(Context& c, int kernel);

// f(integral_constant<int, i>) for i = B .. K-1, unrolled at compile time
template<int B, int K, typename F>
__device__ __forceinline__ void static_for(F&& f)
{
  if constexpr(B < K) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, K>(f);
  }
}

constexpr int kMaxRedBlocks = 1024;
// assembly workgroup: at most kAsmThreads incidences (one per thread) and kAsmThreads CSR blocks
// (one owner thread per block); the row schedule of the system respects both
constexpr int kAsmThreads = 128;
constexpr uint32_t kMaxWgNodes = 255;   // distinct nodes per staged workgroup (one-byte local indices)

// ---------------------------------------------------------------------------------------------
// mesh / system / fields
// ---------------------------------------------------------------------------------------------
struct Mesh {
  Context* ctx = nullptr;
  int etype = 0, dim = 0, k = 0;
  uint32_t n_nodes = 0, n_elems = 0;
  DevBuf x[3];        // SoA coordinates (double)
  DevBuf xp;          // packed coordinates double4 (x, y, z | 0, 0) for the element gathers
  DevBuf conn;        // uint32 n_elems*k, element-major
};

// recycle space of the RCG solver (rcg.hip): W in the first k columns of V and AW = A W in those of
// AV, the next columns the search directions of the current cycle; E^-1 = (W^T A W)^-1 on the device
constexpr int kRcgSpaces = 2;      // per system (the first and the second pressure strategies)
constexpr int kMaxRecycle = 32;    // recycled vectors
constexpr int kMaxCycle = 96;      // search directions per cycle
constexpr int kDefaultCycle = 40;
struct RcgSpace {
  uint64_t n = 0;
  int k = 0;
  double print[2] = {0., 0.};      // fingerprint of the matrix values W was built on
  DevBuf V, AV, tmp, vec, sc, small, einv, gram;
};

struct Lss {
  Context* ctx = nullptr;
  Mesh* mesh = nullptr;   // null when created from a caller graph
  int neq = 0;
  uint32_t n_nodes = 0;   // block rows (the mesh nodes, less the periodically linked ones)
  // periodic links (System::create with periodic_links_nodes/active): p2m[mesh node] = block row; a
  // linked node shares the row of the end of its chain of links.  Elements are still gathered by mesh
  // node (geometry and fields); only rows, columns and the solution update go through p2m.
  bool periodic = false;
  DevBuf p2m;             // uint32 mesh->n_nodes
  uint64_t n_blocks = 0;
  uint64_t dofs = 0;
  DevBuf row_ptr;       // uint32 n_nodes+1 (starting_indices)
  DevBuf cols;          // uint32 n_blocks   (node_connectivity)
  DevBuf diag_blk;      // uint32 n_nodes: block index of the diagonal
  DevBuf values;        // double n_blocks*neq*neq
  DevBuf values32;      // float mirror (precision form)
  DevBuf values32b;     // block-column scaled FP32 operator fl32(A32 D^-1) of a block-scaled precision solve
  DevBuf bdinv;         // double n_nodes*neq*neq: D^-1, the inverted diagonal blocks of that solve
  bool has_fp32 = false;
  DevBuf rhs, sol;      // double dofs
  // element scatter structures (lss_create from a mesh)
  DevBuf slots;         // uint32 n_elems*k*k   element-to-slot map (block slots)
  DevBuf inc_ptr;       // uint32 n_nodes+1
  DevBuf inc;           // uint32 n_elems*k : e*k + a, sorted by (node, e)
  DevBuf wg_rows;       // uint32 n_wg+1 : row ranges of the assembly workgroups
  uint32_t n_wg = 0;
  // staging maps of the row-owner reduction (see sparsity.hip): per incidence and column local node
  // the LDS row its element block is staged to, per block and pass the first staged row it sums
  int pos_stride = 4;   // bytes per incidence in inc_pos (4 for simplices, 8 for hexahedra)
  int passes = 0;       // staging passes per workgroup (k for simplices, 1 for hexahedra)
  DevBuf inc_pos;       // uint8 n_elems*k*pos_stride
  DevBuf blk_beg;       // uint8 n_blocks*4
  DevBuf blk_acc;       // uint16 n_blocks: first accumulator thread (low byte) and count (high byte)
  // node staging of the simplex kernels: each workgroup copies the records of the distinct nodes its
  // elements touch into LDS once, and its incidences address them by local index
  bool staged = false;  // built (simplices whose workgroups touch at most kMaxWgNodes nodes)
  uint32_t wg_nodes_max = 0;
  DevBuf wg_head;       // uint32 8 per workgroup: rows [r0, r1), incidences [i0, i1), blocks [b0, b1), nodes [n0, n1)
  DevBuf wg_nodes;      // uint32: the distinct nodes of each workgroup, ascending
  DevBuf inc_loc;       // uint32 n_elems*k: local indices (one byte per element node) of the element's nodes
  // Dirichlet conditions applied since the last reset (needed by the matrix-free FP64 residual):
  // per dof the first value (column elimination) and the last value (row / solution)
  std::unordered_map<uint64_t, std::pair<double, double> > bc_host;
  std::vector<uint64_t> bc_order;
  uint32_t bc_n = 0;
  DevBuf bc_dof;        // uint32 bc_n
  DevBuf bc_vlast;      // double bc_n
  DevBuf bc_vfirst;     // double dofs (valid at BC dofs)
  DevBuf bc_mask;       // uint8 n_nodes: bit e set when dof (node, e) carries a condition
  DevBuf bc_newmask;    // uint8 n_nodes: conditions added by the current call
  DevBuf xtilde;        // matrix-free residual input (x with eliminated columns at their values)
  DevBuf q1rec, q1tau;  // Q1 cells: per-element Gauss-point records and tau of the last assembly launch
  // matrix builder semantics: CF3HIP_BUILDER_VBR (TrilinosFEVbrMatrix) or CF3HIP_BUILDER_CRS
  // (TrilinosCrsMatrix: symmetric Dirichlet with columns cached until the next reset)
  int builder = 0;
  std::unordered_map<uint64_t, uint64_t> crs_col;   // CRS: dof -> offset of its cached column in crs_cache
  DevBuf crs_cache;                                 // CRS: cached column values (deg(node) * neq per dof)
  uint64_t crs_used = 0;
  bool tied = false;    // block rows tied (periodicity) since the last reset: the matrix-free
                        // residual of the element operator no longer describes the system
  // Krylov workspace (allocated lazily)
  DevBuf work;          // generic vector workspace
  size_t work_vectors = 0;
  DevBuf ir;            // FP64 FGMRES workspace of the precision form (basis, directions, bookkeeping)
  int work_real_bytes = 0;
  // GMRES + ILU(0): the point pattern's level schedules (built once per system, on the host) and the
  // per-solve point values, factors and Krylov workspace (kept between solves)
  struct {
    bool ready = false;
    uint32_t nlev = 0;
    DevBuf prow, pcol, diag, lorder, lptr;   // point pattern and the factorisation's level schedule
    DevBuf fptr, fpiv, fsrc, fdst;   // update lists of the factorisation
    DevBuf aval, fval;
    DevBuf V, w, z, small, out;
    int m_ws = 0;
    // single-workgroup solver: node-level triangular sweeps in slot-major order (per step the
    // (rows, width, int base, value base) record, the int and value-source tables) and a 64-row
    // sliced-ELL copy of the point pattern (per slice (base, width), columns and value sources);
    // the value tables are gathered from fval / aval at every solve
    uint32_t nsteps = 0, nlow = 0;
    uint64_t nsv = 0, nev = 0;   // entries of sval and eval
    DevBuf smeta, sint, ssrc, sval;
    DevBuf emeta, ecol, esrc, eval;
  } ilu;
  std::unique_ptr<RcgSpace> rcg[kRcgSpaces];   // RCG recycle spaces (kept between solves)
};

struct Fields {
  Mesh* mesh = nullptr;
  DevBuf u[3], p, uadv[3], u1[3], u2[3], u3[3], nu, g[3];
  // packed double4 node records refreshed by every assembly / residual call:
  // st = (u, v, w, p) in 3D, (u, v, p, 0) in 2D; ad = (u_adv x, y, z | 0, nu_eff)
  DevBuf st, ad;
  bool has_g = false;   // the body force holds a nonzero value (its source term is skipped otherwise)
};

// NavierStokesSemiImplicit workspace: a velocity system (neq = dim) and a pressure system (neq = 1)
// on one mesh; contribution lists of the deterministic gathers, the inner-loop vectors by LSS row
struct SemiImplicit {
  Context* ctx = nullptr;
  Mesh* mesh = nullptr;
  Lss* u = nullptr;
  Lss* p = nullptr;
  DevBuf mat_ptr, mat_idx;   // per block slot: its (e*k + a)*k + b contributions in element order
  DevBuf vec_ptr, vec_idx;   // per row: its e*k + a incidences in element order
  DevBuf src;                // per row: the last mesh node of the row (SetSolution)
  DevBuf p2m;                // per mesh node: its row (the identity without periodic links)
  std::vector<uint32_t> p2m_host;   // host copy of p2m (boundary conditions are given by mesh node)
  DevBuf stage;              // element outputs, n_elems * k * k * 4 doubles
  DevBuf A, Uv, minv, aup;   // rows x dim: a, u, 1 / lumped mass, scratch
  DevBuf Pv, dps;            // rows: p, delta_p_sum
  DevBuf pvals0;             // the pressure matrix before its Dirichlet conditions
  bool initialized = false;
  // options of NavierStokesSemiImplicit (NavierStokesSemiImplicit.cpp:309-317)
  bool body_force = false;   // enable_body_force: the body_force field enters the velocity rhs
  bool rcg = false;          // pressure_rcg_solve: first and later pressure solves by two RCG spaces
};

// batched FP32 assembly of many systems (serving): a device table of per-system arguments and one
// persistent launch captured in a hipGraph on a private stream
struct Batch {
  Context* ctx = nullptr;
  uint32_t max_cases = 0;
  hipStream_t stream = nullptr;
  DevBuf table;
  void* host_table = nullptr;   // pinned mirror
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

// ---------------------------------------------------------------------------------------------
// entry points between translation units
// ---------------------------------------------------------------------------------------------
void build_element_structures(Lss& L);                   // sparsity.hip
void build_graph_structures(Lss& L);                     // sparsity.hip (diag blocks)
void assemble(Lss& L, Fields& F, const cf3hip_ns_params& prm);   // assembly.hip
void mesh_pack_coords(Mesh& M);                                  // assembly.hip
size_t batch_table_bytes(uint32_t max_cases);                    // assembly.hip
void batch_assemble(Batch& B, uint32_t n, Lss* const* L, Fields* const* F, const cf3hip_ns_params* prm);

// lss.hip
void lss_fill(Context& c, double* p, size_t n, double v);
void lss_fill32(Context& c, float* p, size_t n, float v);
void lss_round_fp32(Lss& L);   // values32 = fl32(values)
void lss_copy(Context& c, double* dst, const double* src, size_t n);
void lss_fill8(Context& c, uint8_t* p, size_t n, uint8_t v);
void lss_clear_conditions(Lss& L);
// FP32 values, FP64 vectors and accumulation: y = A32 x
void lss_spmv_mixed(Lss& L, const double* x, double* y, const double* stop = nullptr);
// matrix-free FP64 residual r = b' - A' x of the NS system (assembly.hip)
void ns_residual(Lss& L, Fields& F, const cf3hip_ns_params& prm, const double* x, double* r);
void ns_solve(Lss& L, Fields& F, const cf3hip_ns_params& nsp, const cf3hip_solver_params& prm,
              cf3hip_solve_stats& st);
// stop (optional, device): a Krylov loop's stop flag; the product is skipped once it is set
void lss_spmv(Lss& L, double alpha, const double* x, double beta, double* y, const double* stop = nullptr);
void lss_spmv32(Lss& L, float alpha, const float* x, float beta, float* y, const double* stop = nullptr);
void lss_dirichlet(Lss& L, uint32_t nbc, const uint32_t* nodes, const uint32_t* eqs, const double* values);
void lss_crs_clone(Lss& src, Lss& dst);                 // TrilinosCrsMatrix::clone_to
void lss_read_native(Lss& L, const char* path);          // TrilinosCrsMatrix::read_native
void lss_tie(Lss& L, uint32_t npairs, const uint32_t* to, const uint32_t* from, bool system);
// noalloc.hip: the residual without device allocation (scratch address per device set at context creation)
void noalloc_init(int device);
double lss_residual_noalloc(Lss& L);
void lss_point_csr(Lss& L, uint32_t* row_ptr_dev, uint32_t* col_dev, double* val_dev);
void lss_diagonal(Lss& L, const double* d_dev, int mode);            // mode 0 set, 1 add
void lss_column_zero(Lss& L, uint32_t col, int ieq, double* out_dev);
void lss_axpby(Context& c, size_t n, double alpha, const double* x, double beta, double* y);

// krylov.hip
void lss_solve(Lss& L, const cf3hip_solver_params& prm, cf3hip_solve_stats& st);
void lss_solve_precision(Lss& L, const cf3hip_solver_params& prm, cf3hip_solve_stats& st);
double lss_residual(Lss& L);
void rcg_solve(Lss& L, const cf3hip_solver_params& prm, cf3hip_solve_stats& st);   // rcg.hip

// nodes.hip
void ns_linearize(Fields& F);
void si_create(SemiImplicit& S, Lss& U, Lss& P);                                            // semi_implicit.hip
void si_initial(SemiImplicit& S, Fields& F, const cf3hip_ns_params& prm);
void si_velocity_assemble(SemiImplicit& S, Fields& F, const cf3hip_ns_params& prm);
void si_action(SemiImplicit& S, Fields& F, const cf3hip_ns_params& prm, int which);
void si_set_solution(SemiImplicit& S, Fields& F);
void si_correct(SemiImplicit& S, double dt);
void si_update(SemiImplicit& S, Fields& F);
void si_zero_dofs(SemiImplicit& S, const std::vector<uint32_t>& dofs, double* v);
void ns_update(Lss& L, Fields& F);

// cells of the Gauss-point lane kernel (assembly_hex.hip): Q1 hexahedra and quadrilaterals and P1
// prisms, which always take the generic expression
inline bool is_cell(int etype) { return etype == CF3HIP_HEXA3D || etype == CF3HIP_QUAD2D || etype == CF3HIP_PRISM3D; }
// lanes per incidence of that kernel: one per Gauss point / column node, 8 on prisms (6 used)
inline int cell_lanes(int etype, int k) { return etype == CF3HIP_PRISM3D ? 8 : k; }

inline unsigned grid_for(size_t n, unsigned block)
{
  size_t g = (n + block - 1) / block;
  if(g == 0) g = 1;
  return (unsigned)g;
}

}  // namespace cf3hip

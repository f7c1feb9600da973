// C++ host layer over the cf3hip C ABI (include/cf3hip.h), mirroring the reference interfaces of the
// Navier-Stokes hot path so that a cf3 component (or any C++ caller) drives it with the same calls:
//
//   lss::Matrix            cf3/math/LSS/Matrix.hpp:35-180 (values addressed column first)
//   lss::Vector            cf3/math/LSS/Vector.hpp:36-173
//   lss::BlockAccumulator  cf3/math/LSS/BlockAccumulator.hpp:28-74 (row-major mat, node indices)
//   lss::SolutionStrategy  cf3/math/LSS/SolutionStrategy.hpp:35-59
//   lss::System            cf3/math/LSS/System.cpp (create, dirichlet, solve, reset)
//   ufem::NavierStokesAssembly  the Assembly action of plugins/UFEM/src/UFEM/NavierStokes.hpp:34-118
//                          with its options (use_specializations, supg_type, theta, density, dt)
//
// Errors become the reference exception types (common/BasicExceptions.hpp): SetupError, BadValue,
// NotImplemented, FailedAssertion.  Device memory is owned by the C handles and released by the
// destructors; no device pointer crosses this interface except through Matrix::apply's own vectors.
// Header-only; link with libcf3hip.so.
#pragma once

#include <cstdint>
#include <fstream>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "cf3hip.h"

namespace cf3hip {

// ---------------------------------------------------------------------------------------------
// exceptions
// ---------------------------------------------------------------------------------------------
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
struct SetupError : Error { using Error::Error; };
struct BadValue : Error { using Error::Error; };
struct NotImplemented : Error { using Error::Error; };
struct FailedAssertion : Error { using Error::Error; };

inline void check(int rc)
{
  if(rc == CF3HIP_OK) return;
  const std::string msg = cf3hip_last_error();
  switch(rc) {
    case CF3HIP_ERR_NOT_CREATED:
    case CF3HIP_ERR_LIMIT: throw SetupError(rc, msg);
    case CF3HIP_ERR_INVALID_ARG:
    case CF3HIP_ERR_ILLEGAL_ENTRY:
    case CF3HIP_ERR_SINGULAR: throw BadValue(rc, msg);
    case CF3HIP_ERR_NOT_IMPLEMENTED: throw NotImplemented(rc, msg);
    default: throw FailedAssertion(rc, msg);
  }
}

// one device and its stream
class Context {
public:
  explicit Context(int device = 0) { check(cf3hip_context_create(device, &h_)); }
  ~Context() { cf3hip_context_destroy(h_); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  cf3hip_context handle() const { return h_; }
  void synchronize() { check(cf3hip_context_synchronize(h_)); }
private:
  cf3hip_context h_ = nullptr;
};

// one LagrangeP1 element region with its coordinates (host node-major) and connectivity
class Mesh {
public:
  Mesh(Context& ctx, int etype, const std::vector<double>& coords, const std::vector<uint32_t>& conn)
  {
    const int k = etype == CF3HIP_TRIAG2D ? 3 : (etype == CF3HIP_HEXA3D ? 8 : 4);
    const int dim = etype == CF3HIP_TRIAG2D || etype == CF3HIP_QUAD2D ? 2 : 3;
    if(coords.size() % dim || conn.size() % k) throw BadValue(CF3HIP_ERR_INVALID_ARG, "mesh array sizes");
    n_nodes_ = (uint32_t)(coords.size() / dim);
    check(cf3hip_mesh_create(ctx.handle(), etype, n_nodes_, (uint32_t)(conn.size() / k), coords.data(), conn.data(),
                             &h_));
  }
  ~Mesh() { cf3hip_mesh_destroy(h_); }
  Mesh(const Mesh&) = delete;
  Mesh& operator=(const Mesh&) = delete;
  cf3hip_mesh handle() const { return h_; }
  uint32_t n_nodes() const { return n_nodes_; }
private:
  cf3hip_mesh h_ = nullptr;
  uint32_t n_nodes_ = 0;
};

namespace lss {

// BlockAccumulator.hpp:28-74: dense element block (row-major), rhs / sol parts and the node indices
struct BlockAccumulator {
  std::vector<double> mat, rhs, sol;
  std::vector<uint32_t> indices;
  uint32_t neq = 0;
  void resize(uint32_t nb, uint32_t neq_)
  {
    neq = neq_;
    indices.assign(nb, 0);
    mat.assign((size_t)nb * neq * nb * neq, 0.);
    rhs.assign((size_t)nb * neq, 0.);
    sol.assign((size_t)nb * neq, 0.);
  }
  void reset(double v = 0.)
  {
    for(double& x : mat) x = v;
    for(double& x : rhs) x = v;
    for(double& x : sol) x = v;
  }
  uint32_t size() const { return (uint32_t)indices.size(); }
};

class System;

// Vector.hpp:36-173, the rhs or the solution of a System
class Vector {
public:
  Vector(System& s, int which) : sys_(s), which_(which) {}
  const std::string solvertype() const { return "cf3hip"; }
  void set_value(uint32_t irow, double v);
  void add_value(uint32_t irow, double v);
  void get_value(uint32_t irow, double& v);
  void set_value(uint32_t iblockrow, uint32_t ieq, double v);
  void add_value(uint32_t iblockrow, uint32_t ieq, double v);
  void get_value(uint32_t iblockrow, uint32_t ieq, double& v);
  void set_rhs_values(const BlockAccumulator& b) { block(b.indices, const_cast<double*>(b.rhs.data()), 0); }
  void add_rhs_values(const BlockAccumulator& b) { block(b.indices, const_cast<double*>(b.rhs.data()), 1); }
  void get_rhs_values(BlockAccumulator& b) { block(b.indices, b.rhs.data(), 2); }
  void set_sol_values(const BlockAccumulator& b) { block(b.indices, const_cast<double*>(b.sol.data()), 0); }
  void add_sol_values(const BlockAccumulator& b) { block(b.indices, const_cast<double*>(b.sol.data()), 1); }
  void get_sol_values(BlockAccumulator& b) { block(b.indices, b.sol.data(), 2); }
  void reset(double v = 0.);
  // get / set: node-major rows of neq values (the multi_array of the reference)
  void get(std::vector<double>& data);
  void set(const std::vector<double>& data);
  void assign(const Vector& source);
  void update(const Vector& source, double alpha = 1.);
  void scale(double alpha);
  void sync() {}
  bool is_created() const;
  uint32_t neq() const;
  uint32_t blockrow_size() const;
  void debug_data(std::vector<double>& values) { get(values); }
  int which() const { return which_; }
  System& system() const { return sys_; }
private:
  void block(const std::vector<uint32_t>& idx, double* vals, int mode);
  System& sys_;
  int which_;
};

// Matrix.hpp:35-180 (column-first scalar access, node-block storage)
class Matrix {
public:
  explicit Matrix(System& s) : sys_(s) {}
  const std::string solvertype() const { return "cf3hip"; }
  void set_value(uint32_t icol, uint32_t irow, double v);
  void add_value(uint32_t icol, uint32_t irow, double v);
  void get_value(uint32_t icol, uint32_t irow, double& v);
  void set_values(const BlockAccumulator& b);
  void add_values(const BlockAccumulator& b);
  void get_values(BlockAccumulator& b);
  void set_row(uint32_t iblockrow, uint32_t ieq, double diagval, double offdiagval);
  void get_column_and_replace_to_zero(uint32_t iblockcol, uint32_t ieq, std::vector<double>& values);
  void symmetric_dirichlet(uint32_t blockrow, uint32_t ieq, double value, Vector& rhs);
  // TrilinosFEVbrMatrix::tie_blockrow_pairs (TrilinosFEVbrMatrix.cpp:558-590): BadValue when the two
  // rows hold different block columns
  void tie_blockrow_pairs(uint32_t iblockrow_to, uint32_t iblockrow_from);
  void set_diagonal(const std::vector<double>& diag);
  void add_diagonal(const std::vector<double>& diag);
  void get_diagonal(std::vector<double>& diag);
  void reset(double v = 0.);
  // y = alpha A x + beta y for vectors of the same system (TrilinosDetail.cpp:380-411)
  void apply(Vector& y, const Vector& x, double alpha = 1., double beta = 0.);
  void debug_data(std::vector<uint32_t>& row_indices, std::vector<uint32_t>& col_indices, std::vector<double>& values);
  // the two Trilinos builders' type names (LSSAction matrix_builder)
  std::string type_name() const;
  std::string name() const { return "Matrix"; }
  // Matrix::print (TrilinosFEVbrMatrix.cpp:704-733, TrilinosCrsMatrix.cpp:589-625): one line
  // "column -row value" per stored scalar entry, point rows in order, then the summary lines; with a file
  // name (:736-743, :629-636) the Tecplot header lines first
  void print(std::ostream& stream);
  void print(const std::string& filename, std::ios_base::openmode mode = std::ios_base::out);
  // Epetra's own Print format is not reproduced: the same lines as print
  void print_native(std::ostream& stream) { print(stream); }
  // TrilinosCrsMatrix::clone_to / read_native (TrilinosCrsMatrix.cpp:518-546); NotImplemented for the
  // block builder (TrilinosFEVbrMatrix.cpp:798-808)
  void clone_to(Matrix& other);
  void read_native(const std::string& path);
  bool is_created() const;
  uint32_t neq() const;
  uint32_t blockrow_size() const;
  uint32_t blockcol_size() const { return blockrow_size(); }
private:
  System& sys_;
};

// SolutionStrategy.hpp:35-59 with the Krylov options of the build (BiCGStab + Jacobi, GMRES(m) + ILU0;
// FP64 or the precision form)
class SolutionStrategy {
public:
  explicit SolutionStrategy(System& s) : sys_(s)
  {
    params.method = CF3HIP_BICGSTAB_JACOBI;
    params.restart = 30;
    params.max_iter = 5000;
    params.tol = 1e-10;
    params.precision = CF3HIP_FP64;
    params.inner_tol = 1e-2;
    params.max_outer = 0;
    params.block_scale = 1;
    params.fp32_vectors = 0;
    params.recycle = 0;
    params.rcg_space = 0;
  }
  // RCGStrategy (RCGStrategy.cpp:59-171): recycling CG with `recycle` vectors kept between solves,
  // cycles of `restart` directions, in recycle space `space` of the system
  void use_rcg(int recycle, int restart, int space = 0)
  {
    params.method = CF3HIP_RCG_JACOBI;
    params.recycle = recycle;
    params.restart = restart;
    params.rcg_space = space;
  }
  cf3hip_solver_params params;
  cf3hip_solve_stats stats = {};
  // solve the system into its solution vector; returns whether the tolerance was reached (the
  // reference logs an unconverged Belos solve instead of throwing)
  bool solve();
  double compute_residual();
private:
  System& sys_;
};

// System.cpp: matrix, rhs, solution and strategy of one linear system
class System {
public:
  // System::create(cp, neq, node_connectivity, starting_indices) (System.cpp:82-110)
  System(Context& ctx, uint32_t neq, const std::vector<uint32_t>& node_connectivity,
         const std::vector<uint32_t>& starting_indices)
      : ctx_(ctx), matrix_(*this), rhs_(*this, CF3HIP_RHS), sol_(*this, CF3HIP_SOLUTION), strategy_(*this)
  {
    if(starting_indices.empty()) throw BadValue(CF3HIP_ERR_INVALID_ARG, "empty graph");
    check(cf3hip_lss_create_from_graph(ctx.handle(), (uint32_t)(starting_indices.size() - 1), (int)neq,
                                       starting_indices.data(), node_connectivity.data(), &h_));
  }
  // the system of a mesh: the sparsity of SparsityBuilder.cpp and the element scatter structures
  System(Context& ctx, Mesh& mesh, uint32_t neq)
      : ctx_(ctx), matrix_(*this), rhs_(*this, CF3HIP_RHS), sol_(*this, CF3HIP_SOLUTION), strategy_(*this)
  {
    check(cf3hip_lss_create(ctx.handle(), mesh.handle(), (int)neq, &h_));
  }
  // System::create with periodic links (System.cpp:82-99): node i with periodic_links_active[i] shares
  // the block row of periodic_links_nodes[i] (the end of its chain of links).  As in the reference
  // (TrilinosFEVbrMatrix's m_p2m), every index given to the matrix, vector and dirichlet calls of this
  // system is a mesh node, converted here; whole-vector get/set/diagonal calls are in block-row order.
  System(Context& ctx, Mesh& mesh, uint32_t neq, const std::vector<uint32_t>& periodic_links_nodes,
         const std::vector<bool>& periodic_links_active)
      : ctx_(ctx), matrix_(*this), rhs_(*this, CF3HIP_RHS), sol_(*this, CF3HIP_SOLUTION), strategy_(*this)
  {
    if(periodic_links_nodes.size() != mesh.n_nodes() || periodic_links_active.size() != mesh.n_nodes())
      throw BadValue(CF3HIP_ERR_INVALID_ARG, "periodic link lists need one entry per mesh node");
    std::vector<uint8_t> active(periodic_links_active.begin(), periodic_links_active.end());
    check(cf3hip_lss_create_periodic(ctx.handle(), mesh.handle(), (int)neq, periodic_links_nodes.data(),
                                     active.data(), &h_));
    p2m_.resize(mesh.n_nodes());
    check(cf3hip_lss_get_p2m(h_, p2m_.data()));
  }
  // block row of mesh node i (the identity without periodic links)
  uint32_t row(uint32_t i) const
  {
    if(p2m_.empty()) return i;
    if(i >= p2m_.size()) throw BadValue(CF3HIP_ERR_INVALID_ARG, "node index outside the mesh");
    return p2m_[i];
  }
  std::vector<uint32_t> rows(const std::vector<uint32_t>& nodes) const
  {
    std::vector<uint32_t> r(nodes.size());
    for(size_t i = 0; i < nodes.size(); ++i) r[i] = row(nodes[i]);
    return r;
  }
  // scalar index node * neq + eq of a mesh node converted to its block row
  uint32_t scalar_row(uint32_t i) const
  {
    const uint32_t q = neq();
    return row(i / q) * q + i % q;
  }
  ~System() { cf3hip_lss_destroy(h_); }
  System(const System&) = delete;
  System& operator=(const System&) = delete;

  Matrix& matrix() { return matrix_; }
  Vector& rhs() { return rhs_; }
  Vector& solution() { return sol_; }
  SolutionStrategy& solution_strategy() { return strategy_; }
  bool is_created() const { return h_ != nullptr; }
  // System::dirichlet (System.cpp:240-255)
  void dirichlet(uint32_t iblockrow, uint32_t ieq, double value, bool preserve_symmetry = true)
  {
    if(preserve_symmetry) {
      matrix_.symmetric_dirichlet(iblockrow, ieq, value, rhs_);
    } else {
      matrix_.set_row(iblockrow, ieq, 1., 0.);
      rhs_.set_value(iblockrow, ieq, value);
    }
    sol_.set_value(iblockrow, ieq, value);
  }
  // the ordered conditions of one BoundaryConditions pass in one call (symmetric form)
  void dirichlet(const std::vector<uint32_t>& nodes, const std::vector<uint32_t>& eqs, const std::vector<double>& values)
  {
    const std::vector<uint32_t> r = rows(nodes);
    check(cf3hip_lss_dirichlet(h_, (uint32_t)r.size(), r.data(), eqs.data(), values.data()));
  }
  // System::periodicity (System.cpp:259-282): tie the two block rows, move the rhs of row from into row
  // to, and give both solution entries their mean
  void periodicity(uint32_t iblockrow_to, uint32_t iblockrow_from)
  {
    const uint32_t t = row(iblockrow_to), f = row(iblockrow_from);
    check(cf3hip_lss_periodicity(h_, 1, &t, &f));
  }
  bool solve() { return strategy_.solve(); }
  // LSSAction matrix_builder: "cf3.math.LSS.TrilinosFEVbrMatrix" (default) or
  // "cf3.math.LSS.TrilinosCrsMatrix" semantics (before the first condition)
  void set_matrix_builder(const std::string& builder)
  {
    int b;
    if(builder == "cf3.math.LSS.TrilinosFEVbrMatrix" || builder == "HipFEVbrMatrix") b = CF3HIP_BUILDER_VBR;
    else if(builder == "cf3.math.LSS.TrilinosCrsMatrix" || builder == "HipCrsMatrix") b = CF3HIP_BUILDER_CRS;
    else throw BadValue(CF3HIP_ERR_INVALID_ARG, "unknown matrix builder " + builder);
    check(cf3hip_lss_set_builder(h_, b));
  }
  // mesh nodes indexed by the matrix calls (block rows without periodic links)
  uint32_t n_mesh_nodes() const { return p2m_.empty() ? n_nodes() : (uint32_t)p2m_.size(); }
  void reset(double v = 0.)
  {
    matrix_.reset(v);
    rhs_.reset(v);
    sol_.reset(v);
  }
  cf3hip_lss handle() const { return h_; }
  Context& context() const { return ctx_; }
  uint32_t n_nodes() const
  {
    uint32_t n = 0;
    int neq = 0;
    uint64_t nb = 0;
    check(cf3hip_lss_info(h_, &n, &neq, &nb));
    return n;
  }
  uint32_t neq() const
  {
    uint32_t n = 0;
    int neq = 0;
    uint64_t nb = 0;
    check(cf3hip_lss_info(h_, &n, &neq, &nb));
    return (uint32_t)neq;
  }
private:
  Context& ctx_;
  cf3hip_lss h_ = nullptr;
  Matrix matrix_;
  Vector rhs_, sol_;
  SolutionStrategy strategy_;
  std::vector<uint32_t> p2m_;   // empty: no periodic links
};

// ---- Vector ----
inline void Vector::set_value(uint32_t irow, double v) { set_value(irow / neq(), irow % neq(), v); }
inline void Vector::add_value(uint32_t irow, double v) { add_value(irow / neq(), irow % neq(), v); }
inline void Vector::get_value(uint32_t irow, double& v) { get_value(irow / neq(), irow % neq(), v); }
inline void Vector::set_value(uint32_t iblockrow, uint32_t ieq, double v)
{
  check(cf3hip_vector_set_value(sys_.handle(), which_, sys_.row(iblockrow), ieq, v));
}
inline void Vector::add_value(uint32_t iblockrow, uint32_t ieq, double v)
{
  check(cf3hip_vector_add_value(sys_.handle(), which_, sys_.row(iblockrow), ieq, v));
}
inline void Vector::get_value(uint32_t iblockrow, uint32_t ieq, double& v)
{
  check(cf3hip_vector_get_value(sys_.handle(), which_, sys_.row(iblockrow), ieq, &v));
}
inline void Vector::block(const std::vector<uint32_t>& nodes, double* vals, int mode)
{
  const std::vector<uint32_t> idx = sys_.rows(nodes);
  const uint32_t nb = (uint32_t)idx.size();
  if(mode == 0) check(cf3hip_vector_set_values(sys_.handle(), which_, nb, idx.data(), vals));
  else if(mode == 1) check(cf3hip_vector_add_values(sys_.handle(), which_, nb, idx.data(), vals));
  else check(cf3hip_vector_get_values(sys_.handle(), which_, nb, idx.data(), vals));
}
inline void Vector::reset(double v) { check(cf3hip_vector_reset(sys_.handle(), which_, v)); }
inline void Vector::get(std::vector<double>& data)
{
  data.resize((size_t)blockrow_size() * neq());
  check(cf3hip_vector_get(sys_.handle(), which_, data.data()));
}
inline void Vector::set(const std::vector<double>& data)
{
  if(data.size() != (size_t)blockrow_size() * neq()) throw BadValue(CF3HIP_ERR_INVALID_ARG, "vector size");
  check(cf3hip_vector_set(sys_.handle(), which_, data.data()));
}
inline void Vector::assign(const Vector& source)
{
  if(&source.sys_ != &sys_)
    throw SetupError(CF3HIP_ERR_NOT_CREATED, "assign needs a vector of the same system");
  check(cf3hip_vector_assign(sys_.handle(), which_, source.which_));
}
inline void Vector::update(const Vector& source, double alpha)
{
  if(&source.sys_ != &sys_)
    throw SetupError(CF3HIP_ERR_NOT_CREATED, "update needs a vector of the same system");
  check(cf3hip_vector_update(sys_.handle(), which_, source.which_, alpha));
}
inline void Vector::scale(double alpha) { check(cf3hip_vector_scale(sys_.handle(), which_, alpha)); }
inline bool Vector::is_created() const { return sys_.is_created(); }
inline uint32_t Vector::neq() const { return sys_.neq(); }
inline uint32_t Vector::blockrow_size() const { return sys_.n_nodes(); }

// ---- Matrix ----
inline void Matrix::set_value(uint32_t icol, uint32_t irow, double v)
{
  check(cf3hip_matrix_set_value(sys_.handle(), sys_.scalar_row(icol), sys_.scalar_row(irow), v));
}
inline void Matrix::add_value(uint32_t icol, uint32_t irow, double v)
{
  check(cf3hip_matrix_add_value(sys_.handle(), sys_.scalar_row(icol), sys_.scalar_row(irow), v));
}
inline void Matrix::get_value(uint32_t icol, uint32_t irow, double& v)
{
  check(cf3hip_matrix_get_value(sys_.handle(), sys_.scalar_row(icol), sys_.scalar_row(irow), &v));
}
inline void Matrix::set_values(const BlockAccumulator& b)
{
  const std::vector<uint32_t> r = sys_.rows(b.indices);
  check(cf3hip_matrix_set_values(sys_.handle(), b.size(), r.data(), b.mat.data()));
}
inline void Matrix::add_values(const BlockAccumulator& b)
{
  const std::vector<uint32_t> r = sys_.rows(b.indices);
  check(cf3hip_matrix_add_values(sys_.handle(), b.size(), r.data(), b.mat.data()));
}
inline void Matrix::get_values(BlockAccumulator& b)
{
  const std::vector<uint32_t> r = sys_.rows(b.indices);
  check(cf3hip_matrix_get_block_values(sys_.handle(), b.size(), r.data(), b.mat.data()));
}
inline void Matrix::set_row(uint32_t iblockrow, uint32_t ieq, double diagval, double offdiagval)
{
  check(cf3hip_matrix_set_row(sys_.handle(), sys_.row(iblockrow), ieq, diagval, offdiagval));
}
inline void Matrix::tie_blockrow_pairs(uint32_t iblockrow_to, uint32_t iblockrow_from)
{
  const uint32_t t = sys_.row(iblockrow_to), f = sys_.row(iblockrow_from);
  check(cf3hip_matrix_tie_blockrow_pairs(sys_.handle(), 1, &t, &f));
}
inline void Matrix::get_column_and_replace_to_zero(uint32_t iblockcol, uint32_t ieq, std::vector<double>& values)
{
  values.resize((size_t)blockcol_size() * neq());
  check(cf3hip_matrix_get_column_and_replace_to_zero(sys_.handle(), sys_.row(iblockcol), ieq, values.data()));
}
inline void Matrix::symmetric_dirichlet(uint32_t blockrow, uint32_t ieq, double value, Vector& rhs)
{
  if(&rhs.system() != &sys_ || rhs.which() != CF3HIP_RHS)
    throw SetupError(CF3HIP_ERR_NOT_CREATED, "symmetric_dirichlet acts on the rhs of the same system");
  const uint32_t r = sys_.row(blockrow);
  check(cf3hip_lss_dirichlet(sys_.handle(), 1, &r, &ieq, &value));
}
inline void Matrix::set_diagonal(const std::vector<double>& diag)
{
  if(diag.size() != (size_t)blockrow_size() * neq()) throw BadValue(CF3HIP_ERR_INVALID_ARG, "diagonal size");
  check(cf3hip_matrix_set_diagonal(sys_.handle(), diag.data()));
}
inline void Matrix::add_diagonal(const std::vector<double>& diag)
{
  if(diag.size() != (size_t)blockrow_size() * neq()) throw BadValue(CF3HIP_ERR_INVALID_ARG, "diagonal size");
  check(cf3hip_matrix_add_diagonal(sys_.handle(), diag.data()));
}
inline void Matrix::get_diagonal(std::vector<double>& diag)
{
  diag.resize((size_t)blockrow_size() * neq());
  check(cf3hip_matrix_get_diagonal(sys_.handle(), diag.data()));
}
inline void Matrix::reset(double v) { check(cf3hip_matrix_reset(sys_.handle(), v)); }
inline void Matrix::apply(Vector& y, const Vector& x, double alpha, double beta)
{
  if(&y.system() != &sys_ || &x.system() != &sys_)
    throw SetupError(CF3HIP_ERR_NOT_CREATED, "apply takes vectors of the same system");
  double *xd = nullptr, *yd = nullptr;
  check(cf3hip_vector_device_ptr(sys_.handle(), x.which(), &xd));
  check(cf3hip_vector_device_ptr(sys_.handle(), y.which(), &yd));
  if(xd == yd) throw BadValue(CF3HIP_ERR_INVALID_ARG, "apply needs distinct x and y");
  check(cf3hip_matrix_apply(sys_.handle(), alpha, xd, beta, yd));
}
inline void Matrix::debug_data(std::vector<uint32_t>& row_indices, std::vector<uint32_t>& col_indices,
                               std::vector<double>& values)
{
  uint32_t n = 0;
  int neq = 0;
  uint64_t nb = 0;
  check(cf3hip_lss_info(sys_.handle(), &n, &neq, &nb));
  row_indices.resize((size_t)n * neq + 1);
  col_indices.resize((size_t)nb * neq * neq);
  values.resize((size_t)nb * neq * neq);
  check(cf3hip_matrix_debug_data(sys_.handle(), row_indices.data(), col_indices.data(), values.data()));
}
inline std::string Matrix::type_name() const
{
  int b = 0;
  check(cf3hip_lss_get_builder(sys_.handle(), &b));
  return b == CF3HIP_BUILDER_CRS ? "HipCrsMatrix" : "HipFEVbrMatrix";
}
inline void Matrix::print(std::ostream& stream)
{
  if(!is_created()) {
    stream << name() << " of type " << type_name() << "::is_created() is false, nothing is printed.";
    return;
  }
  std::vector<uint32_t> rp, ci;
  std::vector<double> v;
  debug_data(rp, ci, v);
  const uint32_t q = neq(), nrow = (uint32_t)rp.size() - 1;
  // scalar index of a point row or column: the last mesh node of its block row (the reference's m2p)
  std::vector<uint32_t> m2p(nrow / q);
  for(uint32_t r = 0; r < nrow / q; ++r) m2p[r] = r;
  for(uint32_t i = 0; i < sys_.n_mesh_nodes(); ++i) m2p[sys_.row(i)] = i;
  for(uint32_t r = 0; r < nrow; ++r)
    for(uint32_t k = rp[r]; k < rp[r + 1]; ++k)
      stream << m2p[ci[k] / q] * q + ci[k] % q << " " << -(int64_t)(m2p[r / q] * q + r % q) << " " << v[k] << "\n";
  stream << "# name:                 " << name() << "\n";
  stream << "# type_name:            " << type_name() << "\n";
  stream << "# process:              " << 0 << "\n";
  stream << "# number of equations:  " << q << "\n";
  stream << "# number of rows:       " << nrow << "\n";
  stream << "# number of cols:       " << nrow << "\n";
  stream << "# number of block rows: " << nrow / q << "\n";
  stream << "# number of block cols: " << nrow / q << "\n";
  stream << "# number of entries:    " << rp[nrow] << "\n" << std::flush;
}
inline void Matrix::print(const std::string& filename, std::ios_base::openmode mode)
{
  std::ofstream stream(filename.c_str(), mode);
  stream << "VARIABLES=COL,ROW,VAL\n" << std::flush;
  stream << "ZONE T=\"" << type_name() << "::" << name() << "\"\n" << std::flush;
  print(stream);
}
inline void Matrix::clone_to(Matrix& other) { check(cf3hip_matrix_clone_to(sys_.handle(), other.sys_.handle())); }
inline void Matrix::read_native(const std::string& path) { check(cf3hip_matrix_read_native(sys_.handle(), path.c_str())); }
inline bool Matrix::is_created() const { return sys_.is_created(); }
inline uint32_t Matrix::neq() const { return sys_.neq(); }
inline uint32_t Matrix::blockrow_size() const { return sys_.n_nodes(); }

// ---- SolutionStrategy ----
inline bool SolutionStrategy::solve()
{
  const int rc = cf3hip_lss_solve(sys_.handle(), &params, &stats);
  if(rc == CF3HIP_ERR_NOT_CONVERGED) return false;
  check(rc);
  return stats.converged != 0;
}
inline double SolutionStrategy::compute_residual()
{
  double r = 0.;
  check(cf3hip_lss_residual(sys_.handle(), &r));
  return r;
}

}  // namespace lss

namespace ufem {

// the NavierStokes fields of one mesh (solution u, p; linearized_velocity; viscosity; body force)
class Fields {
public:
  explicit Fields(Mesh& m) { check(cf3hip_fields_create(m.handle(), &h_)); }
  ~Fields() { cf3hip_fields_destroy(h_); }
  Fields(const Fields&) = delete;
  Fields& operator=(const Fields&) = delete;
  void set(int which, const std::vector<double>& host) { check(cf3hip_fields_set(h_, which, host.data())); }
  void get(int which, std::vector<double>& host, uint32_t n_nodes, int ncomp)
  {
    host.resize((size_t)n_nodes * ncomp);
    check(cf3hip_fields_get(h_, which, host.data()));
  }
  // NavierStokes::on_initial_conditions_set (NavierStokes.cpp:194-205)
  void initialize(double nu) { check(cf3hip_fields_initialize(h_, nu)); }
  cf3hip_fields handle() const { return h_; }
private:
  cf3hip_fields h_ = nullptr;
};

// the NavierStokes LSS action (NavierStokes.hpp:34-118, NavierStokes.cpp:103-121): options with the
// reference's names and defaults, and the actions of one time step
class NavierStokesAssembly {
public:
  bool use_specializations = true;   // NavierStokes "use_specializations"
  std::string supg_type = "metric";  // ComputeTau "supg_type": metric, tezduyar, cf2
  double theta = 1.;                 // NavierStokes "theta" (NavierStokes.cpp:85-88)
  double density = 1.2;              // NavierStokesPhysics "density"
  double dt = 0.;                    // LSSActionUnsteady "dt"
  double c1 = 1., c2 = 16., alpha_ps = 1., alpha_su = 1., alpha_bu = 1.;
  double u_ref = 1.;                 // ComputeTau "u_ref" (SUPG.hpp:103), used by supg_type cf2
  int precision = CF3HIP_FP64;       // FP32: element matrices of the precision form

  cf3hip_ns_params params() const
  {
    if(theta < 0. || theta > 1.)
      throw SetupError(CF3HIP_ERR_NOT_CREATED, "Value for theta option is outside of the valid range from 0 to 1.");
    cf3hip_ns_params p;
    p.rho = density;
    p.dt = dt;
    p.theta = theta;
    p.use_specializations = use_specializations ? 1 : 0;
    if(supg_type == "metric") p.supg_type = CF3HIP_SUPG_METRIC;
    else if(supg_type == "tezduyar") p.supg_type = CF3HIP_SUPG_TEZDUYAR;
    else if(supg_type == "cf2") p.supg_type = CF3HIP_SUPG_CF2;
    else throw BadValue(CF3HIP_ERR_INVALID_ARG, "unknown supg_type " + supg_type);
    p.c1 = c1; p.c2 = c2;
    p.alpha_ps = alpha_ps; p.alpha_su = alpha_su; p.alpha_bu = alpha_bu;
    p.precision = precision;
    p.u_ref = u_ref;
    return p;
  }
  // LinearizeU
  void linearize(Fields& f) { check(cf3hip_ns_linearize(f.handle())); }
  // ZeroLSS + Assembly (matrix and rhs; the solution keeps the previous increment)
  void execute(lss::System& s, Fields& f)
  {
    const cf3hip_ns_params p = params();
    check(cf3hip_ns_assemble(s.handle(), f.handle(), &p));
  }
  // SolveLSS of the NS system (the precision form uses the FP64 element operator for its residuals)
  bool solve(lss::System& s, Fields& f)
  {
    const cf3hip_ns_params p = params();
    const int rc = cf3hip_ns_solve(s.handle(), f.handle(), &p, &s.solution_strategy().params,
                                   &s.solution_strategy().stats);
    if(rc == CF3HIP_ERR_NOT_CONVERGED) return false;
    check(rc);
    return s.solution_strategy().stats.converged != 0;
  }
  // UpdateActions
  void update(lss::System& s, Fields& f) { check(cf3hip_ns_update(s.handle(), f.handle())); }
};

// NavierStokesSemiImplicit (ns_semi_implicit/implementation/NavierStokesSemiImplicit.cpp): a velocity
// system (neq = dim) and a pressure system (neq = 1) of one mesh, with the options of the reference
// ("theta" default 1, "nb_iterations" default 2, the ComputeTau options) and its actions.  The
// solution strategies of the two systems set the solver of each solve.
class NavierStokesSemiImplicit {
public:
  double theta = 1.;                 // NavierStokesSemiImplicit "theta"
  int nb_iterations = 2;             // "nb_iterations" (InnerLoop passes)
  std::string supg_type = "metric";
  double dt = 0.;
  double c1 = 1., c2 = 16., alpha_ps = 1., alpha_su = 1., alpha_bu = 1.;
  bool enable_body_force = false;    // "enable_body_force" (NavierStokesSemiImplicit.cpp:313-317)
  bool pressure_rcg_solve = false;   // "pressure_rcg_solve" (:309-311, :532-537)

  NavierStokesSemiImplicit(lss::System& velocity, lss::System& pressure) : u_(velocity), p_(pressure)
  {
    check(cf3hip_si_create(velocity.handle(), pressure.handle(), &h_));
  }
  ~NavierStokesSemiImplicit() { if(h_) cf3hip_si_destroy(h_); }
  NavierStokesSemiImplicit(const NavierStokesSemiImplicit&) = delete;
  NavierStokesSemiImplicit& operator=(const NavierStokesSemiImplicit&) = delete;

  cf3hip_ns_params params() const
  {
    NavierStokesAssembly a;
    a.theta = theta; a.dt = dt; a.supg_type = supg_type; a.density = 1.;
    a.c1 = c1; a.c2 = c2; a.alpha_ps = alpha_ps; a.alpha_su = alpha_su; a.alpha_bu = alpha_bu;
    return a.params();
  }
  // the initial-condition actions: PressureAssembly, ZeroVelocitySystem, MassMatrixAssembly,
  // SetupInnerLoopData
  void initial(Fields& f)
  {
    apply_options();
    const cf3hip_ns_params p = params();
    check(cf3hip_si_initial(h_, f.handle(), &p));
  }
  // one time step; conditions are the new values at mesh nodes (velocity: nodes x dim, row-major).
  // Returns whether every solve reached its tolerance.
  bool step(Fields& f, const std::vector<uint32_t>& u_nodes, const std::vector<double>& u_values,
            const std::vector<uint32_t>& p_nodes, const std::vector<double>& p_values)
  {
    apply_options();
    const cf3hip_ns_params p = params();
    stats.assign(2 * nb_iterations, cf3hip_solve_stats{});
    const int rc = cf3hip_si_step(h_, f.handle(), &p, nb_iterations, &u_.solution_strategy().params,
                                  &p_.solution_strategy().params, (uint32_t)u_nodes.size(), u_nodes.data(),
                                  u_values.data(), (uint32_t)p_nodes.size(), p_nodes.data(), p_values.data(),
                                  stats.data());
    if(rc == CF3HIP_ERR_NOT_CONVERGED) return false;
    check(rc);
    return true;
  }
  std::vector<cf3hip_solve_stats> stats;   // velocity and pressure solve of each inner pass
  cf3hip_si handle() const { return h_; }
private:
  void apply_options()
  {
    const cf3hip_si_options o = {enable_body_force ? 1 : 0, pressure_rcg_solve ? 1 : 0};
    check(cf3hip_si_set_options(h_, &o));
  }
  lss::System& u_;
  lss::System& p_;
  cf3hip_si h_ = nullptr;
};

}  // namespace ufem
}  // namespace cf3hip

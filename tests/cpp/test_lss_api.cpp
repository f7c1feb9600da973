// C++ checks of the host layer (coolfluid3_amd/host/cf3hip_lss.hpp), written like the reference's
// LSS unit tests: test/math/utest-lss-symmetric-dirichlet.cpp (exact values), utest-lss-atomic.cpp
// (set/add/get, BlockAccumulator, set_row, diagonal, column replace, vector operations, solve) and
// utest-lss-vector.cpp (apply), plus the UFEM assembly through the C++ action compared with the
// oracle.  Usage: test_lss_api --no-device (exception mapping only) | test_lss_api (needs a GPU).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "cf3hip_lss.hpp"
#include "oracle.h"

using namespace cf3hip;

static int g_fail = 0;
#define CHECK(cond)                                                                                \
  do {                                                                                             \
    if(!(cond)) {                                                                                  \
      std::printf("CHECK failed %s:%d: %s\n", __FILE__, __LINE__, #cond);                         \
      ++g_fail;                                                                                    \
    }                                                                                              \
  } while(0)

template<typename Ex, typename F>
static bool throws(F&& f)
{
  try {
    f();
  } catch(const Ex&) {
    return true;
  } catch(...) {
    return false;
  }
  return false;
}

// 4-node chain 0-1-2-3, one equation per node (utest-lss-symmetric-dirichlet.cpp:146-210)
static void symmetric_dirichlet(Context& ctx, const char* builder)
{
  const std::vector<uint32_t> start = {0, 2, 5, 8, 10};
  const std::vector<uint32_t> conn = {0, 1, 0, 1, 2, 1, 2, 3, 2, 3};
  lss::System s(ctx, 1, conn, start);
  s.set_matrix_builder(builder);   // the reference runs this test once per matrix_builder
  for(uint32_t i = 0; i < 4; ++i) s.matrix().set_row(i, 0, 2., 1.);
  s.dirichlet(1, 0, 10.);
  const double expect[4][4] = {{2, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 2, 1}, {0, 0, 1, 2}};
  for(uint32_t r = 0; r < 4; ++r)
    for(uint32_t c = 0; c < 4; ++c) {
      if((r > c ? r - c : c - r) > 1) continue;
      double v = -1.;
      s.matrix().get_value(c, r, v);
      CHECK(v == expect[r][c]);
    }
  std::vector<double> rhs;
  s.rhs().get(rhs);
  CHECK(rhs[0] == -10. && rhs[1] == 10. && rhs[2] == -10. && rhs[3] == 0.);
  double sv = 0.;
  s.solution().get_value(1, 0, sv);
  CHECK(sv == 10.);
  // an entry outside the graph is illegal (TrilinosFEVbrMatrix.cpp:175)
  CHECK(throws<BadValue>([&] { double v; s.matrix().get_value(3, 0, v); }));
  // rows 0 and 3 hold different block columns: they cannot be tied (TrilinosFEVbrMatrix.cpp:575-577)
  CHECK(throws<BadValue>([&] { s.matrix().tie_blockrow_pairs(0, 3); }));
}

// The two builders differ when a condition is applied again before a reset: TrilinosCrsMatrix subtracts
// its cached column again (TrilinosCrsMatrix.cpp:392-398), TrilinosFEVbrMatrix finds the column already
// zero.  Also: print format (TrilinosFEVbrMatrix.cpp:704-743), clone_to and read_native (CRS only) and
// get_column_and_replace_to_zero (VBR only).
static void matrix_builders(Context& ctx)
{
  const std::vector<uint32_t> start = {0, 2, 5, 8, 10};
  const std::vector<uint32_t> conn = {0, 1, 0, 1, 2, 1, 2, 3, 2, 3};
  for(int crs = 0; crs < 2; ++crs) {
    lss::System s(ctx, 1, conn, start);
    s.set_matrix_builder(crs ? "cf3.math.LSS.TrilinosCrsMatrix" : "cf3.math.LSS.TrilinosFEVbrMatrix");
    CHECK(s.matrix().type_name() == (crs ? "HipCrsMatrix" : "HipFEVbrMatrix"));
    for(uint32_t i = 0; i < 4; ++i) s.matrix().set_row(i, 0, 2., 1.);
    s.dirichlet(1, 0, 10.);
    s.dirichlet(1, 0, 10.);
    std::vector<double> rhs;
    s.rhs().get(rhs);
    if(crs) CHECK(rhs[0] == -20. && rhs[1] == 10. && rhs[2] == -20. && rhs[3] == 0.);
    else CHECK(rhs[0] == -10. && rhs[1] == 10. && rhs[2] == -10. && rhs[3] == 0.);
    // print: one "column -row value" line per stored entry (10), then nine summary lines
    std::ostringstream out;
    s.matrix().print(out);
    std::istringstream in(out.str());
    std::string line;
    std::vector<std::string> lines;
    while(std::getline(in, line)) lines.push_back(line);
    CHECK(lines.size() == 19);
    CHECK(lines[0] == "0 0 2");              // row 0, column 0
    CHECK(lines[1] == "1 0 0");              // row 0, column 1: eliminated
    CHECK(lines[2] == "0 -1 0" && lines[3] == "1 -1 1" && lines[4] == "2 -1 0");
    CHECK(lines[10] == "# name:                 Matrix");
    CHECK(lines[11] == std::string("# type_name:            ") + (crs ? "HipCrsMatrix" : "HipFEVbrMatrix"));
    CHECK(lines[18] == "# number of entries:    10");
    lss::System t(ctx, 1, conn, start);
    t.set_matrix_builder(crs ? "cf3.math.LSS.TrilinosCrsMatrix" : "cf3.math.LSS.TrilinosFEVbrMatrix");
    std::vector<double> col;
    if(crs) {
      s.matrix().clone_to(t.matrix());
      double a = 0., b = 0.;
      s.matrix().get_value(2, 3, a);
      t.matrix().get_value(2, 3, b);
      CHECK(a == 1. && b == 1.);
      CHECK(throws<NotImplemented>([&] { s.matrix().get_column_and_replace_to_zero(2, 0, col); }));
    } else {
      CHECK(throws<NotImplemented>([&] { s.matrix().clone_to(t.matrix()); }));
      CHECK(throws<NotImplemented>([&] { s.matrix().read_native("/nonexistent.mtx"); }));
    }
    CHECK(throws<BadValue>([&] { s.set_matrix_builder("cf3.math.LSS.TrilinosCrsMatrix"); }));   // after a condition
  }
}

// System::periodicity on a dense 3-node graph (System.cpp:259-282, TrilinosFEVbrMatrix.cpp:558-590):
// A = [1 2 3; 4 5 6; 7 8 9], rhs = (1, 2, 3), sol = (10, 20, 30); tying row 0 to row 2 gives
// A = [1 0 -1; 4 5 6; 0 10 20], rhs = (0, 2, 4), sol = (20, 20, 20)
static void periodicity(Context& ctx)
{
  const std::vector<uint32_t> start = {0, 3, 6, 9};
  const std::vector<uint32_t> conn = {0, 1, 2, 0, 1, 2, 0, 1, 2};
  lss::System s(ctx, 1, conn, start);
  for(uint32_t r = 0; r < 3; ++r) {
    for(uint32_t c = 0; c < 3; ++c) s.matrix().set_value(c, r, 3. * r + c + 1.);
    s.rhs().set_value(r, 0, r + 1.);
    s.solution().set_value(r, 0, 10. * (r + 1.));
  }
  s.periodicity(2, 0);
  const double expect[3][3] = {{1, 0, -1}, {4, 5, 6}, {0, 10, 20}};
  for(uint32_t r = 0; r < 3; ++r)
    for(uint32_t c = 0; c < 3; ++c) {
      double v = -1.;
      s.matrix().get_value(c, r, v);
      CHECK(v == expect[r][c]);
    }
  std::vector<double> rhs, sol;
  s.rhs().get(rhs);
  s.solution().get(sol);
  CHECK(rhs[0] == 0. && rhs[1] == 2. && rhs[2] == 4.);
  CHECK(sol[0] == 20. && sol[1] == 20. && sol[2] == 20.);
}

// 1D chain of n nodes with neq = 2 (utest-lss-atomic.cpp style)
static void atomic_operations(Context& ctx)
{
  const uint32_t n = 6, neq = 2;
  std::vector<uint32_t> start(1, 0), conn;
  for(uint32_t i = 0; i < n; ++i) {
    for(int d = -1; d <= 1; ++d)
      if((int)i + d >= 0 && (int)i + d < (int)n) conn.push_back(i + d);
    start.push_back((uint32_t)conn.size());
  }
  lss::System s(ctx, neq, conn, start);
  lss::Matrix& A = s.matrix();
  CHECK(A.is_created() && A.neq() == neq && A.blockrow_size() == n && A.blockcol_size() == n);
  A.reset(0.);
  // scalar access: column first
  A.set_value(3, 2, 1.5);
  A.add_value(3, 2, 0.25);
  double v = 0.;
  A.get_value(3, 2, v);
  CHECK(v == 1.75);
  // BlockAccumulator on nodes (1, 2)
  lss::BlockAccumulator ba;
  ba.resize(2, neq);
  ba.indices[0] = 1;
  ba.indices[1] = 2;
  for(size_t i = 0; i < ba.mat.size(); ++i) ba.mat[i] = (double)i;
  A.set_values(ba);
  A.add_values(ba);
  lss::BlockAccumulator back;
  back.resize(2, neq);
  back.indices = ba.indices;
  A.get_values(back);
  for(size_t i = 0; i < ba.mat.size(); ++i) CHECK(back.mat[i] == 2. * ba.mat[i]);
  // set_row: row (node 4, eq 1) = offdiag everywhere in the graph, diag on the diagonal
  A.set_row(4, 1, 7., -1.);
  A.get_value(9, 9, v);
  CHECK(v == 7.);
  A.get_value(8, 9, v);
  CHECK(v == -1.);
  A.get_value(6, 9, v);
  CHECK(v == -1.);
  // diagonal
  std::vector<double> d(n * neq, 1.), dg;
  A.set_diagonal(d);
  A.add_diagonal(d);
  A.get_diagonal(dg);
  for(double x : dg) CHECK(x == 2.);
  // column (node 2, eq 0) out and zeroed
  A.set_value(4, 3, 3.);   // row 3 (node 1, eq 1), column 4 (node 2, eq 0)
  std::vector<double> col;
  A.get_column_and_replace_to_zero(2, 0, col);
  CHECK(col.size() == n * neq);
  CHECK(col[3] == 3.);
  A.get_value(4, 3, v);
  CHECK(v == 0.);
  CHECK(col[0] == 0. && col[10] == 0.);   // nodes 0 and 5 are not neighbours of node 2
  // vectors
  lss::Vector& b = s.rhs();
  lss::Vector& x = s.solution();
  b.reset(1.);
  b.set_value(5, 4.);                     // dof 5 = (node 2, eq 1)
  b.add_value(2, 1, 1.);
  b.get_value(2, 1, v);
  CHECK(v == 5.);
  lss::BlockAccumulator bv;
  bv.resize(1, neq);
  bv.indices[0] = 3;
  bv.rhs[0] = 2.;
  bv.rhs[1] = 3.;
  b.add_rhs_values(bv);
  b.get_rhs_values(bv);
  CHECK(bv.rhs[0] == 3. && bv.rhs[1] == 4.);
  x.assign(b);
  x.update(b, -0.5);
  x.scale(4.);
  std::vector<double> xv, bvv;
  x.get(xv);
  b.get(bvv);
  for(size_t i = 0; i < xv.size(); ++i) CHECK(xv[i] == 2. * bvv[i]);
  // apply: y = 2 A x + 1 y against a host product over the point-CSR debug data
  std::vector<uint32_t> rows, cols;
  std::vector<double> vals;
  A.debug_data(rows, cols, vals);
  std::vector<double> y0 = bvv;
  A.apply(b, x, 2., 1.);
  std::vector<double> y;
  b.get(y);
  for(uint32_t r = 0; r < n * neq; ++r) {
    double acc = 0.;
    for(uint32_t q = rows[r]; q < rows[r + 1]; ++q) acc += vals[q] * xv[cols[q]];
    CHECK(std::fabs(y[r] - (2. * acc + y0[r])) <= 1e-13 * (std::fabs(y[r]) + 1.));
  }
}

// symmetric positive definite chain: -x_{i-1} + 4 x_i - x_{i+1} = f, solved to 1e-10
static void solve_chain(Context& ctx)
{
  const uint32_t n = 200;
  std::vector<uint32_t> start(1, 0), conn;
  for(uint32_t i = 0; i < n; ++i) {
    for(int d = -1; d <= 1; ++d)
      if((int)i + d >= 0 && (int)i + d < (int)n) conn.push_back(i + d);
    start.push_back((uint32_t)conn.size());
  }
  lss::System s(ctx, 1, conn, start);
  s.reset(0.);
  std::vector<double> ref(n), f(n, 0.);
  for(uint32_t i = 0; i < n; ++i) ref[i] = std::sin(0.05 * i) + 0.1 * i;
  for(uint32_t i = 0; i < n; ++i) {
    s.matrix().set_value(i, i, 4.);
    f[i] += 4. * ref[i];
    if(i > 0) { s.matrix().set_value(i - 1, i, -1.); f[i] -= ref[i - 1]; }
    if(i + 1 < n) { s.matrix().set_value(i + 1, i, -1.); f[i] -= ref[i + 1]; }
  }
  s.rhs().set(f);
  s.solution().reset(0.);
  s.solution_strategy().params.tol = 1e-12;
  CHECK(s.solve());
  CHECK(s.solution_strategy().compute_residual() <= 1e-12);
  std::vector<double> xs;
  s.solution().get(xs);
  double err = 0.;
  for(uint32_t i = 0; i < n; ++i) err = std::fmax(err, std::fabs(xs[i] - ref[i]));
  CHECK(err <= 1e-9);
}

// UFEM assembly through the C++ action on a 3^3 Kuhn cube, against the oracle's FP64 assembly
static void ns_assembly(Context& ctx)
{
  const uint32_t c = 3, N = (c + 1) * (c + 1) * (c + 1), E = 6 * c * c * c;
  std::vector<double> X((size_t)N * 3);
  std::vector<uint32_t> conn((size_t)E * 4);
  orc_mesh_box_tetras(c, c, c, 1., 1., 1., X.data(), conn.data());
  uint64_t st = 12345;
  std::vector<double> u((size_t)N * 3), p(N, 0.), nu(N, 1e-2);
  for(double& x : u) x = orc_uniform(&st, -1., 1.);
  Mesh mesh(ctx, CF3HIP_TETRA3D, X, conn);
  ufem::Fields f(mesh);
  f.set(CF3HIP_FIELD_U, u);
  f.set(CF3HIP_FIELD_P, p);
  f.initialize(1e-2);
  lss::System s(ctx, mesh, 4);
  ufem::NavierStokesAssembly ns;
  ns.density = 1.;
  ns.dt = 0.01;
  ns.linearize(f);
  ns.execute(s, f);
  // oracle
  std::vector<uint32_t> start(N + 1), cols(64 * N), slots((size_t)E * 16);
  const int64_t nb = orc_build_sparsity(4, N, E, conn.data(), start.data(), cols.data(), (int64_t)cols.size(), slots.data());
  CHECK(nb > 0);
  orc_ns_params prm = {1., 0.01, 1., 1, ORC_SUPG_METRIC, 1., 16., 1., 1., 1.};
  std::vector<double> vref((size_t)nb * 16), rref((size_t)N * 4);
  CHECK(orc_ns_assemble(ORC_TETRA3D, &prm, N, E, X.data(), conn.data(), u.data(), p.data(), u.data(), nu.data(),
                        nullptr, start.data(), cols.data(), slots.data(), vref.data(), rref.data(), 1) == 0);
  std::vector<uint32_t> rows, ci;
  std::vector<double> vals;
  s.matrix().debug_data(rows, ci, vals);
  // compare through scalar access on the diagonal blocks and the full rhs
  double vmax = 0., dmax = 0.;
  for(uint32_t i = 0; i < N; ++i)
    for(uint32_t q = start[i]; q < start[i + 1]; ++q)
      for(int a = 0; a < 4; ++a)
        for(int b2 = 0; b2 < 4; ++b2) {
          double g = 0.;
          s.matrix().get_value(cols[q] * 4 + b2, i * 4 + a, g);
          const double r = vref[(size_t)q * 16 + a * 4 + b2];
          vmax = std::fmax(vmax, std::fabs(r));
          dmax = std::fmax(dmax, std::fabs(g - r));
        }
  CHECK(dmax <= 1e-12 * vmax);
  std::vector<double> rhs;
  s.rhs().get(rhs);
  double rmax = 0., rd = 0.;
  for(size_t i = 0; i < rhs.size(); ++i) {
    rmax = std::fmax(rmax, std::fabs(rref[i]));
    rd = std::fmax(rd, std::fabs(rhs[i] - rref[i]));
  }
  CHECK(rd <= 1e-12 * rmax);
  // option errors as in the reference
  ns.theta = 1.5;
  CHECK(throws<SetupError>([&] { ns.execute(s, f); }));
  ns.theta = 1.;
  ns.supg_type = "unknown";
  CHECK(throws<BadValue>([&] { ns.execute(s, f); }));
}

// periodic links (System::create with periodic_links_nodes / active): a 4 x 4 triangle mesh of the unit
// square with the right edge linked to the left one; indices are mesh nodes, converted by the host layer
static void periodic_system(Context& ctx)
{
  const uint32_t c = 4, N = (c + 1) * (c + 1), E = 2 * c * c;
  std::vector<double> X((size_t)N * 2);
  std::vector<uint32_t> conn((size_t)E * 3);
  orc_mesh_rect_tris(c, c, 1., 1., X.data(), conn.data());
  std::vector<uint32_t> links(N, 0);
  std::vector<bool> active(N, false);
  for(uint32_t i = 0; i < N; ++i)
    if(std::fabs(X[2 * i] - 1.) < 1e-12)
      for(uint32_t j = 0; j < N; ++j)
        if(std::fabs(X[2 * j]) < 1e-12 && std::fabs(X[2 * j + 1] - X[2 * i + 1]) < 1e-12) {
          links[i] = j;
          active[i] = true;
        }
  uint64_t st = 777;
  std::vector<double> u((size_t)N * 2), p(N, 0.), nu(N, 1e-2);
  for(uint32_t i = 0; i < N; ++i) {   // periodic fields: a linked node carries its master's values
    const uint32_t m = active[i] ? links[i] : i;
    if(m == i) { u[2 * i] = orc_uniform(&st, -1., 1.); u[2 * i + 1] = orc_uniform(&st, -1., 1.); }
  }
  for(uint32_t i = 0; i < N; ++i)
    if(active[i]) { u[2 * i] = u[2 * links[i]]; u[2 * i + 1] = u[2 * links[i] + 1]; }
  Mesh mesh(ctx, CF3HIP_TRIAG2D, X, conn);
  lss::System s(ctx, mesh, 3, links, active);
  CHECK(s.n_nodes() == N - (c + 1));
  uint32_t slave = 0;
  while(!active[slave]) ++slave;
  const uint32_t master = links[slave];
  CHECK(s.row(slave) == s.row(master));
  // a value set through the linked node is the master's
  s.rhs().reset(0.);
  s.rhs().set_value(slave, 1, 2.5);
  double v = 0.;
  s.rhs().get_value(master, 1, v);
  CHECK(v == 2.5);
  // UFEM assembly through the C++ action against the oracle's assembly through the same map
  ufem::Fields f(mesh);
  f.set(CF3HIP_FIELD_U, u);
  f.set(CF3HIP_FIELD_P, p);
  f.initialize(1e-2);
  ufem::NavierStokesAssembly ns;
  ns.density = 1.;
  ns.dt = 0.01;
  ns.use_specializations = false;
  ns.linearize(f);
  ns.execute(s, f);
  std::vector<uint32_t> p2m(N), dconn(conn.size());
  for(uint32_t i = 0; i < N; ++i) p2m[i] = s.row(i);
  for(size_t i = 0; i < conn.size(); ++i) dconn[i] = p2m[conn[i]];
  const uint32_t R = s.n_nodes();
  std::vector<uint32_t> start(R + 1), cols(64 * R), slots((size_t)E * 9);
  const int64_t nb = orc_build_sparsity(3, R, E, dconn.data(), start.data(), cols.data(), (int64_t)cols.size(), slots.data());
  CHECK(nb > 0);
  orc_ns_params prm = {1., 0.01, 1., 0, ORC_SUPG_METRIC, 1., 16., 1., 1., 1.};
  std::vector<double> vref((size_t)nb * 9), rref((size_t)R * 3);
  CHECK(orc_ns_assemble_p2m(ORC_TRIAG2D, &prm, R, E, X.data(), conn.data(), p2m.data(), u.data(), p.data(), u.data(),
                            nu.data(), nullptr, start.data(), slots.data(), vref.data(), rref.data()) == 0);
  // compare through mesh-node scalar access: the free nodes are their own rows' representatives
  std::vector<uint32_t> rep(R);
  for(uint32_t i = 0; i < N; ++i)
    if(!active[i]) rep[p2m[i]] = i;
  double vmax = 0., dmax = 0.;
  for(uint32_t r = 0; r < R; ++r)
    for(uint32_t q = start[r]; q < start[r + 1]; ++q)
      for(int a = 0; a < 3; ++a)
        for(int b2 = 0; b2 < 3; ++b2) {
          double g = 0.;
          s.matrix().get_value(rep[cols[q]] * 3 + b2, rep[r] * 3 + a, g);
          const double ref = vref[(size_t)q * 9 + a * 3 + b2];
          vmax = std::fmax(vmax, std::fabs(ref));
          dmax = std::fmax(dmax, std::fabs(g - ref));
        }
  CHECK(dmax <= 1e-12 * vmax);
  std::vector<double> rhs;
  s.rhs().get(rhs);
  double rmax = 0., rd = 0.;
  for(size_t i = 0; i < rhs.size(); ++i) {
    rmax = std::fmax(rmax, std::fabs(rref[i]));
    rd = std::fmax(rd, std::fabs(rhs[i] - rref[i]));
  }
  CHECK(rd <= 1e-12 * rmax);
  // cycles of links are refused
  std::vector<uint32_t> bad(N, 0);
  std::vector<bool> all(N, true);
  bad[0] = 1;
  CHECK(throws<BadValue>([&] { lss::System t(ctx, mesh, 3, bad, all); }));
}

// NavierStokesSemiImplicit through the host layer: a 6 x 6 lid-driven cavity of quadrilaterals; the
// pressure matrix and, after a step, the velocity matrix against the oracle's element actions; the
// boundary velocities and the pinned pressure take their imposed values
static void semi_implicit(Context& ctx)
{
  const uint32_t c = 6, N = (c + 1) * (c + 1), E = c * c;
  std::vector<double> X((size_t)N * 2);
  std::vector<uint32_t> conn((size_t)E * 4);
  for(uint32_t j = 0; j <= c; ++j)
    for(uint32_t i = 0; i <= c; ++i) { X[2 * (j * (c + 1) + i)] = double(i) / c; X[2 * (j * (c + 1) + i) + 1] = double(j) / c; }
  for(uint32_t j = 0; j < c; ++j)
    for(uint32_t i = 0; i < c; ++i) {
      const uint32_t n0 = j * (c + 1) + i, e = j * c + i;
      conn[4 * e] = n0; conn[4 * e + 1] = n0 + 1; conn[4 * e + 2] = n0 + c + 2; conn[4 * e + 3] = n0 + c + 1;
    }
  uint64_t st = 777;
  std::vector<double> u((size_t)N * 2), p(N, 0.), nu(N, 1e-2);
  for(double& x : u) x = orc_uniform(&st, -0.2, 0.2);
  Mesh mesh(ctx, CF3HIP_QUAD2D, X, conn);
  ufem::Fields f(mesh);
  f.set(CF3HIP_FIELD_U, u);
  f.set(CF3HIP_FIELD_P, p);
  f.initialize(1e-2);
  lss::System us(ctx, mesh, 2), ps(ctx, mesh, 1);
  us.solution_strategy().params.tol = 1e-13;
  ps.solution_strategy().params.tol = 1e-13;
  ufem::NavierStokesSemiImplicit si(us, ps);
  si.dt = 0.01;
  si.initial(f);
  std::vector<uint32_t> start(N + 1), cols(16 * N), slots((size_t)E * 16);
  const int64_t nb = orc_build_sparsity(4, N, E, conn.data(), start.data(), cols.data(), (int64_t)cols.size(), slots.data());
  CHECK(nb > 0);
  orc_ns_params prm = {1., 0.01, 1., 0, ORC_SUPG_METRIC, 1., 16., 1., 1., 1.};
  std::vector<double> pref(nb, 0.), pdev(nb);
  CHECK(orc_si_pressure_matrix(ORC_QUAD2D, &prm, E, X.data(), conn.data(), u.data(), nu.data(), slots.data(),
                               pref.data()) == 0);
  check(cf3hip_matrix_get_values(ps.handle(), pdev.data()));
  double vmax = 0., dmax = 0.;
  for(int64_t q = 0; q < nb; ++q) { vmax = std::fmax(vmax, std::fabs(pref[q])); dmax = std::fmax(dmax, std::fabs(pdev[q] - pref[q])); }
  CHECK(dmax <= 1e-12 * vmax);
  // walls (no slip) and the lid (u = 1, taking the top corners); p = 0 at node 0
  std::vector<uint32_t> un, pn = {0};
  std::vector<double> uv, pv = {0.};
  for(uint32_t n = 0; n < N; ++n) {
    const uint32_t i = n % (c + 1), j = n / (c + 1);
    if(j == c) { un.push_back(n); uv.push_back(1.); uv.push_back(0.); }
    else if(i == 0 || i == c || j == 0) { un.push_back(n); uv.push_back(0.); uv.push_back(0.); }
  }
  CHECK(si.step(f, un, uv, pn, pv));
  CHECK(si.stats.size() == 4 && si.stats[1].iterations > 0);
  std::vector<double> unew, pnew, uadv;
  f.get(CF3HIP_FIELD_U, unew, N, 2);
  f.get(CF3HIP_FIELD_P, pnew, N, 1);
  f.get(CF3HIP_FIELD_UADV, uadv, N, 2);
  double bmax = 0.;
  for(size_t b = 0; b < un.size(); ++b)
    for(int d = 0; d < 2; ++d) bmax = std::fmax(bmax, std::fabs(unew[2 * un[b] + d] - uv[2 * b + d]));
  CHECK(bmax <= 1e-12 && std::fabs(pnew[0]) <= 1e-12);
  // the velocity matrix of the step, before its conditions, from the same advection velocity
  const cf3hip_ns_params sip = si.params();
  check(cf3hip_si_velocity_assemble(si.handle(), f.handle(), &sip));
  std::vector<double> vref((size_t)nb * 4, 0.), vdev((size_t)nb * 4);
  CHECK(orc_si_velocity_matrix(ORC_QUAD2D, &prm, E, X.data(), conn.data(), uadv.data(), nu.data(), slots.data(),
                               vref.data()) == 0);
  check(cf3hip_matrix_get_values(us.handle(), vdev.data()));
  vmax = dmax = 0.;
  for(size_t q = 0; q < vref.size(); ++q) { vmax = std::fmax(vmax, std::fabs(vref[q])); dmax = std::fmax(dmax, std::fabs(vdev[q] - vref[q])); }
  CHECK(dmax <= 1e-12 * vmax);
  si.theta = 2.;
  CHECK(throws<SetupError>([&] { si.initial(f); }));
}

int main(int argc, char** argv)
{
  if(argc > 1 && std::strcmp(argv[1], "--no-device") == 0) {
    // without a GPU: creating a context fails with a mapped exception, never a crash
    bool mapped = false;
    try {
      Context ctx(0);
    } catch(const Error& e) {
      mapped = e.code != 0;
    }
    CHECK(mapped);
    CHECK(throws<SetupError>([] { check(CF3HIP_ERR_NOT_CREATED); }));
    CHECK(throws<BadValue>([] { check(CF3HIP_ERR_ILLEGAL_ENTRY); }));
    CHECK(throws<NotImplemented>([] { check(CF3HIP_ERR_NOT_IMPLEMENTED); }));
    CHECK(throws<FailedAssertion>([] { check(CF3HIP_ERR_HIP); }));
  } else {
    Context ctx(0);
    std::printf("symmetric_dirichlet\n");
    std::fflush(stdout);
    symmetric_dirichlet(ctx, "cf3.math.LSS.TrilinosFEVbrMatrix");
    symmetric_dirichlet(ctx, "cf3.math.LSS.TrilinosCrsMatrix");
    std::printf("matrix_builders\n");
    std::fflush(stdout);
    matrix_builders(ctx);
    std::printf("periodicity\n");
    std::fflush(stdout);
    periodicity(ctx);
    std::printf("atomic_operations\n");
    std::fflush(stdout);
    atomic_operations(ctx);
    std::printf("solve_chain\n");
    std::fflush(stdout);
    solve_chain(ctx);
    std::printf("ns_assembly\n");
    std::fflush(stdout);
    ns_assembly(ctx);
    std::printf("periodic_system\n");
    std::fflush(stdout);
    periodic_system(ctx);
    std::printf("semi_implicit\n");
    std::fflush(stdout);
    semi_implicit(ctx);
  }
  std::printf("%s: %d failed checks\n", g_fail ? "FAIL" : "OK", g_fail);
  return g_fail ? 1 : 0;
}

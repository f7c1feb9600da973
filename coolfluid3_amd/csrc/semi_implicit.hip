// UFEM NavierStokesSemiImplicit on the device (gfx950): the element actions of
// plugins/UFEM/src/UFEM/ns_semi_implicit/MatrixAssembly.hpp and PressureSystem.cpp, and the time step of
// ns_semi_implicit/implementation/NavierStokesSemiImplicit.cpp (LinearizeU, velocity assembly,
// SetSolution, InnerLoop, Update) on a velocity system (neq = dim) and a pressure system (neq = 1)
// that share one mesh and its periodic links.  Triangles, quadrilaterals, tetrahedra, hexahedra and
// prisms (ns_semi_implicit/implementation/{Triags,Quads,Tetras,Hexas,Prisms}MatrixAssembly.cpp).
//
// Each element action runs in two passes.  k_si_elem computes one element per thread and writes its
// element vector or matrix to a staging buffer; k_si_gather then gives every matrix block slot (or
// vector row) one thread that adds its contributions in element order, so the sums are deterministic
// and need no atomics.  The contribution lists are the element-to-slot map of the velocity system (the
// two systems have the same node graph) and the rows p2m[conn], sorted stably once at set-up.
#include "assembly_common.hpp"

#include <algorithm>
#include <vector>

namespace cf3hip {

namespace {

#pragma clang fp contract(off)

// ------------------------------------------------------------------------------------------------
// element types: shape functions, Gauss rules (cf3/mesh/LagrangeP1/Triag2D.cpp, Quad2D.cpp, Tetra3D.cpp,
// Hexa3D.cpp, Prism.cpp; cf3/mesh/Integrators/GaussImplementation.hpp), IntegralOrder
// (MatrixAssembly.hpp:28-47: 1 for triangles and tetrahedra, 2 otherwise)
// ------------------------------------------------------------------------------------------------
struct Tri {
  static constexpr int K = 3, D = 2, ORDER = 1;
  __device__ static int npts(int order) { return order == 1 ? 1 : 3; }
  __device__ static void point(int order, int q, double xi[2], double& w)
  {
    // order 1: centroid, w 1/2; order 2: the edge midpoints (1/2, 0), (1/2, 1/2), (0, 1/2), w 1/6
    if(order == 1) { xi[0] = xi[1] = 0.3333333333333333333333333; w = 0.5; return; }
    xi[0] = q == 2 ? 0.0 : 0.5;
    xi[1] = q == 0 ? 0.0 : 0.5;
    w = 1.0 / 6.0;
  }
  __device__ static void shape(const double xi[2], double N[3], double dN[2][3])
  {
    N[0] = 1. - xi[0] - xi[1]; N[1] = xi[0]; N[2] = xi[1];
    dN[0][0] = -1.; dN[0][1] = 1.; dN[0][2] = 0.;
    dN[1][0] = -1.; dN[1][1] = 0.; dN[1][2] = 1.;
  }
};

struct Quad {
  static constexpr int K = 4, D = 2, ORDER = 2;
  __device__ static int npts(int order) { return order == 1 ? 1 : 4; }
  __device__ static void point(int order, int q, double xi[2], double& w)
  {
    // order 1: centre, w 4; order 2: (x,x), (-x,x), (x,-x), (-x,-x), w 1
    if(order == 1) { xi[0] = xi[1] = 0.; w = 4.; return; }
    const double x = 0.5773502691896257645091488;
    xi[0] = (q & 1) ? -x : x;
    xi[1] = (q & 2) ? -x : x;
    w = 1.;
  }
  __device__ static void shape(const double xi[2], double N[4], double dN[2][4])
  {
    const double s = xi[0], t = xi[1];
    N[0] = 0.25 * (1. - s) * (1. - t); N[1] = 0.25 * (1. + s) * (1. - t);
    N[2] = 0.25 * (1. + s) * (1. + t); N[3] = 0.25 * (1. - s) * (1. + t);
    dN[0][0] = 0.25 * (-1. + t); dN[0][1] = 0.25 * (1. - t); dN[0][2] = 0.25 * (1. + t); dN[0][3] = 0.25 * (-1. - t);
    dN[1][0] = 0.25 * (-1. + s); dN[1][1] = 0.25 * (-1. - s); dN[1][2] = 0.25 * (1. + s); dN[1][3] = 0.25 * (1. - s);
  }
};

struct Tet {
  static constexpr int K = 4, D = 3, ORDER = 1;
  __device__ static int npts(int order) { return order == 1 ? 1 : 4; }
  __device__ static void point(int order, int q, double xi[3], double& w)
  {
    // order 1: centroid, w 1/6; order 2: (mu,mu,mu) and its three permutations with 1 - 3 mu, w 1/24
    if(order == 1) { xi[0] = xi[1] = xi[2] = 0.25; w = 0.1666666666666666666666667; return; }
    const double mu = 0.13819660112501051517954131656;
    xi[0] = q == 1 ? 1. - 3. * mu : mu;
    xi[1] = q == 2 ? 1. - 3. * mu : mu;
    xi[2] = q == 3 ? 1. - 3. * mu : mu;
    w = 0.04166666666666666666666666666666666667;
  }
  __device__ static void shape(const double xi[3], double N[4], double dN[3][4])
  {
    N[0] = 1. - xi[0] - xi[1] - xi[2]; N[1] = xi[0]; N[2] = xi[1]; N[3] = xi[2];
    for(int l = 0; l < 3; ++l)
      for(int m = 0; m < 4; ++m) dN[l][m] = m == 0 ? -1. : (m == l + 1 ? 1. : 0.);
  }
};

struct Hex {
  static constexpr int K = 8, D = 3, ORDER = 2;
  __device__ static int npts(int order) { return order == 1 ? 1 : 8; }
  __device__ static void point(int order, int q, double xi[3], double& w)
  {
    // order 1: centre, w 8; order 2: the eight points in the order of GaussImplementation.hpp:615-676
    if(order == 1) { xi[0] = xi[1] = xi[2] = 0.; w = 8.; return; }
    const double x = 0.5773502691896257645091488;
    const signed char s[8][3] = {{1,1,1},{-1,1,1},{1,-1,1},{1,1,-1},{1,-1,-1},{-1,1,-1},{-1,-1,1},{-1,-1,-1}};
    for(int d = 0; d < 3; ++d) xi[d] = s[q][d] * x;
    w = 1.;
  }
  __device__ static void shape(const double xi[3], double N[8], double dN[3][8])
  {
    // node m at (sx, sy, sz): (-,-,-), (+,-,-), (+,+,-), (-,+,-), then the same with z = +1
    for(int m = 0; m < 8; ++m) {
      const double sx = ((m ^ (m >> 1)) & 1) ? 1. : -1., sy = (m & 2) ? 1. : -1., sz = (m & 4) ? 1. : -1.;
      const double a = 1. + sx * xi[0], b = 1. + sy * xi[1], c = 1. + sz * xi[2];
      N[m] = 0.125 * a * b * c;
      dN[0][m] = 0.125 * sx * b * c;
      dN[1][m] = 0.125 * sy * a * c;
      dN[2][m] = 0.125 * sz * a * b;
    }
  }
};

struct Prism {
  static constexpr int K = 6, D = 3, ORDER = 2;
  __device__ static int npts(int order) { return order == 1 ? 1 : 6; }
  __device__ static void point(int order, int q, double xi[3], double& w)
  {
    // order 1: (1/3, 1/3, 0), w 1; order 2: the triangle's edge midpoints (1/2, 0), (1/2, 1/2), (0, 1/2)
    // at zeta = +1/sqrt(3), then at zeta = -1/sqrt(3), w 1/6 (GaussImplementation.hpp:430-494)
    if(order == 1) { xi[0] = xi[1] = 0.3333333333333333333333333; xi[2] = 0.; w = 1.; return; }
    const int t = q < 3 ? q : q - 3;
    xi[0] = t == 2 ? 0. : 0.5;
    xi[1] = t == 0 ? 0. : 0.5;
    xi[2] = q < 3 ? 0.5773502691896257645091488 : -0.5773502691896257645091488;
    w = 1. / 6.;
  }
  __device__ static void shape(const double xi[3], double N[6], double dN[3][6])
  {
    // the triangle (0,0), (1,0), (0,1) at zeta = -1 (nodes 0-2), then at zeta = +1 (Prism.cpp:24-58)
    const double lo = 0.5 * (1. - xi[2]), hi = 0.5 * (1. + xi[2]);
    const double t[3] = {1. - xi[0] - xi[1], xi[0], xi[1]};
    const double tx[3] = {-1., 1., 0.}, ty[3] = {-1., 0., 1.};
    for(int m = 0; m < 3; ++m) {
      N[m] = lo * t[m]; N[m + 3] = hi * t[m];
      dN[0][m] = lo * tx[m]; dN[0][m + 3] = hi * tx[m];
      dN[1][m] = lo * ty[m]; dN[1][m + 3] = hi * ty[m];
      dN[2][m] = -0.5 * t[m]; dN[2][m + 3] = 0.5 * t[m];
    }
  }
};

// inverse Jacobian (J(l, d) = dx_d / dxi_l) and det J at xi
template<class E>
__device__ __forceinline__ void inverse_jacobian(const double X[E::K][E::D], const double xi[E::D], double N[E::K],
                                                 double dN[E::D][E::K], double Ji[E::D][E::D], double& det)
{
  constexpr int D = E::D;
  E::shape(xi, N, dN);
  double J[D][D];
  for(int l = 0; l < D; ++l)
    for(int d = 0; d < D; ++d) {
      double s = 0.;
      for(int m = 0; m < E::K; ++m) s += dN[l][m] * X[m][d];
      J[l][d] = s;
    }
  if constexpr(D == 2) {
    det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    const double inv = 1. / det;
    Ji[0][0] = J[1][1] * inv; Ji[0][1] = -J[0][1] * inv;
    Ji[1][0] = -J[1][0] * inv; Ji[1][1] = J[0][0] * inv;
  } else {
    const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    const double inv = 1. / det;
    Ji[0][0] = c00 * inv;
    Ji[1][0] = c01 * inv;
    Ji[2][0] = c02 * inv;
    Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * inv;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * inv;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * inv;
    Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * inv;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * inv;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * inv;
  }
}

// N, physical gradients G = J^-1 dN and det J at xi
template<class E>
__device__ __forceinline__ void point_data(const double X[E::K][E::D], const double xi[E::D], double N[E::K],
                                           double G[E::D][E::K], double& det)
{
  constexpr int D = E::D;
  double dN[D][E::K], Ji[D][D];
  inverse_jacobian<E>(X, xi, N, dN, Ji, det);
  for(int d = 0; d < D; ++d)
    for(int m = 0; m < E::K; ++m) {
      double s = 0.;
      for(int l = 0; l < D; ++l) s += Ji[d][l] * dN[l][m];
      G[d][m] = s;
    }
}

struct TauP { int type; double dt, c1, c2, a_ps, a_su, a_bu; };

// ComputeTauImpl (SUPG.hpp:107-219), METRIC or TEZDUYAR, at the element centre with nodal velocity u
// and element viscosity |mean(nu)|
template<class E>
__device__ void tau(const TauP& t, const double X[E::K][E::D], const double u[E::K][E::D], const double nu[E::K],
                    double& tau_ps, double& tau_su, double& tau_bulk)
{
  constexpr int D = E::D, K = E::K;
  double s = 0.;
  for(int m = 0; m < K; ++m) s += nu[m];
  const double enu = fabs(s / K);
  double xi[D], w, N[K], dN[D][K], Ji[D][D], det;
  E::point(1, 0, xi, w);
  inverse_jacobian<E>(X, xi, N, dN, Ji, det);
  double uc[D];
  for(int d = 0; d < D; ++d) {
    double v = 0.;
    for(int m = 0; m < K; ++m) v += N[m] * u[m][d];
    uc[d] = v;
  }
  if(t.type == CF3HIP_SUPG_TEZDUYAR) {
    double h = 1e10;
    for(int i = 0; i < K; ++i)
      for(int j = 0; j < K; ++j)
        if(i != j) {
          double h2 = 0.;
          for(int d = 0; d < D; ++d) { const double df = X[i][d] - X[j][d]; h2 += df * df; }
          h = fmin(h, h2);
        }
    h = sqrt(h);
    double um2 = 0.;
    for(int d = 0; d < D; ++d) um2 += uc[d] * uc[d];
    const double umag = sqrt(um2);
    tau_su = 1. / ((2. * umag) / h + 2. / t.dt + (4. * enu) / (h * h));
    tau_ps = tau_su;
    tau_bulk = tau_su * umag * umag;
  } else {
    double g[D][D];
    for(int i = 0; i < D; ++i)
      for(int j = 0; j < D; ++j) {
        double v = 0.;
        for(int l = 0; l < D; ++l) v += Ji[l][i] * Ji[l][j];
        g[i][j] = v;
      }
    double ugu = 0., gsq = 0., tr = 0.;
    for(int j = 0; j < D; ++j) {
      double v = 0.;
      for(int i = 0; i < D; ++i) v += uc[i] * g[i][j];
      ugu += v * uc[j];
      tr += g[j][j];
      for(int i = 0; i < D; ++i) gsq += g[i][j] * g[i][j];
    }
    const double adv2 = fabs(ugu);
    const double diff = enu * enu * gsq;
    tau_su = 1. / sqrt((4. * t.c1 * t.c1 / (t.dt * t.dt)) + adv2 + t.c2 * diff);
    tau_ps = tau_su;
    const double su_std = 1. / sqrt((4. / (t.dt * t.dt)) + adv2 + 16. * diff);
    tau_bulk = (1. / su_std) / tr;
  }
  tau_ps *= t.a_ps;
  tau_su *= t.a_su;
  tau_bulk *= t.a_bu;
}

enum { OP_MASS = 0, OP_PMAT = 1, OP_VMAT = 2, OP_URHS = 3, OP_PRHS = 4, OP_AUP = 5 };

// element outputs per element: MASS K, PMAT K*K, VMAT K*K*D*D (block (a,b) row-major D x D),
// URHS/AUP K*D (node-major), PRHS K
struct SiArgs {
  uint32_t n_elems;
  const uint32_t* conn;
  const uint32_t* p2m;
  const double* x[3];
  const double* u[3];     // velocity field (tau of PRHS and of the pressure matrix)
  const double* ua[3];    // advection velocity
  const double* nu;
  const double* a;        // LSS vectors (rows x D / rows)
  const double* U;
  const double* da;
  const double* P;
  const double* dps;
  const double* dp;
  const double* gf[3];     // body force (enable_body_force), null otherwise
  double theta;
  TauP t;
  double* stage;
};

template<class E>
__device__ __forceinline__ void interp(const double N[E::K], const double f[E::K][E::D], double out[E::D])
{
  for(int d = 0; d < E::D; ++d) {
    double v = 0.;
    for(int m = 0; m < E::K; ++m) v += N[m] * f[m][d];
    out[d] = v;
  }
}

template<class E, int OP>
__global__ void __launch_bounds__(256) k_si_elem(SiArgs g)
{
  constexpr int K = E::K, D = E::D;
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if(e >= g.n_elems) return;
  uint32_t cn[K], rw[K];
  double X[K][D], ua[K][D], uf[K][D], nu[K];
  for(int m = 0; m < K; ++m) {
    cn[m] = g.conn[(uint64_t)e * K + m];
    rw[m] = g.p2m[cn[m]];
    for(int d = 0; d < D; ++d) {
      X[m][d] = g.x[d][cn[m]];
      ua[m][d] = g.ua[d][cn[m]];
      uf[m][d] = g.u[d][cn[m]];
    }
    nu[m] = g.nu[cn[m]];
  }
  const double dt = g.t.dt;
  double tau_ps = 0., tau_su = 0., tau_bulk = 0.;
  if(OP == OP_PMAT || OP == OP_PRHS) tau<E>(g.t, X, uf, nu, tau_ps, tau_su, tau_bulk);
  else if(OP != OP_MASS) tau<E>(g.t, X, ua, nu, tau_ps, tau_su, tau_bulk);
  double xi[D], w, N[K], G[D][K], det;

  if constexpr(OP == OP_MASS || OP == OP_PMAT) {
    // MassMatrixAssembly: row sums of int N^T N; assemble_pp: theta (tau_ps + dt) int G^T G (order 2)
    double A[K][K];
    for(int a = 0; a < K; ++a)
      for(int b = 0; b < K; ++b) A[a][b] = 0.;
    for(int q = 0; q < E::npts(2); ++q) {
      E::point(2, q, xi, w);
      point_data<E>(X, xi, N, G, det);
      const double wq = w * det;
      for(int a = 0; a < K; ++a)
        for(int b = 0; b < K; ++b) {
          if(OP == OP_MASS) {
            A[a][b] += (wq * N[a]) * N[b];
          } else {
            double gg = 0.;
            for(int d = 0; d < D; ++d) gg += G[d][a] * G[d][b];
            A[a][b] += wq * gg;
          }
        }
    }
    double* out = g.stage + (uint64_t)e * (OP == OP_MASS ? K : K * K);
    if(OP == OP_MASS) {
      for(int a = 0; a < K; ++a) {
        double s = 0.;
        for(int b = 0; b < K; ++b) s += A[a][b];
        out[a] = s;
      }
    } else {
      const double c = g.theta * (tau_ps + dt);
      for(int a = 0; a < K; ++a)
        for(int b = 0; b < K; ++b) out[a * K + b] = c * A[a][b];
    }
  } else if constexpr(OP == OP_VMAT) {
    // VelocityAssembly (MatrixAssembly.hpp:49-146): T + theta dt M
    double T[K][K], L[K][K], B[D][K][K];
    for(int a = 0; a < K; ++a)
      for(int b = 0; b < K; ++b) {
        T[a][b] = 0.; L[a][b] = 0.;
        for(int d = 0; d < D; ++d) B[d][a][b] = 0.;
      }
    for(int q = 0; q < E::npts(2); ++q) {   // mass with the second-order rule
      E::point(2, q, xi, w);
      point_data<E>(X, xi, N, G, det);
      const double wq = w * det;
      for(int a = 0; a < K; ++a)
        for(int b = 0; b < K; ++b) T[a][b] += (wq * N[a]) * N[b];
    }
    for(int q = 0; q < E::npts(E::ORDER); ++q) {   // viscous and bulk diagonal blocks
      E::point(E::ORDER, q, xi, w);
      point_data<E>(X, xi, N, G, det);
      double nuq = 0.;
      for(int m = 0; m < K; ++m) nuq += N[m] * nu[m];
      const double wq = w * det, bulk = wq * tau_bulk, wn = wq * nuq;
      for(int a = 0; a < K; ++a)
        for(int b = 0; b < K; ++b) {
          double gg = 0.;
          for(int d = 0; d < D; ++d) {
            gg += G[d][a] * G[d][b];
            B[d][a][b] += (bulk * G[d][a]) * G[d][b];
          }
          L[a][b] += wn * gg;
        }
    }
    E::point(1, 0, xi, w);   // off-diagonal bulk blocks with the one-point rule
    point_data<E>(X, xi, N, G, det);
    const double bulk = w * det * tau_bulk, tdt = g.theta * dt;
    double* out = g.stage + (uint64_t)e * K * K * D * D;
    for(int a = 0; a < K; ++a)
      for(int b = 0; b < K; ++b) {
        double* o = out + (a * K + b) * D * D;
        for(int r = 0; r < D; ++r)
          for(int c = 0; c < D; ++c)
            o[r * D + c] = r == c ? T[a][b] + tdt * (L[a][b] + B[r][a][b]) : tdt * ((bulk * G[r][a]) * G[c][b]);
      }
  } else if constexpr(OP == OP_URHS) {
    // VelocityRHS (MatrixAssembly.hpp:206-330) with a, dt (1 - theta) a - u_vec, (1 - theta) dps - p_vec
    double av[K][D], dtamu[K][D], ppdp[K];
    for(int m = 0; m < K; ++m) {
      const uint64_t r = rw[m];
      for(int d = 0; d < D; ++d) {
        av[m][d] = g.a[r * D + d];
        dtamu[m][d] = dt * (1. - g.theta) * av[m][d] - g.U[r * D + d];
      }
      ppdp[m] = (1. - g.theta) * g.dps[r] - g.P[r];
    }
    double res[D][K];
    for(int d = 0; d < D; ++d)
      for(int a = 0; a < K; ++a) res[d][a] = 0.;
    // simplices: pass 0 = mass-type terms with the second-order rule, pass 1 = the rest (order 1);
    // Q1 elements: all terms in one second-order loop
    for(int pass = 0; pass < (E::ORDER == 2 ? 1 : 2); ++pass) {
      const bool mass_terms = E::ORDER == 2 || pass == 0;
      const bool other_terms = E::ORDER == 2 || pass == 1;
      const int order = (E::ORDER == 2 || pass == 0) ? 2 : 1;
      for(int q = 0; q < E::npts(order); ++q) {
        E::point(order, q, xi, w);
        point_data<E>(X, xi, N, G, det);
        const double wq = w * det;
        double uq[D];
        interp<E>(N, ua, uq);
        double adv[K], npa[K];
        for(int m = 0; m < K; ++m) {
          double v = 0.;
          for(int d = 0; d < D; ++d) v += uq[d] * G[d][m];
          adv[m] = v;
          npa[m] = tau_su * v + N[m];
        }
        double f = 0.;
        for(int i = 0; i < D; ++i)
          for(int m = 0; m < K; ++m) f += G[i][m] * dtamu[m][i];
        f *= wq;
        double nuq = 0., np = 0.;
        for(int m = 0; m < K; ++m) { nuq += N[m] * nu[m]; np += N[m] * ppdp[m]; }
        const double b = -wq * np, f_bulk = tau_bulk * f, f_skew = 0.5 * f;
        for(int i = 0; i < D; ++i) {
          double na = 0., ad = 0., gp = 0., dv[D];
          for(int l = 0; l < D; ++l) dv[l] = 0.;
          for(int m = 0; m < K; ++m) {
            na += N[m] * av[m][i];
            ad += adv[m] * dtamu[m][i];
            gp += G[i][m] * ppdp[m];
            for(int l = 0; l < D; ++l) dv[l] += G[l][m] * dtamu[m][i];
          }
          const double c = -wq * na, ee = wq * ad, a_ = (wq * tau_su) * gp, wv = wq * nuq;
          for(int a = 0; a < K; ++a) {
            double v = 0.;
            if(other_terms) {
              double gd = 0.;
              for(int l = 0; l < D; ++l) gd += G[l][a] * (wv * dv[l]);
              v += adv[a] * a_ + G[i][a] * (b + f_bulk) + gd;
            }
            if(mass_terms) v += npa[a] * (c + ee + f_skew * uq[i]);
            res[i][a] += v;
          }
        }
      }
    }
    if(g.gf[0]) {
      // the body-force overload (MatrixAssembly.hpp:339-363): N^T g_i(x_q) w_q over the second-order rule,
      // added after the terms above
      double gn[K][D];
      for(int m = 0; m < K; ++m)
        for(int d = 0; d < D; ++d) gn[m][d] = g.gf[d][cn[m]];
      for(int q = 0; q < E::npts(2); ++q) {
        E::point(2, q, xi, w);
        point_data<E>(X, xi, N, G, det);
        const double wq = w * det;
        double gq[D];
        interp<E>(N, gn, gq);
        for(int i = 0; i < D; ++i)
          for(int a = 0; a < K; ++a) res[i][a] += N[a] * gq[i] * wq;
      }
    }
    double* out = g.stage + (uint64_t)e * K * D;
    for(int a = 0; a < K; ++a)
      for(int d = 0; d < D; ++d) out[a * D + d] = res[d][a];
  } else if constexpr(OP == OP_PRHS) {
    // PressureRHS (MatrixAssembly.hpp:148-204)
    double upda[K][D], apda[K][D], pv[K];
    for(int m = 0; m < K; ++m) {
      const uint64_t r = rw[m];
      for(int d = 0; d < D; ++d) {
        upda[m][d] = g.U[r * D + d] + dt * g.da[r * D + d];
        apda[m][d] = g.a[r * D + d] + g.da[r * D + d];
      }
      pv[m] = g.P[r];
    }
    double res[K], App[K][K];
    for(int a = 0; a < K; ++a) {
      res[a] = 0.;
      for(int b = 0; b < K; ++b) App[a][b] = 0.;
    }
    for(int q = 0; q < E::npts(E::ORDER); ++q) {
      E::point(E::ORDER, q, xi, w);
      point_data<E>(X, xi, N, G, det);
      const double wq = w * det, tw = wq * tau_ps;
      double uq[D];
      interp<E>(N, ua, uq);
      double adv[K];
      for(int m = 0; m < K; ++m) {
        double v = 0.;
        for(int d = 0; d < D; ++d) v += uq[d] * G[d][m];
        adv[m] = tau_ps * v;
      }
      for(int a = 0; a < K; ++a)
        for(int b = 0; b < K; ++b) {
          double gg = 0.;
          for(int d = 0; d < D; ++d) gg += G[d][a] * G[d][b];
          App[a][b] += tw * gg;
        }
      for(int i = 0; i < D; ++i) {
        double ga = 0., av = 0., na = 0.;
        for(int m = 0; m < K; ++m) { ga += G[i][m] * upda[m][i]; av += adv[m] * upda[m][i]; na += N[m] * apda[m][i]; }
        const double a_ = wq * ga, b = wq * av, c = tw * na;
        for(int m = 0; m < K; ++m) res[m] -= a_ * (N[m] + adv[m] * 0.5) + (b + c) * G[i][m];
      }
    }
    double* out = g.stage + (uint64_t)e * K;
    for(int a = 0; a < K; ++a) {
      double s = 0.;
      for(int b = 0; b < K; ++b) s += App[a][b] * pv[b];
      out[a] = res[a] - s;
    }
  } else {
    // ApplyAup (MatrixAssembly.hpp:332-376)
    double dp[K];
    for(int m = 0; m < K; ++m) dp[m] = g.dp[rw[m]];
    double res[D][K];
    for(int d = 0; d < D; ++d)
      for(int a = 0; a < K; ++a) res[d][a] = 0.;
    for(int q = 0; q < E::npts(E::ORDER); ++q) {
      E::point(E::ORDER, q, xi, w);
      point_data<E>(X, xi, N, G, det);
      const double wq = g.theta * w * det, ws = wq * tau_su;
      double uq[D];
      interp<E>(N, ua, uq);
      double nd = 0.;
      for(int m = 0; m < K; ++m) nd += N[m] * dp[m];
      const double b = -wq * nd;
      for(int i = 0; i < D; ++i) {
        double a_ = 0.;
        for(int m = 0; m < K; ++m) a_ += G[i][m] * dp[m];
        for(int m = 0; m < K; ++m) {
          double v = 0.;
          for(int d = 0; d < D; ++d) v += uq[d] * G[d][m];
          res[i][m] += a_ * (ws * v) + b * G[i][m];
        }
      }
    }
    double* out = g.stage + (uint64_t)e * K * D;
    for(int a = 0; a < K; ++a)
      for(int d = 0; d < D; ++d) out[a * D + d] = res[d][a];
  }
}

// VelocityAssembly on Q1 cells without the element staging buffer: the row-owner form of the velocity
// system's own scatter maps (one staging pass, K lanes per incidence, as in the Q1 Navier-Stokes kernel).
// A pre-pass (k_si_ctr_q1, one thread per element) writes each element's tau_bulk, w det J tau_bulk at
// the centre and the gradients there (the one-point rule of the off-diagonal bulk blocks).  In the
// row-owner kernel lane q of an incidence (element e, local node a) computes the data of Gauss point q of
// the second-order rule (w det J, w det J nu_q, N and physical gradients at q) into LDS; lane b then forms
// the block (a, b) from those records with the arithmetic of k_si_elem's OP_VMAT, point by point in the
// same order.  The blocks are summed in LDS by the block's accumulator threads in a fixed order and every
// block of the workgroup's rows is written once: no atomics.
struct SiMaps {
  const uint32_t* wg_rows;
  const uint32_t* row_ptr;
  const uint32_t* inc_ptr;
  const uint32_t* inc;
  const uint8_t* inc_pos;
  const uint8_t* blk_beg;
  const uint16_t* blk_acc;
  double* values;
};

template<class E>
__global__ __launch_bounds__(256) void k_si_ctr_q1(SiArgs g, double* __restrict__ ctr)
{
  constexpr int K = E::K, D = E::D, kC = 2 + D * K;
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if(e >= g.n_elems) return;
  double X[K][D], ua[K][D], nu[K];
  for(int n = 0; n < K; ++n) {
    const uint32_t c = g.conn[(uint64_t)e * K + n];
    for(int d = 0; d < D; ++d) { X[n][d] = g.x[d][c]; ua[n][d] = g.ua[d][c]; }
    nu[n] = g.nu[c];
  }
  double tau_ps, tau_su, tau_bulk;
  tau<E>(g.t, X, ua, nu, tau_ps, tau_su, tau_bulk);
  double xi[D], w, N[K], G[D][K], det;
  E::point(1, 0, xi, w);
  point_data<E>(X, xi, N, G, det);
  double* o = ctr + (uint64_t)e * kC;
  o[0] = w * det * tau_bulk;
  o[1] = tau_bulk;
  for(int d = 0; d < D; ++d)
    for(int n = 0; n < K; ++n) o[2 + d * K + n] = G[d][n];
}

template<class E>
__global__ __launch_bounds__(asmk::kThreads) void k_si_vmat_q1(SiArgs g, SiMaps m, const double* __restrict__ ctrs)
{
  using namespace asmk;
  constexpr int K = E::K, D = E::D, NN = D * D, kLanes = K, kSlots = kThreads / K;
  constexpr int kPt = 2 + K + D * K;        // w det, w det nu, N[K], G[D][K]
  constexpr int kC = 2 + D * K;             // pre-pass: w det tau_bulk at the centre, tau_bulk, Gc[D][K]
  constexpr int kSlotW = K * kPt + 1;
  constexpr int LDW = lds_row<double, NN>();
  constexpr size_t kRecBytes = sizeof(double) * kSlots * kSlotW;
  constexpr size_t kBlkBytes = sizeof(double) * kThreads * LDW;
  __shared__ __align__(16) unsigned char s_raw[kRecBytes > kBlkBytes ? kRecBytes : kBlkBytes];
  double* s_rec = reinterpret_cast<double*>(s_raw);   // point records, then
  double* s_blk = reinterpret_cast<double*>(s_raw);   // staged blocks
  __shared__ uint16_t s_acc[kThreads];

  const uint32_t wg = xcd_order(blockIdx.x, gridDim.x);
  const uint32_t r0 = m.wg_rows[wg], r1 = m.wg_rows[wg + 1];
  const uint32_t i0 = m.inc_ptr[r0];
  const uint32_t ninc = m.inc_ptr[r1] - i0;
  const uint32_t blk0 = m.row_ptr[r0];
  const uint32_t nblk = m.row_ptr[r1] - blk0;
  const int t = threadIdx.x;
  const int slot = t / kLanes, lane = t % kLanes;
  if((uint32_t)t < nblk) s_acc[t] = m.blk_acc[blk0 + t];
  const bool active = (uint32_t)slot < ninc;
  int a = 0;
  uint32_t pos = 0, e = 0;
  if(active) {
    const uint32_t ea = m.inc[i0 + slot];
    e = ea / K;
    a = (int)(ea % K);
    pos = m.inc_pos[(uint64_t)(i0 + slot) * kLanes + lane];
    double X[K][D], nu[K];
    for(int n = 0; n < K; ++n) {
      const uint32_t c = g.conn[(uint64_t)e * K + n];
      for(int d = 0; d < D; ++d) X[n][d] = g.x[d][c];
      nu[n] = g.nu[c];
    }
    double xi[D], w, N[K], G[D][K], det;
    E::point(2, lane, xi, w);
    point_data<E>(X, xi, N, G, det);
    double nuq = 0.;
    for(int n = 0; n < K; ++n) nuq += N[n] * nu[n];
    const double wq = w * det;
    double* rec = s_rec + (size_t)slot * kSlotW + lane * kPt;
    rec[0] = wq;
    rec[1] = wq * nuq;
    for(int n = 0; n < K; ++n) rec[2 + n] = N[n];
    for(int d = 0; d < D; ++d)
      for(int n = 0; n < K; ++n) rec[2 + K + d * K + n] = G[d][n];
  }
  __syncthreads();
  // the run of staged rows each accumulator thread sums
  const uint32_t alast = s_acc[nblk - 1];
  const uint32_t n_acc = (alast & 255u) + (alast >> 8);
  const bool owner = (uint32_t)t < n_acc;
  uint32_t jb = 0, je = 0, h = 0, c = 1;
  if(owner) {
    uint32_t lo = 0, hi = nblk;
    while(hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if((s_acc[mid] & 255u) <= (uint32_t)t) lo = mid; else hi = mid;
    }
    h = (uint32_t)t - (s_acc[lo] & 255u);
    c = s_acc[lo] >> 8;
    jb = m.blk_beg[(uint64_t)(blk0 + lo) * 4];
    je = lo + 1 < nblk ? m.blk_beg[(uint64_t)(blk0 + lo + 1) * 4] : ninc * kLanes;
  }
  // block (a, b = lane): T + theta dt (L + B) on the diagonal, the one-point bulk term off it
  double M[NN];
  if(active) {
    const int b = lane;
    const double* recs = s_rec + (size_t)slot * kSlotW;
    const double* ctr = ctrs + (uint64_t)e * kC;
    const double tau_bulk = ctr[1];
    double T = 0., L = 0., B[D];
    for(int d = 0; d < D; ++d) B[d] = 0.;
    for(int q = 0; q < K; ++q) {   // Q1: the second-order rule has K points, the order of k_si_elem
      const double* r = recs + q * kPt;
      const double wq = r[0], wn = r[1];
      const double* N = r + 2;
      const double* G = r + 2 + K;
      T += (wq * N[a]) * N[b];
      const double bulk = wq * tau_bulk;
      double gg = 0.;
      for(int d = 0; d < D; ++d) {
        gg += G[d * K + a] * G[d * K + b];
        B[d] += (bulk * G[d * K + a]) * G[d * K + b];
      }
      L += wn * gg;
    }
    const double bulk1 = ctr[0], tdt = g.theta * g.t.dt;
    const double* Gc = ctr + 2;
    for(int r = 0; r < D; ++r)
      for(int cc = 0; cc < D; ++cc)
        M[r * D + cc] = r == cc ? T + tdt * (L + B[r]) : tdt * ((bulk1 * Gc[r * K + a]) * Gc[cc * K + b]);
  }
  __syncthreads();   // records are no longer read
  if(active) {
    double* d = s_blk + pos * LDW;
    for(int v = 0; v < NN; ++v) d[v] = M[v];
  }
  __syncthreads();
  double acc[NN];
  for(int v = 0; v < NN; ++v) acc[v] = 0.;
  if(owner)
    for(uint32_t j = jb + h; j < je; j += c) {
      double v[NN];
      lds_load_row<double, NN>(s_blk + j * LDW, v);
      for(int q = 0; q < NN; ++q) acc[q] += v[q];
    }
  __syncthreads();
  if(owner) {
    double* d = s_blk + t * LDW;
    for(int q = 0; q < NN; ++q) d[q] = acc[q];
  }
  __syncthreads();
  double* dst = m.values + (uint64_t)blk0 * NN;
  for(uint32_t o = t; o < nblk * NN; o += kThreads) {
    const uint32_t sb = o / NN, v = o % NN;
    const uint32_t ac = s_acc[sb], f = ac & 255u, na = ac >> 8;
    double sum = s_blk[f * LDW + v];
    for(uint32_t r = 1; r < na; ++r) sum += s_blk[(f + r) * LDW + v];
    dst[o] = sum;
  }
}

// out[r * width + c] = (add ? out : 0) + sum over j in [ptr[r], ptr[r+1]) of stage[idx[j] * width + c],
// in list (element) order
__global__ void __launch_bounds__(256) k_si_gather(uint32_t n, int width, const uint32_t* __restrict__ ptr,
                                                   const uint32_t* __restrict__ idx,
                                                   const double* __restrict__ stage, double* __restrict__ out,
                                                   int add)
{
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(t >= (uint64_t)n * width) return;
  const uint32_t r = (uint32_t)(t / width);
  const int c = (int)(t - (uint64_t)r * width);
  double s = add ? out[t] : 0.;
  for(uint32_t j = ptr[r]; j < ptr[r + 1]; ++j) s += stage[(uint64_t)idx[j] * width + c];
  out[t] = s;
}

// SetSolution: row r takes the value of the last mesh node of that row (a node loop in node order)
struct Comp3 { const double* c[3]; };

__global__ void k_si_set_solution(uint32_t rows, int dim, const uint32_t* __restrict__ src, Comp3 f,
                                  double* __restrict__ out)
{
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if(r >= rows) return;
  for(int d = 0; d < dim; ++d) out[(uint64_t)r * dim + d] = f.c[d][src[r]];
}

// delta_a = delta_a - Ml^-1 (Aup delta_p); a += delta_a; U += dt delta_a  (InnerLoop::execute :162-175)
__global__ void k_si_correct_u(uint64_t n, double dt, const double* __restrict__ minv, const double* __restrict__ aup,
                               double* __restrict__ da, double* __restrict__ a, double* __restrict__ U)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const double d = -1. * (minv[i] * aup[i]) + 1. * da[i];
  da[i] = d;
  a[i] += d;
  U[i] += dt * d;
}

// p += dp; delta_p_sum += dp
__global__ void k_si_correct_p(uint64_t n, const double* __restrict__ dp, double* __restrict__ P,
                               double* __restrict__ dps)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  P[i] += dp[i];
  dps[i] += dp[i];
}

__global__ void k_si_reciprocal(uint64_t n, const double* __restrict__ x, double* __restrict__ y)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) y[i] = 1. / x[i];
}

// zero the given dofs of v
__global__ void k_si_zero_dofs(uint32_t n, const uint32_t* __restrict__ dofs, double* __restrict__ v)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) v[dofs[i]] = 0.;
}

// Update: u3 = u2, u2 = u1, u1 = u, u = U(p2m), p = P(p2m)
// one velocity component d (u3 = u2, u2 = u1, u1 = u, u = U(p2m)); the pressure with comp = -1 (p = P(p2m))
__global__ void k_si_update(uint32_t n, int dim, int d, const uint32_t* __restrict__ p2m, double* __restrict__ u,
                            double* __restrict__ u1, double* __restrict__ u2, double* __restrict__ u3,
                            const double* __restrict__ src)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const uint64_t r = p2m[i];
  if(d < 0) { u[i] = src[r]; return; }
  u3[i] = u2[i]; u2[i] = u1[i]; u1[i] = u[i];
  u[i] = src[r * dim + d];
}

void stable_lists(const std::vector<uint32_t>& keys, uint32_t n_keys, std::vector<uint32_t>& ptr,
                  std::vector<uint32_t>& idx)
{
  ptr.assign(n_keys + 1, 0);
  for(uint32_t k : keys) ++ptr[k + 1];
  for(uint32_t i = 0; i < n_keys; ++i) ptr[i + 1] += ptr[i];
  std::vector<uint32_t> fill(ptr.begin(), ptr.end() - 1);
  idx.assign(keys.size(), 0);
  for(size_t j = 0; j < keys.size(); ++j) idx[fill[keys[j]]++] = (uint32_t)j;
}

}  // namespace

// ------------------------------------------------------------------------------------------------
void si_create(SemiImplicit& S, Lss& U, Lss& P)
{
  if(!U.mesh || U.mesh != P.mesh)
    CF3_THROW(CF3HIP_ERR_INVALID_ARG, "the velocity and pressure systems must be created from one mesh");
  Mesh& M = *U.mesh;
  if(U.neq != M.dim || P.neq != 1)
    CF3_THROW(CF3HIP_ERR_INVALID_ARG, "velocity system needs neq = dim, pressure system neq = 1");
  if(U.periodic != P.periodic || U.n_nodes != P.n_nodes || U.n_blocks != P.n_blocks)
    CF3_THROW(CF3HIP_ERR_INVALID_ARG, "the two systems must share the node graph and periodic links");
  Context& c = *M.ctx;
  S.ctx = &c;
  S.mesh = &M;
  S.u = &U;
  S.p = &P;
  const uint32_t E = M.n_elems, K = M.k, rows = U.n_nodes;
  std::vector<uint32_t> p2m(M.n_nodes), conn((size_t)E * K), slots((size_t)E * K * K);
  if(U.periodic) HIP_OK(hipMemcpy(p2m.data(), U.p2m.ptr, sizeof(uint32_t) * M.n_nodes, hipMemcpyDeviceToHost));
  else for(uint32_t i = 0; i < M.n_nodes; ++i) p2m[i] = i;
  HIP_OK(hipMemcpy(conn.data(), M.conn.ptr, sizeof(uint32_t) * conn.size(), hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(slots.data(), U.slots.ptr, sizeof(uint32_t) * slots.size(), hipMemcpyDeviceToHost));
  std::vector<uint32_t> rowkeys(conn.size()), ptr, idx;
  for(size_t j = 0; j < conn.size(); ++j) rowkeys[j] = p2m[conn[j]];
  auto up = [&](DevBuf& b, const std::vector<uint32_t>& h) {
    b.alloc(&c, sizeof(uint32_t) * std::max<size_t>(h.size(), 1));
    HIP_OK(hipMemcpy(b.ptr, h.data(), sizeof(uint32_t) * h.size(), hipMemcpyHostToDevice));
  };
  stable_lists(slots, (uint32_t)U.n_blocks, ptr, idx);
  up(S.mat_ptr, ptr);
  up(S.mat_idx, idx);
  stable_lists(rowkeys, rows, ptr, idx);
  up(S.vec_ptr, ptr);
  up(S.vec_idx, idx);
  std::vector<uint32_t> src(rows, 0);
  for(uint32_t i = 0; i < M.n_nodes; ++i) src[p2m[i]] = i;   // node order: the last node of a row wins
  up(S.src, src);
  up(S.p2m, p2m);
  S.p2m_host = p2m;
  S.stage.alloc(&c, sizeof(double) * (size_t)E * K * K * M.dim * M.dim);
  const size_t vd = (size_t)rows * M.dim;
  for(DevBuf* b : {&S.A, &S.Uv, &S.minv, &S.aup}) { b->alloc(&c, sizeof(double) * vd); lss_fill(c, b->as<double>(), vd, 0.); }
  for(DevBuf* b : {&S.Pv, &S.dps}) { b->alloc(&c, sizeof(double) * rows); lss_fill(c, b->as<double>(), rows, 0.); }
  S.pvals0.alloc(&c, sizeof(double) * P.n_blocks);
  HIP_OK(hipStreamSynchronize(c.stream));
}

namespace {

SiArgs si_args(SemiImplicit& S, Fields& F, const cf3hip_ns_params& prm)
{
  Mesh& M = *S.mesh;
  SiArgs g{};
  g.n_elems = M.n_elems;
  g.conn = M.conn.as<uint32_t>();
  g.p2m = S.p2m.as<uint32_t>();
  for(int d = 0; d < M.dim; ++d) {
    g.x[d] = M.x[d].as<double>();
    g.u[d] = F.u[d].as<double>();
    g.ua[d] = F.uadv[d].as<double>();
  }
  g.nu = F.nu.as<double>();
  g.a = S.A.as<double>();
  g.U = S.Uv.as<double>();
  g.da = S.u->sol.as<double>();
  g.P = S.Pv.as<double>();
  g.dps = S.dps.as<double>();
  g.dp = S.p->sol.as<double>();
  for(int d = 0; d < M.dim; ++d) g.gf[d] = S.body_force ? F.g[d].as<double>() : nullptr;
  g.theta = prm.theta;
  if(prm.supg_type != CF3HIP_SUPG_METRIC && prm.supg_type != CF3HIP_SUPG_TEZDUYAR)
    CF3_THROW(CF3HIP_ERR_NOT_IMPLEMENTED, "the semi-implicit actions take supg_type metric or tezduyar");
  g.t = TauP{prm.supg_type, prm.dt, prm.c1, prm.c2, prm.alpha_ps, prm.alpha_su, prm.alpha_bu};
  g.stage = S.stage.as<double>();
  return g;
}

template<int OP>
void launch_elem(SemiImplicit& S, const SiArgs& g)
{
  const dim3 grid(grid_for(g.n_elems, 256));
  hipStream_t st = S.ctx->stream;
  switch(S.mesh->etype) {
    case CF3HIP_TRIAG2D: hipLaunchKernelGGL((k_si_elem<Tri, OP>), grid, dim3(256), 0, st, g); break;
    case CF3HIP_QUAD2D: hipLaunchKernelGGL((k_si_elem<Quad, OP>), grid, dim3(256), 0, st, g); break;
    case CF3HIP_TETRA3D: hipLaunchKernelGGL((k_si_elem<Tet, OP>), grid, dim3(256), 0, st, g); break;
    case CF3HIP_PRISM3D: hipLaunchKernelGGL((k_si_elem<Prism, OP>), grid, dim3(256), 0, st, g); break;
    default: hipLaunchKernelGGL((k_si_elem<Hex, OP>), grid, dim3(256), 0, st, g); break;
  }
  HIP_OK(hipGetLastError());
}

void gather(SemiImplicit& S, bool matrix, int width, uint32_t n, double* out, bool add)
{
  const uint64_t t = (uint64_t)n * width;
  hipLaunchKernelGGL(k_si_gather, dim3(grid_for(t, 256)), dim3(256), 0, S.ctx->stream, n, width,
                     (matrix ? S.mat_ptr : S.vec_ptr).as<uint32_t>(), (matrix ? S.mat_idx : S.vec_idx).as<uint32_t>(),
                     S.stage.as<double>(), out, add ? 1 : 0);
  HIP_OK(hipGetLastError());
}

void check_fields(SemiImplicit& S, Fields& F)
{
  if(F.mesh != S.mesh) CF3_THROW(CF3HIP_ERR_INVALID_ARG, "fields belong to another mesh");
}

}  // namespace

void si_initial(SemiImplicit& S, Fields& F, const cf3hip_ns_params& prm)
{
  check_fields(S, F);
  Lss &U = *S.u, &P = *S.p;
  Context& c = *S.ctx;
  const SiArgs g = si_args(S, F, prm);
  // PressureAssembly: p_lss->reset(0) then the simplified pressure matrix (PressureSystem.cpp:120-176)
  lss_clear_conditions(P);
  lss_fill(c, P.rhs.as<double>(), P.dofs, 0.);
  lss_fill(c, P.sol.as<double>(), P.dofs, 0.);
  launch_elem<OP_PMAT>(S, g);
  gather(S, true, 1, (uint32_t)P.n_blocks, P.values.as<double>(), false);
  lss_copy(c, S.pvals0.as<double>(), P.values.as<double>(), P.n_blocks);
  // ZeroVelocitySystem + MassMatrixAssembly into the velocity rhs + SetupInnerLoopData (reciprocal)
  lss_clear_conditions(U);
  lss_fill(c, U.values.as<double>(), U.n_blocks * U.neq * U.neq, 0.);
  lss_fill(c, U.sol.as<double>(), U.dofs, 0.);
  launch_elem<OP_MASS>(S, g);
  gather(S, false, 1, U.n_nodes, S.aup.as<double>(), false);   // per row; the same for every component
  // rhs[row * dim + i] = lumped(row)
  std::vector<double> lumped(U.n_nodes), rhs(U.dofs);
  HIP_OK(hipMemcpyAsync(lumped.data(), S.aup.ptr, sizeof(double) * U.n_nodes, hipMemcpyDeviceToHost, c.stream));
  HIP_OK(hipStreamSynchronize(c.stream));
  for(uint32_t r = 0; r < U.n_nodes; ++r)
    for(int d = 0; d < U.neq; ++d) rhs[(size_t)r * U.neq + d] = lumped[r];
  HIP_OK(hipMemcpyAsync(U.rhs.ptr, rhs.data(), sizeof(double) * U.dofs, hipMemcpyHostToDevice, c.stream));
  hipLaunchKernelGGL(k_si_reciprocal, dim3(grid_for(U.dofs, 256)), dim3(256), 0, c.stream, U.dofs,
                     U.rhs.as<double>(), S.minv.as<double>());
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(c.stream));
  S.initialized = true;
}

void si_velocity_assemble(SemiImplicit& S, Fields& F, const cf3hip_ns_params& prm)
{
  check_fields(S, F);
  Lss& U = *S.u;
  Context& c = *S.ctx;
  // ZeroLSS (matrix, rhs and solution) + VelocityAssembly
  lss_clear_conditions(U);
  lss_fill(c, U.rhs.as<double>(), U.dofs, 0.);
  lss_fill(c, U.sol.as<double>(), U.dofs, 0.);
  const SiArgs g = si_args(S, F, prm);
  const int et = S.mesh->etype;
  if((et == CF3HIP_QUAD2D || et == CF3HIP_HEXA3D) && U.passes == 1 && U.pos_stride == (int)S.mesh->k && U.n_wg) {
    const SiMaps mp{U.wg_rows.as<uint32_t>(), U.row_ptr.as<uint32_t>(), U.inc_ptr.as<uint32_t>(), U.inc.as<uint32_t>(),
                    U.inc_pos.as<uint8_t>(), U.blk_beg.as<uint8_t>(), U.blk_acc.as<uint16_t>(), U.values.as<double>()};
    // the per-element centre records go to the element staging buffer of the two-pass path (unused here;
    // 2 + D K doubles per element against its K^2 D^2)
    double* ctr = S.stage.as<double>();
    const dim3 ge(grid_for(S.mesh->n_elems, 256));
    timer_begin(c, CF3HIP_KERNEL_ASSEMBLY);
    if(et == CF3HIP_QUAD2D) {
      hipLaunchKernelGGL(k_si_ctr_q1<Quad>, ge, dim3(256), 0, c.stream, g, ctr);
      hipLaunchKernelGGL(k_si_vmat_q1<Quad>, dim3(U.n_wg), dim3(asmk::kThreads), 0, c.stream, g, mp, ctr);
    } else {
      hipLaunchKernelGGL(k_si_ctr_q1<Hex>, ge, dim3(256), 0, c.stream, g, ctr);
      hipLaunchKernelGGL(k_si_vmat_q1<Hex>, dim3(U.n_wg), dim3(asmk::kThreads), 0, c.stream, g, mp, ctr);
    }
    HIP_OK(hipGetLastError());
    timer_end(c, CF3HIP_KERNEL_ASSEMBLY);
    return;
  }
  launch_elem<OP_VMAT>(S, g);
  gather(S, true, U.neq * U.neq, (uint32_t)U.n_blocks, U.values.as<double>(), false);
}

void si_action(SemiImplicit& S, Fields& F, const cf3hip_ns_params& prm, int which)
{
  check_fields(S, F);
  const SiArgs g = si_args(S, F, prm);
  if(which == CF3HIP_SI_URHS) {
    launch_elem<OP_URHS>(S, g);
    gather(S, false, S.u->neq, S.u->n_nodes, S.u->rhs.as<double>(), true);
  } else if(which == CF3HIP_SI_PRHS) {
    launch_elem<OP_PRHS>(S, g);
    gather(S, false, 1, S.p->n_nodes, S.p->rhs.as<double>(), true);
  } else if(which == CF3HIP_SI_AUP) {
    launch_elem<OP_AUP>(S, g);
    gather(S, false, S.u->neq, S.u->n_nodes, S.u->rhs.as<double>(), true);
  } else {
    CF3_THROW(CF3HIP_ERR_INVALID_ARG, "unknown semi-implicit action");
  }
}

void si_set_solution(SemiImplicit& S, Fields& F)
{
  check_fields(S, F);
  Context& c = *S.ctx;
  const uint32_t rows = S.u->n_nodes;
  const int dim = S.mesh->dim;
  Comp3 fu{{F.u[0].as<double>(), F.u[1].as<double>(), dim == 3 ? F.u[2].as<double>() : nullptr}};
  Comp3 fp{{F.p.as<double>(), nullptr, nullptr}};
  hipLaunchKernelGGL(k_si_set_solution, dim3(grid_for(rows, 256)), dim3(256), 0, c.stream, rows, dim,
                     S.src.as<uint32_t>(), fu, S.u->sol.as<double>());
  hipLaunchKernelGGL(k_si_set_solution, dim3(grid_for(rows, 256)), dim3(256), 0, c.stream, rows, 1,
                     S.src.as<uint32_t>(), fp, S.p->sol.as<double>());
  HIP_OK(hipGetLastError());
}

void si_correct(SemiImplicit& S, double dt)
{
  Context& c = *S.ctx;
  Lss &U = *S.u, &P = *S.p;
  hipLaunchKernelGGL(k_si_correct_u, dim3(grid_for(U.dofs, 256)), dim3(256), 0, c.stream, (uint64_t)U.dofs, dt,
                     S.minv.as<double>(), U.rhs.as<double>(), U.sol.as<double>(), S.A.as<double>(),
                     S.Uv.as<double>());
  hipLaunchKernelGGL(k_si_correct_p, dim3(grid_for(P.dofs, 256)), dim3(256), 0, c.stream, (uint64_t)P.dofs,
                     P.sol.as<double>(), S.Pv.as<double>(), S.dps.as<double>());
  HIP_OK(hipGetLastError());
}

void si_update(SemiImplicit& S, Fields& F)
{
  check_fields(S, F);
  Mesh& M = *S.mesh;
  const dim3 grid(grid_for(M.n_nodes, 256));
  for(int d = 0; d < M.dim; ++d)
    hipLaunchKernelGGL(k_si_update, grid, dim3(256), 0, S.ctx->stream, M.n_nodes, M.dim, d, S.p2m.as<uint32_t>(),
                       F.u[d].as<double>(), F.u1[d].as<double>(), F.u2[d].as<double>(), F.u3[d].as<double>(),
                       S.Uv.as<double>());
  hipLaunchKernelGGL(k_si_update, grid, dim3(256), 0, S.ctx->stream, M.n_nodes, M.dim, -1, S.p2m.as<uint32_t>(),
                     F.p.as<double>(), nullptr, nullptr, nullptr, S.Pv.as<double>());
  HIP_OK(hipGetLastError());
}

void si_zero_dofs(SemiImplicit& S, const std::vector<uint32_t>& dofs, double* v)
{
  if(dofs.empty()) return;
  Context& c = *S.ctx;
  DevBuf d;
  d.alloc(&c, sizeof(uint32_t) * dofs.size());
  HIP_OK(hipMemcpyAsync(d.ptr, dofs.data(), sizeof(uint32_t) * dofs.size(), hipMemcpyHostToDevice, c.stream));
  hipLaunchKernelGGL(k_si_zero_dofs, dim3(grid_for(dofs.size(), 256)), dim3(256), 0, c.stream,
                     (uint32_t)dofs.size(), d.as<uint32_t>(), v);
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(c.stream));
}

}  // namespace cf3hip

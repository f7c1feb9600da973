// UFEM Navier-Stokes element assembly fused with a deterministic, atomic-free scatter into the LSS
// block-CSR matrix (gfx950).
//
// Work decomposition (row owner): a 256-thread workgroup owns a contiguous range of node rows and
// all their incidences (element e, local node a).  Each thread takes one incidence, computes the
// element quantities it needs from coalesced SoA node data, then produces the element row block
// M(a, b) = invdt*T(a,b) + theta*A'(a,b) one column node b at a time.  Per pass b the blocks are
// staged in LDS and the owner threads of every output value sum the contributions of their slot,
// walking the precomputed per-slot contribution list (sorted by (b, q)): a sorted segmented
// reduction with a fixed order, so results are deterministic and no floating-point atomics are used.
// Every CSR value and rhs entry is written exactly once, coalesced.
//
// Element mathematics (reference files relative to the coolfluid3 tree):
//   specialised P1 triangle / tetrahedron: plugins/UFEM/src/UFEM/NavierStokesSpecializations.hpp:38-303
//     (incl. the vk for wk substitution of lines 224-226 and 236-238 and tau from the velocity u)
//   generic expression (use_specializations = false): NavierStokesAssembly.hpp:44-66, order-2 Gauss
//   tau METRIC / TEZDUYAR: plugins/UFEM/src/UFEM/SUPG.hpp:141-219 (gij = J^-T J^-1 = Gram matrix of
//     the shape-function gradients of nodes 1..d for affine elements)
//   combine: NavierStokesAssembly.hpp:69-71 (rhs = -A x + a before the p rows are divided by theta)
#include <type_traits>

#include "internal.hpp"

namespace cf3hip {

namespace {

constexpr int kThreads = 256;   // one incidence and at most one owned CSR block per thread

struct NodeData {
  const double* x[3];
  const double* u[3];
  const double* ua[3];
  const double* p;
  const double* nu;
  const double* g[3];
};

struct AsmArgs {
  NodeData nd;
  const uint32_t* conn;
  const uint32_t* wg_rows;
  const uint32_t* row_ptr;
  const uint32_t* inc_ptr;
  const uint32_t* inc;
  const uint32_t* contrib_ptr;
  const uint16_t* contrib;
  void* values;
  double* rhs;
  const double* xt;    // matrix-free residual: x with constrained columns set to their first value
  double* res;         // matrix-free residual output
  double rho, dt, theta, invdt;
  int supg_type;
  double c1, c2, a_ps, a_su, a_bu;
  int* status;
};

template<int B, int K, typename F>
__device__ inline void static_for(F&& f)
{
  if constexpr(B < K) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, K>(f);
  }
}

// ---------------------------------------------------------------------------------------------
// Specialised P1 simplex (D = 2 triangle, D = 3 tetrahedron) in arithmetic type R
// ---------------------------------------------------------------------------------------------
template<int D, typename R>
struct SpecP1 {
  static constexpr int K = D + 1, NEQ = D + 1, NN = NEQ * NEQ;
  R n[K][D];        // reference face normals (area vectors for D = 3, rotated edges for D = 2)
  R na[D];          // normal of the row node a
  R V;
  R tps;
  R ua[D];          // mean advection velocity u_avg
  R xb[K][NEQ];     // element nodal (u.., p) for -A x
  int a;
  R u_ni, cskew, cpres, time_ii, tp_dV, ts_dV, nu_dV, tb_dV, ni_ut;

  // geometry in type T: normals and volume
  template<typename T>
  __device__ static void geometry(const T X[K][D], T nrm[K][D], T& vol)
  {
    if constexpr(D == 2) {
      nrm[0][0] = X[1][1] - X[2][1]; nrm[0][1] = X[2][0] - X[1][0];
      nrm[1][0] = X[2][1] - X[0][1]; nrm[1][1] = X[0][0] - X[2][0];
      nrm[2][0] = X[0][1] - X[1][1]; nrm[2][1] = X[1][0] - X[0][0];
      vol = T(0.5) * ((X[1][0] - X[0][0]) * (X[2][1] - X[0][1]) - (X[2][0] - X[0][0]) * (X[1][1] - X[0][1]));
    } else {
      // NavierStokesSpecializations.hpp:162-173
      nrm[0][0] = ((X[3][1]-X[1][1])*(X[2][2]-X[1][2]) - (X[2][1]-X[1][1])*(X[3][2]-X[1][2])) / T(2);
      nrm[0][1] = ((X[3][2]-X[1][2])*(X[2][0]-X[1][0]) - (X[2][2]-X[1][2])*(X[3][0]-X[1][0])) / T(2);
      nrm[0][2] = ((X[3][0]-X[1][0])*(X[2][1]-X[1][1]) - (X[2][0]-X[1][0])*(X[3][1]-X[1][1])) / T(2);
      nrm[1][0] = ((X[2][1]-X[0][1])*(X[3][2]-X[0][2]) - (X[3][1]-X[0][1])*(X[2][2]-X[0][2])) / T(2);
      nrm[1][1] = ((X[2][2]-X[0][2])*(X[3][0]-X[0][0]) - (X[3][2]-X[0][2])*(X[2][0]-X[0][0])) / T(2);
      nrm[1][2] = ((X[2][0]-X[0][0])*(X[3][1]-X[0][1]) - (X[3][0]-X[0][0])*(X[2][1]-X[0][1])) / T(2);
      nrm[2][0] = ((X[3][1]-X[0][1])*(X[1][2]-X[0][2]) - (X[1][1]-X[0][1])*(X[3][2]-X[0][2])) / T(2);
      nrm[2][1] = ((X[3][2]-X[0][2])*(X[1][0]-X[0][0]) - (X[1][2]-X[0][2])*(X[3][0]-X[0][0])) / T(2);
      nrm[2][2] = ((X[3][0]-X[0][0])*(X[1][1]-X[0][1]) - (X[1][0]-X[0][0])*(X[3][1]-X[0][1])) / T(2);
      nrm[3][0] = ((X[1][1]-X[0][1])*(X[2][2]-X[0][2]) - (X[2][1]-X[0][1])*(X[1][2]-X[0][2])) / T(2);
      nrm[3][1] = ((X[1][2]-X[0][2])*(X[2][0]-X[0][0]) - (X[2][2]-X[0][2])*(X[1][0]-X[0][0])) / T(2);
      nrm[3][2] = ((X[1][0]-X[0][0])*(X[2][1]-X[0][1]) - (X[2][0]-X[0][0])*(X[1][1]-X[0][1])) / T(2);
      // det J = 2 (x3 - x0) . n_3 ; V = det / 6
      vol = (T(2) * ((X[3][0]-X[0][0])*nrm[3][0] + (X[3][1]-X[0][1])*nrm[3][1] + (X[3][2]-X[0][2])*nrm[3][2])) / T(6);
    }
  }

  // METRIC tau from the velocity at the centroid (SUPG.hpp:176-190); gij = n_i.n_l / (D^2 V^2)
  template<typename T>
  __device__ static void tau_metric(const T nrm[K][D], T vol, const T uc[D], T nu, T dt, T& t_ps, T& t_su, T& t_bu)
  {
    const T s = T(1) / (T(D) * vol);
    T gn[D][D];   // grad N of nodes 1..D
    for(int i = 0; i < D; ++i)
      for(int d = 0; d < D; ++d) gn[i][d] = nrm[i + 1][d] * s;
    T w[D];
    for(int d = 0; d < D; ++d) {
      T acc = T(0);
      for(int i = 0; i < D; ++i) acc += uc[i] * gn[i][d];
      w[d] = acc;
    }
    T adv = T(0);
    for(int d = 0; d < D; ++d) adv += w[d] * w[d];
    T gsq = T(0), tr = T(0);
    for(int i = 0; i < D; ++i)
      for(int l = 0; l < D; ++l) {
        T gil = T(0);
        for(int d = 0; d < D; ++d) gil += gn[i][d] * gn[l][d];
        gsq += gil * gil;
        if(i == l) tr += gil;
      }
    const T diff = nu * nu * gsq;
    const T sq = sqrt(T(4) / (dt * dt) + adv + T(16) * diff);   // c1 = 1, c2 = 16 on the specialised path
    t_su = T(1) / sq;
    t_ps = t_su;
    t_bu = sq / tr;
  }

  __device__ void prepare(const AsmArgs& g, uint32_t e, int a_)
  {
    a = a_;
    uint32_t cn[K];
#pragma unroll
    for(int m = 0; m < K; ++m) cn[m] = g.conn[(uint64_t)e * K + m];
    double X[K][D];
    double uc[D];
    double nus = 0.;
    R uas[D];
    for(int d = 0; d < D; ++d) { uc[d] = 0.; uas[d] = R(0); }
#pragma unroll
    for(int m = 0; m < K; ++m) {
      const uint32_t nd = cn[m];
#pragma unroll
      for(int d = 0; d < D; ++d) {
        X[m][d] = g.nd.x[d][nd];
        const double um = g.nd.u[d][nd];
        uc[d] += um;
        uas[d] += (R)g.nd.ua[d][nd];
        xb[m][d] = (R)um;
      }
      xb[m][D] = (R)g.nd.p[nd];
      nus += g.nd.nu[nd];
    }
    for(int d = 0; d < D; ++d) {
      uc[d] = uc[d] / K;
      ua[d] = uas[d] / R(K);
    }
    const double nu_mean = fabs(nus / K);
    const R nu = (R)nu_mean;
    R tsu, tbu;
    bool fp64_geom = sizeof(R) == sizeof(double);
    if constexpr(sizeof(R) != sizeof(double)) {
      R Xr[K][D];
      for(int m = 0; m < K; ++m)
        for(int d = 0; d < D; ++d) Xr[m][d] = (R)X[m][d];
      geometry<R>(Xr, n, V);
      // strongly skewed elements (volume small against the edge lengths): geometry and tau in FP64
      R l2 = R(0);
      for(int i = 0; i < K; ++i)
        for(int j = i + 1; j < K; ++j)
          for(int d = 0; d < D; ++d) { const R df = Xr[i][d] - Xr[j][d]; l2 += df * df; }
      const R lm = sqrt(l2 / R(K * (K - 1) / 2));
      R lp = lm;
      for(int d = 1; d < D; ++d) lp *= lm;
      if(fabs(V) < R(0.05) * lp) fp64_geom = true;
    }
    if(fp64_geom) {
      double n64[K][D], V64, tp, ts, tb;
      geometry<double>(X, n64, V64);
      tau_metric<double>(n64, V64, uc, nu_mean, g.dt, tp, ts, tb);
      for(int m = 0; m < K; ++m)
        for(int d = 0; d < D; ++d) n[m][d] = (R)n64[m][d];
      V = (R)V64; tps = (R)tp; tsu = (R)ts; tbu = (R)tb;
      if(!(V64 > 0.)) g.status[0] = 1;   // every writer stores the same flag value
    } else {
      R ucr[D];
      for(int d = 0; d < D; ++d) ucr[d] = (R)uc[d];
      tau_metric<R>(n, V, ucr, nu, (R)g.dt, tps, tsu, tbu);
      if(!(V > R(0))) g.status[0] = 1;
    }
    // normal of the row node, selected without dynamic register indexing
#pragma unroll
    for(int d = 0; d < D; ++d) na[d] = n[0][d];
#pragma unroll
    for(int m = 1; m < K; ++m)
      if(m == a)
#pragma unroll
        for(int d = 0; d < D; ++d) na[d] = n[m][d];
    // per-row (node a) coefficients
    const R c0 = D == 2 ? R(1) / R(6) : R(1) / R(12);
    const R rdV = R(1) / (R(D * D) * V);
    ts_dV = tsu * rdV;
    tp_dV = tps * rdV;
    nu_dV = nu * rdV;
    tb_dV = tbu * rdV;
    u_ni = R(0);
    for(int d = 0; d < D; ++d) u_ni += ua[d] * na[d];
    cskew = R(0.5) * c0 + R(0.5) * ts_dV * u_ni;
    cpres = c0 / (R)g.rho + ts_dV / (R)g.rho * u_ni;
    time_ii = tsu * c0 * u_ni;
    // n_i . u~ with the tetrahedron substitution of vk for wk (NavierStokesSpecializations.hpp:236-238)
    ni_ut = na[0] * ua[0] + na[1] * ua[1];
    if constexpr(D == 3) ni_ut += na[2] * ua[1];
  }

  // block (a, B): M = invdt*T + theta*A' into m[NN] (row-major); rhs -= A x_B
  template<int B, typename VR>
  __device__ void block(const AsmArgs& g, VR* m, R* rhs) const
  {
    const R c0 = D == 2 ? R(1) / R(6) : R(1) / R(12);
    const R* nj = n[B];
    R uknj = R(0), ninj = R(0);
    for(int d = 0; d < D; ++d) { uknj += ua[d] * nj[d]; ninj += na[d] * nj[d]; }
    const R conv = c0 * uknj + ts_dV * uknj * u_ni;
    const R uk_row[3] = {ua[0], ua[1], ua[D == 3 ? 1 : 0]};   // W rows use vk (lines 224-226)
    R A[NEQ][NEQ];
    for(int r = 0; r < D; ++r) {
      for(int c = 0; c < D; ++c) {
        R v = cskew * uk_row[r] * nj[c] + (nu_dV / R(3) + tb_dV) * na[r] * nj[c];
        if(r == c) v += conv + nu_dV * ninj;
        A[r][c] = v;
      }
      A[r][D] = nj[r] * cpres;
    }
    for(int c = 0; c < D; ++c)
      A[D][c] = tp_dV * uknj * na[c] + R(0.5) * tp_dV * ni_ut * nj[c] + c0 * nj[c];
    A[D][D] = tp_dV / (R)g.rho * ninj;
    const R mass = (D == 2 ? V / R(12) : V / R(20)) * (a == B ? R(2) : R(1));
    const R tdiag = mass + time_ii;
    const R invdt = (R)g.invdt, theta = (R)g.theta;
    for(int r = 0; r < NEQ; ++r) {
      R s = R(0);
      for(int c = 0; c < NEQ; ++c) s += A[r][c] * xb[B][c];
      rhs[r] -= s;
    }
    for(int r = 0; r < D; ++r)
      for(int c = 0; c < NEQ; ++c)
        m[r * NEQ + c] = (VR)((r == c ? invdt * tdiag : R(0)) + theta * A[r][c]);
    for(int c = 0; c < D; ++c)
      m[D * NEQ + c] = (VR)(invdt * (tps * c0 * na[c]) + theta * (A[D][c] / theta));
    m[D * NEQ + D] = (VR)(theta * (A[D][D] / theta));
  }

  __device__ void rhs_source(const AsmArgs&, R*) const {}
};

// ---------------------------------------------------------------------------------------------
// Generic expression for P1 simplices (use_specializations = false): NavierStokesAssembly.hpp:44-66
// with tau from u_adv and the order-2 Gauss rule; gradients constant over the element.
// ---------------------------------------------------------------------------------------------
template<int D>
struct GenericP1 {
  typedef double R;
  static constexpr int K = D + 1, NEQ = D + 1, NN = NEQ * NEQ, NQ = D == 2 ? 3 : 4;
  R G[D][K];         // physical gradients
  R Ga[D];           // gradient of the row node a
  R detw;            // |J| (weights applied per point)
  R ua_n[K][D];      // nodal advection velocity
  R nu_n[K];
  R g_n[K][D];       // nodal body force
  R xb[K][NEQ];
  R tps, tsu, tbu;
  int a;

  __device__ static void gauss(int q, R xi[D], R& w)
  {
    if constexpr(D == 2) {
      xi[0] = q == 2 ? 0.0 : 0.5;
      xi[1] = q == 0 ? 0.0 : 0.5;
      w = 1.0 / 6.0;
    } else {
      const R mu = 0.13819660112501051517954131656;
      xi[0] = q == 1 ? 1. - 3. * mu : mu;
      xi[1] = q == 2 ? 1. - 3. * mu : mu;
      xi[2] = q == 3 ? 1. - 3. * mu : mu;
      w = 0.04166666666666666666666666666666666667;
    }
  }

  __device__ static void shape(const R xi[D], R N[K])
  {
    R s = 1.;
    #pragma unroll
    for(int d = 0; d < D; ++d) { N[d + 1] = xi[d]; s -= xi[d]; }
    N[0] = s;
  }

  __device__ R select_a(const R v[K]) const
  {
    R r = v[0];
#pragma unroll
    for(int m = 1; m < K; ++m) if(m == a) r = v[m];
    return r;
  }

  __device__ void prepare(const AsmArgs& g, uint32_t e, int a_)
  {
    a = a_;
    uint32_t cn[K];
#pragma unroll
    for(int m = 0; m < K; ++m) cn[m] = g.conn[(uint64_t)e * K + m];
    R X[K][D];
    R nus = 0.;
#pragma unroll
    for(int m = 0; m < K; ++m) {
      const uint32_t nd = cn[m];
#pragma unroll
      for(int d = 0; d < D; ++d) {
        X[m][d] = g.nd.x[d][nd];
        ua_n[m][d] = g.nd.ua[d][nd];
        g_n[m][d] = g.nd.g[d][nd];
        xb[m][d] = g.nd.u[d][nd];
      }
      xb[m][D] = g.nd.p[nd];
      nu_n[m] = g.nd.nu[nd];
      nus += nu_n[m];
    }
    // J(ksi, x) = x_{ksi+1} - x_0, inverse by the adjugate
    R J[D][D];
    #pragma unroll
    for(int i = 0; i < D; ++i)
      #pragma unroll
      for(int d = 0; d < D; ++d) J[i][d] = X[i + 1][d] - X[0][d];
    R Ji[D][D], det;
    if constexpr(D == 2) {
      det = J[0][0] * J[1][1] - J[1][0] * J[0][1];
      const R inv = 1. / det;
      Ji[0][0] = J[1][1] * inv; Ji[0][1] = -J[0][1] * inv;
      Ji[1][0] = -J[1][0] * inv; Ji[1][1] = J[0][0] * inv;
    } else {
      const R c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
      const R c10 = J[2][1] * J[0][2] - J[2][2] * J[0][1];
      const R c20 = J[0][1] * J[1][2] - J[0][2] * J[1][1];
      det = c00 * J[0][0] + c10 * J[1][0] + c20 * J[2][0];
      const R inv = 1. / det;
      Ji[0][0] = c00 * inv; Ji[0][1] = c10 * inv; Ji[0][2] = c20 * inv;
      Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * inv;
      Ji[1][1] = (J[2][2] * J[0][0] - J[2][0] * J[0][2]) * inv;
      Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * inv;
      Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * inv;
      Ji[2][1] = (J[2][0] * J[0][1] - J[2][1] * J[0][0]) * inv;
      Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * inv;
    }
    if(!(det > 0.)) g.status[0] = 1;
    detw = det;
    // grad N = Jinv * dN/dksi ; dN/dksi of node 0 = -1, node m+1 = e_m
    #pragma unroll
    for(int d = 0; d < D; ++d) {
      R s = 0.;
      #pragma unroll
      for(int l = 0; l < D; ++l) { G[d][l + 1] = Ji[d][l]; s -= Ji[d][l]; }
      G[d][0] = s;
    }
    #pragma unroll
    for(int d = 0; d < D; ++d) Ga[d] = select_a(G[d]);
    // tau from u_adv at the centroid (NavierStokesAssembly.hpp:54, SUPG.hpp:137-219)
    const R nu_e = fabs(nus / K);
    R uc[D];
    #pragma unroll
    for(int d = 0; d < D; ++d) {
      R s = 0.;
      #pragma unroll
      for(int m = 0; m < K; ++m) s += ua_n[m][d];
      uc[d] = s / K;
    }
    if(g.supg_type == CF3HIP_SUPG_METRIC) {
      R gij[D][D];
      #pragma unroll
      for(int i = 0; i < D; ++i)
        #pragma unroll
        for(int j = 0; j < D; ++j) {
          R s = 0.;
          #pragma unroll
          for(int l = 0; l < D; ++l) s += Ji[l][i] * Ji[l][j];
          gij[i][j] = s;
        }
      R ugu = 0., gsq = 0., tr = 0.;
      #pragma unroll
      for(int j = 0; j < D; ++j) {
        R s = 0.;
        #pragma unroll
        for(int i = 0; i < D; ++i) s += uc[i] * gij[i][j];
        ugu += s * uc[j];
        tr += gij[j][j];
        #pragma unroll
        for(int i = 0; i < D; ++i) gsq += gij[i][j] * gij[i][j];
      }
      const R diff = nu_e * nu_e * gsq;
      tsu = 1. / sqrt(4. * g.c1 * g.c1 / (g.dt * g.dt) + fabs(ugu) + g.c2 * diff);
      tps = tsu;
      tbu = sqrt(4. / (g.dt * g.dt) + fabs(ugu) + 16. * diff) / tr;
    } else {
      R h = 1e10;
      #pragma unroll
      for(int i = 0; i < K; ++i)
        #pragma unroll
        for(int j = 0; j < K; ++j)
          if(i != j) {
            R s = 0.;
            #pragma unroll
            for(int d = 0; d < D; ++d) { const R df = X[i][d] - X[j][d]; s += df * df; }
            h = fmin(h, s);
          }
      h = sqrt(h);
      R um = 0.;
      #pragma unroll
      for(int d = 0; d < D; ++d) um += uc[d] * uc[d];
      um = sqrt(um);
      tsu = 1. / ((2. * um) / h + 2. / g.dt + (4. * nu_e) / (h * h));
      tps = tsu;
      tbu = tsu * um * um;
    }
    tps *= g.a_ps; tsu *= g.a_su; tbu *= g.a_bu;
  }

  template<int B, typename VR>
  __device__ void block(const AsmArgs& g, VR* m, R* rhs) const
  {
    R A[NEQ][NEQ], T[NEQ][NEQ];
    #pragma unroll
    for(int r = 0; r < NEQ; ++r)
      #pragma unroll
      for(int c = 0; c < NEQ; ++c) { A[r][c] = 0.; T[r][c] = 0.; }
    R gtg = 0.;
    #pragma unroll
    for(int d = 0; d < D; ++d) gtg += Ga[d] * G[d][B];
    #pragma unroll
    for(int q = 0; q < NQ; ++q) {
      R xi[D], wq, N[K];
      gauss(q, xi, wq);
      shape(xi, N);
      const R Na = select_a(N);
      const R w = wq * detw;
      R ua[D], nuq = 0.;
      #pragma unroll
      for(int d = 0; d < D; ++d) {
        R s = 0.;
        #pragma unroll
        for(int mm = 0; mm < K; ++mm) s += N[mm] * ua_n[mm][d];
        ua[d] = s;
      }
      #pragma unroll
      for(int mm = 0; mm < K; ++mm) nuq += N[mm] * nu_n[mm];
      R La = 0., Lb = 0.;
      #pragma unroll
      for(int d = 0; d < D; ++d) { La += ua[d] * Ga[d]; Lb += ua[d] * G[d][B]; }
      const R SUa = Na + tsu * La;
      const R Lps = 0.5 * tps * La;
      #pragma unroll
      for(int i = 0; i < D; ++i) {
        A[D][i] += w * ((Na + Lps) * G[i][B] + (tps * Ga[i]) * Lb);
        A[i][i] += w * (nuq * gtg + SUa * Lb);
        A[i][D] += w * (SUa * G[i][B] / g.rho);
        #pragma unroll
        for(int j = 0; j < D; ++j) A[i][j] += w * ((tbu * Ga[i] + 0.5 * ua[i] * SUa) * G[j][B]);
        T[D][i] += w * (tps * Ga[i] * N[B]);
        T[i][i] += w * (SUa * N[B]);
      }
      A[D][D] += w * ((tps * gtg) / g.rho);
    }
    #pragma unroll
    for(int r = 0; r < NEQ; ++r) {
      R s = 0.;
      #pragma unroll
      for(int c = 0; c < NEQ; ++c) s += A[r][c] * xb[B][c];
      rhs[r] -= s;
    }
    #pragma unroll
    for(int r = 0; r < NEQ; ++r)
      #pragma unroll
      for(int c = 0; c < NEQ; ++c) {
        const R ar = r == D ? A[r][c] / g.theta : A[r][c];
        m[r * NEQ + c] = (VR)(g.invdt * T[r][c] + g.theta * ar);
      }
  }

  // _a[u[_i]] += (N + tau_su u_adv grad N)^T g  (NavierStokesAssembly.hpp:65)
  __device__ void rhs_source(const AsmArgs& g, R* rhs) const
  {
    #pragma unroll
    for(int q = 0; q < NQ; ++q) {
      R xi[D], wq, N[K];
      gauss(q, xi, wq);
      shape(xi, N);
      const R Na = select_a(N);
      const R w = wq * detw;
      R ua[D], gq[D];
      #pragma unroll
      for(int d = 0; d < D; ++d) {
        R s = 0., sg = 0.;
        #pragma unroll
        for(int mm = 0; mm < K; ++mm) { s += N[mm] * ua_n[mm][d]; sg += N[mm] * g_n[mm][d]; }
        ua[d] = s; gq[d] = sg;
      }
      R La = 0.;
      #pragma unroll
      for(int d = 0; d < D; ++d) La += ua[d] * Ga[d];
      const R SUa = Na + tsu * La;
      #pragma unroll
      for(int i = 0; i < D; ++i) rhs[i] += w * (SUa * gq[i]);
    }
  }
};

// ---------------------------------------------------------------------------------------------
// Row-owner assembly kernel
// ---------------------------------------------------------------------------------------------
template<class Op, typename R, typename VR>
__global__ __launch_bounds__(kThreads) void k_assemble(AsmArgs g)
{
  constexpr int K = Op::K, NEQ = Op::NEQ, NN = NEQ * NEQ;
  constexpr int LDW = NN + 1;   // padded row of the staging tile
  __shared__ VR s_blk[kThreads * LDW];
  __shared__ R s_rhs[kThreads * NEQ];
  __shared__ uint32_t s_rowblk[kThreads + 1];
  __shared__ uint32_t s_rowinc[kThreads + 1];

  const uint32_t r0 = g.wg_rows[blockIdx.x], r1 = g.wg_rows[blockIdx.x + 1];
  const uint32_t nrows = r1 - r0;
  const uint32_t i0 = g.inc_ptr[r0];
  const uint32_t ninc = g.inc_ptr[r1] - i0;
  const uint32_t blk0 = g.row_ptr[r0];
  const uint32_t nblk = g.row_ptr[r1] - blk0;
  const int t = threadIdx.x;
  for(uint32_t j = t; j <= nrows; j += kThreads) {
    s_rowblk[j] = g.row_ptr[r0 + j] - blk0;
    s_rowinc[j] = g.inc_ptr[r0 + j] - i0;
  }
  __syncthreads();

  // owner of block t of this workgroup's contiguous block range
  const bool owner = (uint32_t)t < nblk;
  R acc[NN];
#pragma unroll
  for(int v = 0; v < NN; ++v) acc[v] = R(0);
  uint32_t cur = 0, cend = 0, qb = 0;
  if(owner) {
    uint32_t lo = 0, hi = nrows;        // row j with s_rowblk[j] <= t < s_rowblk[j+1]
    while(hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if(s_rowblk[mid] <= (uint32_t)t) lo = mid; else hi = mid;
    }
    qb = s_rowinc[lo];
    cur = g.contrib_ptr[blk0 + t];
    cend = g.contrib_ptr[blk0 + t + 1];
  }

  const bool active = (uint32_t)t < ninc;
  Op op;
  R rhs[NEQ];
#pragma unroll
  for(int r = 0; r < NEQ; ++r) rhs[r] = R(0);
  if(active) {
    const uint32_t ea = g.inc[i0 + t];
    op.prepare(g, ea / K, (int)(ea % K));
  }
  static_for<0, K>([&](auto bc) {
    constexpr int B = decltype(bc)::value;
    VR mb[NN];
    if(active) op.template block<B, VR>(g, mb, rhs);
    __syncthreads();
    if(active)
#pragma unroll
      for(int v = 0; v < NN; ++v) s_blk[t * LDW + v] = mb[v];
    __syncthreads();
    if(owner) {
      while(cur < cend) {
        const uint32_t c = g.contrib[cur];
        if((int)(c & 15u) != B) break;
        const VR* src = s_blk + (qb + (c >> 4)) * LDW;
#pragma unroll
        for(int v = 0; v < NN; ++v) acc[v] += (R)src[v];
        ++cur;
      }
    }
  });
  if(owner) {
    VR* dst = static_cast<VR*>(g.values) + (uint64_t)(blk0 + t) * NN;
#pragma unroll
    for(int v = 0; v < NN; ++v) dst[v] = (VR)acc[v];
  }
  if(active) {
    op.rhs_source(g, rhs);
#pragma unroll
    for(int r = 0; r < NEQ; ++r) s_rhs[t * NEQ + r] = rhs[r];
  }
  __syncthreads();
  if(!g.rhs) return;   // the FP32 matrix pass leaves the FP64 right-hand side alone
  for(uint32_t o = t; o < nrows * NEQ; o += kThreads) {
    const uint32_t row = o / NEQ, r = o % NEQ;
    R s = R(0);
    for(uint32_t q = s_rowinc[row]; q < s_rowinc[row + 1]; ++q) s += s_rhs[q * NEQ + r];
    g.rhs[(uint64_t)(r0 + row) * NEQ + r] = (double)s;
  }
}

// Matrix-free FP64 residual of the assembled, Dirichlet-constrained system: per node row
// r_i = sum over incident elements of (b_e(a) - M_e(a, :) x~_e), with x~ = x except on constrained dofs
// where it holds the eliminated value (the symmetric elimination moves A[:, c] v_c to the rhs).
// Constrained rows are fixed afterwards (r = v - x).  No matrix is read or written.
template<class Op>
__global__ __launch_bounds__(kThreads) void k_residual(AsmArgs g)
{
  constexpr int K = Op::K, NEQ = Op::NEQ, NN = NEQ * NEQ;
  __shared__ double s_res[kThreads * NEQ];
  __shared__ uint32_t s_rowinc[kThreads + 1];
  const uint32_t r0 = g.wg_rows[blockIdx.x], r1 = g.wg_rows[blockIdx.x + 1];
  const uint32_t nrows = r1 - r0;
  const uint32_t i0 = g.inc_ptr[r0];
  const uint32_t ninc = g.inc_ptr[r1] - i0;
  const int t = threadIdx.x;
  for(uint32_t j = t; j <= nrows; j += kThreads) s_rowinc[j] = g.inc_ptr[r0 + j] - i0;
  const bool active = (uint32_t)t < ninc;
  Op op;
  double rhs[NEQ], y[NEQ], xt[K][NEQ];
#pragma unroll
  for(int r = 0; r < NEQ; ++r) { rhs[r] = 0.; y[r] = 0.; }
  if(active) {
    const uint32_t ea = g.inc[i0 + t];
    const uint32_t e = ea / K;
    op.prepare(g, e, (int)(ea % K));
#pragma unroll
    for(int m = 0; m < K; ++m) {
      const uint32_t nd = g.conn[(uint64_t)e * K + m];
#pragma unroll
      for(int c = 0; c < NEQ; ++c) xt[m][c] = g.xt[(uint64_t)nd * NEQ + c];
    }
  }
  static_for<0, K>([&](auto bc) {
    constexpr int B = decltype(bc)::value;
    if(active) {
      double mb[NN];
      op.template block<B, double>(g, mb, rhs);
#pragma unroll
      for(int r = 0; r < NEQ; ++r)
#pragma unroll
        for(int c = 0; c < NEQ; ++c) y[r] += mb[r * NEQ + c] * xt[B][c];
    }
  });
  if(active) {
    op.rhs_source(g, rhs);
#pragma unroll
    for(int r = 0; r < NEQ; ++r) s_res[t * NEQ + r] = rhs[r] - y[r];
  }
  __syncthreads();
  for(uint32_t o = t; o < nrows * NEQ; o += kThreads) {
    const uint32_t row = o / NEQ, r = o % NEQ;
    double s = 0.;
    for(uint32_t q = s_rowinc[row]; q < s_rowinc[row + 1]; ++q) s += s_res[q * NEQ + r];
    g.res[(uint64_t)(r0 + row) * NEQ + r] = s;
  }
}

__global__ void k_xtilde(uint64_t dofs, int neq, const uint8_t* __restrict__ mask, const double* __restrict__ vfirst,
                         const double* __restrict__ x, double* __restrict__ xt)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= dofs) return;
  const uint8_t m = mask ? mask[i / neq] : 0;
  xt[i] = ((m >> (i % neq)) & 1) ? vfirst[i] : x[i];
}

__global__ void k_bc_residual(uint32_t n, const uint32_t* __restrict__ dof, const double* __restrict__ vlast,
                              const double* __restrict__ x, double* __restrict__ r)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) r[dof[i]] = vlast[i] - x[dof[i]];
}

template<class Op, typename R, typename VR>
void launch(const AsmArgs& g, const Lss& L)
{
  Context& c = *L.ctx;
  timer_begin(c, CF3HIP_KERNEL_ASSEMBLY);
  hipLaunchKernelGGL((k_assemble<Op, R, VR>), dim3(L.n_wg), dim3(kThreads), 0, c.stream, g);
  HIP_OK(hipGetLastError());
  timer_end(c, CF3HIP_KERNEL_ASSEMBLY);
}

__global__ void k_clear_flag(int* f) { f[0] = 0; }

}  // namespace

namespace {

AsmArgs make_args(Lss& L, Fields& F, const cf3hip_ns_params& prm)
{
  if(!L.mesh) CF3_THROW(CF3HIP_ERR_NOT_CREATED, "system was created without a mesh: no element scatter map");
  if(prm.theta < 0. || prm.theta > 1.)
    CF3_THROW(CF3HIP_ERR_NOT_CREATED, "Value for theta option is outside of the valid range from 0 to 1.");
  if(!(prm.dt > 0.)) CF3_THROW(CF3HIP_ERR_INVALID_ARG, "dt must be positive");
  const Mesh& M = *L.mesh;
  if(F.mesh != L.mesh) CF3_THROW(CF3HIP_ERR_INVALID_ARG, "fields belong to another mesh");
  if(L.neq != M.dim + 1) CF3_THROW(CF3HIP_ERR_INVALID_ARG, "the NS system needs neq = dim + 1");
  const bool spec = prm.use_specializations && M.etype != CF3HIP_HEXA3D;
  if(!spec && prm.supg_type == CF3HIP_SUPG_CF2)
    CF3_THROW(CF3HIP_ERR_NOT_IMPLEMENTED,
              "supg_type cf2 divides by sum_n u.grad(N_n), identically zero by the partition of unity");
  if(M.etype == CF3HIP_HEXA3D)
    CF3_THROW(CF3HIP_ERR_NOT_IMPLEMENTED, "hexahedra assembly is provided by the Q1 kernel");
  AsmArgs g = {};
  for(int d = 0; d < 3; ++d) {
    g.nd.x[d] = d < M.dim ? M.x[d].as<double>() : nullptr;
    g.nd.u[d] = d < M.dim ? F.u[d].as<double>() : nullptr;
    g.nd.ua[d] = d < M.dim ? F.uadv[d].as<double>() : nullptr;
    g.nd.g[d] = d < M.dim ? F.g[d].as<double>() : nullptr;
  }
  g.nd.p = F.p.as<double>();
  g.nd.nu = F.nu.as<double>();
  g.conn = M.conn.as<uint32_t>();
  g.wg_rows = L.wg_rows.as<uint32_t>();
  g.row_ptr = L.row_ptr.as<uint32_t>();
  g.inc_ptr = L.inc_ptr.as<uint32_t>();
  g.inc = L.inc.as<uint32_t>();
  g.contrib_ptr = L.contrib_ptr.as<uint32_t>();
  g.contrib = L.contrib.as<uint16_t>();
  g.rho = prm.rho; g.dt = prm.dt; g.theta = prm.theta; g.invdt = 1. / prm.dt;
  g.supg_type = prm.supg_type;
  g.c1 = prm.c1; g.c2 = prm.c2; g.a_ps = prm.alpha_ps; g.a_su = prm.alpha_su; g.a_bu = prm.alpha_bu;
  g.status = L.ctx->status_flag();
  return g;
}

void check_status(Context& c)
{
  if(c.async_mode) return;
  int st = 0;
  HIP_OK(hipMemcpyAsync(&st, c.status_flag(), sizeof(int), hipMemcpyDeviceToHost, c.stream));
  HIP_OK(hipStreamSynchronize(c.stream));
  if(st) CF3_THROW(CF3HIP_ERR_SINGULAR, "non-positive element Jacobian determinant");
}

template<class Op>
void launch_residual(const AsmArgs& g, const Lss& L)
{
  Context& c = *L.ctx;
  timer_begin(c, CF3HIP_KERNEL_RESIDUAL);
  hipLaunchKernelGGL((k_residual<Op>), dim3(L.n_wg), dim3(kThreads), 0, c.stream, g);
  HIP_OK(hipGetLastError());
  timer_end(c, CF3HIP_KERNEL_RESIDUAL);
}

}  // namespace

void assemble(Lss& L, Fields& F, const cf3hip_ns_params& prm)
{
  AsmArgs g = make_args(L, F, prm);
  const Mesh& M = *L.mesh;
  if(prm.precision == CF3HIP_FP32 && !L.has_fp32)
    CF3_THROW(CF3HIP_ERR_NOT_CREATED, "FP32 assembly needs cf3hip_lss_enable_fp32");
  const bool spec = prm.use_specializations && M.etype != CF3HIP_HEXA3D;
  Context& c = *L.ctx;
  const bool fp32 = prm.precision == CF3HIP_FP32;
  // ZeroLSS: every value and rhs entry is overwritten by the kernel; the conditions of the previous
  // step are forgotten
  lss_clear_conditions(L);
  // the right-hand side is formed by the FP64 pass only (FP64 outer residual of the precision form)
  g.rhs = fp32 ? nullptr : L.rhs.as<double>();
  // element-validity flag: cleared per call in synchronous mode; in asynchronous mode it accumulates
  // and is reported by cf3hip_context_synchronize
  if(!c.async_mode) hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, c.stream, g.status);
  g.values = fp32 ? L.values32.ptr : L.values.ptr;

  if(spec && M.etype == CF3HIP_TETRA3D) {
    if(fp32) launch<SpecP1<3, float>, float, float>(g, L);
    else launch<SpecP1<3, double>, double, double>(g, L);
  } else if(spec && M.etype == CF3HIP_TRIAG2D) {
    if(fp32) launch<SpecP1<2, float>, float, float>(g, L);
    else launch<SpecP1<2, double>, double, double>(g, L);
  } else if(M.etype == CF3HIP_TETRA3D) {
    if(fp32) launch<GenericP1<3>, double, float>(g, L);
    else launch<GenericP1<3>, double, double>(g, L);
  } else {
    if(fp32) launch<GenericP1<2>, double, float>(g, L);
    else launch<GenericP1<2>, double, double>(g, L);
  }
  check_status(c);
}

void ns_residual(Lss& L, Fields& F, const cf3hip_ns_params& prm, const double* x, double* r)
{
  AsmArgs g = make_args(L, F, prm);
  const Mesh& M = *L.mesh;
  Context& c = *L.ctx;
  if(L.xtilde.bytes < sizeof(double) * L.dofs) L.xtilde.alloc(&c, sizeof(double) * L.dofs);
  hipLaunchKernelGGL(k_xtilde, dim3(grid_for(L.dofs, 256)), dim3(256), 0, c.stream, L.dofs, L.neq,
                     L.bc_n ? L.bc_mask.as<uint8_t>() : (const uint8_t*)nullptr,
                     L.bc_n ? L.bc_vfirst.as<double>() : (const double*)nullptr, x, L.xtilde.as<double>());
  g.xt = L.xtilde.as<double>();
  g.res = r;
  const bool spec = prm.use_specializations && M.etype != CF3HIP_HEXA3D;
  if(spec && M.etype == CF3HIP_TETRA3D) launch_residual<SpecP1<3, double> >(g, L);
  else if(spec && M.etype == CF3HIP_TRIAG2D) launch_residual<SpecP1<2, double> >(g, L);
  else if(M.etype == CF3HIP_TETRA3D) launch_residual<GenericP1<3> >(g, L);
  else launch_residual<GenericP1<2> >(g, L);
  if(L.bc_n)
    hipLaunchKernelGGL(k_bc_residual, dim3(grid_for(L.bc_n, 256)), dim3(256), 0, c.stream, L.bc_n,
                       L.bc_dof.as<uint32_t>(), L.bc_vlast.as<double>(), x, r);
  HIP_OK(hipGetLastError());
}

}  // namespace cf3hip

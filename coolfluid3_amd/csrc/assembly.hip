// UFEM Navier-Stokes element assembly fused with a deterministic, atomic-free scatter into the LSS
// block-CSR matrix (gfx950).
//
// Work decomposition (row owner): a kAsmThreads-thread workgroup owns a contiguous range of node rows
// and all their incidences (element e, local node a).  Each thread takes one incidence, gathers the
// element's nodes from packed node records (double4: coordinates, state (u.., p), advection velocity
// and viscosity), computes the element quantities it needs, then produces the element row block
// M(a, b) = invdt*T(a,b) + theta*A'(a,b) one column node b at a time.  Per pass b the blocks are
// staged in LDS at rows sorted by target CSR block (precomputed staging map); the accumulator threads
// of every CSR block (one, or a few for blocks with many contributions) sum its run of rows, and the
// partial sums are added in thread order when the block is written: a sorted segmented reduction
// with a fixed order, so results are deterministic and no floating-point atomics are used.  Every
// CSR value and rhs entry is written exactly once, the values with coalesced stores.
//
// Element mathematics (reference files relative to the coolfluid3 tree):
//   specialised P1 triangle / tetrahedron: plugins/UFEM/src/UFEM/NavierStokesSpecializations.hpp:38-303
//     (incl. the vk for wk substitution of lines 224-226 and 236-238 and tau from the velocity u)
//   generic expression (use_specializations = false): NavierStokesAssembly.hpp:44-66, order-2 Gauss
//   tau METRIC / TEZDUYAR: plugins/UFEM/src/UFEM/SUPG.hpp:141-219 (gij = J^-T J^-1 = Gram matrix of
//     the shape-function gradients of nodes 1..d for affine elements)
//   combine: NavierStokesAssembly.hpp:69-71 (rhs = -A x + a before the p rows are divided by theta;
//     theta * (A_p / theta) is formed as (theta * (1/theta)) * A_p, equal to within one rounding)
#include <algorithm>

#include "assembly_common.hpp"

// Multiply-adds are fused only within one source expression (a * b + c), never across statements:
// the rounding then follows the source, so a kernel body inlined into different kernels (the
// single-system and the batched persistent assembly) yields the same bits.
#pragma clang fp contract(on)

namespace cf3hip {

using namespace asmk;

namespace {

// ---------------------------------------------------------------------------------------------
// Specialised P1 simplex (D = 2 triangle, D = 3 tetrahedron) in arithmetic type R
// ---------------------------------------------------------------------------------------------
template<int D, typename R>
struct SpecP1 {
  static constexpr int K = D + 1, NEQ = D + 1, NN = NEQ * NEQ;
  uint32_t cn[K];
  R n[K][D];        // reference face normals (area vectors for D = 3, rotated edges for D = 2)
  R ua[D];          // mean advection velocity u_avg
  R na[D];          // normal of the row node a
  R Q[D];           // velocity-row coefficient: skew-symmetric advection + stress diffusion + bulk
  R cc, nudV, cpres, tpdV, pn, pp, tdo, tdd, tpc;
  int a;

  // geometry in type T: normals and volume
  template<typename T>
  __device__ __forceinline__ static void geometry(const T X[K][D], T nrm[K][D], T& vol)
  {
    if constexpr(D == 2) {
      nrm[0][0] = X[1][1] - X[2][1]; nrm[0][1] = X[2][0] - X[1][0];
      nrm[1][0] = X[2][1] - X[0][1]; nrm[1][1] = X[0][0] - X[2][0];
      nrm[2][0] = X[0][1] - X[1][1]; nrm[2][1] = X[1][0] - X[0][0];
      vol = T(0.5) * ((X[1][0] - X[0][0]) * (X[2][1] - X[0][1]) - (X[2][0] - X[0][0]) * (X[1][1] - X[0][1]));
    } else {
      // NavierStokesSpecializations.hpp:162-173
      nrm[0][0] = T(0.5) * ((X[3][1]-X[1][1])*(X[2][2]-X[1][2]) - (X[2][1]-X[1][1])*(X[3][2]-X[1][2]));
      nrm[0][1] = T(0.5) * ((X[3][2]-X[1][2])*(X[2][0]-X[1][0]) - (X[2][2]-X[1][2])*(X[3][0]-X[1][0]));
      nrm[0][2] = T(0.5) * ((X[3][0]-X[1][0])*(X[2][1]-X[1][1]) - (X[2][0]-X[1][0])*(X[3][1]-X[1][1]));
      nrm[1][0] = T(0.5) * ((X[2][1]-X[0][1])*(X[3][2]-X[0][2]) - (X[3][1]-X[0][1])*(X[2][2]-X[0][2]));
      nrm[1][1] = T(0.5) * ((X[2][2]-X[0][2])*(X[3][0]-X[0][0]) - (X[3][2]-X[0][2])*(X[2][0]-X[0][0]));
      nrm[1][2] = T(0.5) * ((X[2][0]-X[0][0])*(X[3][1]-X[0][1]) - (X[3][0]-X[0][0])*(X[2][1]-X[0][1]));
      nrm[2][0] = T(0.5) * ((X[3][1]-X[0][1])*(X[1][2]-X[0][2]) - (X[1][1]-X[0][1])*(X[3][2]-X[0][2]));
      nrm[2][1] = T(0.5) * ((X[3][2]-X[0][2])*(X[1][0]-X[0][0]) - (X[1][2]-X[0][2])*(X[3][0]-X[0][0]));
      nrm[2][2] = T(0.5) * ((X[3][0]-X[0][0])*(X[1][1]-X[0][1]) - (X[1][0]-X[0][0])*(X[3][1]-X[0][1]));
      nrm[3][0] = T(0.5) * ((X[1][1]-X[0][1])*(X[2][2]-X[0][2]) - (X[2][1]-X[0][1])*(X[1][2]-X[0][2]));
      nrm[3][1] = T(0.5) * ((X[1][2]-X[0][2])*(X[2][0]-X[0][0]) - (X[2][2]-X[0][2])*(X[1][0]-X[0][0]));
      nrm[3][2] = T(0.5) * ((X[1][0]-X[0][0])*(X[2][1]-X[0][1]) - (X[2][0]-X[0][0])*(X[1][1]-X[0][1]));
      // det J = 2 (x3 - x0) . n_3 ; V = det / 6
      vol = (T(2) * ((X[3][0]-X[0][0])*nrm[3][0] + (X[3][1]-X[0][1])*nrm[3][1] + (X[3][2]-X[0][2])*nrm[3][2]))
            * T(1.0 / 6.0);
    }
  }

  // METRIC tau from the velocity at the centroid (SUPG.hpp:176-190); gij = n_i.n_l / (D^2 V^2).
  // s = 1 / (D V).  c1 = 1, c2 = 16, alpha = 1 on the specialised path.
  template<typename T>
  __device__ __forceinline__ static void tau_metric(const T nrm[K][D], T s, const T uc[D], T nu, T f4dt2, T& t_su, T& t_bu)
  {
    T gn[D][D];   // grad N of nodes 1..D
    for(int i = 0; i < D; ++i)
      for(int d = 0; d < D; ++d) gn[i][d] = nrm[i + 1][d] * s;
    T adv = T(0);
    for(int d = 0; d < D; ++d) {
      T acc = T(0);
      for(int i = 0; i < D; ++i) acc += uc[i] * gn[i][d];
      adv += acc * acc;
    }
    T gsq = T(0), tr = T(0);
    for(int i = 0; i < D; ++i)
      for(int l = i; l < D; ++l) {
        T gil = T(0);
        for(int d = 0; d < D; ++d) gil += gn[i][d] * gn[l][d];
        gsq += (i == l ? T(1) : T(2)) * gil * gil;
        if(i == l) tr += gil;
      }
    const T sq = sqrt(f4dt2 + adv + T(16) * (nu * nu * gsq));
    t_su = T(1) / sq;
    t_bu = sq / tr;
  }

  __device__ __forceinline__ void prepare(const AsmArgs& g, uint32_t e, int a_)
  {
    load_conn<K>(g.conn, e, cn);
    prepare_nodes(g, g.xp, g.st, g.ad, a_);
  }

  // element quantities from the packed node records xp/st/ad (global, or the workgroup's LDS copy),
  // cn[] indexing them
  __device__ __forceinline__ void prepare_nodes(const AsmArgs& g, const double4* xpr, const double4* str,
                                                const double4* adr, int a_)
  {
    a = a_;
    double X[K][D];
    double uc[D];
    double nus = 0.;
    R uas[D];
    for(int d = 0; d < D; ++d) { uc[d] = 0.; uas[d] = R(0); }
#pragma unroll
    for(int m = 0; m < K; ++m) {
      double xp[4], st[4], ad[4];
      unpack(xpr[cn[m]], xp);
      unpack(str[cn[m]], st);
      unpack(adr[cn[m]], ad);
#pragma unroll
      for(int d = 0; d < D; ++d) {
        X[m][d] = xp[d];
        uc[d] += st[d];
        uas[d] += (R)ad[d];
      }
      nus += ad[3];
    }
    for(int d = 0; d < D; ++d) {
      uc[d] = uc[d] * (1.0 / K);
      ua[d] = uas[d] * R(1.0 / K);
    }
    const double nu_mean = fabs(nus * (1.0 / K));
    const R nu = (R)nu_mean;
    R V = R(0), invV, tsu, tbu;
    bool fp64_geom = sizeof(R) == sizeof(double);
    if constexpr(sizeof(R) != sizeof(double)) {
      // coordinates relative to node 0, formed in FP64: the FP32 differences of absolute coordinates
      // would lose log2(1/h) bits (the normals and the volume are translation invariant)
      R Xr[K][D];
      for(int m = 0; m < K; ++m)
        for(int d = 0; d < D; ++d) Xr[m][d] = (R)(X[m][d] - X[0][d]);
      geometry<R>(Xr, n, V);
      // strongly skewed elements: geometry and tau in FP64.  Criterion from the face normals already
      // formed, S = sum |n_i|^2: tetrahedra V^2 < 1.4e-3 S^(3/2) (regular Kuhn tetrahedron 1.5e-2;
      // aspect ratio (longest edge / smallest height) 15 and more falls below ~7e-4, aspect ratio 8
      // and less stays above ~2.7e-3); triangles V < S / 240 (rotated edges: |V| < 0.0125 l_rms^2).
      R S = R(0);
      for(int m = 0; m < K; ++m)
        for(int d = 0; d < D; ++d) S += n[m][d] * n[m][d];
      if constexpr(D == 3) {
        if(V * V < R(1.4e-3) * S * sqrt(S)) fp64_geom = true;
      } else {
        if(fabs(V) < S * R(1.0 / 240.0)) fp64_geom = true;
      }
    }
    if(fp64_geom) {
      double n64[K][D], V64, ts, tb;
      geometry<double>(X, n64, V64);
      const double iv = 1. / V64;
      tau_metric<double>(n64, iv * (1.0 / D), uc, nu_mean, g.f4dt2, ts, tb);
      for(int m = 0; m < K; ++m)
        for(int d = 0; d < D; ++d) n[m][d] = (R)n64[m][d];
      V = (R)V64; invV = (R)iv; tsu = (R)ts; tbu = (R)tb;
      if(!(V64 > 0.)) g.status[0] = 1;   // every writer stores the same flag value
      if constexpr(D == 3 && sizeof(R) == sizeof(double)) {
        // the element terms take the reference's volume (NavierStokesSpecializations.hpp:156), tau above
        // det J of the Jacobian, as the reference's ComputeTau does (SUPG.hpp:178, ElementData.hpp:302-314)
        const double Vr = tet_volume_ref(X);
        V = Vr;
        invV = 1. / Vr;
      }
    } else {
      R ucr[D];
      for(int d = 0; d < D; ++d) ucr[d] = (R)uc[d];
      invV = R(1) / V;
      tau_metric<R>(n, invV * R(1.0 / D), ucr, nu, (R)g.f4dt2, tsu, tbu);
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
    // per-row (node a) coefficients (NavierStokesSpecializations.hpp:66-145 / 196-301)
    const R c0 = R(D == 2 ? 1.0 / 6.0 : 1.0 / 12.0);
    const R rdV = invV * R(1.0 / (D * D));
    const R tsdV = tsu * rdV;
    nudV = nu * rdV;
    const R tbdV = tbu * rdV;
    R u_ni = R(0);
    for(int d = 0; d < D; ++d) u_ni += ua[d] * na[d];
    const R cskew = R(0.5) * c0 + R(0.5) * tsdV * u_ni;
    cc = c0 + tsdV * u_ni;
    const R inv_rho = (R)g.inv_rho;
    cpres = cc * inv_rho;
    const R time_ii = tsu * c0 * u_ni;
    // n_i . u~ with the tetrahedron substitution of vk for wk (lines 236-238)
    R ni_ut = na[0] * ua[0] + na[1] * ua[1];
    if constexpr(D == 3) ni_ut += na[2] * ua[1];
    const R visc = nudV * R(1.0 / 3.0) + tbdV;
#pragma unroll
    for(int r = 0; r < D; ++r) {
      const R uk = ua[(D == 3 && r == 2) ? 1 : r];   // W rows use vk (lines 224-226)
      Q[r] = cskew * uk + visc * na[r];
    }
    tpdV = tsdV;   // tau_ps = tau_su
    pn = R(0.5) * tpdV * ni_ut + c0;
    pp = tpdV * inv_rho;
    const R m0 = V * R(D == 2 ? 1.0 / 12.0 : 1.0 / 20.0);
    const R invdt = (R)g.invdt;
    tdo = invdt * (m0 + time_ii);
    tdd = invdt * (R(2) * m0 + time_ii);
    tpc = invdt * tsu * c0;
  }

  // block (a, B): M = invdt*T + theta*A' into m[NN] (row-major); with RHS, rhs -= A x_B
  template<int B, bool RHS, typename VR>
  __device__ __forceinline__ void block(const AsmArgs& g, VR* m, R* rhs, const R* xb) const
  {
    const R* nj = n[B];
    R uknj = R(0), ninj = R(0);
#pragma unroll
    for(int d = 0; d < D; ++d) { uknj += ua[d] * nj[d]; ninj += na[d] * nj[d]; }
    const R dsum = uknj * cc + nudV * ninj;   // convection (std + SUPG) + diffusion n_a.n_b
    const R td = a == B ? tdd : tdo;
    const R th = (R)g.theta, thq = (R)g.thq;
#pragma unroll
    for(int r = 0; r < D; ++r) {
      R Ar[NEQ];
#pragma unroll
      for(int c = 0; c < D; ++c) Ar[c] = Q[r] * nj[c] + (r == c ? dsum : R(0));
      Ar[D] = cpres * nj[r];
      if constexpr(RHS) {
        R s = R(0);
#pragma unroll
        for(int c = 0; c < NEQ; ++c) s += Ar[c] * xb[c];
        rhs[r] -= s;
      }
#pragma unroll
      for(int c = 0; c < D; ++c) m[r * NEQ + c] = (VR)(th * Ar[c] + (r == c ? td : R(0)));
      m[r * NEQ + D] = (VR)(th * Ar[D]);
    }
    {
      R Ar[NEQ];
#pragma unroll
      for(int c = 0; c < D; ++c) Ar[c] = uknj * (tpdV * na[c]) + pn * nj[c];
      Ar[D] = pp * ninj;
      if constexpr(RHS) {
        R s = R(0);
#pragma unroll
        for(int c = 0; c < NEQ; ++c) s += Ar[c] * xb[c];
        rhs[D] -= s;
      }
#pragma unroll
      for(int c = 0; c < D; ++c) m[D * NEQ + c] = (VR)(tpc * na[c] + thq * Ar[c]);
      m[D * NEQ + D] = (VR)(thq * Ar[D]);
    }
  }

  __device__ __forceinline__ void rhs_source(const AsmArgs&, R*) const {}
};

// ---------------------------------------------------------------------------------------------
// Generic expression for P1 simplices (use_specializations = false): NavierStokesAssembly.hpp:44-66
// with tau from u_adv and the order-2 Gauss rule; gradients constant over the element.
// ---------------------------------------------------------------------------------------------
template<int D>
struct GenericP1 {
  typedef double R;
  static constexpr int K = D + 1, NEQ = D + 1, NN = NEQ * NEQ, NQ = D == 2 ? 3 : 4;
  uint32_t cn[K];
  R G[D][K];         // physical gradients
  R Ga[D];           // gradient of the row node a
  R detw;            // |J| (weights applied per point)
  R ua_n[K][D];      // nodal advection velocity
  R nu_n[K];
  R tps, tsu, tbu;
  R oh[K];           // one-hot of the row node a
  int a;

  __device__ __forceinline__ static void gauss(int q, R xi[D], R& w)
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

  __device__ __forceinline__ static void shape(const R xi[D], R N[K])
  {
    R s = 1.;
#pragma unroll
    for(int d = 0; d < D; ++d) { N[d + 1] = xi[d]; s -= xi[d]; }
    N[0] = s;
  }

  // v[a] as a one-hot weighted sum: a select chain over a local array is turned into a dynamically
  // indexed (scratch) load by the optimiser
  __device__ __forceinline__ R select_a(const R v[K]) const
  {
    R r = 0.;
#pragma unroll
    for(int m = 0; m < K; ++m) r += oh[m] * v[m];
    return r;
  }

  __device__ __forceinline__ void prepare(const AsmArgs& g, uint32_t e, int a_)
  {
    a = a_;
#pragma unroll
    for(int m = 0; m < K; ++m) oh[m] = m == a ? 1. : 0.;
    load_conn<K>(g.conn, e, cn);
    R X[K][D];
    R nus = 0.;
#pragma unroll
    for(int m = 0; m < K; ++m) {
      double xp[4], ad[4];
      unpack(g.xp[cn[m]], xp);
      unpack(g.ad[cn[m]], ad);
#pragma unroll
      for(int d = 0; d < D; ++d) {
        X[m][d] = xp[d];
        ua_n[m][d] = ad[d];
      }
      nu_n[m] = ad[3];
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
    const R nu_e = fabs(nus * (1.0 / K));
    R uc[D];
#pragma unroll
    for(int d = 0; d < D; ++d) {
      R s = 0.;
#pragma unroll
      for(int m = 0; m < K; ++m) s += ua_n[m][d];
      uc[d] = s * (1.0 / K);
    }
    if(g.supg_type == CF3HIP_SUPG_CF2) {
      // volume as the reference's etype::volume: Tetra3D's 24-term expansion, Triag2D's 0.5 det J
      double vol;
      if constexpr(D == 3) {
        double Xv[4][3];
#pragma unroll
        for(int m = 0; m < K; ++m)
#pragma unroll
          for(int d = 0; d < D; ++d) Xv[m][d] = X[m][d];
        vol = tet_volume_ref(Xv);
      } else {
        vol = 0.5 * det;
      }
      R um2 = 0., s = 0.;
#pragma unroll
      for(int d = 0; d < D; ++d) um2 += uc[d] * uc[d];
#pragma unroll
      for(int m = 0; m < K; ++m) {
        R un = 0.;
#pragma unroll
        for(int d = 0; d < D; ++d) un += uc[d] * G[d][m];
        s += un;
      }
      if(!tau_cf2(D, vol, nu_e, sqrt(um2), s, g.dt, g.u_ref, tps, tsu, tbu)) g.status[0] = 2;
    } else if(g.supg_type == CF3HIP_SUPG_METRIC) {
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
      tsu = 1. / sqrt(g.c1 * g.c1 * g.f4dt2 + fabs(ugu) + g.c2 * diff);
      tps = tsu;
      tbu = sqrt(g.f4dt2 + fabs(ugu) + 16. * diff) / tr;
    } else {
      R h = 1e10;
#pragma unroll
      for(int i = 0; i < K; ++i)
#pragma unroll
        for(int j = i + 1; j < K; ++j) {
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

  template<int B, bool RHS, typename VR>
  __device__ __forceinline__ void block(const AsmArgs& g, VR* m, R* rhs, const R* xb) const
  {
    R A[NEQ][NEQ];
#pragma unroll
    for(int r = 0; r < NEQ; ++r)
#pragma unroll
      for(int c = 0; c < NEQ; ++c) A[r][c] = 0.;
    R Tuu = 0., Tp[D];   // T is diagonal in the velocity rows; T[p][u_i] in the pressure row
#pragma unroll
    for(int d = 0; d < D; ++d) Tp[d] = 0.;
    R gtg = 0.;
#pragma unroll
    for(int d = 0; d < D; ++d) gtg += Ga[d] * G[d][B];
    const R inv_rho = g.inv_rho;
#pragma unroll
    for(int q = 0; q < NQ; ++q) {
      R xi[D], wq, N[K];
      gauss(q, xi, wq);
      shape(xi, N);
      const R Na = select_a(N);
      const R w = wq * detw;
      R uq[D], nuq = 0.;
#pragma unroll
      for(int d = 0; d < D; ++d) {
        R s = 0.;
#pragma unroll
        for(int mm = 0; mm < K; ++mm) s += N[mm] * ua_n[mm][d];
        uq[d] = s;
      }
#pragma unroll
      for(int mm = 0; mm < K; ++mm) nuq += N[mm] * nu_n[mm];
      R La = 0., Lb = 0.;
#pragma unroll
      for(int d = 0; d < D; ++d) { La += uq[d] * Ga[d]; Lb += uq[d] * G[d][B]; }
      const R SUa = Na + tsu * La;
      const R Lps = 0.5 * tps * La;
#pragma unroll
      for(int i = 0; i < D; ++i) {
        A[D][i] += w * ((Na + Lps) * G[i][B] + (tps * Ga[i]) * Lb);
        A[i][i] += w * (nuq * gtg + SUa * Lb);
        A[i][D] += w * (SUa * G[i][B] * inv_rho);
#pragma unroll
        for(int j = 0; j < D; ++j) A[i][j] += w * ((tbu * Ga[i] + 0.5 * uq[i] * SUa) * G[j][B]);
        Tp[i] += w * (tps * Ga[i] * N[B]);
      }
      Tuu += w * (SUa * N[B]);
      A[D][D] += w * ((tps * gtg) * inv_rho);
    }
    if constexpr(RHS) {
#pragma unroll
      for(int r = 0; r < NEQ; ++r) {
        R s = 0.;
#pragma unroll
        for(int c = 0; c < NEQ; ++c) s += A[r][c] * xb[c];
        rhs[r] -= s;
      }
    }
    const R invdt = g.invdt, th = g.theta, thq = g.thq;
#pragma unroll
    for(int r = 0; r < D; ++r)
#pragma unroll
      for(int c = 0; c < NEQ; ++c)
        m[r * NEQ + c] = (VR)((r == c ? invdt * Tuu : 0.) + th * A[r][c]);
#pragma unroll
    for(int c = 0; c < D; ++c) m[D * NEQ + c] = (VR)(invdt * Tp[c] + thq * A[D][c]);
    m[D * NEQ + D] = (VR)(thq * A[D][D]);
  }

  // _a[u[_i]] += (N + tau_su u_adv grad N)^T g  (NavierStokesAssembly.hpp:65)
  __device__ __forceinline__ void rhs_source(const AsmArgs& g, R* rhs) const
  {
    R g_n[K][D];
#pragma unroll
    for(int m = 0; m < K; ++m)
#pragma unroll
      for(int d = 0; d < D; ++d) g_n[m][d] = g.g[d][cn[m]];
#pragma unroll
    for(int q = 0; q < NQ; ++q) {
      R xi[D], wq, N[K];
      gauss(q, xi, wq);
      shape(xi, N);
      const R Na = select_a(N);
      const R w = wq * detw;
      R uq[D], gq[D];
#pragma unroll
      for(int d = 0; d < D; ++d) {
        R s = 0., sg = 0.;
#pragma unroll
        for(int mm = 0; mm < K; ++mm) { s += N[mm] * ua_n[mm][d]; sg += N[mm] * g_n[mm][d]; }
        uq[d] = s; gq[d] = sg;
      }
      R La = 0.;
#pragma unroll
      for(int d = 0; d < D; ++d) La += uq[d] * Ga[d];
      const R SUa = Na + tsu * La;
#pragma unroll
      for(int i = 0; i < D; ++i) rhs[i] += w * (SUa * gq[i]);
    }
  }
};

// ---------------------------------------------------------------------------------------------
// Row-owner assembly kernel.  RHS = false (FP32 matrix pass of the precision form) skips the
// right-hand side, which is formed by the FP64 pass.
// ---------------------------------------------------------------------------------------------
// the work of one assembly workgroup (row range wg of the system described by g)
// STAGED (specialised simplices with node staging): the workgroup's range bounds come from one
// 32-byte header, the records of its distinct nodes are copied into LDS once, and every element
// gather and nodal state read of the passes is an LDS read (three dependent global loads per
// workgroup instead of five, and one global read per distinct node instead of one per incidence).
template<class Op, typename R, typename VR, bool RHS, bool STAGED = false>
__device__ __forceinline__ void assemble_wg(const AsmArgs& g, uint32_t wg)
{
  constexpr int K = Op::K, NEQ = Op::NEQ, NN = NEQ * NEQ;
  static_assert(K <= 4, "one staging byte per column local node in a 32-bit word");
  constexpr int LDW = lds_row<VR, NN>();
  constexpr size_t kBlkBytes = sizeof(VR) * kThreads * LDW;
  constexpr size_t kRhsBytes = sizeof(R) * kThreads * NEQ;
  __shared__ __align__(16) unsigned char s_raw[kBlkBytes > kRhsBytes ? kBlkBytes : kRhsBytes];
  VR* s_blk = reinterpret_cast<VR*>(s_raw);
  R* s_rhs = reinterpret_cast<R*>(s_raw);   // reused after the values are written
  __shared__ uint32_t s_rowinc[kThreads + 1];
  __shared__ uint16_t s_acc[kThreads];

  uint32_t r0, r1, i0, ninc, blk0, nblk, n0 = 0, nn = 0;
  if constexpr(STAGED) {
    const uint4 h0 = reinterpret_cast<const uint4*>(g.wg_head)[2 * wg];
    const uint4 h1 = reinterpret_cast<const uint4*>(g.wg_head)[2 * wg + 1];
    r0 = h0.x; r1 = h0.y; i0 = h0.z; ninc = h0.w - h0.z;
    blk0 = h1.x; nblk = h1.y - h1.x; n0 = h1.z; nn = h1.w - h1.z;
  } else {
    r0 = g.wg_rows[wg]; r1 = g.wg_rows[wg + 1];
    i0 = g.inc_ptr[r0]; ninc = g.inc_ptr[r1] - i0;
    blk0 = g.row_ptr[r0]; nblk = g.row_ptr[r1] - blk0;
  }
  const uint32_t nrows = r1 - r0;
  const int t = threadIdx.x;
  // node staging area (dynamic LDS): xp records [0, nn), st [nn, 2 nn), ad [2 nn, 3 nn)
  extern __shared__ double4 s_nodes[];
  if constexpr(STAGED)
    for(uint32_t j = t; j < nn; j += kThreads) {
      const uint32_t nd = g.wg_nodes[n0 + j];
      s_nodes[j] = g.xp[nd];
      s_nodes[nn + j] = g.st[nd];
      s_nodes[2 * nn + j] = g.ad[nd];
    }
  if constexpr(RHS)
    for(uint32_t j = t; j <= nrows; j += kThreads) s_rowinc[j] = g.inc_ptr[r0 + j] - i0;
  if((uint32_t)t < nblk) s_acc[t] = g.blk_acc[blk0 + t];

  const bool active = (uint32_t)t < ninc;
  uint32_t posw = 0;
  Op op;
  R rhs[NEQ];
#pragma unroll
  for(int r = 0; r < NEQ; ++r) rhs[r] = R(0);
  if constexpr(STAGED) {
    uint32_t ea = 0, loc = 0;
    if(active) {
      posw = reinterpret_cast<const uint32_t*>(g.inc_pos)[i0 + t];
      ea = g.inc[i0 + t];
      loc = g.inc_loc[i0 + t];
    }
    __syncthreads();   // node records staged
    if(active) {
#pragma unroll
      for(int m = 0; m < K; ++m) op.cn[m] = (loc >> (8 * m)) & 255u;
      op.prepare_nodes(g, s_nodes, s_nodes + nn, s_nodes + 2 * nn, (int)(ea % K));
    }
  } else {
    if(active) {
      posw = reinterpret_cast<const uint32_t*>(g.inc_pos)[i0 + t];
      const uint32_t ea = g.inc[i0 + t];
      op.prepare(g, ea / K, (int)(ea % K));
    }
    __syncthreads();
  }
  // accumulator thread t: block bs of this workgroup, rank h of c; rows [beg_p + h, end_p) step c
  const uint32_t alast = s_acc[nblk - 1];
  const uint32_t n_acc = (alast & 255u) + (alast >> 8);
  const bool owner = (uint32_t)t < n_acc;
  uint32_t begw = 0, endw = 0, h = 0, c = 1;
  if(owner) {
    uint32_t lo = 0, hi = nblk;
    while(hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if((s_acc[mid] & 255u) <= (uint32_t)t) lo = mid; else hi = mid;
    }
    h = (uint32_t)t - (s_acc[lo] & 255u);
    c = s_acc[lo] >> 8;
    const uint32_t* bb = reinterpret_cast<const uint32_t*>(g.blk_beg);
    begw = bb[blk0 + lo];
    endw = lo + 1 < nblk ? bb[blk0 + lo + 1] : ninc * 0x01010101u;
  }
  R acc[NN];
#pragma unroll
  for(int v = 0; v < NN; ++v) acc[v] = R(0);

  static_for<0, K>([&](auto bc) {
    constexpr int B = decltype(bc)::value;
    if(active) {
      R xb[NEQ];
      if constexpr(RHS) load_state<NEQ, R>(STAGED ? s_nodes + nn : g.st, op.cn[B], xb);
      op.template block<B, RHS, VR>(g, s_blk + ((posw >> (8 * B)) & 255u) * LDW, rhs, xb);
    }
    __syncthreads();
    if(owner) {
      const uint32_t je = (endw >> (8 * B)) & 255u;
      for(uint32_t j = ((begw >> (8 * B)) & 255u) + h; j < je; j += c) {
        VR v[NN];
        lds_load_row<VR, NN>(s_blk + j * LDW, v);
#pragma unroll
        for(int q = 0; q < NN; ++q) acc[q] += (R)v[q];
      }
    }
    __syncthreads();
  });
  // partial sums to LDS row t, then each block = sum of its accumulator rows in thread order,
  // stored with consecutive lanes on consecutive values
  if(owner) {
    VR* d = s_blk + t * LDW;
#pragma unroll
    for(int q = 0; q < NN; ++q) d[q] = (VR)acc[q];
  }
  __syncthreads();
  {
    VR* dst = static_cast<VR*>(g.values) + (uint64_t)blk0 * NN;
    for(uint32_t o = t; o < nblk * NN; o += kThreads) {
      const uint32_t sb = o / NN, v = o % NN;
      const uint32_t a = s_acc[sb], f = a & 255u, na = a >> 8;
      R sum = (R)s_blk[f * LDW + v];
      for(uint32_t r = 1; r < na; ++r) sum += (R)s_blk[(f + r) * LDW + v];
      dst[o] = (VR)sum;
    }
  }
  if constexpr(RHS) {
    __syncthreads();
    if(active) {
      if(g.body_force) op.rhs_source(g, rhs);
#pragma unroll
      for(int r = 0; r < NEQ; ++r) s_rhs[t * NEQ + r] = rhs[r];
    }
    __syncthreads();
    for(uint32_t o = t; o < nrows * NEQ; o += kThreads) {
      const uint32_t row = o / NEQ, r = o % NEQ;
      const uint32_t qe = s_rowinc[row + 1];
      uint32_t q = s_rowinc[row];
      // four interleaved partial sums (fixed order) shorten the LDS latency chain
      R s0 = R(0), s1 = R(0), s2 = R(0), s3 = R(0);
      for(; q + 4 <= qe; q += 4) {
        s0 += s_rhs[q * NEQ + r];
        s1 += s_rhs[(q + 1) * NEQ + r];
        s2 += s_rhs[(q + 2) * NEQ + r];
        s3 += s_rhs[(q + 3) * NEQ + r];
      }
      for(; q < qe; ++q) s0 += s_rhs[q * NEQ + r];
      g.rhs[(uint64_t)(r0 + row) * NEQ + r] = (double)((s0 + s1) + (s2 + s3));
    }
  }
}

// FP32 kernels: at most 96 VGPRs, so that 5 waves fit per SIMD (104 would allow 4)
template<class Op, typename R, typename VR, bool RHS, bool STAGED>
__global__ __launch_bounds__(kThreads, sizeof(R) == sizeof(float) ? 5 : 1) void k_assemble(AsmArgs g)
{
  assemble_wg<Op, R, VR, RHS, STAGED>(g, xcd_order(blockIdx.x, gridDim.x));
}

// ---------------------------------------------------------------------------------------------
// Batched FP32 assembly of many tetrahedral systems (serving): one persistent launch whose
// workgroups walk the concatenated workgroup lists of the systems in a device table.
// ---------------------------------------------------------------------------------------------
struct PackArgs {
  const double* u[3];
  const double* ua[3];
  const double* p;
  const double* nu;
  double4* st;
  double4* ad;
};

struct BatchEntry {
  AsmArgs g;
  PackArgs pk;
  uint32_t wg_begin, node_begin, n_nodes, pad;
};

struct BatchHeader {
  uint32_t n_cases, total_wg, total_nodes, pad;
};

__device__ __forceinline__ uint32_t batch_find(const BatchEntry* e, uint32_t n, uint32_t key, bool by_wg)
{
  uint32_t lo = 0, hi = n;   // last entry with begin <= key
  while(hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if((by_wg ? e[mid].wg_begin : e[mid].node_begin) <= key) lo = mid; else hi = mid;
  }
  return lo;
}

__global__ void k_batch_pack(const BatchHeader* __restrict__ hd, const BatchEntry* __restrict__ e)
{
  const uint32_t total = hd->total_nodes, n = hd->n_cases;
  for(uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < total; j += gridDim.x * blockDim.x) {
    const uint32_t c = batch_find(e, n, j, false);
    const uint32_t i = j - e[c].node_begin;
    const PackArgs& a = e[c].pk;
    a.st[i] = make_double4(a.u[0][i], a.u[1][i], a.u[2][i], a.p[i]);
    a.ad[i] = make_double4(a.ua[0][i], a.ua[1][i], a.ua[2][i], a.nu[i]);
  }
}

__global__ __launch_bounds__(kThreads) void k_batch_assemble(const BatchHeader* __restrict__ hd,
                                                              const BatchEntry* __restrict__ e)
{
  const uint32_t total = hd->total_wg, n = hd->n_cases;
  for(uint32_t task = blockIdx.x; task < total; task += gridDim.x) {
    __syncthreads();   // the previous task's LDS reads are done
    const uint32_t c = batch_find(e, n, task, true);
    assemble_wg<SpecP1<3, float>, float, float, false>(e[c].g, task - e[c].wg_begin);
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
  const uint32_t wg = xcd_order(blockIdx.x, gridDim.x);
  const uint32_t r0 = g.wg_rows[wg], r1 = g.wg_rows[wg + 1];
  const uint32_t nrows = r1 - r0;
  const uint32_t i0 = g.inc_ptr[r0];
  const uint32_t ninc = g.inc_ptr[r1] - i0;
  const int t = threadIdx.x;
  for(uint32_t j = t; j <= nrows; j += kThreads) s_rowinc[j] = g.inc_ptr[r0 + j] - i0;
  const bool active = (uint32_t)t < ninc;
  Op op;
  double rhs[NEQ], y[NEQ];
#pragma unroll
  for(int r = 0; r < NEQ; ++r) { rhs[r] = 0.; y[r] = 0.; }
  if(active) {
    const uint32_t ea = g.inc[i0 + t];
    op.prepare(g, ea / K, (int)(ea % K));
    static_for<0, K>([&](auto bc) {
      constexpr int B = decltype(bc)::value;
      double xb[NEQ], xt[NEQ], mb[NN];
      load_state<NEQ, double>(g, op.cn[B], xb);
#pragma unroll
      for(int c = 0; c < NEQ; ++c) xt[c] = g.xt[(uint64_t)op.cn[B] * NEQ + c];
      op.template block<B, true, double>(g, mb, rhs, xb);
#pragma unroll
      for(int r = 0; r < NEQ; ++r)
#pragma unroll
        for(int c = 0; c < NEQ; ++c) y[r] += mb[r * NEQ + c] * xt[c];
    });
    if(g.body_force) op.rhs_source(g, rhs);
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

// packed node records
__global__ void k_pack_coords(uint32_t n, int dim, const double* __restrict__ x, const double* __restrict__ y,
                              const double* __restrict__ z, double4* __restrict__ xp)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  xp[i] = make_double4(x[i], y[i], dim == 3 ? z[i] : 0., 0.);
}

template<int D>
__global__ void k_pack_fields(uint32_t n, PackArgs a)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  if constexpr(D == 3) {
    a.st[i] = make_double4(a.u[0][i], a.u[1][i], a.u[2][i], a.p[i]);
    a.ad[i] = make_double4(a.ua[0][i], a.ua[1][i], a.ua[2][i], a.nu[i]);
  } else {
    a.st[i] = make_double4(a.u[0][i], a.u[1][i], a.p[i], 0.);
    a.ad[i] = make_double4(a.ua[0][i], a.ua[1][i], 0., a.nu[i]);
  }
}

void pack_fields(Fields& F)
{
  const Mesh& M = *F.mesh;
  Context& c = *M.ctx;
  PackArgs a = {};
  for(int d = 0; d < M.dim; ++d) {
    a.u[d] = F.u[d].as<double>();
    a.ua[d] = F.uadv[d].as<double>();
  }
  a.p = F.p.as<double>();
  a.nu = F.nu.as<double>();
  a.st = F.st.as<double4>();
  a.ad = F.ad.as<double4>();
  if(M.dim == 3) hipLaunchKernelGGL(k_pack_fields<3>, dim3(grid_for(M.n_nodes, 256)), dim3(256), 0, c.stream, M.n_nodes, a);
  else hipLaunchKernelGGL(k_pack_fields<2>, dim3(grid_for(M.n_nodes, 256)), dim3(256), 0, c.stream, M.n_nodes, a);
  HIP_OK(hipGetLastError());
}

template<class Op, typename R, typename VR, bool RHS>
void launch(const AsmArgs& g, const Lss& L)
{
  Context& c = *L.ctx;
  timer_begin(c, CF3HIP_KERNEL_ASSEMBLY);
  hipLaunchKernelGGL((k_assemble<Op, R, VR, RHS, false>), dim3(L.n_wg), dim3(kThreads), 0, c.stream, g);
  HIP_OK(hipGetLastError());
  timer_end(c, CF3HIP_KERNEL_ASSEMBLY);
}

// specialised simplex kernels: the FP64 kernel is node-staged when the system has the staging
// structures (measured at 64^3 tetrahedra: 0.437-0.448 ms against 0.451-0.463 ms unstaged, 144
// instead of 164 VGPRs).  The FP32 kernel stays unstaged: it already runs 5 waves per SIMD, and the
// staging copy and its barrier cost more than the gathers it saves (0.260 ms against 0.237-0.249 ms).
template<class Op, typename R, typename VR, bool RHS>
void launch_spec(const AsmArgs& g, const Lss& L)
{
  if constexpr(sizeof(R) != sizeof(double)) {
    launch<Op, R, VR, RHS>(g, L);
  } else {
    if(!L.staged) {
      launch<Op, R, VR, RHS>(g, L);
      return;
    }
    Context& c = *L.ctx;
    const size_t lds = sizeof(double4) * 3 * (size_t)L.wg_nodes_max;
    timer_begin(c, CF3HIP_KERNEL_ASSEMBLY);
    hipLaunchKernelGGL((k_assemble<Op, R, VR, RHS, true>), dim3(L.n_wg), dim3(kThreads), lds, c.stream, g);
    HIP_OK(hipGetLastError());
    timer_end(c, CF3HIP_KERNEL_ASSEMBLY);
  }
}

__global__ void k_clear_flag(int* f) { f[0] = 0; }

AsmArgs make_args(Lss& L, Fields& F, const cf3hip_ns_params& prm, bool pack = true)
{
  if(!L.mesh) CF3_THROW(CF3HIP_ERR_NOT_CREATED, "system was created without a mesh: no element scatter map");
  if(prm.theta < 0. || prm.theta > 1.)
    CF3_THROW(CF3HIP_ERR_NOT_CREATED, "Value for theta option is outside of the valid range from 0 to 1.");
  if(!(prm.dt > 0.)) CF3_THROW(CF3HIP_ERR_INVALID_ARG, "dt must be positive");
  const Mesh& M = *L.mesh;
  if(F.mesh != L.mesh) CF3_THROW(CF3HIP_ERR_INVALID_ARG, "fields belong to another mesh");
  if(L.neq != M.dim + 1) CF3_THROW(CF3HIP_ERR_INVALID_ARG, "the NS system needs neq = dim + 1");
  const bool spec = prm.use_specializations && !is_cell(M.etype);
  if(!spec && prm.supg_type == CF3HIP_SUPG_CF2 && !(prm.u_ref > 0.))
    CF3_THROW(CF3HIP_ERR_INVALID_ARG, "supg_type cf2 needs a positive u_ref (SUPG.hpp:196)");
  AsmArgs g = {};
  g.body_force = F.has_g ? 1 : 0;
  g.xp = M.xp.as<double4>();
  g.st = F.st.as<double4>();
  g.ad = F.ad.as<double4>();
  for(int d = 0; d < 3; ++d) g.g[d] = d < M.dim ? F.g[d].as<double>() : nullptr;
  g.conn = M.conn.as<uint32_t>();
  g.wg_rows = L.wg_rows.as<uint32_t>();
  g.row_ptr = L.row_ptr.as<uint32_t>();
  g.inc_ptr = L.inc_ptr.as<uint32_t>();
  g.inc = L.inc.as<uint32_t>();
  g.inc_pos = L.inc_pos.as<uint8_t>();
  g.blk_beg = L.blk_beg.as<uint8_t>();
  g.blk_acc = L.blk_acc.as<uint16_t>();
  if(L.staged) {
    g.wg_head = L.wg_head.as<uint32_t>();
    g.wg_nodes = L.wg_nodes.as<uint32_t>();
    g.inc_loc = L.inc_loc.as<uint32_t>();
  }
  g.rho = prm.rho; g.inv_rho = 1. / prm.rho;
  g.dt = prm.dt; g.theta = prm.theta; g.thq = prm.theta * (1. / prm.theta);
  g.invdt = 1. / prm.dt; g.f4dt2 = 4. / (prm.dt * prm.dt);
  g.supg_type = prm.supg_type;
  g.c1 = prm.c1; g.c2 = prm.c2; g.a_ps = prm.alpha_ps; g.a_su = prm.alpha_su; g.a_bu = prm.alpha_bu;
  g.u_ref = prm.u_ref;
  g.status = L.ctx->status_flag();
  if(pack) pack_fields(F);
  return g;
}

void check_status(Context& c)
{
  if(c.async_mode) return;
  int st = 0;
  HIP_OK(hipMemcpyAsync(&st, c.status_flag(), sizeof(int), hipMemcpyDeviceToHost, c.stream));
  HIP_OK(hipStreamSynchronize(c.stream));
  if(st == 2) CF3_THROW(CF3HIP_ERR_INVALID_ARG, "ComputeTau cf2: element Reynolds number u_ref he / (2 nu) not positive");
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

void mesh_pack_coords(Mesh& M)
{
  Context& c = *M.ctx;
  if(!M.xp.ptr) M.xp.alloc(&c, sizeof(double4) * M.n_nodes);
  hipLaunchKernelGGL(k_pack_coords, dim3(grid_for(M.n_nodes, 256)), dim3(256), 0, c.stream, M.n_nodes, M.dim,
                     M.x[0].as<double>(), M.x[1].as<double>(), M.dim == 3 ? M.x[2].as<double>() : nullptr,
                     M.xp.as<double4>());
  HIP_OK(hipGetLastError());
}

void assemble(Lss& L, Fields& F, const cf3hip_ns_params& prm)
{
  AsmArgs g = make_args(L, F, prm);
  const Mesh& M = *L.mesh;
  if(prm.precision == CF3HIP_FP32 && !L.has_fp32)
    CF3_THROW(CF3HIP_ERR_NOT_CREATED, "FP32 assembly needs cf3hip_lss_enable_fp32");
  const bool spec = prm.use_specializations && !is_cell(M.etype);
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

  if(is_cell(M.etype)) {
    launch_hex(g, L, fp32 ? 1 : 0);
  } else if(spec && M.etype == CF3HIP_TETRA3D) {
    if(fp32) launch_spec<SpecP1<3, float>, float, float, false>(g, L);
    else launch_spec<SpecP1<3, double>, double, double, true>(g, L);
  } else if(spec && M.etype == CF3HIP_TRIAG2D) {
    if(fp32) launch_spec<SpecP1<2, float>, float, float, false>(g, L);
    else launch_spec<SpecP1<2, double>, double, double, true>(g, L);
  } else if(M.etype == CF3HIP_TETRA3D) {
    if(fp32) launch<GenericP1<3>, double, float, false>(g, L);
    else launch<GenericP1<3>, double, double, true>(g, L);
  } else {
    if(fp32) launch<GenericP1<2>, double, float, false>(g, L);
    else launch<GenericP1<2>, double, double, true>(g, L);
  }
  check_status(c);
}

void ns_residual(Lss& L, Fields& F, const cf3hip_ns_params& prm, const double* x, double* r)
{
  if(L.periodic)
    CF3_THROW(CF3HIP_ERR_NOT_IMPLEMENTED, "the matrix-free residual (precision form) takes systems without periodic links");
  AsmArgs g = make_args(L, F, prm);
  const Mesh& M = *L.mesh;
  Context& c = *L.ctx;
  if(L.xtilde.bytes < sizeof(double) * L.dofs) L.xtilde.alloc(&c, sizeof(double) * L.dofs);
  hipLaunchKernelGGL(k_xtilde, dim3(grid_for(L.dofs, 256)), dim3(256), 0, c.stream, L.dofs, L.neq,
                     L.bc_n ? L.bc_mask.as<uint8_t>() : (const uint8_t*)nullptr,
                     L.bc_n ? L.bc_vfirst.as<double>() : (const double*)nullptr, x, L.xtilde.as<double>());
  g.xt = L.xtilde.as<double>();
  g.res = r;
  const bool spec = prm.use_specializations && !is_cell(M.etype);
  if(is_cell(M.etype)) launch_hex(g, L, 2);
  else if(spec && M.etype == CF3HIP_TETRA3D) launch_residual<SpecP1<3, double> >(g, L);
  else if(spec && M.etype == CF3HIP_TRIAG2D) launch_residual<SpecP1<2, double> >(g, L);
  else if(M.etype == CF3HIP_TETRA3D) launch_residual<GenericP1<3> >(g, L);
  else launch_residual<GenericP1<2> >(g, L);
  if(L.bc_n)
    hipLaunchKernelGGL(k_bc_residual, dim3(grid_for(L.bc_n, 256)), dim3(256), 0, c.stream, L.bc_n,
                       L.bc_dof.as<uint32_t>(), L.bc_vlast.as<double>(), x, r);
  HIP_OK(hipGetLastError());
}

// ---------------------------------------------------------------------------------------------
// batched FP32 assembly (serving): device table + persistent launch captured once in a hipGraph
// ---------------------------------------------------------------------------------------------
namespace {
constexpr unsigned kBatchPackGrid = 1024;
constexpr unsigned kBatchAsmGrid = 2048;   // persistent workgroups (8 per CU)
constexpr size_t kBatchHeaderBytes = 256;
}  // namespace

size_t batch_table_bytes(uint32_t max_cases) { return kBatchHeaderBytes + sizeof(BatchEntry) * max_cases; }

void batch_assemble(Batch& B, uint32_t n, Lss* const* L, Fields* const* F, const cf3hip_ns_params* prm)
{
  if(n > B.max_cases) CF3_THROW(CF3HIP_ERR_LIMIT, "more systems than the batch was created for");
  BatchHeader* hd = reinterpret_cast<BatchHeader*>(B.host_table);
  BatchEntry* ent = reinterpret_cast<BatchEntry*>(static_cast<char*>(B.host_table) + kBatchHeaderBytes);
  uint32_t wg = 0, nodes = 0;
  for(uint32_t i = 0; i < n; ++i) {
    Lss& Li = *L[i];
    Fields& Fi = *F[i];
    if(!Li.mesh || Li.mesh->etype != CF3HIP_TETRA3D || !prm[i].use_specializations)
      CF3_THROW(CF3HIP_ERR_NOT_IMPLEMENTED, "batched assembly takes tetrahedral systems with the specialised kernel");
    if(!Li.has_fp32) CF3_THROW(CF3HIP_ERR_NOT_CREATED, "batched FP32 assembly needs cf3hip_lss_enable_fp32");
    if(Li.periodic) CF3_THROW(CF3HIP_ERR_NOT_IMPLEMENTED, "batched assembly takes systems without periodic links");
    if((uint64_t)wg + Li.n_wg > 0xffffffffull || (uint64_t)nodes + Li.n_nodes > 0xffffffffull)
      CF3_THROW(CF3HIP_ERR_LIMIT, "batch too large");
    // ZeroLSS: forget the conditions of the previous step (mask cleared on the batch stream)
    Li.bc_host.clear();
    Li.bc_order.clear();
    Li.bc_n = 0;
    HIP_OK(hipStreamSynchronize(Li.ctx->stream));
    if(Li.bc_mask.ptr) HIP_OK(hipMemsetAsync(Li.bc_mask.ptr, 0, Li.n_nodes, B.stream));
    HIP_OK(hipMemsetAsync(Li.ctx->status_flag(), 0, sizeof(int), B.stream));   // element-validity flag
    AsmArgs g = make_args(Li, Fi, prm[i], false);
    g.values = Li.values32.ptr;
    g.rhs = nullptr;
    ent[i].g = g;
    PackArgs& pk = ent[i].pk;
    for(int d = 0; d < 3; ++d) {
      pk.u[d] = Fi.u[d].as<double>();
      pk.ua[d] = Fi.uadv[d].as<double>();
    }
    pk.p = Fi.p.as<double>();
    pk.nu = Fi.nu.as<double>();
    pk.st = Fi.st.as<double4>();
    pk.ad = Fi.ad.as<double4>();
    ent[i].wg_begin = wg;
    ent[i].node_begin = nodes;
    ent[i].n_nodes = Li.n_nodes;
    wg += Li.n_wg;
    nodes += Li.n_nodes;
  }
  hd->n_cases = n;
  hd->total_wg = n ? wg : 0;
  hd->total_nodes = n ? nodes : 0;
  HIP_OK(hipMemcpyAsync(B.table.ptr, B.host_table, kBatchHeaderBytes + sizeof(BatchEntry) * (n ? n : 1),
                        hipMemcpyHostToDevice, B.stream));
  if(!B.exec) {
    // the two launches read the table from device memory, so one capture serves every later batch
    const BatchHeader* dh = B.table.as<BatchHeader>();
    const BatchEntry* de = reinterpret_cast<const BatchEntry*>(B.table.as<char>() + kBatchHeaderBytes);
    hipGraph_t graph = nullptr;
    HIP_OK(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
    hipLaunchKernelGGL(k_batch_pack, dim3(kBatchPackGrid), dim3(256), 0, B.stream, dh, de);
    hipLaunchKernelGGL(k_batch_assemble, dim3(kBatchAsmGrid), dim3(kThreads), 0, B.stream, dh, de);
    HIP_OK(hipStreamEndCapture(B.stream, &graph));
    HIP_OK(hipGraphInstantiate(&B.exec, graph, nullptr, nullptr, 0));
    B.graph = graph;
  }
  HIP_OK(hipGraphLaunch(B.exec, B.stream));
  HIP_OK(hipStreamSynchronize(B.stream));
  for(uint32_t i = 0; i < n; ++i) check_status(*L[i]->ctx);
}

}  // namespace cf3hip

// Cell assembly (Q1 hexahedra and quadrilaterals, P1 prisms): the generic Navier-Stokes expression,
// which these cells always take (plugins/UFEM/src/UFEM/ns_implementation/NavierStokesHexas.cpp:20-25,
// NavierStokesPrisms.cpp:20-26), with the second-order Gauss rule of the cell (ElementIntegration.hpp:
// 70-84, 247: order 2 for linear shape functions): 2^D points on Q1 cells, 6 on prisms.
//
// Element expression: plugins/UFEM/src/UFEM/NavierStokesAssembly.hpp:44-66 (tau from u_adv, :54);
// tau METRIC / TEZDUYAR at the element centre: SUPG.hpp:141-219; Q1 shape functions and gradients:
// cf3/mesh/LagrangeP1/Hexa.cpp:24-89 (node order :103-118) and Quad.cpp (compute_value,
// compute_gradient), prism shape functions cf3/mesh/LagrangeP1/Prism.cpp:24-58; Gauss rule:
// GaussImplementation.hpp:29-35 (x = 1/sqrt(3), w = 1), :566-600 (quad point order), :615-676
// (hexahedron point order) and :456-494 (prism: the triangle's edge midpoints at zeta = +x, then -x,
// w = 1/6); combine: NavierStokesAssembly.hpp:69-71.
//
// Work decomposition: a pre-pass (k_q1_rec, one thread per element and Gauss point; k_q1_tau, one thread
// per element) computes each element's Gauss-point records (J^-1, det J, u_adv, nu at each of its NQ
// points) and its tau once, into device buffers of the system.  The row-owner kernel (k_q1, one staging
// pass) then gives each incidence (element e, local node a) of its rows kLanes threads (2^D on Q1
// cells, 8 on prisms, whose lanes 6 and 7 idle): lane q reads the record of Gauss point q, forms the
// row-node terms and stores them in LDS; lane b then integrates the element block M(a, b) over the
// NQ points.
// Blocks are staged in LDS at rows sorted by target CSR block and summed by the block's accumulator
// threads in a fixed order: deterministic, no floating-point atomics.  (The element geometry is thus
// computed once per element rather than once per incidence, i.e. K times.)
//
// Per Gauss point the element block terms are (w = w_q det J_q, G = grad N, L = u_adv . G,
// SU = N + tau_su L):
//   A[p][u_i]   += w ((N_a + tau_ps L_a / 2) G_b,i + tau_ps G_a,i L_b)
//   A[p][p]     += w tau_ps G_a.G_b / rho
//   A[u_i][u_i] += w (nu G_a.G_b + SU_a L_b)
//   A[u_i][p]   += w SU_a G_b,i / rho
//   A[u_i][u_j] += w (tau_bulk G_a,i + u_i SU_a / 2) G_b,j
//   T[p][u_i]   += w tau_ps G_a,i N_b,   T[u_i][u_i] += w SU_a N_b,   a[u_i] += w SU_a g_i
#include "assembly_common.hpp"

namespace cf3hip {
namespace asmk {

namespace {

constexpr double kGauss = 0.5773502691896257645091488;

enum { MODE_FP64 = 0, MODE_FP32 = 1, MODE_RESIDUAL = 2 };

// reference coordinate signs of the Q1 nodes: (-,-,-), (+,-,-), (+,+,-), (-,+,-), then z = +
__device__ __forceinline__ double node_sx(int m) { return ((m ^ (m >> 1)) & 1) ? 1. : -1.; }
__device__ __forceinline__ double node_sy(int m) { return (m & 2) ? 1. : -1.; }
__device__ __forceinline__ double node_sz(int m) { return (m & 4) ? 1. : -1.; }

// Cell types.  Each gives D, its nodes K, its Gauss points NQ (= K here), the lanes per incidence, the
// Gauss points and weights, the shape functions through a per-node record (Node) computed once per
// node, and the shape-function gradients at the point where tau is evaluated (the one-point rule's
// point, SUPG.hpp:145), where every N_m = 1/K.
template<int D_>
struct Q1 {
  static constexpr int D = D_, K = 1 << D, NQ = K, NEQ = D + 1, NN = NEQ * NEQ;
  static constexpr int kLanes = K;                    // threads per incidence: Gauss point q, then column node b
  static constexpr int kSlots = kThreads / kLanes;    // incidences per workgroup
  static constexpr int kRec = D * D + 2 * D + 3;      // doubles per (incidence, Gauss point) record
  static constexpr int kRecSlot = kLanes * kRec + 2;  // doubles per incidence (padded: slots on distinct LDS banks)
  static constexpr int kQRec = D == 3 ? 16 : 8;       // doubles per (element, Gauss point) in the record buffer:
                                                      // J^-1 (D x D), det J, u_adv (D), nu, padding
  static constexpr double kInvK = 1.0 / K;
  static constexpr int kUnroll = 2;                   // Gauss-point loop of the block integration
  // Gauss point q, times 1/sqrt(3): hexahedra (+,+,+), (-,+,+), (+,-,+), (+,+,-), (+,-,-), (-,+,-), (-,-,+),
  // (-,-,-); quadrilaterals (+,+), (-,+), (+,-), (-,-); weight 1
  __device__ static void gauss(int q, double x[3])
  {
    if constexpr(D == 3) {
      x[0] = ((0xE2 >> q) & 1) ? -kGauss : kGauss;
      x[1] = ((0xD4 >> q) & 1) ? -kGauss : kGauss;
      x[2] = ((0xB8 >> q) & 1) ? -kGauss : kGauss;
    } else {
      x[0] = (q & 1) ? -kGauss : kGauss;
      x[1] = (q & 2) ? -kGauss : kGauss;
      x[2] = 0.;
    }
  }
  __device__ static double weight(int) { return 1.; }
  struct Node { double s[3]; };
  __device__ static Node node(int m) { return Node{{node_sx(m), node_sy(m), node_sz(m)}}; }
  // N_m and dN_m/d(xi, eta[, zeta]) at x
  __device__ static void shape(const Node& n, const double x[3], double& N, double dN[3])
  {
    const double* s = n.s;
    if constexpr(D == 3) {
      const double ax = 1. + s[0] * x[0], by = 1. + s[1] * x[1], cz = 1. + s[2] * x[2];
      N = ax * by * cz * 0.125;
      dN[0] = s[0] * 0.125 * by * cz;
      dN[1] = s[1] * 0.125 * ax * cz;
      dN[2] = s[2] * 0.125 * ax * by;
    } else {
      const double ax = 1. + s[0] * x[0], by = 1. + s[1] * x[1];
      N = ax * by * 0.25;
      dN[0] = s[0] * 0.25 * by;
      dN[1] = s[1] * 0.25 * ax;
      dN[2] = 0.;
    }
  }
  // dN_m/dxi_k at the centre: s_k / 2^D
  __device__ static double centre_grad(int m, int k) { return (k == 0 ? node_sx(m) : k == 1 ? node_sy(m) : node_sz(m)) * kInvK; }
  // the reference's etype::volume from det J at the centre and the nodes: Hexa3D 8 det J
  // (Hexa3D.cpp:317-320), Quad2D half the cross product of the diagonals
  __device__ static double volume(double detc, const double X[][D])
  {
    if constexpr(D == 3) return 8. * detc;
    else return 0.5 * ((X[2][0] - X[0][0]) * (X[3][1] - X[1][1]) - (X[2][1] - X[0][1]) * (X[3][0] - X[1][0]));
  }
};

// P1 prism (Prism3D): the triangle (0,0), (1,0), (0,1) at zeta = -1 (nodes 0-2), then at zeta = +1
// (nodes 3-5); N_m = t_(m mod 3)(xi, eta) (1 -+ zeta) / 2
struct Prism {
  static constexpr int D = 3, K = 6, NQ = 6, NEQ = 4, NN = 16;
  static constexpr int kLanes = 8;
  static constexpr int kSlots = kThreads / kLanes;
  static constexpr int kRec = D * D + 2 * D + 3;
  static constexpr int kRecSlot = kLanes * kRec + 2;
  static constexpr int kQRec = 16;
  static constexpr double kInvK = 1.0 / K;
  static constexpr int kUnroll = 1;
  __device__ static void gauss(int q, double x[3])
  {
    const int t = q < 3 ? q : q - 3;
    x[0] = t == 2 ? 0. : 0.5;
    x[1] = t == 0 ? 0. : 0.5;
    x[2] = q < 3 ? kGauss : -kGauss;
  }
  __device__ static double weight(int) { return 1. / 6.; }
  struct Node { int t; double dtx, dty, sz; };   // triangle node, its xi / eta derivatives, side of zeta
  __device__ static Node node(int m)
  {
    const int t = m < 3 ? m : m - 3;
    return Node{t, t == 0 ? -1. : (t == 1 ? 1. : 0.), t == 0 ? -1. : (t == 2 ? 1. : 0.), m < 3 ? -1. : 1.};
  }
  __device__ static void shape(const Node& n, const double x[3], double& N, double dN[3])
  {
    const double tv = n.t == 0 ? 1. - x[0] - x[1] : (n.t == 1 ? x[0] : x[1]);
    const double zf = 0.5 * (1. + n.sz * x[2]);
    N = tv * zf;
    dN[0] = n.dtx * zf;
    dN[1] = n.dty * zf;
    dN[2] = 0.5 * n.sz * tv;
  }
  // at (1/3, 1/3, 0): dN/dxi = dt/dxi / 2, dN/dzeta = +-1/6
  __device__ static double centre_grad(int m, int k)
  {
    const Node n = node(m);
    return k == 0 ? 0.5 * n.dtx : (k == 1 ? 0.5 * n.dty : n.sz / 6.);
  }
  // Prism3D::volume: det J at (1/3, 1/3, 0) (Prism3D.cpp:163-170)
  __device__ static double volume(double detc, const double (*)[3]) { return detc; }
};

// connectivity of element e (K node ids) with the widest aligned loads
template<int K>
__device__ __forceinline__ void load_conn(const uint32_t* conn, uint64_t e, uint32_t cn[K])
{
  if constexpr(K % 4 == 0) {
#pragma unroll
    for(int h = 0; h < K / 4; ++h) {
      const uint4 c = reinterpret_cast<const uint4*>(conn)[(K / 4) * e + h];
      cn[4 * h] = c.x; cn[4 * h + 1] = c.y; cn[4 * h + 2] = c.z; cn[4 * h + 3] = c.w;
    }
  } else {
#pragma unroll
    for(int h = 0; h < K / 2; ++h) {
      const uint2 c = reinterpret_cast<const uint2*>(conn)[(K / 2) * e + h];
      cn[2 * h] = c.x; cn[2 * h + 1] = c.y;
    }
  }
}

// inverse and determinant of a D x D matrix (row-major); 3 x 3 by cofactor expansion along column 0
template<int D>
__device__ __forceinline__ double inverse(const double J[D * D], double Ji[D * D])
{
  if constexpr(D == 3) {
    const double c00 = J[4] * J[8] - J[5] * J[7];
    const double c10 = J[7] * J[2] - J[8] * J[1];
    const double c20 = J[1] * J[5] - J[2] * J[4];
    const double det = c00 * J[0] + c10 * J[3] + c20 * J[6];
    const double inv = 1. / det;
    Ji[0] = c00 * inv; Ji[1] = c10 * inv; Ji[2] = c20 * inv;
    Ji[3] = (J[5] * J[6] - J[3] * J[8]) * inv;
    Ji[4] = (J[8] * J[0] - J[6] * J[2]) * inv;
    Ji[5] = (J[2] * J[3] - J[0] * J[5]) * inv;
    Ji[6] = (J[3] * J[7] - J[4] * J[6]) * inv;
    Ji[7] = (J[6] * J[1] - J[7] * J[0]) * inv;
    Ji[8] = (J[0] * J[4] - J[1] * J[3]) * inv;
    return det;
  } else {
    const double det = J[0] * J[3] - J[2] * J[1];
    const double inv = 1. / det;
    Ji[0] = J[3] * inv; Ji[1] = -J[1] * inv;
    Ji[2] = -J[2] * inv; Ji[3] = J[0] * inv;
    return det;
  }
}

// Element pre-pass, records: one thread per (element e, Gauss point q), the 2^D threads of an element
// adjacent; each streams the element's nodes once and writes the record of its point (J^-1, det J,
// u_adv, nu).
template<class Q>
__global__ __launch_bounds__(256) void k_q1_rec(uint32_t n_elems, AsmArgs g)
{
  constexpr int D = Q::D, K = Q::K, NQ = Q::NQ, kQRec = Q::kQRec;
  const uint64_t t = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  const uint64_t e = t / NQ;
  const int q = (int)(t % NQ);
  if(e >= n_elems) return;
  uint32_t cn[K];
  load_conn<K>(g.conn, e, cn);
  double xq[3];
  Q::gauss(q, xq);
  double J[D * D], uq[D], nuq = 0.;
#pragma unroll
  for(int v = 0; v < D * D; ++v) J[v] = 0.;
#pragma unroll
  for(int d = 0; d < D; ++d) uq[d] = 0.;
#pragma unroll
  for(int m = 0; m < K; ++m) {
    double xp[4], ad[4];
    unpack(g.xp[cn[m]], xp);
    unpack(g.ad[cn[m]], ad);
    double Nm, dN[3];
    Q::shape(Q::node(m), xq, Nm, dN);
#pragma unroll
    for(int k = 0; k < D; ++k)
#pragma unroll
      for(int d = 0; d < D; ++d) J[k * D + d] += dN[k] * xp[d];
#pragma unroll
    for(int d = 0; d < D; ++d) uq[d] += Nm * ad[d];
    nuq += Nm * ad[3];
  }
  double Ji[D * D];
  const double det = inverse<D>(J, Ji);
  if(!(det > 0.)) g.status[0] = 1;   // every writer stores the same flag value
  double r[kQRec];
#pragma unroll
  for(int v = 0; v < D * D; ++v) r[v] = Ji[v];
  r[D * D] = det;
#pragma unroll
  for(int d = 0; d < D; ++d) r[D * D + 1 + d] = uq[d];
  r[D * D + 1 + D] = nuq;
#pragma unroll
  for(int v = D * D + 2 + D; v < kQRec; ++v) r[v] = 0.;
  double2* dst = reinterpret_cast<double2*>(g.q1rec + (e * NQ + q) * kQRec);
#pragma unroll
  for(int v = 0; v < kQRec / 2; ++v) dst[v] = make_double2(r[2 * v], r[2 * v + 1]);
}

// Element pre-pass, tau: one thread per element, tau at the element centre (SUPG.hpp:141-219) from J,
// u_adv and nu there (every N_m = 1/K there; the gradients from Q::centre_grad).
template<class Q, int SUPG>
__global__ __launch_bounds__(256) void k_q1_tau(uint32_t n_elems, AsmArgs g)
{
  constexpr int D = Q::D, K = Q::K;
  const uint64_t e = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  if(e >= n_elems) return;
  uint32_t cn[K];
  load_conn<K>(g.conn, e, cn);
  double Jc[D * D], uc[D], nus = 0.;
#pragma unroll
  for(int v = 0; v < D * D; ++v) Jc[v] = 0.;
#pragma unroll
  for(int d = 0; d < D; ++d) uc[d] = 0.;
  double X[SUPG == CF3HIP_SUPG_METRIC ? 1 : K][D];
#pragma unroll
  for(int m = 0; m < K; ++m) {
    double xp[4], ad[4];
    unpack(g.xp[cn[m]], xp);
    unpack(g.ad[cn[m]], ad);
#pragma unroll
    for(int k = 0; k < D; ++k)
#pragma unroll
      for(int d = 0; d < D; ++d) Jc[k * D + d] += Q::centre_grad(m, k) * xp[d];
#pragma unroll
    for(int d = 0; d < D; ++d) uc[d] += Q::kInvK * ad[d];
    nus += ad[3];
    if constexpr(SUPG != CF3HIP_SUPG_METRIC) {
#pragma unroll
      for(int d = 0; d < D; ++d) X[m][d] = xp[d];
    }
  }
  const double nu_e = fabs(nus * Q::kInvK);
  double tps, tsu, tbu;
  if constexpr(SUPG == CF3HIP_SUPG_METRIC) {
    double Jci[D * D];
    inverse<D>(Jc, Jci);
    double gij[D * D];
#pragma unroll
    for(int i = 0; i < D; ++i)
#pragma unroll
      for(int j = 0; j < D; ++j) {
        double s = 0.;
#pragma unroll
        for(int k = 0; k < D; ++k) s += Jci[k * D + i] * Jci[k * D + j];
        gij[i * D + j] = s;
      }
    double ugu = 0., gsq = 0., tr = 0.;
#pragma unroll
    for(int j = 0; j < D; ++j) {
      double s = 0.;
#pragma unroll
      for(int i = 0; i < D; ++i) s += uc[i] * gij[i * D + j];
      ugu += s * uc[j];
      tr += gij[j * (D + 1)];
    }
#pragma unroll
    for(int v = 0; v < D * D; ++v) gsq += gij[v] * gij[v];
    const double diff = nu_e * nu_e * gsq;
    tsu = 1. / sqrt(g.c1 * g.c1 * g.f4dt2 + fabs(ugu) + g.c2 * diff);
    tps = tsu;
    tbu = sqrt(g.f4dt2 + fabs(ugu) + 16. * diff) / tr;
  } else if constexpr(SUPG == CF3HIP_SUPG_CF2) {
    // volume as the reference's etype::volume
    double Jci[D * D];
    const double detc = inverse<D>(Jc, Jci);
    const double vol = Q::volume(detc, X);
    // sum over the nodes of u_c . grad N_n at the centre, grad N_n = J^-1 dN_n/dxi
    double um2 = 0., sum = 0.;
#pragma unroll
    for(int d = 0; d < D; ++d) um2 += uc[d] * uc[d];
#pragma unroll
    for(int m = 0; m < K; ++m) {
      double un = 0.;
#pragma unroll
      for(int d = 0; d < D; ++d) {
        double gdn = 0.;
#pragma unroll
        for(int l = 0; l < D; ++l) gdn += Jci[d * D + l] * Q::centre_grad(m, l);
        un += uc[d] * gdn;
      }
      sum += un;
    }
    if(!tau_cf2(D, vol, nu_e, sqrt(um2), sum, g.dt, g.u_ref, tps, tsu, tbu)) g.status[0] = 2;
  } else {
    double hmin2 = 1e20;
#pragma unroll
    for(int i = 0; i < K; ++i)
#pragma unroll
      for(int j = i + 1; j < K; ++j) {
        double s = 0.;
#pragma unroll
        for(int d = 0; d < D; ++d) { const double df = X[i][d] - X[j][d]; s += df * df; }
        hmin2 = fmin(hmin2, s);
      }
    const double h = sqrt(hmin2);
    double um2 = 0.;
#pragma unroll
    for(int d = 0; d < D; ++d) um2 += uc[d] * uc[d];
    const double um = sqrt(um2);
    tsu = 1. / ((2. * um) / h + 2. / g.dt + (4. * nu_e) / (h * h));
    tps = tsu;
    tbu = tsu * um * um;
  }
  g.q1tau[e] = make_double4(tps * g.a_ps, tsu * g.a_su, tbu * g.a_bu, 0.);
}

template<class Q, int MODE>
__global__ __launch_bounds__(kThreads, 4) void k_q1(AsmArgs g)
{
  constexpr int D = Q::D, K = Q::K, NQ = Q::NQ, NEQ = Q::NEQ, NN = Q::NN, kLanes = Q::kLanes, kRec = Q::kRec;
  constexpr int kRecSlot = Q::kRecSlot;
  typedef typename std::conditional<MODE == MODE_FP32, float, double>::type VR;
  constexpr int LDW = lds_row<VR, NN>();
  constexpr size_t kRecBytes = sizeof(double) * Q::kSlots * kRecSlot;
  constexpr size_t kBlkBytes = sizeof(VR) * kThreads * LDW;
  constexpr size_t kRhsBytes = sizeof(double) * kThreads * NEQ;
  constexpr size_t kRaw = kRecBytes > kBlkBytes ? (kRecBytes > kRhsBytes ? kRecBytes : kRhsBytes)
                                                 : (kBlkBytes > kRhsBytes ? kBlkBytes : kRhsBytes);
  __shared__ __align__(16) unsigned char s_raw[kRaw];
  double* s_rec = reinterpret_cast<double*>(s_raw);   // Gauss point records, then
  VR* s_blk = reinterpret_cast<VR*>(s_raw);           // staged blocks, then
  double* s_rhs = reinterpret_cast<double*>(s_raw);   // rhs / residual parts
  __shared__ uint32_t s_rowinc[kThreads + 1];
  __shared__ uint16_t s_acc[kThreads];

  const uint32_t wg = xcd_order(blockIdx.x, gridDim.x);
  const uint32_t r0 = g.wg_rows[wg], r1 = g.wg_rows[wg + 1];
  const uint32_t nrows = r1 - r0;
  const uint32_t i0 = g.inc_ptr[r0];
  const uint32_t ninc = g.inc_ptr[r1] - i0;
  const uint32_t blk0 = g.row_ptr[r0];
  const uint32_t nblk = g.row_ptr[r1] - blk0;
  const int t = threadIdx.x;
  const int slot = t / kLanes, lane = t % kLanes;
  for(uint32_t j = t; j <= nrows; j += kThreads) s_rowinc[j] = g.inc_ptr[r0 + j] - i0;
  if(MODE != MODE_RESIDUAL && (uint32_t)t < nblk) s_acc[t] = g.blk_acc[blk0 + t];

  const bool active = (uint32_t)slot < ninc;
  const bool node_lane = active && lane < K;   // lanes K .. kLanes - 1 of a prism incidence idle
  uint64_t e = 0;
  int a = 0;
  uint32_t nb = 0, pos = 0;   // column node b = lane of the incidence's element, staging row of its block
  double tps = 0., tsu = 0., tbu = 0.;
  double rhs_part[NEQ];
#pragma unroll
  for(int r = 0; r < NEQ; ++r) rhs_part[r] = 0.;
  if(active) {
    const uint32_t ea = g.inc[i0 + slot];
    e = ea / K;
    a = (int)(ea % K);
    const double4 tt = g.q1tau[e];
    tps = tt.x; tsu = tt.y; tbu = tt.z;
  }
  if(node_lane) {
    nb = g.conn[e * K + lane];
    if(MODE != MODE_RESIDUAL) pos = g.inc_pos[(uint64_t)(i0 + slot) * kLanes + lane];
  }
  if(active && lane < NQ) {
    // ---- the pre-pass record of Gauss point q = lane ----
    double xq[3];
    Q::gauss(lane, xq);
    double rq[Q::kQRec];
    {
      const double2* src = reinterpret_cast<const double2*>(g.q1rec + (e * NQ + lane) * Q::kQRec);
#pragma unroll
      for(int v = 0; v < Q::kQRec / 2; ++v) { const double2 w2 = src[v]; rq[2 * v] = w2.x; rq[2 * v + 1] = w2.y; }
    }
    double Ji[D * D], uq[D];
#pragma unroll
    for(int v = 0; v < D * D; ++v) Ji[v] = rq[v];
    const double det = rq[D * D];
#pragma unroll
    for(int d = 0; d < D; ++d) uq[d] = rq[D * D + 1 + d];
    const double nuq = rq[D * D + 1 + D];
    // ---- Gauss point record of lane q ----
    const double w = Q::weight(lane) * det;
    double Na, dNa[3];
    Q::shape(Q::node(a), xq, Na, dNa);
    double Ga[D], La = 0.;
#pragma unroll
    for(int d = 0; d < D; ++d) {
      double s = 0.;
#pragma unroll
      for(int l = 0; l < D; ++l) s += Ji[d * D + l] * dNa[l];
      Ga[d] = s;
      La += uq[d] * Ga[d];
    }
    const double SUa = Na + tsu * La;
    const double delta = w * SUa;
    double* rec = s_rec + (size_t)slot * kRecSlot + lane * kRec;
#pragma unroll
    for(int v = 0; v < D * D; ++v) rec[v] = Ji[v];
#pragma unroll
    for(int d = 0; d < D; ++d) { rec[D * D + d] = uq[d]; rec[D * D + D + d] = w * Ga[d]; }
    rec[D * D + 2 * D] = nuq;
    rec[D * D + 2 * D + 1] = delta;
    rec[D * D + 2 * D + 2] = w * (Na + 0.5 * tps * La);
  }
  __syncthreads();

  // the run of staged rows each accumulator thread sums (its global loads overlap the integration below)
  uint32_t jb = 0, je = 0, h = 0, c = 1;
  bool owner = false;
  if constexpr(MODE != MODE_RESIDUAL) {
    const uint32_t alast = s_acc[nblk - 1];
    const uint32_t n_acc = (alast & 255u) + (alast >> 8);
    owner = (uint32_t)t < n_acc;
    if(owner) {
      uint32_t lo = 0, hi = nblk;
      while(hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if((s_acc[mid] & 255u) <= (uint32_t)t) lo = mid; else hi = mid;
      }
      h = (uint32_t)t - (s_acc[lo] & 255u);
      c = s_acc[lo] >> 8;
      jb = g.blk_beg[(uint64_t)(blk0 + lo) * 4];
      je = lo + 1 < nblk ? g.blk_beg[(uint64_t)(blk0 + lo + 1) * 4] : ninc * K;
    }
  }
  // ---- block (a, b = lane): integrate over the NQ Gauss points ----
  constexpr int RU = D * D, RW = D * D + D, RNU = D * D + 2 * D;   // record offsets: u_adv, w G_a, nu
  double M[NN];
  double xb[NEQ];
  if(node_lane) {
    const int b = lane;
    const typename Q::Node nd_b = Q::node(b);
    double Ap[D], Xp[D], Aup[D], Tp[D], Auu[D * D];
#pragma unroll
    for(int d = 0; d < D; ++d) { Ap[d] = 0.; Xp[d] = 0.; Aup[d] = 0.; Tp[d] = 0.; }
#pragma unroll
    for(int v = 0; v < D * D; ++v) Auu[v] = 0.;
    double App = 0., Add = 0., Tuu = 0.;
    const double* recs = s_rec + (size_t)slot * kRecSlot;
#pragma unroll Q::kUnroll   // full unrolling makes the scheduler hoist the LDS reads and spill
    for(int q = 0; q < NQ; ++q) {
      double r[kRec];
      lds_load_row<double, kRec>(recs + q * kRec, r);
      double xq[3];
      Q::gauss(q, xq);
      double Nb, dNb[3];
      Q::shape(nd_b, xq, Nb, dNb);
      double Gb[D];
#pragma unroll
      for(int d = 0; d < D; ++d) {
        double s = 0.;
#pragma unroll
        for(int l = 0; l < D; ++l) s += r[d * D + l] * dNb[l];
        Gb[d] = s;
      }
      double Lb = 0., gw = 0.;   // u_adv . G_b and w G_a . G_b
#pragma unroll
      for(int d = 0; d < D; ++d) { Lb += r[RU + d] * Gb[d]; gw += r[RW + d] * Gb[d]; }
      const double delta = r[RNU + 1], alpha = r[RNU + 2];
      App += gw;
      Add += r[RNU] * gw + delta * Lb;
      Tuu += delta * Nb;
#pragma unroll
      for(int i = 0; i < D; ++i) {
        Ap[i] += alpha * Gb[i];
        Xp[i] += r[RW + i] * Lb;
        Aup[i] += delta * Gb[i];
        Tp[i] += r[RW + i] * Nb;
        const double ci = tbu * r[RW + i] + 0.5 * r[RU + i] * delta;
#pragma unroll
        for(int j = 0; j < D; ++j) Auu[i * D + j] += ci * Gb[j];
      }
    }
    // element matrices of the block, node-major (u, v, [w,] p)
    double A[NN];
#pragma unroll
    for(int i = 0; i < D; ++i) {
#pragma unroll
      for(int j = 0; j < D; ++j) A[i * NEQ + j] = Auu[i * D + j] + (i == j ? Add : 0.);
      A[i * NEQ + D] = Aup[i] * g.inv_rho;
      A[D * NEQ + i] = Ap[i] + tps * Xp[i];
    }
    A[D * NEQ + D] = tps * App * g.inv_rho;
    load_state<NEQ, double>(g, nb, xb);
    if constexpr(MODE != MODE_FP32) {
#pragma unroll
      for(int r = 0; r < NEQ; ++r) {
        double s = 0.;
#pragma unroll
        for(int c = 0; c < NEQ; ++c) s += A[r * NEQ + c] * xb[c];
        rhs_part[r] -= s;
      }
      // body force source (NavierStokesAssembly.hpp:65), sum_q w SU_a(q) g(q) with g(q) = sum_b N_b(q) g_b:
      // lane b adds T_uu(a, b) g_b
      if(g.body_force)
#pragma unroll
        for(int d = 0; d < D; ++d) rhs_part[d] += Tuu * g.g[d][nb];
    }
    const double invdt = g.invdt, th = g.theta, thq = g.thq;
#pragma unroll
    for(int i = 0; i < D; ++i) {
#pragma unroll
      for(int j = 0; j < NEQ; ++j) M[i * NEQ + j] = (i == j ? invdt * Tuu : 0.) + th * A[i * NEQ + j];
      M[D * NEQ + i] = invdt * (tps * Tp[i]) + thq * A[D * NEQ + i];
    }
    M[D * NEQ + D] = thq * A[D * NEQ + D];
  }

  if constexpr(MODE == MODE_RESIDUAL) {
    // r_a = sum_b (b_e(a) - M(a, b) x~_b): lane part with x~ of node b
    if(node_lane) {
      double xt[NEQ];
#pragma unroll
      for(int c = 0; c < NEQ; ++c) xt[c] = g.xt[(uint64_t)nb * NEQ + c];
#pragma unroll
      for(int r = 0; r < NEQ; ++r) {
        double s = 0.;
#pragma unroll
        for(int c = 0; c < NEQ; ++c) s += M[r * NEQ + c] * xt[c];
        rhs_part[r] -= s;
      }
    }
    __syncthreads();   // records are no longer read
  } else {
    // ---- stage the blocks at their sorted rows; accumulator threads sum their runs ----
    __syncthreads();   // records are no longer read
    if(node_lane) {
      VR* d = s_blk + pos * LDW;
#pragma unroll
      for(int v = 0; v < NN; ++v) d[v] = (VR)M[v];
    }
    __syncthreads();
    VR acc[NN];
#pragma unroll
    for(int v = 0; v < NN; ++v) acc[v] = VR(0);
    if(owner) {
      for(uint32_t j = jb + h; j < je; j += c) {
        VR v[NN];
        lds_load_row<VR, NN>(s_blk + j * LDW, v);
#pragma unroll
        for(int q = 0; q < NN; ++q) acc[q] += v[q];
      }
    }
    __syncthreads();
    if(owner) {
      VR* d = s_blk + t * LDW;
#pragma unroll
      for(int q = 0; q < NN; ++q) d[q] = acc[q];
    }
    __syncthreads();
    VR* dst = static_cast<VR*>(g.values) + (uint64_t)blk0 * NN;
    for(uint32_t o = t; o < nblk * NN; o += kThreads) {
      const uint32_t sb = o / NN, v = o % NN;
      const uint32_t ac = s_acc[sb], f = ac & 255u, na = ac >> 8;
      VR sum = s_blk[f * LDW + v];
      for(uint32_t r = 1; r < na; ++r) sum += s_blk[(f + r) * LDW + v];
      dst[o] = sum;
    }
    if constexpr(MODE == MODE_FP32) return;
    __syncthreads();
  }
  // ---- rhs (or residual) of each row: sum over the lanes of its incidences ----
  if(active)
#pragma unroll
    for(int r = 0; r < NEQ; ++r) s_rhs[t * NEQ + r] = rhs_part[r];
  __syncthreads();
  double* out = MODE == MODE_RESIDUAL ? g.res : g.rhs;
  for(uint32_t o = t; o < nrows * NEQ; o += kThreads) {
    const uint32_t row = o / NEQ, r = o % NEQ;
    const uint32_t qe = s_rowinc[row + 1] * kLanes;
    uint32_t q = s_rowinc[row] * kLanes;
    double s0 = 0., s1 = 0., s2 = 0., s3 = 0.;
    for(; q + 4 <= qe; q += 4) {
      s0 += s_rhs[q * NEQ + r];
      s1 += s_rhs[(q + 1) * NEQ + r];
      s2 += s_rhs[(q + 2) * NEQ + r];
      s3 += s_rhs[(q + 3) * NEQ + r];
    }
    for(; q < qe; ++q) s0 += s_rhs[q * NEQ + r];
    out[(uint64_t)(r0 + row) * NEQ + r] = (s0 + s1) + (s2 + s3);
  }
}

template<class Q, int MODE>
void launch_mode(const AsmArgs& g, const Lss& L)
{
  Context& c = *L.ctx;
  const int timer = MODE == MODE_RESIDUAL ? CF3HIP_KERNEL_RESIDUAL : CF3HIP_KERNEL_ASSEMBLY;
  const unsigned gp = (unsigned)(((uint64_t)L.mesh->n_elems * Q::NQ + 255) / 256);
  const unsigned ge = (unsigned)((L.mesh->n_elems + 255) / 256);
  timer_begin(c, timer);
  hipLaunchKernelGGL((k_q1_rec<Q>), dim3(gp), dim3(256), 0, c.stream, L.mesh->n_elems, g);
  if(g.supg_type == CF3HIP_SUPG_TEZDUYAR)
    hipLaunchKernelGGL((k_q1_tau<Q, CF3HIP_SUPG_TEZDUYAR>), dim3(ge), dim3(256), 0, c.stream, L.mesh->n_elems, g);
  else if(g.supg_type == CF3HIP_SUPG_CF2)
    hipLaunchKernelGGL((k_q1_tau<Q, CF3HIP_SUPG_CF2>), dim3(ge), dim3(256), 0, c.stream, L.mesh->n_elems, g);
  else
    hipLaunchKernelGGL((k_q1_tau<Q, CF3HIP_SUPG_METRIC>), dim3(ge), dim3(256), 0, c.stream, L.mesh->n_elems, g);
  hipLaunchKernelGGL((k_q1<Q, MODE>), dim3(L.n_wg), dim3(kThreads), 0, c.stream, g);
  HIP_OK(hipGetLastError());
  timer_end(c, timer);
}

template<class Q>
void launch_cell(AsmArgs g, Lss& L, int mode)
{
  if(L.passes != 1 || L.pos_stride != Q::kLanes)
    CF3_THROW(CF3HIP_ERR_INVALID_ARG, "system has no cell staging maps");
  // record buffers of the pre-pass (allocated at the first cell launch of the system)
  const size_t nrec = (size_t)L.mesh->n_elems * Q::NQ * Q::kQRec;
  if(L.q1rec.bytes < sizeof(double) * nrec) L.q1rec.alloc(L.ctx, sizeof(double) * nrec);
  if(L.q1tau.bytes < sizeof(double4) * L.mesh->n_elems) L.q1tau.alloc(L.ctx, sizeof(double4) * L.mesh->n_elems);
  g.q1rec = L.q1rec.as<double>();
  g.q1tau = L.q1tau.as<double4>();
  if(mode == MODE_FP64) launch_mode<Q, MODE_FP64>(g, L);
  else if(mode == MODE_FP32) launch_mode<Q, MODE_FP32>(g, L);
  else launch_mode<Q, MODE_RESIDUAL>(g, L);
}

}  // namespace

void launch_hex(AsmArgs g, Lss& L, int mode)
{
  if(L.mesh->etype == CF3HIP_QUAD2D) launch_cell<Q1<2> >(g, L, mode);
  else if(L.mesh->etype == CF3HIP_PRISM3D) launch_cell<Prism>(g, L, mode);
  else launch_cell<Q1<3> >(g, L, mode);
}

}  // namespace asmk
}  // namespace cf3hip

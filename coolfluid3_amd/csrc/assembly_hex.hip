// Q1 hexahedron assembly: the generic Navier-Stokes expression, which hexahedra always take
// (plugins/UFEM/src/UFEM/ns_implementation/NavierStokesHexas.cpp:20-25), with the 8-point Gauss rule.
//
// Element expression: plugins/UFEM/src/UFEM/NavierStokesAssembly.hpp:44-66 (tau from u_adv, :54);
// tau METRIC / TEZDUYAR at the element centre: SUPG.hpp:141-219; Q1 shape functions and gradients:
// cf3/mesh/LagrangeP1/Hexa.cpp:24-89 (node order :103-118); Gauss rule: GaussImplementation.hpp:29-35
// (x = 1/sqrt(3), w = 1) and :615-676 (point order); combine: NavierStokesAssembly.hpp:69-71.
//
// Work decomposition (row owner, one staging pass): a workgroup owns a range of node rows; each of its
// incidences (element e, local node a) gets kLanes = 8 threads.  Lane q first computes the quantities
// of Gauss point q that do not depend on the column node (J^-1, u_adv, nu, the row-node terms) and
// stores them in LDS; lane b then integrates the element block M(a, b) over the 8 points.  Blocks are
// staged in LDS at rows sorted by target CSR block and summed by the block's accumulator threads in a
// fixed order: deterministic, no floating-point atomics.
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

constexpr int HK = 8, HNEQ = 4, HNN = 16;
constexpr int kLanes = HK;                  // threads per incidence: Gauss point q, then column node b
constexpr int kSlots = kThreads / kLanes;   // incidences per workgroup
constexpr int kRec = 18;                    // doubles per (incidence, Gauss point) record
constexpr int kRecSlot = kLanes * kRec + 2; // doubles per incidence (padded: slots on distinct LDS banks)
constexpr double kGauss = 0.5773502691896257645091488;

enum { MODE_FP64 = 0, MODE_FP32 = 1, MODE_RESIDUAL = 2 };

// reference coordinate signs of the Q1 nodes: (-,-,-), (+,-,-), (+,+,-), (-,+,-), then z = +
__device__ __forceinline__ double node_sx(int m) { return ((m ^ (m >> 1)) & 1) ? 1. : -1.; }
__device__ __forceinline__ double node_sy(int m) { return (m & 2) ? 1. : -1.; }
__device__ __forceinline__ double node_sz(int m) { return (m & 4) ? 1. : -1.; }
// Gauss point q: (+,+,+), (-,+,+), (+,-,+), (+,+,-), (+,-,-), (-,+,-), (-,-,+), (-,-,-) times 1/sqrt(3)
__device__ __forceinline__ void gauss_point(int q, double& x, double& y, double& z)
{
  x = ((0xE2 >> q) & 1) ? -kGauss : kGauss;
  y = ((0xD4 >> q) & 1) ? -kGauss : kGauss;
  z = ((0xB8 >> q) & 1) ? -kGauss : kGauss;
}

// N_m and dN_m/d(xi, eta, zeta) at (x, y, z) for node signs (sx, sy, sz)
__device__ __forceinline__ void shape(double sx, double sy, double sz, double x, double y, double z, double& N,
                                      double dN[3])
{
  const double ax = 1. + sx * x, by = 1. + sy * y, cz = 1. + sz * z;
  N = ax * by * cz * 0.125;
  dN[0] = sx * 0.125 * by * cz;
  dN[1] = sy * 0.125 * ax * cz;
  dN[2] = sz * 0.125 * ax * by;
}

// inverse and determinant of a 3x3 matrix (row-major), cofactor expansion along column 0
__device__ __forceinline__ double inverse3(const double J[9], double Ji[9])
{
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
}

template<int MODE, int SUPG>
__global__ __launch_bounds__(kThreads) void k_hexa(AsmArgs g)
{
  typedef typename std::conditional<MODE == MODE_FP32, float, double>::type VR;
  constexpr int LDW = lds_row<VR, HNN>();
  constexpr size_t kRecBytes = sizeof(double) * kSlots * kRecSlot;
  constexpr size_t kBlkBytes = sizeof(VR) * kThreads * LDW;
  constexpr size_t kRhsBytes = sizeof(double) * kThreads * HNEQ;
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
  uint32_t cn[HK];
  uint64_t e = 0;
  int a = 0;
  double tps = 0., tsu = 0., tbu = 0.;
  double rhs_part[HNEQ] = {0., 0., 0., 0.};
  if(active) {
    const uint32_t ea = g.inc[i0 + slot];
    e = ea / HK;
    a = (int)(ea % HK);
    const uint4 c0 = reinterpret_cast<const uint4*>(g.conn)[2 * (uint64_t)e];
    const uint4 c1 = reinterpret_cast<const uint4*>(g.conn)[2 * (uint64_t)e + 1];
    cn[0] = c0.x; cn[1] = c0.y; cn[2] = c0.z; cn[3] = c0.w;
    cn[4] = c1.x; cn[5] = c1.y; cn[6] = c1.z; cn[7] = c1.w;
    // ---- stream over the nodes: J at Gauss point q = lane and at the centre, u_adv and nu ----
    double qx, qy, qz;
    gauss_point(lane, qx, qy, qz);
    double J[9], Jc[9], uq[3] = {0., 0., 0.}, uc[3] = {0., 0., 0.}, nuq = 0., nus = 0.;
    double gq[3] = {0., 0., 0.};
#pragma unroll
    for(int v = 0; v < 9; ++v) { J[v] = 0.; Jc[v] = 0.; }
    double hmin2 = 1e20;
    double Xall[SUPG == CF3HIP_SUPG_TEZDUYAR ? HK : 1][3];
#pragma unroll
    for(int m = 0; m < HK; ++m) {
      double xp[4], ad[4];
      unpack(g.xp[cn[m]], xp);
      unpack(g.ad[cn[m]], ad);
      const double sx = node_sx(m), sy = node_sy(m), sz = node_sz(m);
      double Nm, dN[3];
      shape(sx, sy, sz, qx, qy, qz, Nm, dN);
      const double dc[3] = {sx * 0.125, sy * 0.125, sz * 0.125};   // gradients at the centre
#pragma unroll
      for(int k = 0; k < 3; ++k)
#pragma unroll
        for(int d = 0; d < 3; ++d) {
          J[k * 3 + d] += dN[k] * xp[d];
          Jc[k * 3 + d] += dc[k] * xp[d];
        }
#pragma unroll
      for(int d = 0; d < 3; ++d) {
        uq[d] += Nm * ad[d];
        uc[d] += 0.125 * ad[d];
      }
      nuq += Nm * ad[3];
      nus += ad[3];
      if(MODE != MODE_FP32 && g.body_force) {
#pragma unroll
        for(int d = 0; d < 3; ++d) gq[d] += Nm * g.g[d][cn[m]];
      }
      if constexpr(SUPG == CF3HIP_SUPG_TEZDUYAR) {
#pragma unroll
        for(int d = 0; d < 3; ++d) Xall[m][d] = xp[d];
      }
    }
    // ---- tau at the centre (SUPG.hpp:141-219), identical on the 8 lanes ----
    const double nu_e = fabs(nus * 0.125);
    if constexpr(SUPG == CF3HIP_SUPG_METRIC) {
      double Jci[9];
      inverse3(Jc, Jci);
      double gij[9];
#pragma unroll
      for(int i = 0; i < 3; ++i)
#pragma unroll
        for(int j = 0; j < 3; ++j)
          gij[i * 3 + j] = Jci[i] * Jci[j] + Jci[3 + i] * Jci[3 + j] + Jci[6 + i] * Jci[6 + j];
      double ugu = 0., gsq = 0., tr = 0.;
#pragma unroll
      for(int j = 0; j < 3; ++j) {
        ugu += (uc[0] * gij[j] + uc[1] * gij[3 + j] + uc[2] * gij[6 + j]) * uc[j];
        tr += gij[j * 4];
      }
#pragma unroll
      for(int v = 0; v < 9; ++v) gsq += gij[v] * gij[v];
      const double diff = nu_e * nu_e * gsq;
      tsu = 1. / sqrt(g.c1 * g.c1 * g.f4dt2 + fabs(ugu) + g.c2 * diff);
      tps = tsu;
      tbu = sqrt(g.f4dt2 + fabs(ugu) + 16. * diff) / tr;
    } else {
#pragma unroll
      for(int i = 0; i < HK; ++i)
#pragma unroll
        for(int j = i + 1; j < HK; ++j) {
          double s = 0.;
#pragma unroll
          for(int d = 0; d < 3; ++d) { const double df = Xall[i][d] - Xall[j][d]; s += df * df; }
          hmin2 = fmin(hmin2, s);
        }
      const double h = sqrt(hmin2);
      const double um = sqrt(uc[0] * uc[0] + uc[1] * uc[1] + uc[2] * uc[2]);
      tsu = 1. / ((2. * um) / h + 2. / g.dt + (4. * nu_e) / (h * h));
      tps = tsu;
      tbu = tsu * um * um;
    }
    tps *= g.a_ps; tsu *= g.a_su; tbu *= g.a_bu;
    // ---- Gauss point record of lane q ----
    double Ji[9];
    const double det = inverse3(J, Ji);
    if(!(det > 0.)) g.status[0] = 1;   // every writer stores the same flag value
    const double w = det;               // Gauss weight 1
    double Na, dNa[3];
    shape(node_sx(a), node_sy(a), node_sz(a), qx, qy, qz, Na, dNa);
    double Ga[3], La = 0.;
#pragma unroll
    for(int d = 0; d < 3; ++d) {
      Ga[d] = Ji[d * 3] * dNa[0] + Ji[d * 3 + 1] * dNa[1] + Ji[d * 3 + 2] * dNa[2];
      La += uq[d] * Ga[d];
    }
    const double SUa = Na + tsu * La;
    const double delta = w * SUa;
    double* rec = s_rec + (size_t)slot * kRecSlot + lane * kRec;
#pragma unroll
    for(int v = 0; v < 9; ++v) rec[v] = Ji[v];
#pragma unroll
    for(int d = 0; d < 3; ++d) { rec[9 + d] = uq[d]; rec[12 + d] = w * Ga[d]; }
    rec[15] = nuq;
    rec[16] = delta;
    rec[17] = w * (Na + 0.5 * tps * La);
    // body force source of point q (NavierStokesAssembly.hpp:65)
    if(MODE != MODE_FP32 && g.body_force)
#pragma unroll
      for(int d = 0; d < 3; ++d) rhs_part[d] += delta * gq[d];
  }
  __syncthreads();

  // ---- block (a, b = lane): integrate over the 8 Gauss points ----
  double M[HNN];
  double xb[HNEQ];
  if(active) {
    const int b = lane;
    const double bx = node_sx(b), by = node_sy(b), bz = node_sz(b);
    double Ap[3] = {0., 0., 0.}, Xp[3] = {0., 0., 0.}, Aup[3] = {0., 0., 0.}, Tp[3] = {0., 0., 0.};
    double Auu[9];
#pragma unroll
    for(int v = 0; v < 9; ++v) Auu[v] = 0.;
    double App = 0., Add = 0., Tuu = 0.;
    const double* recs = s_rec + (size_t)slot * kRecSlot;
#pragma unroll
    for(int q = 0; q < HK; ++q) {
      double r[kRec];
      lds_load_row<double, kRec>(recs + q * kRec, r);
      double qx, qy, qz;
      gauss_point(q, qx, qy, qz);
      double Nb, dNb[3];
      shape(bx, by, bz, qx, qy, qz, Nb, dNb);
      double Gb[3];
#pragma unroll
      for(int d = 0; d < 3; ++d) Gb[d] = r[d * 3] * dNb[0] + r[d * 3 + 1] * dNb[1] + r[d * 3 + 2] * dNb[2];
      const double Lb = r[9] * Gb[0] + r[10] * Gb[1] + r[11] * Gb[2];
      const double gw = r[12] * Gb[0] + r[13] * Gb[1] + r[14] * Gb[2];   // w G_a . G_b
      const double delta = r[16], alpha = r[17];
      App += gw;
      Add += r[15] * gw + delta * Lb;
      Tuu += delta * Nb;
#pragma unroll
      for(int i = 0; i < 3; ++i) {
        Ap[i] += alpha * Gb[i];
        Xp[i] += r[12 + i] * Lb;
        Aup[i] += delta * Gb[i];
        Tp[i] += r[12 + i] * Nb;
        const double ci = tbu * r[12 + i] + 0.5 * r[9 + i] * delta;
#pragma unroll
        for(int j = 0; j < 3; ++j) Auu[i * 3 + j] += ci * Gb[j];
      }
    }
    // element matrices of the block, node-major (u, v, w, p)
    double A[HNN];
#pragma unroll
    for(int i = 0; i < 3; ++i) {
#pragma unroll
      for(int j = 0; j < 3; ++j) A[i * 4 + j] = Auu[i * 3 + j] + (i == j ? Add : 0.);
      A[i * 4 + 3] = Aup[i] * g.inv_rho;
      A[12 + i] = Ap[i] + tps * Xp[i];
    }
    A[15] = tps * App * g.inv_rho;
    const uint32_t nb = g.conn[e * HK + b];   // column node (lane-dependent: not taken from cn[])
    load_state<HNEQ, double>(g, nb, xb);
    if constexpr(MODE != MODE_FP32) {
#pragma unroll
      for(int r = 0; r < HNEQ; ++r)
        rhs_part[r] -= A[r * 4] * xb[0] + A[r * 4 + 1] * xb[1] + A[r * 4 + 2] * xb[2] + A[r * 4 + 3] * xb[3];
    }
    const double invdt = g.invdt, th = g.theta, thq = g.thq;
#pragma unroll
    for(int i = 0; i < 3; ++i) {
#pragma unroll
      for(int j = 0; j < 4; ++j) M[i * 4 + j] = (i == j ? invdt * Tuu : 0.) + th * A[i * 4 + j];
      M[12 + i] = invdt * (tps * Tp[i]) + thq * A[12 + i];
    }
    M[15] = thq * A[15];
  }

  if constexpr(MODE == MODE_RESIDUAL) {
    // r_a = sum_b (b_e(a) - M(a, b) x~_b): lane part with x~ of node b
    if(active) {
      double xt[HNEQ];
      const uint32_t nb = g.conn[e * HK + lane];
#pragma unroll
      for(int c = 0; c < HNEQ; ++c) xt[c] = g.xt[(uint64_t)nb * HNEQ + c];
#pragma unroll
      for(int r = 0; r < HNEQ; ++r)
        rhs_part[r] -= M[r * 4] * xt[0] + M[r * 4 + 1] * xt[1] + M[r * 4 + 2] * xt[2] + M[r * 4 + 3] * xt[3];
    }
    __syncthreads();   // records are no longer read
  } else {
    // ---- stage the blocks at their sorted rows; accumulator threads sum their runs ----
    __syncthreads();   // records are no longer read
    if(active) {
      const uint32_t pos = g.inc_pos[(uint64_t)(i0 + slot) * kLanes + lane];
      VR* d = s_blk + pos * LDW;
#pragma unroll
      for(int v = 0; v < HNN; ++v) d[v] = (VR)M[v];
    }
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
      jb = g.blk_beg[(uint64_t)(blk0 + lo) * 4];
      je = lo + 1 < nblk ? g.blk_beg[(uint64_t)(blk0 + lo + 1) * 4] : ninc * kLanes;
    }
    __syncthreads();
    VR acc[HNN];
#pragma unroll
    for(int v = 0; v < HNN; ++v) acc[v] = VR(0);
    if(owner) {
      for(uint32_t j = jb + h; j < je; j += c) {
        VR v[HNN];
        lds_load_row<VR, HNN>(s_blk + j * LDW, v);
#pragma unroll
        for(int q = 0; q < HNN; ++q) acc[q] += v[q];
      }
    }
    __syncthreads();
    if(owner) {
      VR* d = s_blk + t * LDW;
#pragma unroll
      for(int q = 0; q < HNN; ++q) d[q] = acc[q];
    }
    __syncthreads();
    VR* dst = static_cast<VR*>(g.values) + (uint64_t)blk0 * HNN;
    for(uint32_t o = t; o < nblk * HNN; o += kThreads) {
      const uint32_t sb = o / HNN, v = o % HNN;
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
    for(int r = 0; r < HNEQ; ++r) s_rhs[t * HNEQ + r] = rhs_part[r];
  __syncthreads();
  double* out = MODE == MODE_RESIDUAL ? g.res : g.rhs;
  for(uint32_t o = t; o < nrows * HNEQ; o += kThreads) {
    const uint32_t row = o / HNEQ, r = o % HNEQ;
    const uint32_t qe = s_rowinc[row + 1] * kLanes;
    uint32_t q = s_rowinc[row] * kLanes;
    double s0 = 0., s1 = 0., s2 = 0., s3 = 0.;
    for(; q + 4 <= qe; q += 4) {
      s0 += s_rhs[q * HNEQ + r];
      s1 += s_rhs[(q + 1) * HNEQ + r];
      s2 += s_rhs[(q + 2) * HNEQ + r];
      s3 += s_rhs[(q + 3) * HNEQ + r];
    }
    for(; q < qe; ++q) s0 += s_rhs[q * HNEQ + r];
    out[(uint64_t)(r0 + row) * HNEQ + r] = (s0 + s1) + (s2 + s3);
  }
}

template<int MODE>
void launch_mode(const AsmArgs& g, const Lss& L)
{
  Context& c = *L.ctx;
  const int timer = MODE == MODE_RESIDUAL ? CF3HIP_KERNEL_RESIDUAL : CF3HIP_KERNEL_ASSEMBLY;
  timer_begin(c, timer);
  if(g.supg_type == CF3HIP_SUPG_TEZDUYAR)
    hipLaunchKernelGGL((k_hexa<MODE, CF3HIP_SUPG_TEZDUYAR>), dim3(L.n_wg), dim3(kThreads), 0, c.stream, g);
  else
    hipLaunchKernelGGL((k_hexa<MODE, CF3HIP_SUPG_METRIC>), dim3(L.n_wg), dim3(kThreads), 0, c.stream, g);
  HIP_OK(hipGetLastError());
  timer_end(c, timer);
}

}  // namespace

void launch_hex(const AsmArgs& g, const Lss& L, int mode)
{
  if(L.passes != 1 || L.pos_stride != kLanes) CF3_THROW(CF3HIP_ERR_INVALID_ARG, "system has no hexahedral staging maps");
  if(mode == MODE_FP64) launch_mode<MODE_FP64>(g, L);
  else if(mode == MODE_FP32) launch_mode<MODE_FP32>(g, L);
  else launch_mode<MODE_RESIDUAL>(g, L);
}

}  // namespace asmk
}  // namespace cf3hip

// Definitions shared by the element assembly translation units (assembly.hip, assembly_hex.hip).
#pragma once

#include <type_traits>

#include "internal.hpp"

namespace cf3hip {
namespace asmk {

constexpr int kThreads = kAsmThreads;   // one incidence and at most one owned CSR block per thread

struct AsmArgs {
  const double4* xp;   // packed coordinates (x, y, z | 0, 0)
  const double4* st;   // packed state (u, v, w, p) in 3D, (u, v, p, 0) in 2D
  const double4* ad;   // packed (u_adv x, y, z | 0, nu_eff)
  const double* g[3];  // body force (SoA), generic expression only
  const uint32_t* conn;
  const uint32_t* wg_rows;
  const uint32_t* row_ptr;
  const uint32_t* inc_ptr;
  const uint32_t* inc;
  const uint8_t* inc_pos;   // per incidence: staging row of its block in pass b (byte b)
  const uint8_t* blk_beg;   // per block: first staging row of pass p (byte p)
  const uint16_t* blk_acc;  // per block: first accumulator thread (low byte), accumulator count (high byte)
  const uint32_t* wg_head;  // node staging: per workgroup r0, r1, i0, i1, b0, b1, n0, n1
  const uint32_t* wg_nodes; // node staging: distinct nodes of each workgroup
  const uint32_t* inc_loc;  // node staging: per incidence the local indices of its element's nodes
  void* values;
  double* rhs;
  const double* xt;    // matrix-free residual: x with constrained columns set to their first value
  double* res;         // matrix-free residual output
  double rho, inv_rho, dt, theta, thq, invdt, f4dt2;   // thq = theta * (1 / theta); f4dt2 = 4 / dt^2
  int supg_type;
  int body_force;      // nonzero when the body force field holds a nonzero value
  double c1, c2, a_ps, a_su, a_bu;
  double u_ref;        // CF2 reference velocity
  int* status;         // 1: non-positive Jacobian determinant, 2: CF2 element Reynolds number not positive
  double* q1rec;       // Q1 cells: per element and Gauss point J^-1, det J, u_adv, nu (Q1Rec doubles)
  double4* q1tau;      // Q1 cells: per element (tau_ps, tau_su, tau_bulk, 0)
};

// ComputeTau CF2 (plugins/UFEM/src/UFEM/SUPG.hpp:191-210).  vol: the element volume the reference takes
// (etype::volume); um: |u| at the centroid; s: sum over the nodes of u . grad N_n at the centroid (zero
// up to rounding, since the gradients sum to zero: tau_su then comes out as dt / 2 to within rounding).
// pi is the reference's 3.141592654.  Returns false when Re_e = u_ref he / (2 nu) is not positive (the
// reference's cf3_assert).
__device__ __forceinline__ bool tau_cf2(int dim, double vol, double nu_e, double um, double s, double dt,
                                        double u_ref, double& tps, double& tsu, double& tbu)
{
  const double he = dim == 2 ? sqrt(4. / 3.141592654 * vol) : pow(3. / 4. / 3.141592654 * vol, 1. / 3.);
  const double ree = u_ref * he / (2. * nu_e);
  const double xi = ree < 3. ? ree / 3. : 1.;
  tps = he * xi / (2. * u_ref);
  tbu = he * u_ref / xi;
  tsu = 0.;
  if(um > 1e-10) {
    const double h = 2. * um / s;
    const double tau_adv = h / (2. * um);
    const double tau_time = 0.5 * dt;
    const double tau_diff = h * h / (4. * nu_e);
    tsu = 1. / (1. / tau_adv + 1. / tau_time + 1. / tau_diff);
  }
  return ree > 0.;
}

// Tetra3D::volume (cf3/mesh/LagrangeP1/Tetra3D.cpp:104-128, 205-208): the 24-term expansion of det J
// in absolute coordinates, with the reference's products, sums and final division in their source order
// and no contraction into FMAs.  Its rounding error grows like |x|^3 / V (about 1e-11 relative for
// 64^3 cells of the unit cube): computed any other way, results would differ from the reference's by that.
__device__ __forceinline__ double tet_volume_ref(const double X[4][3])
{
#pragma clang fp contract(off)
  const double x0 = X[0][0], y0 = X[0][1], z0 = X[0][2];
  const double x1 = X[1][0], y1 = X[1][1], z1 = X[1][2];
  const double x2 = X[2][0], y2 = X[2][1], z2 = X[2][2];
  const double x3 = X[3][0], y3 = X[3][1], z3 = X[3][2];
  return (x2*y1*z0 - x3*y1*z0 - x1*y2*z0 + x3*y2*z0 + x1*y3*z0 -
          x2*y3*z0 - x2*y0*z1 + x3*y0*z1 + x0*y2*z1 - x3*y2*z1 -
          x0*y3*z1 + x2*y3*z1 + x1*y0*z2 - x3*y0*z2 - x0*y1*z2 +
          x3*y1*z2 + x0*y3*z2 - x1*y3*z2 - x1*y0*z3 + x2*y0*z3 +
          x0*y1*z3 - x2*y1*z3 - x0*y2*z3 + x1*y2*z3) / 6.;
}

// one LDS row of a staged element block: 16-byte aligned, odd number of 16-byte chunks (bank spread)
template<typename VR, int NN>
constexpr int lds_row()
{
  constexpr int bytes = (NN * (int)sizeof(VR) + 15) / 16 * 16;
  constexpr int padded = (bytes / 16) % 2 == 0 ? bytes + 16 : bytes;
  return padded / (int)sizeof(VR);
}

__device__ __forceinline__ void unpack(const double4& v, double o[4]) { o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w; }

// XCD-aware workgroup order: workgroups are dispatched round-robin over the 8 XCDs of the MI355X, each
// with its own L2.  Workgroup i (the k-th on XCD i mod 8) takes logical workgroup k of that XCD's
// contiguous share, so each L2 serves a compact part of the mesh (node records and maps shared by
// neighbouring row ranges).  A bijection of [0, n); the result of a logical workgroup does not depend
// on where it runs.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_order(uint32_t i, uint32_t n)
{
  const uint32_t q = n / kXcds, r = n % kXcds;
  const uint32_t x = i % kXcds, k = i / kXcds;
  return x * q + (x < r ? x : r) + k;
}

template<int K>
__device__ __forceinline__ void load_conn(const uint32_t* conn, uint32_t e, uint32_t cn[K])
{
  if constexpr(K == 4) {
    const uint4 c = reinterpret_cast<const uint4*>(conn)[e];
    cn[0] = c.x; cn[1] = c.y; cn[2] = c.z; cn[3] = c.w;
  } else {
#pragma unroll
    for(int m = 0; m < K; ++m) cn[m] = conn[(uint64_t)e * K + m];
  }
}

// nodal unknowns (u.., p) of record nd of st (the packed state, or its LDS copy) in arithmetic type R
template<int NEQ, typename R>
__device__ __forceinline__ void load_state(const double4* st, uint32_t nd, R xb[NEQ])
{
  double s[4];
  unpack(st[nd], s);
#pragma unroll
  for(int c = 0; c < NEQ; ++c) xb[c] = (R)s[c];
}
template<int NEQ, typename R>
__device__ __forceinline__ void load_state(const AsmArgs& g, uint32_t nd, R xb[NEQ])
{
  load_state<NEQ, R>(g.st, nd, xb);
}

// LDS row access in 16-byte pieces
template<typename VR, int NN>
__device__ __forceinline__ void lds_load_row(const VR* src, VR v[NN])
{
  if constexpr((NN * sizeof(VR)) % 16 == 0) {
    typedef typename std::conditional<sizeof(VR) == 8, double2, float4>::type V16;
    constexpr int W = 16 / sizeof(VR);
#pragma unroll
    for(int c = 0; c < NN / W; ++c) {
      const V16 x = reinterpret_cast<const V16*>(src)[c];
      if constexpr(W == 2) { v[2 * c] = x.x; v[2 * c + 1] = x.y; }
      else { v[4 * c] = x.x; v[4 * c + 1] = x.y; v[4 * c + 2] = x.z; v[4 * c + 3] = x.w; }
    }
  } else {
#pragma unroll
    for(int c = 0; c < NN; ++c) v[c] = src[c];
  }
}

// mode 0 = FP64 matrix + rhs, 1 = FP32 matrix, 2 = matrix-free residual
void launch_hex(AsmArgs g, Lss& L, int mode);

}  // namespace asmk
}  // namespace cf3hip

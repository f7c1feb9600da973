// Oracle: CPU FP64 restatement of the COOLFluiD 3 UFEM Navier-Stokes hot path.
// TEST INFRASTRUCTURE ONLY (the test oracle).  It is loaded by tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg; the product library never links or calls it.
//
// Reference citations are relative to the coolfluid3 source tree.
//
// Parity status of each part (pins live in tests/test_oracle_*.py):
//   geometry / shape functions / Gauss rules ... pinned by the reference's known-answer fixtures
//                                                (test/mesh/utest-mesh-lagrangep1-*.cpp) and exactness
//   sparsity ................................... pinned by utest-lss-test-matrix.hpp element lists/CSR
//                                                and brute-force pair enumeration
//   specialised P1 kernels ..................... pinned by closed forms and by the generic quadrature
//                                                (agreement / documented differences)
//   generic quadrature (hex) ................... pinned by brute-force dense assembly, patch tests,
//                                                integration exactness
//   tau METRIC ................................. closed form on reference elements
//   symmetric Dirichlet / set_row / apply ...... utest-lss-symmetric-dirichlet.cpp, utest-lss-vector.cpp
//   BiCGStab / GMRES+ILU0 ...................... Trilinos arithmetic is third-party: "parity by
//                                                specification", pinned by dense direct solves
//   semi-implicit actions (orc_si_*) ........... lumped mass and pressure matrix: closed forms on the
//                                                unit square; URHS, PRHS, ApplyAup and the velocity
//                                                diagonal blocks: the implicit generic operator's blocks
//                                                on periodic meshes (tests/test_semi_implicit.py); the
//                                                off-diagonal one-point bulk blocks: symmetry only
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <set>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

typedef double Real;
typedef uint32_t Uint;

inline int enodes(int et) { return et == ORC_TRIAG2D ? 3 : (et == ORC_HEXA3D ? 8 : (et == ORC_PRISM3D ? 6 : 4)); }
inline int edim(int et) { return et == ORC_TRIAG2D || et == ORC_QUAD2D ? 2 : 3; }

// ---------------------------------------------------------------------------------------------
// Shape functions: Triag.cpp:24-41, Tetra.cpp:24-48, Hexa.cpp:24-89, Quad.cpp (compute_value,
// compute_gradient: node order (-1,-1), (1,-1), (1,1), (-1,1)), Prism.cpp:24-58 (the triangle
// (0,0), (1,0), (0,1) at zeta = -1, then at zeta = +1)
// ---------------------------------------------------------------------------------------------
void shape_values(int et, const Real* xi, Real* N)
{
  if(et == ORC_PRISM3D) {
    const Real lo = 0.5 * (1. - xi[2]), hi = 0.5 * (1. + xi[2]), l0 = 1. - xi[0] - xi[1];
    N[0] = lo * l0; N[1] = lo * xi[0]; N[2] = lo * xi[1];
    N[3] = hi * l0; N[4] = hi * xi[0]; N[5] = hi * xi[1];
    return;
  }
  if(et == ORC_QUAD2D) {
    const Real ksi = xi[0], eta = xi[1];
    N[0] = 0.25 * (1.0 - ksi) * (1.0 - eta);
    N[1] = 0.25 * (1.0 + ksi) * (1.0 - eta);
    N[2] = 0.25 * (1.0 + ksi) * (1.0 + eta);
    N[3] = 0.25 * (1.0 - ksi) * (1.0 + eta);
  } else if(et == ORC_TRIAG2D) {
    N[0] = 1. - xi[0] - xi[1];
    N[1] = xi[0];
    N[2] = xi[1];
  } else if(et == ORC_TETRA3D) {
    N[0] = 1. - xi[0] - xi[1] - xi[2];
    N[1] = xi[0];
    N[2] = xi[1];
    N[3] = xi[2];
  } else {
    const Real a1 = (1 + xi[0]), a2 = (1 - xi[0]);
    const Real b1 = (1 + xi[1]), b2 = (1 - xi[1]);
    const Real c1 = (1 + xi[2]), c2 = (1 - xi[2]);
    N[0] = a2*b2*c2; N[1] = a1*b2*c2; N[2] = a1*b1*c2; N[3] = a2*b1*c2;
    N[4] = a2*b2*c1; N[5] = a1*b2*c1; N[6] = a1*b1*c1; N[7] = a2*b1*c1;
    for(int i = 0; i < 8; ++i) N[i] *= 0.125;
  }
}

// dN(xi_dir, node) row-major dim x k
void shape_gradients(int et, const Real* xi, Real* dN)
{
  if(et == ORC_PRISM3D) {
    const Real ksi = xi[0], eta = xi[1], zta = xi[2];
    Real* K = dN; Real* E = dN + 6; Real* Z = dN + 12;
    K[0] = 0.5 * (zta - 1.); E[0] = 0.5 * (zta - 1.); Z[0] = 0.5 * (ksi + eta - 1.);
    K[1] = 0.5 * (1. - zta); E[1] = 0.;              Z[1] = -0.5 * ksi;
    K[2] = 0.;               E[2] = 0.5 * (1. - zta); Z[2] = -0.5 * eta;
    K[3] = -0.5 * (zta + 1.); E[3] = -0.5 * (zta + 1.); Z[3] = 0.5 * (1. - ksi - eta);
    K[4] = 0.5 * (1. + zta); E[4] = 0.;              Z[4] = 0.5 * ksi;
    K[5] = 0.;               E[5] = 0.5 * (1. + zta); Z[5] = 0.5 * eta;
    return;
  }
  if(et == ORC_QUAD2D) {
    const Real ksi = xi[0], eta = xi[1];
    Real* K = dN; Real* E = dN + 4;
    K[0] = 0.25 * (-1. + eta); E[0] = 0.25 * (-1. + ksi);
    K[1] = 0.25 * ( 1. - eta); E[1] = 0.25 * (-1. - ksi);
    K[2] = 0.25 * ( 1. + eta); E[2] = 0.25 * ( 1. + ksi);
    K[3] = 0.25 * (-1. - eta); E[3] = 0.25 * ( 1. - ksi);
  } else if(et == ORC_TRIAG2D) {
    const Real g[6] = {-1., 1., 0., -1., 0., 1.};
    std::memcpy(dN, g, sizeof(g));
  } else if(et == ORC_TETRA3D) {
    const Real g[12] = {-1., 1., 0., 0., -1., 0., 1., 0., -1., 0., 0., 1.};
    std::memcpy(dN, g, sizeof(g));
  } else {
    const Real a1 = (1 + xi[0]), a2 = (1 - xi[0]);
    const Real b1 = (1 + xi[1]), b2 = (1 - xi[1]);
    const Real c1 = (1 + xi[2]), c2 = (1 - xi[2]);
    Real* K = dN; Real* E = dN + 8; Real* Z = dN + 16;
    K[0] = -0.125*b2*c2; E[0] = -0.125*a2*c2; Z[0] = -0.125*a2*b2;
    K[1] =  0.125*b2*c2; E[1] = -0.125*a1*c2; Z[1] = -0.125*a1*b2;
    K[2] =  0.125*b1*c2; E[2] =  0.125*a1*c2; Z[2] = -0.125*a1*b1;
    K[3] = -0.125*b1*c2; E[3] =  0.125*a2*c2; Z[3] = -0.125*a2*b1;
    K[4] = -0.125*b2*c1; E[4] = -0.125*a2*c1; Z[4] =  0.125*a2*b2;
    K[5] =  0.125*b2*c1; E[5] = -0.125*a1*c1; Z[5] =  0.125*a1*b2;
    K[6] =  0.125*b1*c1; E[6] =  0.125*a1*c1; Z[6] =  0.125*a1*b1;
    K[7] = -0.125*b1*c1; E[7] =  0.125*a2*c1; Z[7] =  0.125*a2*b1;
  }
}

// ---------------------------------------------------------------------------------------------
// Jacobians: Triag2D.cpp:134-140, Tetra3D.cpp:131-175, Hexa3D.cpp:246-287 (J(ksi, x), row-major)
// ---------------------------------------------------------------------------------------------
void jacobian(int et, const Real* xi, const Real* X, Real* J)
{
  if(et == ORC_PRISM3D) {
    // Prism3D::compute_jacobian (Prism3D.cpp:93-127): J(k, d) = sum_m dN_m/dxi_k x_m,d, grouped as the
    // bottom / top node differences (which cancel exactly on right prisms)
    const Real ksi = xi[0], eta = xi[1], zta = xi[2];
    const Real lo = -0.5*zta + 0.5, hi = 0.5*zta + 0.5, l0 = -eta - ksi + 1;
    for(int d = 0; d < 3; ++d) {
      const Real x0 = X[0*3+d], x1 = X[1*3+d], x2 = X[2*3+d], x3 = X[3*3+d], x4 = X[4*3+d], x5 = X[5*3+d];
      J[0*3+d] = -x0*lo + x1*lo - x3*hi + x4*hi;
      J[1*3+d] = -x0*lo + x2*lo - x3*hi + x5*hi;
      J[2*3+d] = -0.5*eta*x2 + 0.5*eta*x5 - 0.5*ksi*x1 + 0.5*ksi*x4 - 0.5*x0*l0 + 0.5*x3*l0;
    }
    return;
  }
  if(et == ORC_QUAD2D) {
    // Quad2D::compute_jacobian with the JacobianCoefficients of Quad2D.hpp:71-96
    const Real bx = 0.25*(-X[0] + X[2] + X[4] - X[6]), cx = 0.25*(-X[0] - X[2] + X[4] + X[6]);
    const Real dx = 0.25*( X[0] - X[2] + X[4] - X[6]);
    const Real by = 0.25*(-X[1] + X[3] + X[5] - X[7]), cy = 0.25*(-X[1] - X[3] + X[5] + X[7]);
    const Real dy = 0.25*( X[1] - X[3] + X[5] - X[7]);
    J[0] = bx + dx*xi[1]; J[1] = by + dy*xi[1];
    J[2] = cx + dx*xi[0]; J[3] = cy + dy*xi[0];
  } else if(et == ORC_TRIAG2D) {
    J[0] = X[2] - X[0]; J[1] = X[3] - X[1];
    J[2] = X[4] - X[0]; J[3] = X[5] - X[1];
  } else if(et == ORC_TETRA3D) {
    for(int d = 0; d < 3; ++d) {
      J[0*3+d] = -X[0*3+d] + X[1*3+d];
      J[1*3+d] = -X[0*3+d] + X[2*3+d];
      J[2*3+d] = -X[0*3+d] + X[3*3+d];
    }
  } else {
    const Real xi0 = xi[0], eta = xi[1], zeta = xi[2];
    for(int d = 0; d < 3; ++d) {
      const Real v0 = X[0*3+d], v1 = X[1*3+d], v2 = X[2*3+d], v3 = X[3*3+d];
      const Real v4 = X[4*3+d], v5 = X[5*3+d], v6 = X[6*3+d], v7 = X[7*3+d];
      J[0*3+d] = 0.125 * (v1 + v2 + v5 + v6 - v0 - v3 - v4 - v7 + eta*(v0 + v2 + v4 + v6 - v1 - v3 - v5 - v7) + zeta*(v0 + v3 + v5 + v6 - v1 - v2 - v4 - v7) + eta*zeta*(v1 + v3 + v4 + v6 - v0 - v2 - v5 - v7));
      J[1*3+d] = 0.125 * (v2 + v3 + v6 + v7 - v0 - v1 - v4 - v5 + xi0*(v0 + v2 + v4 + v6 - v1 - v3 - v5 - v7) + zeta*(v0 + v1 + v6 + v7 - v2 - v3 - v4 - v5) + xi0*zeta*(v1 + v3 + v4 + v6 - v0 - v2 - v5 - v7));
      J[2*3+d] = 0.125 * (v4 + v5 + v6 + v7 - v0 - v1 - v2 - v3 + eta*(v0 + v1 + v6 + v7 - v2 - v3 - v4 - v5) + xi0*(v0 + v3 + v5 + v6 - v1 - v2 - v4 - v7) + eta*xi0*(v1 + v3 + v4 + v6 - v0 - v2 - v5 - v7));
    }
  }
}

// adjugate of a 3x3 matrix (Tetra3D.cpp:188-200, Hexa3D.cpp:296-308)
void adjoint3(const Real* J, Real* r)
{
#define Jm(i,j) J[(i)*3+(j)]
  r[0] =  (Jm(1,1)*Jm(2,2) - Jm(1,2)*Jm(2,1));
  r[1] = -(Jm(0,1)*Jm(2,2) - Jm(0,2)*Jm(2,1));
  r[2] =  (Jm(0,1)*Jm(1,2) - Jm(1,1)*Jm(0,2));
  r[3] = -(Jm(1,0)*Jm(2,2) - Jm(1,2)*Jm(2,0));
  r[4] =  (Jm(0,0)*Jm(2,2) - Jm(0,2)*Jm(2,0));
  r[5] = -(Jm(0,0)*Jm(1,2) - Jm(0,2)*Jm(1,0));
  r[6] =  (Jm(1,0)*Jm(2,1) - Jm(1,1)*Jm(2,0));
  r[7] = -(Jm(0,0)*Jm(2,1) - Jm(0,1)*Jm(2,0));
  r[8] =  (Jm(0,0)*Jm(1,1) - Jm(0,1)*Jm(1,0));
#undef Jm
}

// Inverse and determinant of J (ElementData.hpp:302-314 uses Eigen's computeInverseAndDetWithCheck;
// restated from the adjugate formula: det by cofactor expansion along column 0, inv = adj / det).
bool inverse_det(int dim, const Real* J, Real* Jinv, Real& det)
{
  if(dim == 2) {
    det = J[0]*J[3] - J[2]*J[1];
    if(!(std::fabs(det) > 0.)) return false;
    const Real inv = 1. / det;
    Jinv[0] = J[3]*inv; Jinv[1] = -J[1]*inv;
    Jinv[2] = -J[2]*inv; Jinv[3] = J[0]*inv;
    return true;
  }
  // cofactors of column 0
  const Real c00 = J[4]*J[8] - J[5]*J[7];
  const Real c10 = J[7]*J[2] - J[8]*J[1];
  const Real c20 = J[1]*J[5] - J[2]*J[4];
  det = c00*J[0] + c10*J[3] + c20*J[6];
  if(!(std::fabs(det) > 0.)) return false;
  const Real inv = 1. / det;
  Jinv[0] = c00*inv; Jinv[1] = c10*inv; Jinv[2] = c20*inv;
  Jinv[3] = (J[5]*J[6] - J[3]*J[8])*inv;
  Jinv[4] = (J[8]*J[0] - J[6]*J[2])*inv;
  Jinv[5] = (J[2]*J[3] - J[0]*J[5])*inv;
  Jinv[6] = (J[3]*J[7] - J[4]*J[6])*inv;
  Jinv[7] = (J[6]*J[1] - J[7]*J[0])*inv;
  Jinv[8] = (J[0]*J[4] - J[1]*J[3])*inv;
  return true;
}

// Tetra3D::jacobian_determinant (Tetra3D.cpp:104-128)
Real tetra_det(const Real* X)
{
  const Real x0 = X[0], y0 = X[1], z0 = X[2];
  const Real x1 = X[3], y1 = X[4], z1 = X[5];
  const Real x2 = X[6], y2 = X[7], z2 = X[8];
  const Real x3 = X[9], y3 = X[10], z3 = X[11];
  return
      x2*y1*z0 - x3*y1*z0 - x1*y2*z0 + x3*y2*z0 + x1*y3*z0 -
      x2*y3*z0 - x2*y0*z1 + x3*y0*z1 + x0*y2*z1 - x3*y2*z1 -
      x0*y3*z1 + x2*y3*z1 + x1*y0*z2 - x3*y0*z2 - x0*y1*z2 +
      x3*y1*z2 + x0*y3*z2 - x1*y3*z2 - x1*y0*z3 + x2*y0*z3 +
      x0*y1*z3 - x2*y1*z3 - x0*y2*z3 + x1*y2*z3;
}

// Triag2D jacobian_determinant_helper (Triag2D.cpp:44-53)
Real triag_det(const Real* X)
{
  return (X[2] - X[0]) * (X[5] - X[1]) - (X[4] - X[0]) * (X[3] - X[1]);
}

// Hexa3D::jacobian_determinant (Hexa3D.cpp:169-239)
Real hexa_det(const Real* m, const Real* X)
{
#define XC(n) X[(n)*3+0]
#define YC(n) X[(n)*3+1]
#define ZC(n) X[(n)*3+2]
  const Real x0 = XC(0), x1 = XC(1), x2 = XC(2), x3 = XC(3), x4 = XC(4), x5 = XC(5), x6 = XC(6), x7 = XC(7);
  const Real y0 = YC(0), y1 = YC(1), y2 = YC(2), y3 = YC(3), y4 = YC(4), y5 = YC(5), y6 = YC(6), y7 = YC(7);
  const Real z0 = ZC(0), z1 = ZC(1), z2 = ZC(2), z3 = ZC(3), z4 = ZC(4), z5 = ZC(5), z6 = ZC(6), z7 = ZC(7);
#undef XC
#undef YC
#undef ZC
  const Real xi = m[0], eta = m[1], zeta = m[2];
  const Real etaxi = (eta*xi), etazeta = (eta*zeta), xizeta = (xi*zeta);
  const Real f1 = 0.125*(-1.0 + eta + xi - etaxi);
  const Real f2 = 0.125*(-1.0 + eta - xi + etaxi);
  const Real f3 = 0.125*(-1.0 - eta - xi - etaxi);
  const Real f4 = 0.125*(-1.0 - eta + xi + etaxi);
  const Real f5 = 0.125*(+1.0 - eta - xi + etaxi);
  const Real f6 = 0.125*(+1.0 - eta + xi - etaxi);
  const Real f7 = 0.125*(+1.0 + eta + xi + etaxi);
  const Real f8 = 0.125*(+1.0 + eta - xi - etaxi);
  const Real f9  = 0.125*(-1.0 - eta - zeta - etazeta);
  const Real f10 = 0.125*(+1.0 + eta - zeta - etazeta);
  const Real f11 = 0.125*(+1.0 - eta + zeta - etazeta);
  const Real f12 = 0.125*(-1.0 + eta + zeta - etazeta);
  const Real f13 = 0.125*(+1.0 - eta - zeta + etazeta);
  const Real f14 = 0.125*(-1.0 + eta - zeta + etazeta);
  const Real f15 = 0.125*(-1.0 - eta + zeta + etazeta);
  const Real f16 = 0.125*(+1.0 + eta + zeta + etazeta);
  const Real f17 = 0.125*(-1.0 - xi - zeta - xizeta);
  const Real f18 = 0.125*(+1.0 + xi - zeta - xizeta);
  const Real f19 = 0.125*(+1.0 - xi + zeta - xizeta);
  const Real f20 = 0.125*(-1.0 + xi + zeta - xizeta);
  const Real f21 = 0.125*(+1.0 - xi - zeta + xizeta);
  const Real f22 = 0.125*(-1.0 + xi - zeta + xizeta);
  const Real f23 = 0.125*(-1.0 - xi + zeta + xizeta);
  const Real f24 = 0.125*(+1.0 + xi + zeta + xizeta);
  return (f1*z0 + f2*z1 + f3*z2 + f4*z3 + f5*z4 + f6*z5 + f7*z6 + f8*z7)*
         (-((y7*f9 + y2*f10 + y5*f11 + y0*f12 + y1*f13 + y4*f14 + y3*f15 + y6*f16)*
         (x5*f17 + x2*f18 + x7*f19 + x0*f20 + x3*f21 + x4*f22 + x1*f23 + x6*f24)) +
         (x7*f9 + x2*f10 + x5*f11 + x0*f12 + x1*f13 + x4*f14 + x3*f15 + x6*f16)*
         (y5*f17 + y2*f18 + y7*f19 + y0*f20 + y3*f21 + y4*f22 + y1*f23 + y6*f24)) -
         (f1*y0 + f2*y1 + f3*y2 + f4*y3 + f5*y4 + f6*y5 + f7*y6 + f8*y7)*
         (-((z7*f9 + z2*f10 + z5*f11 + z0*f12 + z1*f13 + z4*f14 + z3*f15 + z6*f16)*
         (x5*f17 + x2*f18 + x7*f19 + x0*f20 + x3*f21 + x4*f22 + x1*f23 + x6*f24)) +
         (x7*f9 + x2*f10 + x5*f11 + x0*f12 + x1*f13 + x4*f14 + x3*f15 + x6*f16)*
         (z5*f17 + z2*f18 + z7*f19 + z0*f20 + z3*f21 + z4*f22 + z1*f23 + z6*f24)) +
         (x2*f3 + x7*f8 + x5*f6 + x0*f1 + x4*f5 + x1*f2 + x3*f4 + x6*f7)*
         (-((z7*f9 + z2*f10 + z5*f11 + z0*f12 + z1*f13 + z4*f14 + z3*f15 + z6*f16)*
         (y5*f17 + y2*f18 + y7*f19 + y0*f20 + y3*f21 + y4*f22 + y1*f23 + y6*f24)) +
         (y7*f9 + y2*f10 + y5*f11 + y0*f12 + y1*f13 + y4*f14 + y3*f15 + y6*f16)*
         (z5*f17 + z2*f18 + z7*f19 + z0*f20 + z3*f21 + z4*f22 + z1*f23 + z6*f24));
}

// Element volume: Triag2D.cpp:150-153 (0.5 det), Tetra3D.cpp:205-208 (det/6), Hexa3D.cpp:317-320 (8 det(0))
Real volume(int et, const Real* X)
{
  if(et == ORC_TRIAG2D) return 0.5 * triag_det(X);
  if(et == ORC_QUAD2D)   // Quad2D::volume: half the cross product of the diagonals
    return 0.5 * ((X[4] - X[0]) * (X[7] - X[3]) - (X[5] - X[1]) * (X[6] - X[2]));
  if(et == ORC_TETRA3D) return tetra_det(X) / 6.;
  if(et == ORC_PRISM3D) {   // Prism3D::volume (Prism3D.cpp:163-170): det J at (1/3, 1/3, 0)
    const Real c[3] = {0.3333333333333333333333333, 0.3333333333333333333333333, 0.};
    Real J[9], Ji[9], det;
    jacobian(et, c, X, J);
    inverse_det(3, J, Ji, det);
    return det;
  }
  const Real zero[3] = {0., 0., 0.};
  return 8 * hexa_det(zero, X);
}

// ---------------------------------------------------------------------------------------------
// Gauss rules: GaussImplementation.hpp:29-35 (x=1/sqrt3, w=1), 151-168 (hex 1pt), 170-226 (tri),
// 370-428 (tet), 615-676 (hex 8pt ordering)
// ---------------------------------------------------------------------------------------------
int gauss(int et, int order, Real* xi, Real* w)
{
  if(et == ORC_QUAD2D) {   // GaussImplementation.hpp:131-148 (centre, w 4) and :566-600 (2 x 2 order)
    if(order == 1) { xi[0] = xi[1] = 0.; w[0] = 4.; return 1; }
    const Real x = 0.5773502691896257645091488;
    const Real s[4][2] = {{x, x}, {-x, x}, {x, -x}, {-x, -x}};
    for(int i = 0; i < 4; ++i) { xi[2*i] = s[i][0]; xi[2*i+1] = s[i][1]; w[i] = 1.; }
    return 4;
  }
  if(et == ORC_TRIAG2D) {
    if(order == 1) { xi[0] = xi[1] = 0.3333333333333333333333333; w[0] = 0.5; return 1; }
    const Real c[6] = {0.5, 0.0, 0.5, 0.5, 0.0, 0.5};
    std::memcpy(xi, c, sizeof(c));
    w[0] = w[1] = w[2] = 1.0/6.0;
    return 3;
  }
  if(et == ORC_TETRA3D) {
    if(order == 1) { xi[0] = xi[1] = xi[2] = 0.25; w[0] = 0.1666666666666666666666667; return 1; }
    const Real mu = 0.13819660112501051517954131656;
    const Real c[12] = {mu, mu, mu, 1.-3.*mu, mu, mu, mu, 1.-3.*mu, mu, mu, mu, 1.-3.*mu};
    std::memcpy(xi, c, sizeof(c));
    for(int i = 0; i < 4; ++i) w[i] = 0.04166666666666666666666666666666666667;
    return 4;
  }
  if(et == ORC_PRISM3D) {   // GaussImplementation.hpp:430-494: (1/3, 1/3, 0) w 1; the triangle's edge
                            // midpoints at zeta = +x, then at zeta = -x, w 1/6
    if(order == 1) { xi[0] = xi[1] = 0.3333333333333333333333333; xi[2] = 0.; w[0] = 1.; return 1; }
    const Real x = 0.5773502691896257645091488;
    const Real t[3][2] = {{0.5, 0.0}, {0.5, 0.5}, {0.0, 0.5}};
    for(int q = 0; q < 6; ++q) {
      xi[3*q] = t[q % 3][0]; xi[3*q+1] = t[q % 3][1]; xi[3*q+2] = q < 3 ? x : -x;
      w[q] = 1. / 6.;
    }
    return 6;
  }
  if(order == 1) { xi[0] = xi[1] = xi[2] = 0.; w[0] = 8.; return 1; }
  const Real x = 0.5773502691896257645091488;
  const Real s[8][3] = {{x,x,x},{-x,x,x},{x,-x,x},{x,x,-x},{x,-x,-x},{-x,x,-x},{-x,-x,x},{-x,-x,-x}};
  for(int i = 0; i < 8; ++i) { xi[3*i] = s[i][0]; xi[3*i+1] = s[i][1]; xi[3*i+2] = s[i][2]; w[i] = 1.*1.*1.; }
  return 8;
}

// ---------------------------------------------------------------------------------------------
// Tau (SUPG.hpp:137-219).  u: nodal values (k x dim) of the velocity used for tau.
// ---------------------------------------------------------------------------------------------
struct TauCfg { int type; Real c1, c2, a_ps, a_su, a_bu, u_ref; };

int compute_tau(int et, const TauCfg& cfg, const Real* X, const Real* u, Real element_nu, Real dt,
                Real& tau_ps, Real& tau_su, Real& tau_bulk)
{
  const int k = enodes(et), dim = edim(et);
  Real xi[3], w[1];
  gauss(et, 1, xi, w);
  Real N[8];
  shape_values(et, xi, N);
  // u.compute_values(centroid): u.eval() = N * u_nodal (1 x dim)
  Real uc[3] = {0., 0., 0.};
  for(int d = 0; d < dim; ++d) {
    Real s = 0.;
    for(int n = 0; n < k; ++n) s += N[n] * u[n*dim+d];
    uc[d] = s;
  }
  if(cfg.type == ORC_SUPG_TEZDUYAR) {
    Real h = 1e10;
    for(int i = 0; i < k; ++i)
      for(int j = 0; j < k; ++j)
        if(i != j) {
          Real s = 0.;
          for(int d = 0; d < dim; ++d) { const Real df = X[i*dim+d] - X[j*dim+d]; s += df*df; }
          h = std::min(h, s);
        }
    h = std::sqrt(h);
    Real umag2 = 0.;
    for(int d = 0; d < dim; ++d) umag2 += uc[d]*uc[d];
    const Real umag = std::sqrt(umag2);
    const Real tau_adv_inv = (2.*umag)/h;
    const Real tau_time_inv = 2./dt;
    const Real tau_diff_inv = (4.*element_nu)/(h*h);
    tau_su = 1./(tau_adv_inv + tau_time_inv + tau_diff_inv);
    tau_ps = tau_su;
    tau_bulk = tau_su*umag*umag;
  } else if(cfg.type == ORC_SUPG_METRIC) {
    Real J[9], Ji[9], det;
    jacobian(et, xi, X, J);
    if(!inverse_det(dim, J, Ji, det)) return -2;
    // gij = Jinv^T * Jinv
    Real g[9];
    for(int i = 0; i < dim; ++i)
      for(int j = 0; j < dim; ++j) {
        Real s = 0.;
        for(int l = 0; l < dim; ++l) s += Ji[l*dim+i] * Ji[l*dim+j];
        g[i*dim+j] = s;
      }
    // u * gij * u^T
    Real ugu = 0.;
    for(int j = 0; j < dim; ++j) {
      Real s = 0.;
      for(int i = 0; i < dim; ++i) s += uc[i] * g[i*dim+j];
      ugu += s * uc[j];
    }
    const Real tau_adv_sq = std::fabs(ugu);
    Real gsq = 0., tr = 0.;
    for(int i = 0; i < dim*dim; ++i) gsq += g[i]*g[i];
    for(int i = 0; i < dim; ++i) tr += g[i*dim+i];
    const Real tau_diff = element_nu*element_nu*gsq;
    tau_su = 1. / std::sqrt((4.*cfg.c1*cfg.c1/(dt*dt)) + tau_adv_sq + cfg.c2*tau_diff);
    tau_ps = tau_su;
    const Real tau_su_std = 1. / std::sqrt((4./(dt*dt)) + tau_adv_sq + 16.*tau_diff);
    tau_bulk = (1./tau_su_std) / tr;
  } else if(cfg.type == ORC_SUPG_CF2) {
    // SUPG.hpp:191-210.  he from the element volume with pi as written there (3.141592654)
    const Real vol = volume(et, X);
    const Real he = dim == 2 ? std::sqrt(4./3.141592654*vol) : std::pow(3./4./3.141592654*vol, 1./3.);
    const Real ree = cfg.u_ref*he/(2.*element_nu);
    if(!(ree > 0.)) return -3;   // cf3_assert(ree > 0.)
    const Real xi_ = ree < 3. ? ree/3. : 1.;
    tau_ps = he*xi_/(2.*cfg.u_ref);
    tau_bulk = he*cfg.u_ref/xi_;
    tau_su = 0.;
    Real umag2 = 0.;
    for(int d = 0; d < dim; ++d) umag2 += uc[d]*uc[d];
    const Real umag = std::sqrt(umag2);
    if(umag > 1e-10) {
      // (u.eval() * u.nabla()).sum(): u at the centroid times the shape-function gradients there,
      // summed over the nodes (zero up to rounding: the gradients sum to zero)
      Real J[9], Ji[9], det, dN[24];
      jacobian(et, xi, X, J);
      if(!inverse_det(dim, J, Ji, det)) return -2;
      shape_gradients(et, xi, dN);
      Real s = 0.;
      for(int n = 0; n < k; ++n) {
        Real un = 0.;
        for(int d = 0; d < dim; ++d) {
          Real gdn = 0.;
          for(int l = 0; l < dim; ++l) gdn += Ji[d*dim+l] * dN[l*k+n];
          un += uc[d] * gdn;
        }
        s += un;
      }
      const Real h = 2.*umag / s;
      const Real tau_adv = h/(2.*umag);
      const Real tau_time = 0.5*dt;
      const Real tau_diff = h*h/(4.*element_nu);
      tau_su = 1./(1./tau_adv + 1./tau_time + 1./tau_diff);
    }
  } else {
    return -3;
  }
  tau_ps *= cfg.a_ps;
  tau_su *= cfg.a_su;
  tau_bulk *= cfg.a_bu;
  return 0;
}

// ---------------------------------------------------------------------------------------------
// Specialised P1 triangle: NavierStokesSpecializations.hpp:38-147
// A, T var-major (9 x 9): u at 0..2, v at 3..5, p at 6..8
// ---------------------------------------------------------------------------------------------
int supg_specialized_tri(const Real* nodes, const Real* u, const Real* u_adv, const Real* nu_eff,
                         Real dt, Real rho, Real* A, Real* T, Real* tau3)
{
  const int n9 = 9;
  Real u_avg[2];
  for(int d = 0; d < 2; ++d) u_avg[d] = (u_adv[0*2+d] + u_adv[1*2+d] + u_adv[2*2+d]) / 3.;
  const Real vol = volume(ORC_TRIAG2D, nodes);
  const Real fc = 0.5;
  const Real nu = std::fabs((nu_eff[0] + nu_eff[1] + nu_eff[2]) / 3.);
#define ND(i,d) nodes[(i)*2+(d)]
  Real nrm[3][2];
  nrm[0][0] = ND(1,1) - ND(2,1);
  nrm[0][1] = ND(2,0) - ND(1,0);
  nrm[1][0] = ND(2,1) - ND(0,1);
  nrm[1][1] = ND(0,0) - ND(2,0);
  nrm[2][0] = ND(0,1) - ND(1,1);
  nrm[2][1] = ND(1,0) - ND(0,0);
#undef ND
  Real tau_ps, tau_su, tau_bulk;
  const TauCfg cfg = {ORC_SUPG_METRIC, 1., 16., 1., 1., 1., 1.};
  const int err = compute_tau(ORC_TRIAG2D, cfg, nodes, u, nu, dt, tau_ps, tau_su, tau_bulk);
  if(err) return err;
  tau3[0] = tau_ps; tau3[1] = tau_su; tau3[2] = tau_bulk;
#define AA(r,c) A[(r)*n9+(c)]
#define TT(r,c) T[(r)*n9+(c)]
  for(int i = 0; i < 3; ++i) {
    const int Pi = i + 3*2, Ui = i + 3*0, Vi = Ui + 3;
    const Real u_ni = u_avg[0]*nrm[i][0] + u_avg[1]*nrm[i][1];
    for(int j = 0; j < 3; ++j) {
      const int Pj = j + 3*2, Uj = j + 3*0, Vj = Uj + 3;
      const Real uk = u_avg[0], vk = u_avg[1];
      const Real uknj = uk*nrm[j][0] + vk*nrm[j][1];
      Real val = 1./6.*uknj;
      val += tau_su/(4.*vol)*uknj*u_ni;
      AA(Ui,Uj) += val; AA(Vi,Vj) += val;
      val = tau_ps/(4.*vol)*uknj;
      AA(Pi,Uj) += val*nrm[i][0]; AA(Pi,Vj) += val*nrm[i][1];
      val = fc/6.;
      val += fc*tau_su/(4.*vol)*u_ni;
      AA(Ui,Uj) += val*uk*nrm[j][0]; AA(Ui,Vj) += val*uk*nrm[j][1];
      AA(Vi,Uj) += val*vk*nrm[j][0]; AA(Vi,Vj) += val*vk*nrm[j][1];
      val = fc*tau_ps/(4.*vol);
      AA(Pi,Uj) += val*nrm[i][0]*uk*nrm[j][0]; AA(Pi,Vj) += val*nrm[i][0]*uk*nrm[j][1];
      AA(Pi,Uj) += val*nrm[i][1]*vk*nrm[j][0]; AA(Pi,Vj) += val*nrm[i][1]*vk*nrm[j][1];
      val = nu/(4.*vol);
      AA(Ui,Uj) += val*(4./3.*nrm[i][0]*nrm[j][0] + nrm[i][1]*nrm[j][1]);
      AA(Ui,Vj) += val*nrm[i][0]*nrm[j][1]/3.;
      AA(Vi,Uj) += val*nrm[i][1]*nrm[j][0]/3.;
      AA(Vi,Vj) += val*(nrm[i][0]*nrm[j][0] + 4./3.*nrm[i][1]*nrm[j][1]);
      val = 1./(6.*rho);
      val += tau_su/(4.*rho*vol)*u_ni;
      AA(Ui,Pj) += nrm[j][0]*val; AA(Vi,Pj) += nrm[j][1]*val;
      const Real laplacian = 1./(4.*rho*vol)*(nrm[i][0]*nrm[j][0] + nrm[i][1]*nrm[j][1]);
      AA(Pi,Pj) += tau_ps*laplacian;
      val = 1./6.;
      AA(Pi,Uj) += val*nrm[j][0]; AA(Pi,Vj) += val*nrm[j][1];
      val = tau_bulk/(4.*vol);
      AA(Ui,Uj) += val*nrm[i][0]*nrm[j][0]; AA(Ui,Vj) += val*nrm[i][0]*nrm[j][1];
      AA(Vi,Uj) += val*nrm[i][1]*nrm[j][0]; AA(Vi,Vj) += val*nrm[i][1]*nrm[j][1];
      val = vol/12.*(1.+((i)==(j)?1.:0.));
      val += tau_su/6.*u_ni;
      TT(Ui,Uj) += val; TT(Vi,Vj) += val;
      val = tau_ps/6.;
      TT(Pi,Uj) += val*nrm[i][0]; TT(Pi,Vj) += val*nrm[i][1];
    }
  }
#undef AA
#undef TT
  return 0;
}

// ---------------------------------------------------------------------------------------------
// Specialised P1 tetrahedron: NavierStokesSpecializations.hpp:151-303 (incl. the vk quirks of
// lines 224-226 and 236-238).  A, T var-major 16x16: u 0..3, v 4..7, w 8..11, p 12..15
// ---------------------------------------------------------------------------------------------
int supg_specialized_tet(const Real* nodes, const Real* u, const Real* u_adv, const Real* nu_eff,
                         Real dt, Real rho, Real* A, Real* T, Real* tau3)
{
  const int n16 = 16;
  Real u_avg[3];
  for(int d = 0; d < 3; ++d) u_avg[d] = (u_adv[0*3+d] + u_adv[1*3+d] + u_adv[2*3+d] + u_adv[3*3+d]) / 4.;
  const Real vol = volume(ORC_TETRA3D, nodes);
  const Real fc = 0.5;
  const Real nu = std::fabs((nu_eff[0] + nu_eff[1] + nu_eff[2] + nu_eff[3]) / 4.);
#define ND(i,d) nodes[(i)*3+(d)]
  Real n[4][3];
  n[0][0] = ((ND(3,1)-ND(1,1))*(ND(2,2)-ND(1,2)) - (ND(2,1)-ND(1,1))*(ND(3,2)-ND(1,2))) / 2.0;
  n[0][1] = ((ND(3,2)-ND(1,2))*(ND(2,0)-ND(1,0)) - (ND(2,2)-ND(1,2))*(ND(3,0)-ND(1,0))) / 2.0;
  n[0][2] = ((ND(3,0)-ND(1,0))*(ND(2,1)-ND(1,1)) - (ND(2,0)-ND(1,0))*(ND(3,1)-ND(1,1))) / 2.0;
  n[1][0] = ((ND(2,1)-ND(0,1))*(ND(3,2)-ND(0,2)) - (ND(3,1)-ND(0,1))*(ND(2,2)-ND(0,2))) / 2.0;
  n[1][1] = ((ND(2,2)-ND(0,2))*(ND(3,0)-ND(0,0)) - (ND(3,2)-ND(0,2))*(ND(2,0)-ND(0,0))) / 2.0;
  n[1][2] = ((ND(2,0)-ND(0,0))*(ND(3,1)-ND(0,1)) - (ND(3,0)-ND(0,0))*(ND(2,1)-ND(0,1))) / 2.0;
  n[2][0] = ((ND(3,1)-ND(0,1))*(ND(1,2)-ND(0,2)) - (ND(1,1)-ND(0,1))*(ND(3,2)-ND(0,2))) / 2.0;
  n[2][1] = ((ND(3,2)-ND(0,2))*(ND(1,0)-ND(0,0)) - (ND(1,2)-ND(0,2))*(ND(3,0)-ND(0,0))) / 2.0;
  n[2][2] = ((ND(3,0)-ND(0,0))*(ND(1,1)-ND(0,1)) - (ND(1,0)-ND(0,0))*(ND(3,1)-ND(0,1))) / 2.0;
  n[3][0] = ((ND(1,1)-ND(0,1))*(ND(2,2)-ND(0,2)) - (ND(2,1)-ND(0,1))*(ND(1,2)-ND(0,2))) / 2.0;
  n[3][1] = ((ND(1,2)-ND(0,2))*(ND(2,0)-ND(0,0)) - (ND(2,2)-ND(0,2))*(ND(1,0)-ND(0,0))) / 2.0;
  n[3][2] = ((ND(1,0)-ND(0,0))*(ND(2,1)-ND(0,1)) - (ND(2,0)-ND(0,0))*(ND(1,1)-ND(0,1))) / 2.0;
#undef ND
  Real tau_ps, tau_su, tau_bulk;
  const TauCfg cfg = {ORC_SUPG_METRIC, 1., 16., 1., 1., 1., 1.};
  const int err = compute_tau(ORC_TETRA3D, cfg, nodes, u, nu, dt, tau_ps, tau_su, tau_bulk);
  if(err) return err;
  tau3[0] = tau_ps; tau3[1] = tau_su; tau3[2] = tau_bulk;
#define AA(r,c) A[(r)*n16+(c)]
#define TT(r,c) T[(r)*n16+(c)]
  Real val;
  for(int i = 0; i < 4; ++i) {
    const int Pi = i + 4*3, Ui = i + 4*0, Vi = Ui + 4, Wi = Vi + 4;
    const Real u_ni = u_avg[0]*n[i][0] + u_avg[1]*n[i][1] + u_avg[2]*n[i][2];
    for(int j = 0; j < 4; ++j) {
      const int Pj = j + 4*3, Uj = j + 4*0, Vj = Uj + 4, Wj = Vj + 4;
      const Real uk = u_avg[0], vk = u_avg[1], wk = u_avg[2];
      const Real uknj = uk*n[j][0] + vk*n[j][1] + wk*n[j][2];
      val  = 1./12.*uknj;
      val += tau_su/(9.*vol)*uknj*u_ni;
      AA(Ui,Uj) += val; AA(Vi,Vj) += val; AA(Wi,Wj) += val;
      val = tau_ps/(9.*vol)*uknj;
      AA(Pi,Uj) += val*n[i][0]; AA(Pi,Vj) += val*n[i][1]; AA(Pi,Wj) += val*n[i][2];
      val  = fc/12.;
      val += fc*tau_su/(9.*vol)*u_ni;
      AA(Ui,Uj) += val*uk*n[j][0]; AA(Ui,Vj) += val*uk*n[j][1]; AA(Ui,Wj) += val*uk*n[j][2];
      AA(Vi,Uj) += val*vk*n[j][0]; AA(Vi,Vj) += val*vk*n[j][1]; AA(Vi,Wj) += val*vk*n[j][2];
      AA(Wi,Uj) += val*vk*n[j][0]; AA(Wi,Vj) += val*vk*n[j][1]; AA(Wi,Wj) += val*vk*n[j][2];
      val = fc*tau_ps/(9.*vol);
      AA(Pi,Uj) += val*n[i][0]*uk*n[j][0]; AA(Pi,Vj) += val*n[i][0]*uk*n[j][1]; AA(Pi,Wj) += val*n[i][0]*uk*n[j][2];
      AA(Pi,Uj) += val*n[i][1]*vk*n[j][0]; AA(Pi,Vj) += val*n[i][1]*vk*n[j][1]; AA(Pi,Wj) += val*n[i][1]*vk*n[j][2];
      AA(Pi,Uj) += val*n[i][2]*vk*n[j][0]; AA(Pi,Vj) += val*n[i][2]*vk*n[j][1]; AA(Pi,Wj) += val*n[i][2]*vk*n[j][2];
      val = nu/(9.*vol);
      AA(Ui,Uj) += val*(4./3.*n[i][0]*n[j][0] + n[i][1]*n[j][1] + n[i][2]*n[j][2]);
      AA(Ui,Vj) += val*n[i][0]*n[j][1]/3.;
      AA(Ui,Wj) += val*n[i][0]*n[j][2]/3.;
      AA(Vi,Uj) += val*n[i][1]*n[j][0]/3.;
      AA(Vi,Vj) += val*(n[i][0]*n[j][0] + 4./3.*n[i][1]*n[j][1] + n[i][2]*n[j][2]);
      AA(Vi,Wj) += val*n[i][1]*n[j][2]/3.;
      AA(Wi,Uj) += val*n[i][2]*n[j][0]/3.;
      AA(Wi,Vj) += val*n[i][2]*n[j][1]/3.;
      AA(Wi,Wj) += val*(n[i][0]*n[j][0] + n[i][1]*n[j][1] + 4./3.*n[i][2]*n[j][2]);
      val  = 1./(12.*rho);
      val += tau_su/(9.*rho*vol)*u_ni;
      AA(Ui,Pj) += n[j][0]*val; AA(Vi,Pj) += n[j][1]*val; AA(Wi,Pj) += n[j][2]*val;
      val = tau_ps/(9.*rho*vol)*(n[i][0]*n[j][0] + n[i][1]*n[j][1] + n[i][2]*n[j][2]);
      AA(Pi,Pj) += val;
      val = 1./12.;
      AA(Pi,Uj) += val*n[j][0]; AA(Pi,Vj) += val*n[j][1]; AA(Pi,Wj) += val*n[j][2];
      val = tau_bulk/(9.*vol);
      AA(Ui,Uj) += val*n[i][0]*n[j][0]; AA(Ui,Vj) += val*n[i][0]*n[j][1]; AA(Ui,Wj) += val*n[i][0]*n[j][2];
      AA(Vi,Uj) += val*n[i][1]*n[j][0]; AA(Vi,Vj) += val*n[i][1]*n[j][1]; AA(Vi,Wj) += val*n[i][1]*n[j][2];
      AA(Wi,Uj) += val*n[i][2]*n[j][0]; AA(Wi,Vj) += val*n[i][2]*n[j][1]; AA(Wi,Wj) += val*n[i][2]*n[j][2];
      val  = vol/20.*(1.+ (i == j ? 1. : 0.));
      val += tau_su/12.*u_ni;
      TT(Ui,Uj) += val; TT(Vi,Vj) += val; TT(Wi,Wj) += val;
      val = tau_ps/12.;
      TT(Pi,Uj) += val*n[i][0]; TT(Pi,Vj) += val*n[i][1]; TT(Pi,Wj) += val*n[i][2];
    }
  }
#undef AA
#undef TT
  return 0;
}

// ---------------------------------------------------------------------------------------------
// Generic Proto expression: NavierStokesAssembly.hpp:44-66 with element_quadrature
// (ElementIntegration.hpp:184-257: order 2, weight w_q * detJ_q) and tau from u_adv (:54).
// Matrices var-major: u_i block at i*k, p block at dim*k.
// ---------------------------------------------------------------------------------------------
int generic_element(int et, const TauCfg& cfg, Real dt, Real rho, const Real* X, const Real* u_adv,
                    const Real* nu_eff, const Real* g_force, Real* A, Real* T, Real* a, Real* tau3)
{
  const int k = enodes(et), dim = edim(et), n = k*(dim+1);
  Real mean_nu = 0.;
  for(int i = 0; i < k; ++i) mean_nu += nu_eff[i];
  const Real element_nu = std::fabs(mean_nu / k);
  Real tau_ps, tau_su, tau_bulk;
  const int err = compute_tau(et, cfg, X, u_adv, element_nu, dt, tau_ps, tau_su, tau_bulk);
  if(err) return err;
  tau3[0] = tau_ps; tau3[1] = tau_su; tau3[2] = tau_bulk;

  Real gxi[24], gw[8];
  const int ngp = gauss(et, 2, gxi, gw);
  const int P = dim*k;
  for(int q = 0; q < ngp; ++q) {
    const Real* xi = gxi + q*dim;
    Real N[8], dN[24], J[9], Ji[9], det, G[24];
    shape_values(et, xi, N);
    shape_gradients(et, xi, dN);
    jacobian(et, xi, X, J);
    if(!inverse_det(dim, J, Ji, det)) return -2;
    for(int d = 0; d < dim; ++d)
      for(int m = 0; m < k; ++m) {
        Real s = 0.;
        for(int l = 0; l < dim; ++l) s += Ji[d*dim+l] * dN[l*k+m];
        G[d*k+m] = s;
      }
    Real ua[3] = {0., 0., 0.}, gq[3] = {0., 0., 0.}, nuq = 0.;
    for(int d = 0; d < dim; ++d) {
      Real s = 0., sg = 0.;
      for(int m = 0; m < k; ++m) { s += N[m]*u_adv[m*dim+d]; sg += N[m]*g_force[m*dim+d]; }
      ua[d] = s; gq[d] = sg;
    }
    for(int m = 0; m < k; ++m) nuq += N[m]*nu_eff[m];
    const Real w = gw[q] * det;
    Real L[8], Lps[8], SU[8];
    for(int m = 0; m < k; ++m) {
      Real l = 0., lps = 0., lsu = 0.;
      for(int d = 0; d < dim; ++d) {
        l += ua[d]*G[d*k+m];
        lps += (tau_ps*ua[d])*G[d*k+m];
        lsu += (tau_su*ua[d])*G[d*k+m];
      }
      L[m] = l; Lps[m] = lps*0.5; SU[m] = N[m] + lsu;
    }
    for(int r = 0; r < k; ++r)
      for(int c = 0; c < k; ++c) {
        Real gtg = 0.;
        for(int d = 0; d < dim; ++d) gtg += G[d*k+r]*G[d*k+c];
        // _A(p, u[_i]) (:57)
        for(int i = 0; i < dim; ++i) {
          const Real v = (N[r] + Lps[r]) * G[i*k+c] + ((tau_ps*G[i*k+r]) * L[c]);
          A[(P+r)*n + i*k+c] += w * v;
        }
        // _A(p, p) (:58)
        A[(P+r)*n + P+c] += w * ((tau_ps*gtg) / rho);
        for(int i = 0; i < dim; ++i) {
          // _A(u[_i], u[_i]) (:59)
          A[(i*k+r)*n + i*k+c] += w * (nuq*gtg + SU[r]*L[c]);
          // _A(u[_i], p) (:60)
          A[(i*k+r)*n + P+c] += w * (SU[r]*G[i*k+c] / rho);
          // _A(u[_i], u[_j]) (:61-62)
          for(int j = 0; j < dim; ++j)
            A[(i*k+r)*n + j*k+c] += w * ((tau_bulk*G[i*k+r] + 0.5*ua[i]*SU[r]) * G[j*k+c]);
          // _T(p, u[_i]) (:63)
          T[(P+r)*n + i*k+c] += w * (tau_ps*G[i*k+r]*N[c]);
          // _T(u[_i], u[_i]) (:64)
          T[(i*k+r)*n + i*k+c] += w * (SU[r]*N[c]);
        }
      }
    // _a[u[_i]] (:65)
    for(int i = 0; i < dim; ++i)
      for(int r = 0; r < k; ++r)
        a[i*k+r] += w * (SU[r]*gq[i]);
  }
  return 0;
}

TauCfg tau_cfg(const orc_ns_params* p)
{
  TauCfg c = {p->supg_type, p->c1, p->c2, p->alpha_ps, p->alpha_su, p->alpha_bu, p->u_ref};
  return c;
}

int element_matrices(int et, const orc_ns_params* prm, const Real* X, const Real* u, const Real* u_adv,
                     const Real* nu, const Real* g, Real* A, Real* T, Real* a, Real* tau3)
{
  const int k = enodes(et), dim = edim(et), n = k*(dim+1);
  std::fill(A, A + n*n, 0.);   // _A = _0, _T = _0, _a = _0 (NavierStokesAssembly.hpp:50)
  std::fill(T, T + n*n, 0.);
  std::fill(a, a + n, 0.);
  // use_specializations applies to triangles and tetrahedra only (NavierStokesTriags/Tetras.cpp;
  // hexahedra always take the generic expression, NavierStokesHexas.cpp:20-25)
  if(prm->use_specializations && et == ORC_TETRA3D)
    return supg_specialized_tet(X, u, u_adv, nu, prm->dt, prm->rho, A, T, tau3);
  if(prm->use_specializations && et == ORC_TRIAG2D)
    return supg_specialized_tri(X, u, u_adv, nu, prm->dt, prm->rho, A, T, tau3);
  return generic_element(et, tau_cfg(prm), prm->dt, prm->rho, X, u_adv, nu, g, A, T, a, tau3);
}

// combine + node-major reorder (NavierStokesAssembly.hpp:69-71, BlockAccumulator.hpp:134-184)
void element_system(int et, const orc_ns_params* prm, Real* A, const Real* T, const Real* a,
                    const Real* x_varmajor, Real* M, Real* rhs)
{
  const int k = enodes(et), dim = edim(et), neq = dim+1, n = k*neq;
  // system_rhs += -_A * _x + _a  (before the /theta of the p rows)
  for(int r = 0; r < n; ++r) {
    Real s = 0.;
    for(int c = 0; c < n; ++c) s += (-A[r*n+c]) * x_varmajor[c];
    const int br = (r % k)*neq + r / k;
    rhs[br] = s + a[r];
  }
  // _A(p) = _A(p) / theta
  for(int r = dim*k; r < n; ++r)
    for(int c = 0; c < n; ++c) A[r*n+c] = A[r*n+c] / prm->theta;
  // system_matrix += invdt * _T + theta * _A
  const Real invdt = 1. / prm->dt;
  for(int r = 0; r < n; ++r) {
    const int br = (r % k)*neq + r / k;
    for(int c = 0; c < n; ++c) {
      const int bc = (c % k)*neq + c / k;
      M[br*n+bc] = invdt * T[r*n+c] + prm->theta * A[r*n+c];
    }
  }
}

void gather_element(int et, const Uint* cn, const Real* coords, const Real* u, const Real* p,
                    const Real* u_adv, const Real* nu, const Real* g,
                    Real* X, Real* ue, Real* uae, Real* nue, Real* ge, Real* xe)
{
  const int k = enodes(et), dim = edim(et);
  for(int m = 0; m < k; ++m) {
    const Uint nd = cn[m];
    for(int d = 0; d < dim; ++d) {
      X[m*dim+d] = coords[(size_t)nd*dim+d];
      ue[m*dim+d] = u[(size_t)nd*dim+d];
      uae[m*dim+d] = u_adv[(size_t)nd*dim+d];
      ge[m*dim+d] = g ? g[(size_t)nd*dim+d] : 0.;
      xe[d*k+m] = u[(size_t)nd*dim+d];         // FillRhs: var-major element vector (ElementData.hpp:1228-1248)
    }
    nue[m] = nu[nd];
    xe[dim*k+m] = p[nd];
  }
}

// p2m (optional): matrix block row of each mesh node (periodic links; TrilinosDetail.cpp:147-227).  The
// element is gathered by its mesh nodes; its rhs rows are those of p2m (the slots already are).
// When set, the scatter accumulates |M| and |b| of every element contribution (orc_ns_assemble_abs):
// the scale of each entry's rounding error for parity bars.
bool g_abs_assembly = false;

int assemble_range(int et, const orc_ns_params* prm, const Real* coords, const Uint* conn,
                   const Real* u, const Real* p, const Real* u_adv, const Real* nu, const Real* g,
                   const Uint* slots, Real* values, Real* rhs, const Uint* elems, size_t ne,
                   const Uint* p2m = nullptr)
{
  const int k = enodes(et), dim = edim(et), neq = dim+1, n = k*neq, nn = neq*neq;
  std::vector<Real> A(n*n), T(n*n), a(n), M(n*n), b(n);
  Real X[24], ue[24], uae[24], nue[8], ge[24], xe[32], tau3[3];
  for(size_t ii = 0; ii < ne; ++ii) {
    const size_t e = elems ? elems[ii] : ii;
    const Uint* cn = conn + e*k;
    gather_element(et, cn, coords, u, p, u_adv, nu, g, X, ue, uae, nue, ge, xe);
    const int err = element_matrices(et, prm, X, ue, uae, nue, ge, A.data(), T.data(), a.data(), tau3);
    if(err) return err;
    element_system(et, prm, A.data(), T.data(), a.data(), xe, M.data(), b.data());
    // TrilinosFEVbrMatrix::add_values (TrilinosFEVbrMatrix.cpp:315-355): block (ra, cb) += M block
    for(int ra = 0; ra < k; ++ra) {
      for(int cb = 0; cb < k; ++cb) {
        Real* blk = values + (size_t)slots[(e*k + ra)*k + cb] * nn;
        for(int al = 0; al < neq; ++al)
          for(int be = 0; be < neq; ++be) {
            const Real m = M[(ra*neq+al)*n + cb*neq+be];
            blk[al*neq+be] += g_abs_assembly ? std::fabs(m) : m;
          }
      }
      // TrilinosVector::add_rhs_values (TrilinosVector.cpp:189-203)
      const Uint row = p2m ? p2m[cn[ra]] : cn[ra];
      for(int al = 0; al < neq; ++al)
        rhs[(size_t)row*neq+al] += g_abs_assembly ? std::fabs(b[ra*neq+al]) : b[ra*neq+al];
    }
  }
  return 0;
}

// greedy element colouring (no two elements of a colour share a node) for the threaded baseline
void colour_elements(int k, Uint n_nodes, Uint n_elems, const Uint* conn, std::vector<std::vector<Uint> >& colours)
{
  std::vector<uint64_t> node_mask(n_nodes, 0);
  std::vector<int> ecol(n_elems);
  int ncol = 0;
  for(Uint e = 0; e < n_elems; ++e) {
    uint64_t used = 0;
    for(int m = 0; m < k; ++m) used |= node_mask[conn[(size_t)e*k+m]];
    int c = 0;
    while(c < 64 && (used >> c) & 1ull) ++c;
    if(c == 64) c = 63; // not reached for the meshes used here
    ecol[e] = c;
    ncol = std::max(ncol, c+1);
    for(int m = 0; m < k; ++m) node_mask[conn[(size_t)e*k+m]] |= (1ull << c);
  }
  colours.assign(ncol, std::vector<Uint>());
  for(Uint e = 0; e < n_elems; ++e) colours[ecol[e]].push_back(e);
}

// ---------------------------------------------------------------------------------------------
// Block matrix helpers (row-major neq x neq blocks, node-row contiguous)
// ---------------------------------------------------------------------------------------------
inline int64_t find_block(const Uint* start, const Uint* cols, Uint row, Uint col)
{
  for(Uint i = start[row]; i != start[row+1]; ++i)
    if(cols[i] == col) return i;
  return -1;
}

// Threads of the Krylov kernels (SpMV rows, dot products, vector updates) of the oracle, 1 by default.
// Only the CPU-baseline timing raises it; with several threads a dot product sums fixed chunks in
// order, so results do not depend on the thread count but differ from the one-thread sums in the
// last bits.
int g_krylov_threads = 1;

void spmv_plain(int neq, Uint N, const Uint* start, const Uint* cols, const Real* values, const Real* x, Real* y)
{
  const int nn = neq*neq;
#ifdef _OPENMP
#pragma omp parallel for num_threads(g_krylov_threads) if(g_krylov_threads > 1) schedule(static)
#endif
  for(Uint i = 0; i < N; ++i) {
    for(int al = 0; al < neq; ++al) {
      Real s = 0.;
      for(Uint q = start[i]; q != start[i+1]; ++q) {
        const Real* blk = values + (size_t)q*nn + al*neq;
        const Real* xb = x + (size_t)cols[q]*neq;
        for(int be = 0; be < neq; ++be) s += blk[be] * xb[be];
      }
      y[(size_t)i*neq+al] = s;
    }
  }
}

Real dot(size_t n, const Real* a, const Real* b)
{
  if(g_krylov_threads <= 1) {
    Real s = 0.;
    for(size_t i = 0; i < n; ++i) s += a[i]*b[i];
    return s;
  }
  constexpr int kChunks = 256;
  Real part[kChunks];
#ifdef _OPENMP
#pragma omp parallel for num_threads(g_krylov_threads) schedule(static)
#endif
  for(int c = 0; c < kChunks; ++c) {
    Real s = 0.;
    for(size_t i = n * c / kChunks; i < n * (c + 1) / kChunks; ++i) s += a[i]*b[i];
    part[c] = s;
  }
  Real s = 0.;
  for(int c = 0; c < kChunks; ++c) s += part[c];
  return s;
}

// ILU(0) on the point matrix (IKJ variant) -- Ifpack ILU level 0 is third-party; restated by
// specification.  Point CSR with sorted columns.
struct PointCSR { std::vector<Uint> rp, ci; std::vector<Real> v; std::vector<Uint> diag; int neq = 1; };

void to_point(int neq, Uint N, const Uint* start, const Uint* cols, const Real* values, PointCSR& P)
{
  const size_t n = (size_t)N*neq;
  P.neq = neq;
  P.rp.assign(n+1, 0);
  const size_t nnz = (size_t)start[N]*neq*neq;
  P.ci.resize(nnz); P.v.resize(nnz);
  orc_point_csr(neq, N, start, cols, values, P.rp.data(), P.ci.data(), P.v.data());
  P.diag.resize(n);
  for(size_t r = 0; r < n; ++r)
    for(Uint q = P.rp[r]; q != P.rp[r+1]; ++q) if(P.ci[q] == r) { P.diag[r] = q; break; }
}

void ilu0_factor(PointCSR& P)
{
  const size_t n = P.rp.size()-1;
  std::vector<int64_t> pos(n, -1);
  for(size_t i = 0; i < n; ++i) {
    for(Uint q = P.rp[i]; q != P.rp[i+1]; ++q) pos[P.ci[q]] = q;
    for(Uint q = P.rp[i]; q != P.rp[i+1]; ++q) {
      const Uint kcol = P.ci[q];
      if(kcol >= i) break;
      const Real lik = P.v[q] / P.v[P.diag[kcol]];
      P.v[q] = lik;
      for(Uint r = P.diag[kcol]+1; r != P.rp[kcol+1]; ++r) {
        const int64_t t = pos[P.ci[r]];
        if(t >= 0) P.v[t] -= lik * P.v[r];
      }
    }
    for(Uint q = P.rp[i]; q != P.rp[i+1]; ++q) pos[P.ci[q]] = -1;
  }
}

void ilu0_solve(const PointCSR& P, const Real* b, Real* x)
{
  const size_t n = P.rp.size()-1;
  for(size_t i = 0; i < n; ++i) {
    Real s = b[i];
    for(Uint q = P.rp[i]; q != P.diag[i]; ++q) s -= P.v[q] * x[P.ci[q]];
    x[i] = s;
  }
  // upper rows: the entries of the other nodes' columns first, then those of the row's own node (the
  // block rows of a node are solved together), each in ascending column order
  const Uint neq = (Uint)P.neq;
  for(size_t ii = n; ii-- > 0;) {
    const Uint own = (Uint)(ii / neq) * neq;
    Real s = x[ii];
    for(Uint q = P.diag[ii]+1; q != P.rp[ii+1]; ++q)
      if(P.ci[q] >= own + neq) s -= P.v[q] * x[P.ci[q]];
    for(Uint q = P.diag[ii]+1; q != P.rp[ii+1]; ++q)
      if(P.ci[q] < own + neq) s -= P.v[q] * x[P.ci[q]];
    x[ii] = s / P.v[P.diag[ii]];
  }
}

} // namespace

// =============================================================================================
// C API
// =============================================================================================
extern "C" {

uint64_t orc_splitmix64_next(uint64_t* s)
{
  uint64_t z = (*s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

double orc_uniform(uint64_t* s, double a, double b)
{
  const double u = (double)(orc_splitmix64_next(s) >> 11) * (1.0 / 9007199254740992.0);
  return a + (b - a) * u;
}

void orc_mesh_rect_tris(uint32_t nx, uint32_t ny, double lx, double ly, double* coords, uint32_t* conn)
{
  const Real xs = lx / (Real)nx, ys = ly / (Real)ny;
  for(Uint j = 0; j <= ny; ++j) {
    const Real y = (Real)j * ys;
    for(Uint i = 0; i <= nx; ++i) {
      coords[2*(j*(nx+1)+i)+0] = (Real)i * xs;
      coords[2*(j*(nx+1)+i)+1] = y;
    }
  }
  for(Uint j = 0; j < ny; ++j)
    for(Uint i = 0; i < nx; ++i) {
      const Uint n0 = j*(nx+1)+i, n1 = n0+1, n3 = (j+1)*(nx+1)+i, n2 = n3+1;
      Uint* u = conn + 3*(2*(j*nx+i));
      u[0] = n0; u[1] = n2; u[2] = n3;
      Uint* l = conn + 3*(2*(j*nx+i)+1);
      l[0] = n0; l[1] = n1; l[2] = n2;
    }
}

static void box_nodes(uint32_t nx, uint32_t ny, uint32_t nz, double lx, double ly, double lz, double* coords)
{
  const Real xs = lx / (Real)nx, ys = ly / (Real)ny, zs = lz / (Real)nz;
  for(Uint k = 0; k <= nz; ++k) {
    const Real z = (Real)k * zs;
    for(Uint j = 0; j <= ny; ++j) {
      const Real y = (Real)j * ys;
      for(Uint i = 0; i <= nx; ++i) {
        const size_t nd = ((size_t)k*(ny+1) + j)*(nx+1) + i;
        coords[3*nd+0] = (Real)i * xs;
        coords[3*nd+1] = y;
        coords[3*nd+2] = z;
      }
    }
  }
}

static void hex_nodes(uint32_t nx, uint32_t ny, Uint i, Uint j, Uint k, Uint* h)
{
  auto nid = [&](Uint a, Uint b, Uint c) { return (c*(ny+1) + b)*(nx+1) + a; };
  h[0] = nid(i,j,k); h[1] = nid(i+1,j,k); h[2] = nid(i+1,j+1,k); h[3] = nid(i,j+1,k);
  h[4] = nid(i,j,k+1); h[5] = nid(i+1,j,k+1); h[6] = nid(i+1,j+1,k+1); h[7] = nid(i,j+1,k+1);
}

void orc_mesh_box_hexas(uint32_t nx, uint32_t ny, uint32_t nz, double lx, double ly, double lz, double* coords, uint32_t* conn)
{
  box_nodes(nx, ny, nz, lx, ly, lz, coords);
  for(Uint k = 0; k < nz; ++k)
    for(Uint j = 0; j < ny; ++j)
      for(Uint i = 0; i < nx; ++i)
        hex_nodes(nx, ny, i, j, k, conn + 8*(((size_t)k*ny + j)*nx + i));
}

void orc_mesh_box_tetras(uint32_t nx, uint32_t ny, uint32_t nz, double lx, double ly, double lz, double* coords, uint32_t* conn)
{
  // Kuhn split along the local 0 -> 6 diagonal; the six monotone paths x-y-z, x-z-y, y-x-z, y-z-x,
  // z-x-y, z-y-x with the two middle nodes swapped for odd permutations so that detJ > 0.
  static const int kuhn[6][4] = {{0,1,2,6},{0,5,1,6},{0,2,3,6},{0,3,7,6},{0,4,5,6},{0,7,4,6}};
  box_nodes(nx, ny, nz, lx, ly, lz, coords);
  Uint h[8];
  for(Uint k = 0; k < nz; ++k)
    for(Uint j = 0; j < ny; ++j)
      for(Uint i = 0; i < nx; ++i) {
        hex_nodes(nx, ny, i, j, k, h);
        const size_t e0 = 6*(((size_t)k*ny + j)*nx + i);
        for(int t = 0; t < 6; ++t)
          for(int m = 0; m < 4; ++m) conn[4*(e0+t)+m] = h[kuhn[t][m]];
      }
}

void orc_jitter_structured(int dim, uint32_t nx, uint32_t ny, uint32_t nz, double lx, double ly, double lz,
                           double frac, uint64_t seed, double* coords)
{
  uint64_t s = seed;
  const Real h[3] = {lx/(Real)nx, ly/(Real)ny, dim == 3 ? lz/(Real)nz : 0.};
  const Uint n[3] = {nx, ny, dim == 3 ? nz : 0};
  const Uint nzz = dim == 3 ? nz : 0;
  for(Uint k = 0; k <= nzz; ++k)
    for(Uint j = 0; j <= ny; ++j)
      for(Uint i = 0; i <= nx; ++i) {
        const size_t nd = ((size_t)k*(ny+1) + j)*(nx+1) + i;
        const Uint idx[3] = {i, j, k};
        for(int d = 0; d < dim; ++d) {
          const Real r = orc_uniform(&s, -frac*h[d], frac*h[d]);
          if(idx[d] != 0 && idx[d] != n[d]) coords[dim*nd+d] += r;
        }
      }
}

double orc_tetra_volume(const double* X) { return volume(ORC_TETRA3D, X); }
void orc_tetra_jacobian(const double* X, double* J) { const Real xi[3] = {0,0,0}; jacobian(ORC_TETRA3D, xi, X, J); }
void orc_tetra_jacobian_adjoint(const double* X, double* adj) { Real J[9]; orc_tetra_jacobian(X, J); adjoint3(J, adj); }
double orc_hexa_jacobian_determinant(const double* m, const double* X) { return hexa_det(m, X); }
void orc_hexa_jacobian(const double* m, const double* X, double* J) { jacobian(ORC_HEXA3D, m, X, J); }
void orc_hexa_jacobian_adjoint(const double* m, const double* X, double* adj) { Real J[9]; jacobian(ORC_HEXA3D, m, X, J); adjoint3(J, adj); }
double orc_hexa_volume(const double* X) { return volume(ORC_HEXA3D, X); }
double orc_element_volume(int et, const double* X) { return volume(et, X); }
void orc_element_jacobian(int et, const double* m, const double* X, double* J) { jacobian(et, m, X, J); }
void orc_shape_values(int et, const double* m, double* N) { shape_values(et, m, N); }
void orc_shape_gradients(int et, const double* m, double* dN) { shape_gradients(et, m, dN); }
int orc_gauss_points(int et, int order, double* c, double* w) { return gauss(et, order, c, w); }

int64_t orc_build_sparsity(int k, uint32_t n_nodes, uint32_t n_elems, const uint32_t* conn,
                           uint32_t* start, uint32_t* cols, int64_t cols_cap, uint32_t* slots)
{
  // SparsityBuilder.cpp:157-189 with an identity used-node map (all nodes are used)
  std::vector< std::set<Uint> > sets(n_nodes);
  for(Uint e = 0; e < n_elems; ++e)
    for(int a = 0; a < k; ++a)
      for(int b = 0; b < k; ++b)
        sets[conn[(size_t)e*k+a]].insert(conn[(size_t)e*k+b]);
  start[0] = 0;
  for(Uint i = 1; i <= n_nodes; ++i) start[i] = (Uint)(sets[i-1].size() + start[i-1]);
  if((int64_t)start[n_nodes] > cols_cap) return -1;
  for(Uint i = 0; i < n_nodes; ++i) std::copy(sets[i].begin(), sets[i].end(), cols + start[i]);
  if(slots) {
    for(Uint e = 0; e < n_elems; ++e)
      for(int a = 0; a < k; ++a) {
        const Uint r = conn[(size_t)e*k+a];
        for(int b = 0; b < k; ++b) {
          const Uint* beg = cols + start[r];
          const Uint* end = cols + start[r+1];
          const Uint* it = std::lower_bound(beg, end, conn[(size_t)e*k+b]);
          slots[((size_t)e*k+a)*k+b] = (Uint)(it - cols);
        }
      }
  }
  return start[n_nodes];
}

int orc_ns_element(int et, const orc_ns_params* prm, const double* nodes, const double* u,
                   const double* u_adv, const double* nu, const double* g,
                   double* A, double* T, double* a, double* tau3)
{
  Real t3[3];
  const int err = element_matrices(et, prm, nodes, u, u_adv, nu, g, A, T, a, t3);
  if(tau3) { tau3[0] = t3[0]; tau3[1] = t3[1]; tau3[2] = t3[2]; }
  return err;
}

int orc_ns_element_system(int et, const orc_ns_params* prm, const double* nodes, const double* u,
                          const double* p, const double* u_adv, const double* nu, const double* g,
                          double* M, double* rhs)
{
  const int k = enodes(et), dim = edim(et), n = k*(dim+1);
  std::vector<Real> A(n*n), T(n*n), a(n), xe(n);
  Real t3[3];
  const int err = element_matrices(et, prm, nodes, u, u_adv, nu, g, A.data(), T.data(), a.data(), t3);
  if(err) return err;
  for(int m = 0; m < k; ++m) {
    for(int d = 0; d < dim; ++d) xe[d*k+m] = u[m*dim+d];
    xe[dim*k+m] = p[m];
  }
  element_system(et, prm, A.data(), T.data(), a.data(), xe.data(), M, rhs);
  return 0;
}

int orc_ns_assemble(int et, const orc_ns_params* prm, uint32_t n_nodes, uint32_t n_elems,
                    const double* coords, const uint32_t* conn, const double* u, const double* p,
                    const double* u_adv, const double* nu, const double* g,
                    const uint32_t* start, const uint32_t* cols, const uint32_t* slots,
                    double* values, double* rhs, int nthreads)
{
  if(prm->theta < 0. || prm->theta > 1.) return -4;   // NavierStokesAssembly.hpp:38-40 SetupError
  const int k = enodes(et), neq = edim(et)+1;
  (void)cols;
  // ZeroLSS (ZeroLSS.cpp:56-85): matrix and rhs reset, solution kept
  std::fill(values, values + (size_t)start[n_nodes]*neq*neq, 0.);
  std::fill(rhs, rhs + (size_t)n_nodes*neq, 0.);
  if(nthreads <= 1)
    return assemble_range(et, prm, coords, conn, u, p, u_adv, nu, g, slots, values, rhs, 0, n_elems);
  std::vector< std::vector<Uint> > colours;
  colour_elements(k, n_nodes, n_elems, conn, colours);
  int err = 0;
  for(size_t c = 0; c < colours.size(); ++c) {
    const std::vector<Uint>& el = colours[c];
    const size_t chunks = (size_t)nthreads * 8;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads) schedule(dynamic) reduction(|:err)
#endif
    for(long ch = 0; ch < (long)chunks; ++ch) {
      const size_t b = el.size() * ch / chunks, e = el.size() * (ch+1) / chunks;
      if(e > b) err |= assemble_range(et, prm, coords, conn, u, p, u_adv, nu, g, slots, values, rhs, el.data()+b, e-b);
    }
  }
  return err;
}

int orc_ns_assemble_abs(int et, const orc_ns_params* prm, uint32_t n_nodes, uint32_t n_elems,
                        const double* coords, const uint32_t* conn, const double* u, const double* p,
                        const double* u_adv, const double* nu, const double* g,
                        const uint32_t* start, const uint32_t* cols, const uint32_t* slots,
                        double* values, double* rhs, int nthreads)
{
  g_abs_assembly = true;
  const int err = orc_ns_assemble(et, prm, n_nodes, n_elems, coords, conn, u, p, u_adv, nu, g, start, cols, slots,
                                  values, rhs, nthreads);
  g_abs_assembly = false;
  return err;
}

int orc_ns_assemble_p2m(int et, const orc_ns_params* prm, uint32_t n_rows, uint32_t n_elems,
                        const double* coords, const uint32_t* conn, const uint32_t* p2m, const double* u,
                        const double* p, const double* u_adv, const double* nu, const double* g,
                        const uint32_t* start, const uint32_t* slots, double* values, double* rhs)
{
  if(prm->theta < 0. || prm->theta > 1.) return -4;
  const int neq = edim(et)+1;
  std::fill(values, values + (size_t)start[n_rows]*neq*neq, 0.);
  std::fill(rhs, rhs + (size_t)n_rows*neq, 0.);
  return assemble_range(et, prm, coords, conn, u, p, u_adv, nu, g, slots, values, rhs, 0, n_elems, p2m);
}

void orc_set_row(int neq, const uint32_t* start, const uint32_t* cols, double* values,
                 uint32_t br, uint32_t ieq, double diag, double offdiag)
{
  const int nn = neq*neq;
  for(Uint q = start[br]; q != start[br+1]; ++q) {
    for(int j = 0; j < neq; ++j) values[(size_t)q*nn + ieq*neq + j] = offdiag;
    if(cols[q] == br) values[(size_t)q*nn + ieq*neq + ieq] = diag;
  }
}

// System::periodicity (System.cpp:259-282) with TrilinosFEVbrMatrix::tie_blockrow_pairs
// (TrilinosFEVbrMatrix.cpp:558-590) for each (to, from) pair in order; serial, so every row is owned.
// rhs / sol may be null (the matrix part alone).  Returns -1 when the two rows hold different block
// counts (BadValue of the reference) and -2 when their block columns differ (its cf3_assert).
int orc_tie_blockrow_pairs(int neq, const uint32_t* start, const uint32_t* cols, double* values, double* rhs,
                           double* sol, uint32_t npairs, const uint32_t* to, const uint32_t* from)
{
  const int nn = neq*neq;
  for(Uint p = 0; p < npairs; ++p) {
    const Uint bt = to[p], bf = from[p];
    const Uint nt = start[bt+1] - start[bt], nf = start[bf+1] - start[bf];
    if(nt != nf) return -1;
    int diag = -1, pair = -1;
    for(Uint i = 0; i < nt; ++i) {
      const Uint ct = cols[start[bt]+i], cf = cols[start[bf]+i];
      if(ct != cf) return -2;
      if(cf == bf) diag = (int)i;
      if(ct == bt) pair = (int)i;
      Real* vt = values + (size_t)(start[bt]+i)*nn;
      Real* vf = values + (size_t)(start[bf]+i)*nn;
      for(int k = 0; k < nn; ++k) { vt[k] += vf[k]; vf[k] = 0.; }
    }
    if(diag < 0 || pair < 0) return -2;
    Real* vfd = values + (size_t)(start[bf]+diag)*nn;
    Real* vfp = values + (size_t)(start[bf]+pair)*nn;
    for(int e = 0; e < neq; ++e) { vfd[e*neq+e] = 1.; vfp[e*neq+e] = -1.; }
    Real* vtp = values + (size_t)(start[bt]+pair)*nn;
    Real* vtd = values + (size_t)(start[bt]+diag)*nn;
    for(int k = 0; k < nn; ++k) { vtp[k] += vtd[k]; vtd[k] = 0.; }
    if(rhs)
      for(int e = 0; e < neq; ++e) { rhs[(size_t)bt*neq+e] += rhs[(size_t)bf*neq+e]; rhs[(size_t)bf*neq+e] = 0.; }
    if(sol)
      for(int e = 0; e < neq; ++e) {
        const Real m = 0.5*(sol[(size_t)bt*neq+e] + sol[(size_t)bf*neq+e]);
        sol[(size_t)bf*neq+e] = m;
        sol[(size_t)bt*neq+e] = m;
      }
  }
  return 0;
}

void orc_dirichlet(int neq, uint32_t n_nodes, const uint32_t* start, const uint32_t* cols,
                   double* values, double* rhs, double* sol, uint32_t nbc,
                   const uint32_t* bc_node, const uint32_t* bc_eq, const double* bc_val)
{
  (void)n_nodes;
  const int nn = neq*neq;
  for(Uint b = 0; b < nbc; ++b) {
    const Uint br = bc_node[b], ieq = bc_eq[b];
    const Real value = bc_val[b];
    for(Uint ci = start[br]; ci != start[br+1]; ++ci) {
      const Uint col = cols[ci];  // other row
      if(find_block(start, cols, col, br) < 0) continue;   // row_has_node check
      const int64_t q = find_block(start, cols, col, br);
      Real* blk = values + (size_t)q*nn;
      for(int j = 0; j < neq; ++j) {
        rhs[(size_t)col*neq + j] += -blk[j*neq + ieq] * value;
        blk[j*neq + ieq] = 0;
      }
      if(col == br) {
        for(Uint r = start[br]; r != start[br+1]; ++r) {
          Real* bb = values + (size_t)r*nn;
          for(int j = 0; j < neq; ++j) bb[ieq*neq + j] = 0;
          if(cols[r] == br) bb[ieq*neq + ieq] = 1.;
        }
      }
    }
    rhs[(size_t)br*neq + ieq] = value;
    sol[(size_t)br*neq + ieq] = value;
  }
}

void orc_spmv(int neq, uint32_t n_nodes, const uint32_t* start, const uint32_t* cols, const double* values,
              double alpha, const double* x, double beta, double* y)
{
  const size_t n = (size_t)n_nodes*neq;
  std::vector<Real> old(y, y + n), ax(n);
  spmv_plain(neq, n_nodes, start, cols, values, x, ax.data());
  for(size_t i = 0; i < n; ++i) {
    Real v = ax[i];
    if(alpha != 1.) v *= alpha;
    if(beta != 0.) v = beta*old[i] + 1.*v;
    y[i] = v;
  }
}

void orc_point_csr(int neq, uint32_t N, const uint32_t* start, const uint32_t* cols, const double* values,
                   uint32_t* row_ptr, uint32_t* col_idx, double* vals)
{
  const int nn = neq*neq;
  size_t pos = 0;
  row_ptr[0] = 0;
  for(Uint i = 0; i < N; ++i)
    for(int al = 0; al < neq; ++al) {
      for(Uint q = start[i]; q != start[i+1]; ++q)
        for(int be = 0; be < neq; ++be) {
          col_idx[pos] = cols[q]*neq + be;
          vals[pos] = values[(size_t)q*nn + al*neq + be];
          ++pos;
        }
      row_ptr[(size_t)i*neq + al + 1] = (Uint)pos;
    }
}

void orc_set_krylov_threads(int n) { g_krylov_threads = n < 1 ? 1 : n; }

int orc_bicgstab_jacobi(int neq, uint32_t N, const uint32_t* start, const uint32_t* cols, const double* values,
                        const double* b, double* x, double tol, int maxit, orc_solve_stats* st)
{
  const size_t n = (size_t)N*neq;
  const int nn = neq*neq;
  std::vector<Real> dinv(n), r(n), rh(n), p(n, 0.), v(n, 0.), ph(n), s(n), sh(n), t(n);
  for(Uint i = 0; i < N; ++i) {
    const int64_t q = find_block(start, cols, i, i);
    for(int al = 0; al < neq; ++al) dinv[(size_t)i*neq+al] = 1. / values[(size_t)q*nn + al*neq + al];
  }
  int spmvs = 0;
  spmv_plain(neq, N, start, cols, values, x, r.data()); ++spmvs;
  for(size_t i = 0; i < n; ++i) { r[i] = b[i] - r[i]; rh[i] = r[i]; }
  Real bnorm = std::sqrt(dot(n, b, b));
  if(bnorm == 0.) bnorm = 1.;
  Real rho_prev = 1., alpha = 1., omega = 1.;
  Real rnorm = std::sqrt(dot(n, r.data(), r.data()));
  int it = 0, conv = rnorm <= tol*bnorm;
  while(!conv && it < maxit) {
    ++it;
    const Real rho = dot(n, rh.data(), r.data());
    if(rho == 0.) break;
    const Real beta = (rho/rho_prev)*(alpha/omega);
    for(size_t i = 0; i < n; ++i) { p[i] = r[i] + beta*(p[i] - omega*v[i]); ph[i] = dinv[i]*p[i]; }
    spmv_plain(neq, N, start, cols, values, ph.data(), v.data()); ++spmvs;
    alpha = rho / dot(n, rh.data(), v.data());
    for(size_t i = 0; i < n; ++i) s[i] = r[i] - alpha*v[i];
    const Real snorm = std::sqrt(dot(n, s.data(), s.data()));
    if(snorm <= tol*bnorm) {
      for(size_t i = 0; i < n; ++i) { x[i] += alpha*ph[i]; r[i] = s[i]; }
      conv = 1; break;
    }
    for(size_t i = 0; i < n; ++i) sh[i] = dinv[i]*s[i];
    spmv_plain(neq, N, start, cols, values, sh.data(), t.data()); ++spmvs;
    omega = dot(n, t.data(), s.data()) / dot(n, t.data(), t.data());
    for(size_t i = 0; i < n; ++i) { x[i] += alpha*ph[i] + omega*sh[i]; r[i] = s[i] - omega*t[i]; }
    rho_prev = rho;
    rnorm = std::sqrt(dot(n, r.data(), r.data()));
    if(rnorm <= tol*bnorm) conv = 1;
    if(omega == 0.) break;
  }
  std::vector<Real> ax(n);
  spmv_plain(neq, N, start, cols, values, x, ax.data());
  for(size_t i = 0; i < n; ++i) ax[i] = b[i] - ax[i];
  if(st) {
    st->iterations = it; st->converged = conv; st->spmvs = spmvs + 1;
    st->rel_residual = std::sqrt(dot(n, ax.data(), ax.data())) / bnorm;
  }
  return conv ? 0 : 1;
}

int orc_gmres_ilu0(int neq, uint32_t N, const uint32_t* start, const uint32_t* cols, const double* values,
                   const double* b, double* x, int m, double tol, int maxit, orc_solve_stats* st)
{
  const size_t n = (size_t)N*neq;
  PointCSR P;
  to_point(neq, N, start, cols, values, P);
  PointCSR F = P;
  ilu0_factor(F);
  Real bnorm = std::sqrt(dot(n, b, b));
  if(bnorm == 0.) bnorm = 1.;
  std::vector<Real> V((size_t)(m+1)*n), H((size_t)(m+1)*m), cs(m), sn(m), g(m+1), w(n), z(n), h(m+1), y(m);
  int it = 0, conv = 0, spmvs = 0;
  auto matvec = [&](const Real* in, Real* out) { spmv_plain(neq, N, start, cols, values, in, out); ++spmvs; };
  while(it < maxit) {
    matvec(x, w.data());
    for(size_t i = 0; i < n; ++i) w[i] = b[i] - w[i];
    const Real beta = std::sqrt(dot(n, w.data(), w.data()));
    if(beta <= tol*bnorm) { conv = 1; break; }
    for(size_t i = 0; i < n; ++i) V[i] = w[i] / beta;
    std::fill(g.begin(), g.end(), 0.); g[0] = beta;
    int j = 0;
    for(; j < m && it < maxit; ++j) {
      ++it;
      ilu0_solve(F, &V[(size_t)j*n], z.data());
      matvec(z.data(), w.data());
      // CGS2: two passes of classical Gram-Schmidt
      for(int i = 0; i <= j; ++i) h[i] = 0.;
      for(int pass = 0; pass < 2; ++pass) {
        std::vector<Real> c(j+1);
        for(int i = 0; i <= j; ++i) c[i] = dot(n, &V[(size_t)i*n], w.data());
        for(int i = 0; i <= j; ++i) { h[i] += c[i]; const Real* vi = &V[(size_t)i*n]; for(size_t l = 0; l < n; ++l) w[l] -= c[i]*vi[l]; }
      }
      const Real hn = std::sqrt(dot(n, w.data(), w.data()));
      for(size_t l = 0; l < n; ++l) V[(size_t)(j+1)*n+l] = w[l] / hn;
      h[j+1] = hn;
      for(int i = 0; i < j; ++i) {
        const Real t0 = cs[i]*h[i] + sn[i]*h[i+1];
        h[i+1] = -sn[i]*h[i] + cs[i]*h[i+1];
        h[i] = t0;
      }
      const Real den = std::sqrt(h[j]*h[j] + h[j+1]*h[j+1]);
      cs[j] = h[j]/den; sn[j] = h[j+1]/den;
      h[j] = den; h[j+1] = 0.;
      g[j+1] = -sn[j]*g[j]; g[j] = cs[j]*g[j];
      for(int i = 0; i <= j; ++i) H[(size_t)i*m + j] = h[i];
      if(std::fabs(g[j+1]) <= tol*bnorm) { ++j; conv = 1; break; }
    }
    // back substitution and update x += M^-1 V y
    for(int i = j-1; i >= 0; --i) {
      Real s = g[i];
      for(int l = i+1; l < j; ++l) s -= H[(size_t)i*m + l]*y[l];
      y[i] = s / H[(size_t)i*m + i];
    }
    std::fill(w.begin(), w.end(), 0.);
    for(int i = 0; i < j; ++i) { const Real* vi = &V[(size_t)i*n]; for(size_t l = 0; l < n; ++l) w[l] += y[i]*vi[l]; }
    ilu0_solve(F, w.data(), z.data());
    for(size_t l = 0; l < n; ++l) x[l] += z[l];
    if(conv) break;
  }
  matvec(x, w.data());
  for(size_t i = 0; i < n; ++i) w[i] = b[i] - w[i];
  const Real rel = std::sqrt(dot(n, w.data(), w.data())) / bnorm;
  if(st) { st->iterations = it; st->converged = rel <= tol*1.0000001 || conv; st->spmvs = spmvs; st->rel_residual = rel; }
  return (rel <= tol*10.) ? 0 : 1;
}

void orc_linearize_u(uint32_t n, int dim, const double* u, const double* u1, const double* u2, const double* u3, double* u_adv)
{
  for(size_t i = 0; i < (size_t)n*dim; ++i)
    u_adv[i] = 2.1875*u[i] - 2.1875*u1[i] + 1.3125*u2[i] - 0.3125*u3[i];
}

void orc_update(uint32_t n, int dim, double* u, double* p, double* u1, double* u2, double* u3, const double* sol)
{
  const int neq = dim+1;
  for(size_t i = 0; i < n; ++i) {
    for(int d = 0; d < dim; ++d) {
      const size_t q = i*dim+d;
      u3[q] = u2[q]; u2[q] = u1[q]; u1[q] = u[q];
      u[q] += sol[i*neq+d];
    }
    p[i] += sol[i*neq+dim];
  }
}

} // extern "C"

// ---------------------------------------------------------------------------------------------
// NavierStokesSemiImplicit element actions (plugins/UFEM/src/UFEM/ns_semi_implicit/MatrixAssembly.hpp,
// PressureSystem.cpp:166-176 and implementation/NavierStokesSemiImplicit.cpp).  Element vectors and
// matrices are var-major like the Proto element matrices (component i of node a at i*k + a).
// The velocity system has neq = dim, the pressure system neq = 1; both are indexed by LSS row (p2m).
// ---------------------------------------------------------------------------------------------
namespace {

bool simplex(int et) { return et == ORC_TRIAG2D || et == ORC_TETRA3D; }

// IntegralOrder (MatrixAssembly.hpp:28-47): 1 for triangles and tetrahedra, 2 otherwise
int si_order(int et) { return simplex(et) ? 1 : 2; }

// shape values N, physical gradients G (dim x k) and detJ at mapped point xi
bool point_data(int et, const Real* xi, const Real* X, Real* N, Real* G, Real& det)
{
  const int k = enodes(et), dim = edim(et);
  Real dN[24], J[9], Ji[9];
  shape_values(et, xi, N);
  shape_gradients(et, xi, dN);
  jacobian(et, xi, X, J);
  if(!inverse_det(dim, J, Ji, det)) return false;
  for(int d = 0; d < dim; ++d)
    for(int m = 0; m < k; ++m) {
      Real s = 0.;
      for(int l = 0; l < dim; ++l) s += Ji[d*dim+l] * dN[l*k+m];
      G[d*k+m] = s;
    }
  return true;
}

// value at the point of nodal data f (k x dim, node-major)
void interp(int k, int dim, const Real* N, const Real* f, Real* out)
{
  for(int d = 0; d < dim; ++d) {
    Real s = 0.;
    for(int m = 0; m < k; ++m) s += N[m]*f[m*dim+d];
    out[d] = s;
  }
}

Real element_nu(int k, const Real* nu)
{
  Real s = 0.;
  for(int i = 0; i < k; ++i) s += nu[i];
  return std::fabs(s / k);   // ComputeTauImpl::operator() (SUPG.hpp:110-115)
}

// VelocityAssembly (MatrixAssembly.hpp:49-146): M and T (dim*k)^2, then system += T + theta*dt*M (:458)
int si_velocity_element(int et, const TauCfg& cfg, Real dt, Real theta, const Real* X, const Real* uadv,
                        const Real* nu, Real* S)
{
  const int k = enodes(et), dim = edim(et), n = dim*k;
  Real M[576], T[576];
  std::fill(M, M + n*n, 0.);
  std::fill(T, T + n*n, 0.);
  Real tau_ps, tau_su, tau_bulk;
  if(compute_tau(et, cfg, X, uadv, element_nu(k, nu), dt, tau_ps, tau_su, tau_bulk)) return -2;
  Real gxi[24], gw[8], N[8], G[24], det;
  const int order = si_order(et);
  if(order != 2) {   // mass part with the second-order rule (:66-86)
    const int ngp = gauss(et, 2, gxi, gw);
    for(int q = 0; q < ngp; ++q) {
      if(!point_data(et, gxi + q*dim, X, N, G, det)) return -2;
      const Real w = gw[q]*det;
      for(int i = 0; i < dim; ++i)
        for(int a = 0; a < k; ++a)
          for(int b = 0; b < k; ++b) T[(i*k+a)*n + i*k+b] += (w*N[a])*N[b];
    }
  }
  const int ngp = gauss(et, order, gxi, gw);
  for(int q = 0; q < ngp; ++q) {   // (:88-118)
    if(!point_data(et, gxi + q*dim, X, N, G, det)) return -2;
    Real nuq = 0.;
    for(int m = 0; m < k; ++m) nuq += N[m]*nu[m];
    const Real w = gw[q]*det;
    const Real bulk = w*tau_bulk;
    for(int a = 0; a < k; ++a)
      for(int b = 0; b < k; ++b) {
        Real gtg = 0.;
        for(int d = 0; d < dim; ++d) gtg += G[d*k+a]*G[d*k+b];
        const Real lap = (w*nuq)*gtg;
        for(int i = 0; i < dim; ++i) {
          M[(i*k+a)*n + i*k+b] += lap + (bulk*G[i*k+a])*G[i*k+b];
          if(order == 2) T[(i*k+a)*n + i*k+b] += (w*N[a])*N[b];
        }
      }
  }
  // off-diagonal bulk blocks with the one-point rule (:120-145)
  gauss(et, 1, gxi, gw);
  if(!point_data(et, gxi, X, N, G, det)) return -2;
  const Real bulk = gw[0]*det*tau_bulk;
  for(int i = 0; i < dim; ++i)
    for(int j = i+1; j < dim; ++j)
      for(int a = 0; a < k; ++a)
        for(int b = 0; b < k; ++b) {
          const Real v = (bulk*G[i*k+a])*G[j*k+b];
          M[(i*k+a)*n + j*k+b] += v;
          M[(j*k+b)*n + i*k+a] += v;
        }
  const Real tdt = theta*dt;
  for(int r = 0; r < n*n; ++r) S[r] = T[r] + tdt*M[r];
  return 0;
}

// VelocityRHS (MatrixAssembly.hpp:206-330).  avec, dtamu: k x dim node-major (dt(1-theta) a - u);
// ppdp: k ((1-theta) delta_p_sum - p).  res: dim*k var-major.
int si_velocity_rhs_element(int et, const TauCfg& cfg, Real dt, const Real* X, const Real* uadv, const Real* nu,
                            const Real* avec, const Real* dtamu, const Real* ppdp, Real* res)
{
  const int k = enodes(et), dim = edim(et);
  std::fill(res, res + dim*k, 0.);
  Real tau_ps, tau_su, tau_bulk;
  if(compute_tau(et, cfg, X, uadv, element_nu(k, nu), dt, tau_ps, tau_su, tau_bulk)) return -2;
  Real gxi[24], gw[8], N[8], G[24], det, uq[3], adv[8], npa[8];
  const int order = si_order(et);
  // adv = u.G, N_plus_adv = tau_su adv + N, f = w sum_i G_i . dtamu_i
  auto common = [&](Real w, Real& f) {
    for(int m = 0; m < k; ++m) {
      Real s = 0.;
      for(int d = 0; d < dim; ++d) s += uq[d]*G[d*k+m];
      adv[m] = s;
      npa[m] = tau_su*s + N[m];
    }
    f = 0.;
    for(int i = 0; i < dim; ++i)
      for(int m = 0; m < k; ++m) f += G[i*k+m]*dtamu[m*dim+i];
    f *= w;
  };
  if(order != 2) {   // mass-type terms with the second-order rule (:236-268)
    const int ngp = gauss(et, 2, gxi, gw);
    for(int q = 0; q < ngp; ++q) {
      if(!point_data(et, gxi + q*dim, X, N, G, det)) return -2;
      interp(k, dim, N, uadv, uq);
      const Real w = gw[q]*det;
      Real f;
      common(w, f);
      const Real f_skew = 0.5*f;
      for(int i = 0; i < dim; ++i) {
        Real na = 0., ad = 0.;
        for(int m = 0; m < k; ++m) { na += N[m]*avec[m*dim+i]; ad += adv[m]*dtamu[m*dim+i]; }
        const Real c = -w*na, e = w*ad;
        for(int a = 0; a < k; ++a) res[i*k+a] += npa[a]*(c + e + f_skew*uq[i]);
      }
    }
  }
  const int ngp = gauss(et, order, gxi, gw);
  for(int q = 0; q < ngp; ++q) {   // (:270-326)
    if(!point_data(et, gxi + q*dim, X, N, G, det)) return -2;
    interp(k, dim, N, uadv, uq);
    Real nuq = 0.;
    for(int m = 0; m < k; ++m) nuq += N[m]*nu[m];
    const Real w = gw[q]*det;
    const Real w_visc = w*nuq, tau_su_w = w*tau_su;
    Real np = 0.;
    for(int m = 0; m < k; ++m) np += N[m]*ppdp[m];
    const Real b = -w*np;
    Real f;
    common(w, f);
    const Real f_bulk = tau_bulk*f, f_skew = 0.5*f;
    for(int i = 0; i < dim; ++i) {
      Real gp = 0.;
      for(int m = 0; m < k; ++m) gp += G[i*k+m]*ppdp[m];
      const Real a_ = tau_su_w*gp;
      Real dv[3];
      for(int l = 0; l < dim; ++l) {
        Real s = 0.;
        for(int m = 0; m < k; ++m) s += G[l*k+m]*dtamu[m*dim+i];
        dv[l] = w_visc*s;
      }
      Real c = 0., e = 0.;
      if(order == 2) {
        Real na = 0., ad = 0.;
        for(int m = 0; m < k; ++m) { na += N[m]*avec[m*dim+i]; ad += adv[m]*dtamu[m*dim+i]; }
        c = -w*na; e = w*ad;
      }
      for(int a = 0; a < k; ++a) {
        Real gd = 0.;
        for(int l = 0; l < dim; ++l) gd += G[l*k+a]*dv[l];
        Real v = adv[a]*a_ + G[i*k+a]*(b + f_bulk);
        if(order == 2) v += npa[a]*(c + e + f_skew*uq[i]);
        res[i*k+a] += v + gd;
      }
    }
  }
  return 0;
}

// PressureRHS (MatrixAssembly.hpp:148-204).  tau_ps from u (NavierStokesSemiImplicit.cpp PRHSAssembly),
// the operators from u_adv; uvec, avec, davec k x dim; pvec k.  res: k.
int si_pressure_rhs_element(int et, const TauCfg& cfg, Real dt, const Real* X, const Real* u, const Real* uadv,
                            const Real* nu, const Real* uvec, const Real* avec, const Real* davec,
                            const Real* pvec, Real* res)
{
  const int k = enodes(et), dim = edim(et);
  Real tau_ps, tau_su, tau_bulk;
  if(compute_tau(et, cfg, X, u, element_nu(k, nu), dt, tau_ps, tau_su, tau_bulk)) return -2;
  Real upda[24], apda[24];
  for(int r = 0; r < k*dim; ++r) { upda[r] = uvec[r] + dt*davec[r]; apda[r] = avec[r] + davec[r]; }
  Real App[64];
  std::fill(App, App + k*k, 0.);
  std::fill(res, res + k, 0.);
  Real gxi[24], gw[8], N[8], G[24], det, uq[3], adv[8];
  const int ngp = gauss(et, si_order(et), gxi, gw);
  for(int q = 0; q < ngp; ++q) {
    if(!point_data(et, gxi + q*dim, X, N, G, det)) return -2;
    interp(k, dim, N, uadv, uq);
    const Real w = gw[q]*det, tau_w = w*tau_ps;
    for(int m = 0; m < k; ++m) {
      Real s = 0.;
      for(int d = 0; d < dim; ++d) s += uq[d]*G[d*k+m];
      adv[m] = tau_ps*s;
    }
    for(int a = 0; a < k; ++a)
      for(int b = 0; b < k; ++b) {
        Real gtg = 0.;
        for(int d = 0; d < dim; ++d) gtg += G[d*k+a]*G[d*k+b];
        App[a*k+b] += tau_w*gtg;
      }
    for(int i = 0; i < dim; ++i) {
      Real ga = 0., av = 0., na = 0.;
      for(int m = 0; m < k; ++m) {
        ga += G[i*k+m]*upda[m*dim+i];
        av += adv[m]*upda[m*dim+i];
        na += N[m]*apda[m*dim+i];
      }
      const Real a_ = w*ga, b = w*av, c = tau_w*na;
      for(int m = 0; m < k; ++m) res[m] -= a_*(N[m] + adv[m]*0.5) + (b + c)*G[i*k+m];
    }
  }
  for(int a = 0; a < k; ++a) {
    Real s = 0.;
    for(int b = 0; b < k; ++b) s += App[a*k+b]*pvec[b];
    res[a] -= s;
  }
  return 0;
}

// ApplyAup (MatrixAssembly.hpp:332-376): tau_su from u_adv.  dp: k.  res: dim*k var-major.
int si_apply_aup_element(int et, const TauCfg& cfg, Real dt, Real theta, const Real* X, const Real* uadv,
                         const Real* nu, const Real* dp, Real* res)
{
  const int k = enodes(et), dim = edim(et);
  Real tau_ps, tau_su, tau_bulk;
  if(compute_tau(et, cfg, X, uadv, element_nu(k, nu), dt, tau_ps, tau_su, tau_bulk)) return -2;
  std::fill(res, res + dim*k, 0.);
  Real gxi[24], gw[8], N[8], G[24], det, uq[3], adv[8];
  const int ngp = gauss(et, si_order(et), gxi, gw);
  for(int q = 0; q < ngp; ++q) {
    if(!point_data(et, gxi + q*dim, X, N, G, det)) return -2;
    interp(k, dim, N, uadv, uq);
    const Real w = theta*gw[q]*det;
    const Real ws = w*tau_su;
    for(int m = 0; m < k; ++m) {
      Real s = 0.;
      for(int d = 0; d < dim; ++d) s += uq[d]*G[d*k+m];
      adv[m] = ws*s;
    }
    Real nd = 0.;
    for(int m = 0; m < k; ++m) nd += N[m]*dp[m];
    const Real b = -w*nd;
    for(int i = 0; i < dim; ++i) {
      Real a_ = 0.;
      for(int m = 0; m < k; ++m) a_ += G[i*k+m]*dp[m];
      for(int m = 0; m < k; ++m) res[i*k+m] += a_*adv[m] + b*G[i*k+m];
    }
  }
  return 0;
}

// PressureSystemAssembly::assemble_pp (PressureSystem.cpp:166-176): theta (tau_ps + dt) int G^T G,
// tau_ps from u.  S: k x k.
int si_pressure_matrix_element(int et, const TauCfg& cfg, Real dt, Real theta, const Real* X, const Real* u,
                               const Real* nu, Real* S)
{
  const int k = enodes(et), dim = edim(et);
  Real tau_ps, tau_su, tau_bulk;
  if(compute_tau(et, cfg, X, u, element_nu(k, nu), dt, tau_ps, tau_su, tau_bulk)) return -2;
  Real A[64];
  std::fill(A, A + k*k, 0.);
  Real gxi[24], gw[8], N[8], G[24], det;
  const int ngp = gauss(et, 2, gxi, gw);
  for(int q = 0; q < ngp; ++q) {
    if(!point_data(et, gxi + q*dim, X, N, G, det)) return -2;
    const Real w = gw[q]*det;
    for(int a = 0; a < k; ++a)
      for(int b = 0; b < k; ++b) {
        Real gtg = 0.;
        for(int d = 0; d < dim; ++d) gtg += G[d*k+a]*G[d*k+b];
        A[a*k+b] += w*gtg;
      }
  }
  const Real c = theta*(tau_ps + dt);
  for(int r = 0; r < k*k; ++r) S[r] = c*A[r];
  return 0;
}

// lumped mass matrix (NavierStokesSemiImplicit.cpp MassMatrixAssembly): row sums of int N^T N
int si_lumped_mass_element(int et, const Real* X, Real* m)
{
  const int k = enodes(et), dim = edim(et);
  Real A[64];
  std::fill(A, A + k*k, 0.);
  Real gxi[24], gw[8], N[8], G[24], det;
  const int ngp = gauss(et, 2, gxi, gw);
  for(int q = 0; q < ngp; ++q) {
    if(!point_data(et, gxi + q*dim, X, N, G, det)) return -2;
    const Real w = gw[q]*det;
    for(int a = 0; a < k; ++a)
      for(int b = 0; b < k; ++b) A[a*k+b] += (w*N[a])*N[b];
  }
  for(int a = 0; a < k; ++a) {   // lump(_A): row sum onto the diagonal (EigenTransforms.hpp:536-553)
    Real s = 0.;
    for(int b = 0; b < k; ++b) s += A[a*k+b];
    m[a] = s;
  }
  return 0;
}

// gather k x dim node-major values of mesh-node field f (n x dim) or of LSS vector v (rows x dim) via p2m
void gather(int k, int dim, const uint32_t* c, const uint32_t* p2m, const Real* f, Real* out)
{
  for(int a = 0; a < k; ++a) {
    const size_t r = p2m ? p2m[c[a]] : c[a];
    for(int d = 0; d < dim; ++d) out[a*dim+d] = f[r*dim+d];
  }
}

}  // namespace

extern "C" {

int orc_si_lumped_mass(int et, uint32_t n_elems, const double* coords, const uint32_t* conn,
                       const uint32_t* p2m, double* diag)
{
  const int k = enodes(et), dim = edim(et);
  Real X[24], m[8];
  for(uint32_t e = 0; e < n_elems; ++e) {
    const uint32_t* c = conn + (size_t)e*k;
    gather(k, dim, c, nullptr, coords, X);
    if(si_lumped_mass_element(et, X, m)) return -2;
    for(int a = 0; a < k; ++a)
      for(int i = 0; i < dim; ++i) diag[(size_t)p2m[c[a]]*dim + i] += m[a];
  }
  return 0;
}

int orc_si_velocity_matrix(int et, const orc_ns_params* prm, uint32_t n_elems, const double* coords,
                           const uint32_t* conn, const double* uadv, const double* nu, const uint32_t* slots,
                           double* values)
{
  const int k = enodes(et), dim = edim(et), n = dim*k;
  const TauCfg cfg = tau_cfg(prm);
  Real X[24], ua[24], nn[8], S[576];
  for(uint32_t e = 0; e < n_elems; ++e) {
    const uint32_t* c = conn + (size_t)e*k;
    gather(k, dim, c, nullptr, coords, X);
    gather(k, dim, c, nullptr, uadv, ua);
    gather(k, 1, c, nullptr, nu, nn);
    if(si_velocity_element(et, cfg, prm->dt, prm->theta, X, ua, nn, S)) return -2;
    for(int a = 0; a < k; ++a)
      for(int b = 0; b < k; ++b) {
        double* blk = values + (size_t)slots[((size_t)e*k + a)*k + b]*dim*dim;
        for(int i = 0; i < dim; ++i)
          for(int j = 0; j < dim; ++j) blk[i*dim+j] += S[(i*k+a)*n + j*k+b];
      }
  }
  return 0;
}

int orc_si_pressure_matrix(int et, const orc_ns_params* prm, uint32_t n_elems, const double* coords,
                           const uint32_t* conn, const double* u, const double* nu, const uint32_t* slots,
                           double* values)
{
  const int k = enodes(et), dim = edim(et);
  const TauCfg cfg = tau_cfg(prm);
  Real X[24], uu[24], nn[8], S[64];
  for(uint32_t e = 0; e < n_elems; ++e) {
    const uint32_t* c = conn + (size_t)e*k;
    gather(k, dim, c, nullptr, coords, X);
    gather(k, dim, c, nullptr, u, uu);
    gather(k, 1, c, nullptr, nu, nn);
    if(si_pressure_matrix_element(et, cfg, prm->dt, prm->theta, X, uu, nn, S)) return -2;
    for(int a = 0; a < k; ++a)
      for(int b = 0; b < k; ++b) values[slots[((size_t)e*k + a)*k + b]] += S[a*k+b];
  }
  return 0;
}

// the body-force overload of VelocityRHS (MatrixAssembly.hpp:339-363): after the terms above, every
// point of the second-order rule adds N^T g_i(x_q) w_q to component i, g interpolated from the nodes
static int si_body_force_element(int et, const Real* X, const Real* g, Real* res)
{
  const int k = enodes(et), dim = edim(et);
  Real gxi[24], gw[8], N[8], G[24], det, gq[3];
  const int ngp = gauss(et, 2, gxi, gw);
  for(int q = 0; q < ngp; ++q) {
    if(!point_data(et, gxi + q*dim, X, N, G, det)) return -2;
    interp(k, dim, N, g, gq);
    const Real w = gw[q]*det;
    for(int i = 0; i < dim; ++i)
      for(int a = 0; a < k; ++a) res[i*k+a] += N[a]*gq[i]*w;
  }
  return 0;
}

int orc_si_velocity_rhs(int et, const orc_ns_params* prm, uint32_t n_elems, const double* coords,
                        const uint32_t* conn, const uint32_t* p2m, const double* uadv, const double* nu,
                        const double* a_lss, const double* u_lss, const double* dps_lss, const double* p_lss,
                        double* rhs, const double* body_force)
{
  const int k = enodes(et), dim = edim(et);
  const TauCfg cfg = tau_cfg(prm);
  const Real dt = prm->dt, th = prm->theta;
  Real X[24], ua[24], nn[8], av[24], uv[24], dps[8], pv[8], dtamu[24], ppdp[8], res[24], gn[24];
  for(uint32_t e = 0; e < n_elems; ++e) {
    const uint32_t* c = conn + (size_t)e*k;
    gather(k, dim, c, nullptr, coords, X);
    gather(k, dim, c, nullptr, uadv, ua);
    gather(k, 1, c, nullptr, nu, nn);
    gather(k, dim, c, p2m, a_lss, av);
    gather(k, dim, c, p2m, u_lss, uv);
    gather(k, 1, c, p2m, dps_lss, dps);
    gather(k, 1, c, p2m, p_lss, pv);
    // lit(dt)*(1 - lit(theta))*lit(a) - lit(u_vec), (1 - lit(theta))*lit(delta_p_sum) - lit(p_vec)
    for(int r = 0; r < k*dim; ++r) dtamu[r] = dt*(1. - th)*av[r] - uv[r];
    for(int r = 0; r < k; ++r) ppdp[r] = (1. - th)*dps[r] - pv[r];
    if(si_velocity_rhs_element(et, cfg, dt, X, ua, nn, av, dtamu, ppdp, res)) return -2;
    if(body_force) {
      gather(k, dim, c, nullptr, body_force, gn);
      if(si_body_force_element(et, X, gn, res)) return -2;
    }
    for(int a = 0; a < k; ++a)
      for(int i = 0; i < dim; ++i) rhs[(size_t)p2m[c[a]]*dim + i] += res[i*k+a];
  }
  return 0;
}

int orc_si_pressure_rhs(int et, const orc_ns_params* prm, uint32_t n_elems, const double* coords,
                        const uint32_t* conn, const uint32_t* p2m, const double* u, const double* uadv,
                        const double* nu, const double* u_lss, const double* a_lss, const double* da_lss,
                        const double* p_lss, double* rhs)
{
  const int k = enodes(et), dim = edim(et);
  const TauCfg cfg = tau_cfg(prm);
  Real X[24], uu[24], ua[24], nn[8], uv[24], av[24], dav[24], pv[8], res[8];
  for(uint32_t e = 0; e < n_elems; ++e) {
    const uint32_t* c = conn + (size_t)e*k;
    gather(k, dim, c, nullptr, coords, X);
    gather(k, dim, c, nullptr, u, uu);
    gather(k, dim, c, nullptr, uadv, ua);
    gather(k, 1, c, nullptr, nu, nn);
    gather(k, dim, c, p2m, u_lss, uv);
    gather(k, dim, c, p2m, a_lss, av);
    gather(k, dim, c, p2m, da_lss, dav);
    gather(k, 1, c, p2m, p_lss, pv);
    if(si_pressure_rhs_element(et, cfg, prm->dt, X, uu, ua, nn, uv, av, dav, pv, res)) return -2;
    for(int a = 0; a < k; ++a) rhs[p2m[c[a]]] += res[a];
  }
  return 0;
}

int orc_si_apply_aup(int et, const orc_ns_params* prm, uint32_t n_elems, const double* coords,
                     const uint32_t* conn, const uint32_t* p2m, const double* uadv, const double* nu,
                     const double* dp_lss, double* rhs)
{
  const int k = enodes(et), dim = edim(et);
  const TauCfg cfg = tau_cfg(prm);
  Real X[24], ua[24], nn[8], dp[8], res[24];
  for(uint32_t e = 0; e < n_elems; ++e) {
    const uint32_t* c = conn + (size_t)e*k;
    gather(k, dim, c, nullptr, coords, X);
    gather(k, dim, c, nullptr, uadv, ua);
    gather(k, 1, c, nullptr, nu, nn);
    gather(k, 1, c, p2m, dp_lss, dp);
    if(si_apply_aup_element(et, cfg, prm->dt, prm->theta, X, ua, nn, dp, res)) return -2;
    for(int a = 0; a < k; ++a)
      for(int i = 0; i < dim; ++i) rhs[(size_t)p2m[c[a]]*dim + i] += res[i*k+a];
  }
  return 0;
}

} // extern "C"

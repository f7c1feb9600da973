// Node-wise actions of the UFEM NavierStokes solver (gfx950), SoA fields.
//   LinearizeU    u_adv = 2.1875 u - 2.1875 u1 + 1.3125 u2 - 0.3125 u3   (ns_implementation/NavierStokes.cpp:105)
//   UpdateActions u3 = u2; u2 = u1; u1 = u; u += du; p += dp              (NavierStokes.cpp:172-180)
// Evaluated left to right without contraction into fused multiply-adds, as in the reference expression.
#include "internal.hpp"

namespace cf3hip {

namespace {

#pragma clang fp contract(off)

__global__ void k_linearize(uint32_t n, const double* __restrict__ u, const double* __restrict__ u1,
                            const double* __restrict__ u2, const double* __restrict__ u3, double* __restrict__ ua)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  ua[i] = 2.1875 * u[i] - 2.1875 * u1[i] + 1.3125 * u2[i] - 0.3125 * u3[i];
}

// p2m (null: identity) gives the block row of mesh node i under periodic links
__global__ void k_update(uint32_t n, int dim, double* __restrict__ u, double* __restrict__ u1,
                         double* __restrict__ u2, double* __restrict__ u3, const double* __restrict__ sol,
                         const uint32_t* __restrict__ p2m, int comp)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const uint64_t row = p2m ? p2m[i] : i;
  u3[i] = u2[i];
  u2[i] = u1[i];
  u1[i] = u[i];
  u[i] += sol[row * (dim + 1) + comp];
}

__global__ void k_update_p(uint32_t n, int dim, double* __restrict__ p, const double* __restrict__ sol,
                           const uint32_t* __restrict__ p2m)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  const uint64_t row = p2m ? p2m[i] : i;
  p[i] += sol[row * (dim + 1) + dim];
}

}  // namespace

void ns_linearize(Fields& F)
{
  const Mesh& M = *F.mesh;
  Context& c = *M.ctx;
  for(int d = 0; d < M.dim; ++d)
    hipLaunchKernelGGL(k_linearize, dim3(grid_for(M.n_nodes, 256)), dim3(256), 0, c.stream, M.n_nodes,
                       F.u[d].as<double>(), F.u1[d].as<double>(), F.u2[d].as<double>(), F.u3[d].as<double>(),
                       F.uadv[d].as<double>());
  HIP_OK(hipGetLastError());
}

void ns_update(Lss& L, Fields& F)
{
  const Mesh& M = *F.mesh;
  Context& c = *M.ctx;
  if(L.neq != M.dim + 1 || (L.periodic ? L.mesh != F.mesh : L.n_nodes != M.n_nodes))
    CF3_THROW(CF3HIP_ERR_INVALID_ARG, "system does not match the fields");
  const uint32_t* p2m = L.periodic ? L.p2m.as<uint32_t>() : nullptr;
  for(int d = 0; d < M.dim; ++d)
    hipLaunchKernelGGL(k_update, dim3(grid_for(M.n_nodes, 256)), dim3(256), 0, c.stream, M.n_nodes, M.dim,
                       F.u[d].as<double>(), F.u1[d].as<double>(), F.u2[d].as<double>(), F.u3[d].as<double>(),
                       L.sol.as<double>(), p2m, d);
  hipLaunchKernelGGL(k_update_p, dim3(grid_for(M.n_nodes, 256)), dim3(256), 0, c.stream, M.n_nodes, M.dim,
                     F.p.as<double>(), L.sol.as<double>(), p2m);
  HIP_OK(hipGetLastError());
}

}  // namespace cf3hip

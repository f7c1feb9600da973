// Arnoldi building blocks shared by GMRES(m)+ILU0 (gmres.hip) and the FGMRES outer iteration of the
// precision form (krylov.hip): multi-vector dot products with a fixed summation order, classical
// Gram-Schmidt updates, and the small dense bookkeeping (Givens rotations, back substitution) done by a
// single device thread so that the host only reads the residual estimate.
#pragma once

#include <cmath>

#include "internal.hpp"

namespace cf3hip {
namespace arnoldi {
// internal linkage: every translation unit that includes this header gets its own kernels
namespace {

constexpr int kMD = 256;     // block size of the multi-dot kernels
constexpr int kMaxV = 33;    // basis vectors per multi-dot (restart length <= 32)

// part[l * G + b] = block b's share of V_l . w, l < nv (V holds nv vectors of length n back to back)
__global__ __launch_bounds__(kMD) void k_mdot_part(uint64_t n, int nv, const double* __restrict__ V,
                                                         const double* __restrict__ w, double* __restrict__ part)
{
  __shared__ double sh[kMaxV][kMD / 64];
  for(int l0 = 0; l0 < nv; l0 += 8) {   // eight basis vectors per sweep keep the accumulators in registers
    double acc[8] = {0., 0., 0., 0., 0., 0., 0., 0.};
    const int nl = nv - l0 < 8 ? nv - l0 : 8;
    const uint64_t stride = (uint64_t)gridDim.x * kMD;
    for(uint64_t i = (uint64_t)blockIdx.x * kMD + threadIdx.x; i < n; i += stride) {
      const double wi = w[i];
#pragma unroll
      for(int l = 0; l < 8; ++l)
        if(l < nl) acc[l] += V[(uint64_t)(l0 + l) * n + i] * wi;
    }
#pragma unroll
    for(int l = 0; l < 8; ++l) {
      if(l >= nl) break;
      double x = acc[l];
      for(int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
      if((threadIdx.x & 63) == 0) sh[l0 + l][threadIdx.x >> 6] = x;
    }
  }
  __syncthreads();
  if((int)threadIdx.x < nv) {
    double s = 0.;
    for(int q = 0; q < kMD / 64; ++q) s += sh[threadIdx.x][q];
    part[(uint64_t)threadIdx.x * gridDim.x + blockIdx.x] = s;
  }
}

// out[l] = sum over b of part[l * G + b], in a fixed order
__global__ __launch_bounds__(kMD) void k_mdot_fin(unsigned G, int nv, const double* __restrict__ part,
                                                        double* __restrict__ out)
{
  __shared__ double sh[kMD / 64];
  for(int l = 0; l < nv; ++l) {
    double s = 0.;
    for(unsigned i = threadIdx.x; i < G; i += kMD) s += part[(uint64_t)l * G + i];
    for(int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if(threadIdx.x == 0) out[l] = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
  }
}

// w -= sum_l c[l] V_l
__global__ void k_orth_update(uint64_t n, int nv, const double* __restrict__ V, const double* __restrict__ c,
                                     double* __restrict__ w)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  double s = w[i];
  for(int l = 0; l < nv; ++l) s -= c[l] * V[(uint64_t)l * n + i];
  w[i] = s;
}

// b = a / s[0]
__global__ void k_scale(uint64_t n, const double* __restrict__ a, const double* __restrict__ s,
                               double* __restrict__ b)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) b[i] = a[i] / s[0];
}

// w = b - w
__global__ void k_sub(uint64_t n, const double* __restrict__ b, double* __restrict__ w)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) w[i] = b[i] - w[i];
}

// u = sum_l y[l] V_l
__global__ void k_lincomb(uint64_t n, int nv, const double* __restrict__ V, const double* __restrict__ y,
                                 double* __restrict__ u)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  double s = 0.;
  for(int l = 0; l < nv; ++l) s += y[l] * V[(uint64_t)l * n + i];
  u[i] = s;
}

// x += sum_l y[l] V_l
__global__ void k_add_lincomb(uint64_t n, int nv, const double* __restrict__ V, const double* __restrict__ y,
                                     double* __restrict__ x)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  double s = x[i];
  for(int l = 0; l < nv; ++l) s += y[l] * V[(uint64_t)l * n + i];
  x[i] = s;
}

__global__ void k_add(uint64_t n, const double* __restrict__ z, double* __restrict__ x)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) x[i] += z[i];
}

// small dense bookkeeping of one Arnoldi step (single thread): H column j from the two
// Gram-Schmidt passes, norm, Givens rotations, residual estimate
// layout: [0] beta or |g_{j+1}|, [1] ||w||, cs[m], sn[m], g[m+1], H[(m+1)*m], h1[m+2], h2[m+2], y[m+2]
struct GmresDev {
  int m;
  double* base;
  __host__ __device__ double* cs() const { return base + 2; }
  __host__ __device__ double* sn() const { return base + 2 + m; }
  __host__ __device__ double* g() const { return base + 2 + 2 * m; }
  __host__ __device__ double* H() const { return base + 3 + 3 * m; }
  __host__ __device__ double* h1() const { return H() + (m + 1) * m; }
  __host__ __device__ double* h2() const { return h1() + (m + 2); }
  __host__ __device__ double* y() const { return h2() + (m + 2); }
  static int size(int m) { return 3 + 3 * m + (m + 1) * m + 3 * (m + 2); }
};

__global__ void k_arnoldi_close(int m, double* base, int j, const double* __restrict__ h1,
                                       const double* __restrict__ h2, const double* __restrict__ wn2)
{
  GmresDev G{m, base};
  double h[64];
  for(int i = 0; i <= j; ++i) h[i] = h1[i] + h2[i];
  const double hn = sqrt(wn2[0]);
  h[j + 1] = hn;
  base[1] = hn;
  double* cs = G.cs();
  double* sn = G.sn();
  double* g = G.g();
  for(int i = 0; i < j; ++i) {
    const double t0 = cs[i] * h[i] + sn[i] * h[i + 1];
    h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1];
    h[i] = t0;
  }
  const double den = sqrt(h[j] * h[j] + h[j + 1] * h[j + 1]);
  cs[j] = h[j] / den;
  sn[j] = h[j + 1] / den;
  h[j] = den;
  h[j + 1] = 0.;
  g[j + 1] = -sn[j] * g[j];
  g[j] = cs[j] * g[j];
  for(int i = 0; i <= j; ++i) G.H()[i * m + j] = h[i];
  base[0] = fabs(g[j + 1]);
}

__global__ void k_back_subst(int m, double* base, int j, double* __restrict__ y)
{
  GmresDev G{m, base};
  for(int i = j - 1; i >= 0; --i) {
    double s = G.g()[i];
    for(int l = i + 1; l < j; ++l) s -= G.H()[i * m + l] * y[l];
    y[i] = s / G.H()[i * m + i];
  }
}

__global__ void k_gmres_start(int m, double* base, const double* __restrict__ beta2)
{
  GmresDev G{m, base};
  for(int i = 0; i <= m; ++i) G.g()[i] = 0.;
  G.g()[0] = sqrt(beta2[0]);
  base[0] = G.g()[0];
}

// out[l] = V_l . w for l < nv, on the context stream (partials in the context's reduction buffer)
void multidot(Context& c, uint64_t n, int nv, const double* V, const double* w, double* out)
{
  unsigned G = grid_for(n, kMD);
  if(G > (unsigned)kMaxRedBlocks * 8u / (unsigned)kMaxV) G = (unsigned)kMaxRedBlocks * 8u / (unsigned)kMaxV;
  double* part = c.red_partials.as<double>();
  hipLaunchKernelGGL(k_mdot_part, dim3(G), dim3(kMD), 0, c.stream, n, nv, V, w, part);
  hipLaunchKernelGGL(k_mdot_fin, dim3(1), dim3(kMD), 0, c.stream, G, nv, part, out);
  HIP_OK(hipGetLastError());
}

}  // namespace
}  // namespace arnoldi
}  // namespace cf3hip

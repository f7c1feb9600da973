// LSS matrix/vector kernels on the block-CSR system (gfx950): fill, SpMV (FP64 / FP32),
// symmetric Dirichlet, point-CSR export.
//
// Matrix layout: node row i holds blocks row_ptr[i] .. row_ptr[i+1]-1 (columns sorted), each block
// neq x neq row-major; all values of a node row are contiguous, so a node row streams as one
// coalesced segment.  Semantics: cf3/math/LSS/Trilinos/TrilinosFEVbrMatrix.cpp (reset :661-665,
// symmetric_dirichlet :482-553) and TrilinosDetail.cpp:380-411 (apply y = alpha A x + beta y).
#include <algorithm>
#include <unordered_map>

#include <cctype>
#include <cstdio>
#include <string>

#include "internal.hpp"
#include "spmv4.hpp"

namespace cf3hip {

namespace {

template<typename T>
__global__ void k_fill(T* __restrict__ p, uint64_t n, T v)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for(uint64_t j = i; j < n; j += stride) p[j] = v;
}

// FP32 mirror of FP64 values, each rounded to nearest
__global__ void k_round_fp32(float* __restrict__ dst, const double* __restrict__ src, uint64_t n)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for(uint64_t j = i; j < n; j += stride) dst[j] = (float)src[j];
}

// FP64, neq = 4: 16 lanes per node row (lane layout of spmv4::row_f64)
__global__ __launch_bounds__(256) void k_spmv4_f64(uint32_t N, const uint32_t* __restrict__ row_ptr,
                                                   const uint32_t* __restrict__ cols, const double* __restrict__ vals,
                                                   double alpha, const double* __restrict__ x, double beta,
                                                   double* __restrict__ y, const double* __restrict__ stop)
{
  if(stop && stop[0] != 0.) return;   // Krylov loop already stopped on the device
  const uint32_t row = (blockIdx.x * blockDim.x + threadIdx.x) >> 4;
  const int l = threadIdx.x & 15;
  if(row >= N) return;
  int a;
  bool out;
  const double acc = spmv4::row_f64(row_ptr, cols, vals, x, row, l, a, out);
  if(out) {
    const uint64_t o = (uint64_t)row * 4 + a;
    double r = alpha * acc;
    if(beta != 0.) r += beta * y[o];
    y[o] = r;
  }
}

// FP32, neq = 4: 16 lanes per node row, a lane streams one block row (4 floats, 16 B) of blocks
// 4k + (l >> 2); alpha = l & 3.  Partial sums reduced across the 4 blocks of an iteration.
__global__ __launch_bounds__(256) void k_spmv4_f32(uint32_t N, const uint32_t* __restrict__ row_ptr,
                                                   const uint32_t* __restrict__ cols, const float* __restrict__ vals,
                                                   float alpha, const float* __restrict__ x, float beta,
                                                   float* __restrict__ y, const double* __restrict__ stop)
{
  if(stop && stop[0] != 0.) return;
  const uint32_t row = (blockIdx.x * blockDim.x + threadIdx.x) >> 4;
  const int l = threadIdx.x & 15;
  if(row >= N) return;
  const uint32_t q0 = row_ptr[row], q1 = row_ptr[row + 1];
  const int sub = l >> 2, a = l & 3;
  float acc = 0.f;
  for(uint32_t q = q0 + sub; q < q1; q += 4) {
    const float4 v = *reinterpret_cast<const float4*>(vals + (uint64_t)q * 16 + 4 * a);
    const float4 xv = *reinterpret_cast<const float4*>(x + (uint64_t)cols[q] * 4);
    acc += v.x * xv.x + v.y * xv.y + v.z * xv.z + v.w * xv.w;
  }
  acc += __shfl_xor(acc, 4, 16);
  acc += __shfl_xor(acc, 8, 16);
  if(l < 4) {
    const uint64_t o = (uint64_t)row * 4 + a;
    float r = alpha * acc;
    if(beta != 0.f) r += beta * y[o];
    y[o] = r;
  }
}

// FP32 values with FP64 vectors and FP64 accumulation, neq = 4 (lane layout of spmv4::row_mixed)
__global__ __launch_bounds__(256) void k_spmv4_mixed(uint32_t N, const uint32_t* __restrict__ row_ptr,
                                                     const uint32_t* __restrict__ cols, const float* __restrict__ vals,
                                                     const double* __restrict__ x, double* __restrict__ y,
                                                     const double* __restrict__ stop)
{
  if(stop && stop[0] != 0.) return;
  const uint32_t row = (blockIdx.x * blockDim.x + threadIdx.x) >> 4;
  const int l = threadIdx.x & 15;
  if(row >= N) return;
  int a;
  bool out;
  const double acc = spmv4::row_mixed(row_ptr, cols, vals, x, row, l, a, out);
  if(out) y[(uint64_t)row * 4 + a] = acc;
}

__global__ void k_spmv_mixed_generic(uint32_t N, int neq, const uint32_t* __restrict__ row_ptr,
                                     const uint32_t* __restrict__ cols, const float* __restrict__ vals,
                                     const double* __restrict__ x, double* __restrict__ y,
                                     const double* __restrict__ stop)
{
  if(stop && stop[0] != 0.) return;
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(r >= (uint64_t)N * neq) return;
  const uint32_t i = (uint32_t)(r / neq);
  const int a = (int)(r % neq);
  const int nn = neq * neq;
  double acc = 0.;
  for(uint32_t q = row_ptr[i]; q < row_ptr[i + 1]; ++q) {
    const float* blk = vals + (uint64_t)q * nn + a * neq;
    const double* xb = x + (uint64_t)cols[q] * neq;
    for(int b = 0; b < neq; ++b) acc += (double)blk[b] * xb[b];
  }
  y[r] = acc;
}

// any neq: one thread per scalar row
template<typename T>
__global__ void k_spmv_generic(uint32_t N, int neq, const uint32_t* __restrict__ row_ptr,
                               const uint32_t* __restrict__ cols, const T* __restrict__ vals, T alpha,
                               const T* __restrict__ x, T beta, T* __restrict__ y, const double* __restrict__ stop)
{
  if(stop && stop[0] != 0.) return;
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(r >= (uint64_t)N * neq) return;
  const uint32_t i = (uint32_t)(r / neq);
  const int a = (int)(r % neq);
  const int nn = neq * neq;
  T acc = T(0);
  for(uint32_t q = row_ptr[i]; q < row_ptr[i + 1]; ++q) {
    const T* blk = vals + (uint64_t)q * nn + a * neq;
    const T* xb = x + (uint64_t)cols[q] * neq;
    for(int b = 0; b < neq; ++b) acc += blk[b] * xb[b];
  }
  T res = alpha * acc;
  if(beta != T(0)) res += beta * y[r];
  y[r] = res;
}

// symmetric Dirichlet, column part: every node row m eliminates the columns of BC dofs
// (rhs[m, j] -= A[(m, j), (n, e)] * v_first(n, e); A[(m, j), (n, e)] = 0)
template<typename T>
__global__ void k_dirichlet_columns(uint32_t N, int neq, const uint32_t* __restrict__ row_ptr,
                                    const uint32_t* __restrict__ cols, T* __restrict__ vals,
                                    const uint8_t* __restrict__ mask, const double* __restrict__ vfirst,
                                    double* __restrict__ rhs)
{
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if(m >= N) return;
  const int nn = neq * neq;
  double r[4];
  bool touched = false;
  for(int j = 0; j < neq; ++j) r[j] = 0.;
  for(uint32_t q = row_ptr[m]; q < row_ptr[m + 1]; ++q) {
    const uint32_t n = cols[q];
    const uint8_t mk = mask[n];
    if(!mk) continue;
    if(!touched) { for(int j = 0; j < neq; ++j) r[j] = rhs ? rhs[(uint64_t)m * neq + j] : 0.; touched = true; }
    T* blk = vals + (uint64_t)q * nn;
    for(int e = 0; e < neq; ++e) {
      if(!((mk >> e) & 1)) continue;
      const double v = vfirst[(uint64_t)n * neq + e];
      for(int j = 0; j < neq; ++j) {
        r[j] += -(double)blk[j * neq + e] * v;
        blk[j * neq + e] = T(0);
      }
    }
  }
  if(touched && rhs)
    for(int j = 0; j < neq; ++j) rhs[(uint64_t)m * neq + j] = r[j];
}

// symmetric Dirichlet, row part: BC row -> identity; rhs = sol = last value
template<typename T>
__global__ void k_dirichlet_rows(uint32_t nbc, int neq, const uint32_t* __restrict__ dof, const double* __restrict__ vlast,
                                 const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ cols,
                                 T* __restrict__ vals, double* __restrict__ rhs, double* __restrict__ sol)
{
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if(b >= nbc) return;
  const uint32_t d = dof[b];
  const uint32_t n = d / neq;
  const int e = d % neq;
  const int nn = neq * neq;
  for(uint32_t q = row_ptr[n]; q < row_ptr[n + 1]; ++q) {
    T* blk = vals + (uint64_t)q * nn;
    for(int j = 0; j < neq; ++j) blk[e * neq + j] = T(0);
    if(cols[q] == n) blk[e * neq + e] = T(1);
  }
  if(rhs) rhs[d] = vlast[b];
  if(sol) sol[d] = vlast[b];
}

__global__ void k_scatter_mask(uint32_t nnodes_bc, const uint32_t* __restrict__ node, const uint8_t* __restrict__ m,
                               uint8_t* __restrict__ mask)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < nnodes_bc) mask[node[i]] = m[i];
}

__global__ void k_scatter_values(uint32_t n, const uint32_t* __restrict__ dof, const double* __restrict__ v,
                                 double* __restrict__ dst)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) dst[dof[i]] = v[i];
}

__global__ void k_point_csr(uint32_t N, int neq, const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ cols,
                            const double* __restrict__ vals, uint32_t* __restrict__ prow, uint32_t* __restrict__ pcol,
                            double* __restrict__ pval)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= N) return;
  const uint32_t q0 = row_ptr[i], deg = row_ptr[i + 1] - q0;
  const int nn = neq * neq;
  for(int a = 0; a < neq; ++a) {
    const uint64_t base = (uint64_t)nn * q0 + (uint64_t)a * neq * deg;
    prow[(uint64_t)i * neq + a] = (uint32_t)base;
    for(uint32_t s = 0; s < deg; ++s)
      for(int b = 0; b < neq; ++b) {
        pcol[base + s * neq + b] = cols[q0 + s] * neq + b;
        pval[base + s * neq + b] = vals[(uint64_t)(q0 + s) * nn + a * neq + b];
      }
  }
  if(i == N - 1) prow[(uint64_t)N * neq] = (uint32_t)((uint64_t)nn * row_ptr[N]);
}

}  // namespace

void lss_fill(Context& c, double* p, size_t n, double v)
{
  if(!n) return;
  hipLaunchKernelGGL(k_fill<double>, dim3(std::min<unsigned>(grid_for(n, 256), 4096)), dim3(256), 0, c.stream, p,
                     (uint64_t)n, v);
  HIP_OK(hipGetLastError());
}

__global__ void k_copy(double* __restrict__ dst, const double* __restrict__ src, uint64_t n)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for(uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) dst[j] = src[j];
}

void lss_copy(Context& c, double* dst, const double* src, size_t n)
{
  if(!n || dst == src) return;
  hipLaunchKernelGGL(k_copy, dim3(std::min<unsigned>(grid_for(n, 256), 4096)), dim3(256), 0, c.stream, dst, src,
                     (uint64_t)n);
  HIP_OK(hipGetLastError());
}

void lss_round_fp32(Lss& L)
{
  if(!L.has_fp32) CF3_THROW(CF3HIP_ERR_NOT_CREATED, "FP32 values not enabled");
  const uint64_t n = (uint64_t)L.n_blocks * L.neq * L.neq;
  if(!n) return;
  hipLaunchKernelGGL(k_round_fp32, dim3(std::min<unsigned>(grid_for(n, 256), 8192)), dim3(256), 0, L.ctx->stream,
                     L.values32.as<float>(), L.values.as<double>(), n);
  HIP_OK(hipGetLastError());
}

void lss_fill32(Context& c, float* p, size_t n, float v)
{
  if(!n) return;
  hipLaunchKernelGGL(k_fill<float>, dim3(std::min<unsigned>(grid_for(n, 256), 4096)), dim3(256), 0, c.stream, p,
                     (uint64_t)n, v);
  HIP_OK(hipGetLastError());
}

void lss_spmv(Lss& L, double alpha, const double* x, double beta, double* y, const double* stop)
{
  Context& c = *L.ctx;
  timer_begin(c, CF3HIP_KERNEL_SPMV);
  if(L.neq == 4) {
    hipLaunchKernelGGL(k_spmv4_f64, dim3(grid_for((uint64_t)L.n_nodes * 16, 256)), dim3(256), 0, c.stream, L.n_nodes,
                       L.row_ptr.as<uint32_t>(), L.cols.as<uint32_t>(), L.values.as<double>(), alpha, x, beta, y, stop);
  } else {
    hipLaunchKernelGGL(k_spmv_generic<double>, dim3(grid_for(L.dofs, 256)), dim3(256), 0, c.stream, L.n_nodes, L.neq,
                       L.row_ptr.as<uint32_t>(), L.cols.as<uint32_t>(), L.values.as<double>(), alpha, x, beta, y,
                       stop);
  }
  HIP_OK(hipGetLastError());
  timer_end(c, CF3HIP_KERNEL_SPMV);
}

void lss_spmv32(Lss& L, float alpha, const float* x, float beta, float* y, const double* stop)
{
  Context& c = *L.ctx;
  if(!L.has_fp32) CF3_THROW(CF3HIP_ERR_NOT_CREATED, "FP32 values not enabled");
  timer_begin(c, CF3HIP_KERNEL_SPMV_FP32);
  if(L.neq == 4) {
    hipLaunchKernelGGL(k_spmv4_f32, dim3(grid_for((uint64_t)L.n_nodes * 16, 256)), dim3(256), 0, c.stream, L.n_nodes,
                       L.row_ptr.as<uint32_t>(), L.cols.as<uint32_t>(), L.values32.as<float>(), alpha, x, beta, y, stop);
  } else {
    hipLaunchKernelGGL(k_spmv_generic<float>, dim3(grid_for(L.dofs, 256)), dim3(256), 0, c.stream, L.n_nodes, L.neq,
                       L.row_ptr.as<uint32_t>(), L.cols.as<uint32_t>(), L.values32.as<float>(), alpha, x, beta, y,
                       stop);
  }
  HIP_OK(hipGetLastError());
  timer_end(c, CF3HIP_KERNEL_SPMV_FP32);
}

void lss_spmv_mixed(Lss& L, const double* x, double* y, const double* stop)
{
  Context& c = *L.ctx;
  if(!L.has_fp32) CF3_THROW(CF3HIP_ERR_NOT_CREATED, "FP32 values not enabled");
  timer_begin(c, CF3HIP_KERNEL_SPMV_FP32);
  if(L.neq == 4) {
    hipLaunchKernelGGL(k_spmv4_mixed, dim3(grid_for((uint64_t)L.n_nodes * 16, 256)), dim3(256), 0, c.stream, L.n_nodes,
                       L.row_ptr.as<uint32_t>(), L.cols.as<uint32_t>(), L.values32.as<float>(), x, y, stop);
  } else {
    hipLaunchKernelGGL(k_spmv_mixed_generic, dim3(grid_for(L.dofs, 256)), dim3(256), 0, c.stream, L.n_nodes, L.neq,
                       L.row_ptr.as<uint32_t>(), L.cols.as<uint32_t>(), L.values32.as<float>(), x, y, stop);
  }
  HIP_OK(hipGetLastError());
  timer_end(c, CF3HIP_KERNEL_SPMV_FP32);
}

namespace {

// TrilinosCrsMatrix::symmetric_dirichlet (TrilinosCrsMatrix.cpp:325-402) followed by the solution value of
// System::dirichlet (System.cpp:240-255), the conditions one after the other in one workgroup.  The first
// application of a dof since the reset caches the column entries of the rows of its node's neighbours
// (the dof's own row excluded) and eliminates them into the rhs, then makes the dof's row the identity;
// a later application of the same dof only subtracts the cached column times its value from those rows'
// rhs again.  Both end with rhs = sol = value at the dof.  fresh[b] marks first applications; off[b] is the
// dof's cache offset (deg(node) * neq entries, row order of the node's block columns, then equation).
template<typename V32>
__global__ __launch_bounds__(256) void k_crs_dirichlet(uint32_t n, int neq, const uint32_t* __restrict__ node,
                                                       const uint32_t* __restrict__ eq, const double* __restrict__ val,
                                                       const uint8_t* __restrict__ fresh, const uint64_t* __restrict__ off,
                                                       const uint32_t* __restrict__ row_ptr,
                                                       const uint32_t* __restrict__ cols, double* values,
                                                       V32* values32, double* rhs, double* sol, double* cache)
{
  const uint32_t nn = (uint32_t)(neq * neq);
  for(uint32_t b = 0; b < n; ++b) {
    const uint32_t i = node[b];
    const uint32_t e = eq[b];
    const double v = val[b];
    const uint32_t q0 = row_ptr[i], deg = row_ptr[i + 1] - q0;
    const uint64_t bc = (uint64_t)i * neq + e;
    double* cb = cache + off[b];
    const uint32_t nent = deg * (uint32_t)neq;
    if(fresh[b]) {
      for(uint32_t t = threadIdx.x; t < nent; t += blockDim.x) {
        const uint32_t c = cols[q0 + t / neq], j = t % neq;
        const uint64_t r = (uint64_t)c * neq + j;
        double cv = 0.;
        if(r != bc) {
          // block (c, i) of row c (the graph is structurally symmetric)
          uint32_t lo = row_ptr[c], hi = row_ptr[c + 1];
          while(lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if(cols[mid] < i) lo = mid + 1; else hi = mid;
          }
          if(lo < row_ptr[c + 1] && cols[lo] == i) {
            const uint64_t k = (uint64_t)lo * nn + (uint64_t)j * neq + e;
            cv = values[k];
            rhs[r] -= cv * v;
            values[k] = 0.;
            if(values32) values32[k] = V32(0);
          }
        }
        cb[t] = cv;
      }
      __syncthreads();
      // the dof's own row: 1 on the diagonal, 0 elsewhere
      for(uint32_t t = threadIdx.x; t < nent; t += blockDim.x) {
        const uint32_t q = q0 + t / neq, l = t % neq;
        const uint64_t k = (uint64_t)q * nn + (uint64_t)e * neq + l;
        const double d = (cols[q] == i && l == e) ? 1. : 0.;
        values[k] = d;
        if(values32) values32[k] = (V32)d;
      }
    } else {
      for(uint32_t t = threadIdx.x; t < nent; t += blockDim.x) {
        const uint32_t c = cols[q0 + t / neq], j = t % neq;
        const uint64_t r = (uint64_t)c * neq + j;
        if(r != bc) rhs[r] -= cb[t] * v;
      }
    }
    __syncthreads();
    if(threadIdx.x == 0) {
      rhs[bc] = v;
      sol[bc] = v;
    }
    __syncthreads();
  }
}

// the conditions of one call on a CRS-builder system (host bookkeeping of the column cache)
void crs_dirichlet(Lss& L, uint32_t nbc, const uint32_t* nodes, const uint32_t* eqs, const double* values)
{
  Context& c = *L.ctx;
  const int neq = L.neq;
  std::vector<uint32_t> rp(L.n_nodes + 1);
  HIP_OK(hipMemcpyAsync(rp.data(), L.row_ptr.ptr, sizeof(uint32_t) * (L.n_nodes + 1), hipMemcpyDeviceToHost, c.stream));
  HIP_OK(hipStreamSynchronize(c.stream));
  std::vector<uint8_t> fresh(nbc);
  std::vector<uint64_t> off(nbc);
  uint64_t need = L.crs_used;
  for(uint32_t b = 0; b < nbc; ++b) {
    if(nodes[b] >= L.n_nodes || eqs[b] >= (uint32_t)neq)
      CF3_THROW(CF3HIP_ERR_INVALID_ARG, "Dirichlet condition on a dof outside the system");
    const uint64_t d = (uint64_t)nodes[b] * neq + eqs[b];
    auto it = L.crs_col.find(d);
    if(it == L.crs_col.end()) {
      fresh[b] = 1;
      off[b] = need;
      L.crs_col.emplace(d, need);
      need += (uint64_t)(rp[nodes[b] + 1] - rp[nodes[b]]) * neq;
    } else {
      fresh[b] = 0;
      off[b] = it->second;
    }
  }
  if(L.crs_cache.bytes < sizeof(double) * std::max<uint64_t>(need, 1)) {   // grow, keeping the cached columns
    DevBuf nb;
    nb.alloc(&c, sizeof(double) * std::max<uint64_t>(2 * need, 64));
    if(L.crs_used) HIP_OK(hipMemcpyAsync(nb.ptr, L.crs_cache.ptr, sizeof(double) * L.crs_used, hipMemcpyDeviceToDevice, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    std::swap(L.crs_cache.ptr, nb.ptr);
    std::swap(L.crs_cache.bytes, nb.bytes);
    std::swap(L.crs_cache.ctx, nb.ctx);
  }
  L.crs_used = need;
  DevBuf d_node, d_eq, d_val, d_fresh, d_off;
  d_node.alloc(&c, sizeof(uint32_t) * nbc);
  d_eq.alloc(&c, sizeof(uint32_t) * nbc);
  d_val.alloc(&c, sizeof(double) * nbc);
  d_fresh.alloc(&c, nbc);
  d_off.alloc(&c, sizeof(uint64_t) * nbc);
  HIP_OK(hipMemcpyAsync(d_node.ptr, nodes, sizeof(uint32_t) * nbc, hipMemcpyHostToDevice, c.stream));
  HIP_OK(hipMemcpyAsync(d_eq.ptr, eqs, sizeof(uint32_t) * nbc, hipMemcpyHostToDevice, c.stream));
  HIP_OK(hipMemcpyAsync(d_val.ptr, values, sizeof(double) * nbc, hipMemcpyHostToDevice, c.stream));
  HIP_OK(hipMemcpyAsync(d_fresh.ptr, fresh.data(), nbc, hipMemcpyHostToDevice, c.stream));
  HIP_OK(hipMemcpyAsync(d_off.ptr, off.data(), sizeof(uint64_t) * nbc, hipMemcpyHostToDevice, c.stream));
  hipLaunchKernelGGL(k_crs_dirichlet<float>, dim3(1), dim3(256), 0, c.stream, nbc, neq, d_node.as<uint32_t>(),
                     d_eq.as<uint32_t>(), d_val.as<double>(), d_fresh.as<uint8_t>(), d_off.as<uint64_t>(),
                     L.row_ptr.as<uint32_t>(), L.cols.as<uint32_t>(), L.values.as<double>(),
                     L.has_fp32 ? L.values32.as<float>() : (float*)nullptr, L.rhs.as<double>(), L.sol.as<double>(),
                     L.crs_cache.as<double>());
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(c.stream));
}

}  // namespace

// TrilinosCrsMatrix::clone_to (TrilinosCrsMatrix.cpp:518-538): the values and the cached Dirichlet columns
// into another CRS-builder system of the same graph
void lss_crs_clone(Lss& src, Lss& dst)
{
  Context& c = *src.ctx;
  if(src.builder != CF3HIP_BUILDER_CRS)
    CF3_THROW(CF3HIP_ERR_NOT_IMPLEMENTED, "clone_to is not implemented for the block (VBR) matrix builder");
  if(dst.builder != CF3HIP_BUILDER_CRS)
    CF3_THROW(CF3HIP_ERR_NOT_CREATED, "clone_to needs another system with the CRS matrix builder");
  if(dst.neq != src.neq || dst.n_nodes != src.n_nodes || dst.n_blocks != src.n_blocks)
    CF3_THROW(CF3HIP_ERR_NOT_CREATED, "clone_to needs a system of the same graph");
  std::vector<uint32_t> a(src.n_nodes + 1), b(src.n_nodes + 1), ca(src.n_blocks), cb(src.n_blocks);
  HIP_OK(hipMemcpyAsync(a.data(), src.row_ptr.ptr, sizeof(uint32_t) * a.size(), hipMemcpyDeviceToHost, c.stream));
  HIP_OK(hipMemcpyAsync(ca.data(), src.cols.ptr, sizeof(uint32_t) * ca.size(), hipMemcpyDeviceToHost, c.stream));
  HIP_OK(hipStreamSynchronize(c.stream));
  Context& cd = *dst.ctx;
  HIP_OK(hipMemcpyAsync(b.data(), dst.row_ptr.ptr, sizeof(uint32_t) * b.size(), hipMemcpyDeviceToHost, cd.stream));
  HIP_OK(hipMemcpyAsync(cb.data(), dst.cols.ptr, sizeof(uint32_t) * cb.size(), hipMemcpyDeviceToHost, cd.stream));
  HIP_OK(hipStreamSynchronize(cd.stream));
  if(a != b || ca != cb) CF3_THROW(CF3HIP_ERR_NOT_CREATED, "clone_to needs a system of the same graph");
  const size_t nv = (size_t)src.n_blocks * src.neq * src.neq;
  HIP_OK(hipMemcpyAsync(dst.values.ptr, src.values.ptr, sizeof(double) * nv, hipMemcpyDeviceToDevice, c.stream));
  if(src.has_fp32 && dst.has_fp32)
    HIP_OK(hipMemcpyAsync(dst.values32.ptr, src.values32.ptr, sizeof(float) * nv, hipMemcpyDeviceToDevice, c.stream));
  else if(dst.has_fp32)
    lss_round_fp32(dst);
  dst.crs_col = src.crs_col;
  dst.crs_used = src.crs_used;
  if(src.crs_used) {
    if(dst.crs_cache.bytes < sizeof(double) * src.crs_used) dst.crs_cache.alloc(&cd, sizeof(double) * src.crs_used);
    HIP_OK(hipMemcpyAsync(dst.crs_cache.ptr, src.crs_cache.ptr, sizeof(double) * src.crs_used, hipMemcpyDeviceToDevice,
                          c.stream));
  }
  HIP_OK(hipStreamSynchronize(c.stream));
}

// TrilinosCrsMatrix::read_native (TrilinosCrsMatrix.cpp:540-546, EpetraExt::readEpetraLinearSystem): a
// Matrix Market coordinate file (real, general or symmetric, 1-based indices) of dofs x dofs whose entries
// lie in the system's pattern; the matrix is reset to zero first (entries absent from the file are zero)
void lss_read_native(Lss& L, const char* path)
{
  if(L.builder != CF3HIP_BUILDER_CRS)
    CF3_THROW(CF3HIP_ERR_NOT_IMPLEMENTED, "read_native is not implemented for the block (VBR) matrix builder");
  FILE* f = std::fopen(path, "r");
  if(!f) CF3_THROW(CF3HIP_ERR_INVALID_ARG, std::string("cannot open ") + path);
  char line[512];
  bool header = false, symmetric = false, sized = false;
  uint64_t M = 0, Ncol = 0, nnz = 0, got = 0;
  const int neq = L.neq;
  Context& c = *L.ctx;
  std::vector<uint32_t> rp(L.n_nodes + 1), cl(L.n_blocks);
  HIP_OK(hipMemcpyAsync(rp.data(), L.row_ptr.ptr, sizeof(uint32_t) * rp.size(), hipMemcpyDeviceToHost, c.stream));
  HIP_OK(hipMemcpyAsync(cl.data(), L.cols.ptr, sizeof(uint32_t) * cl.size(), hipMemcpyDeviceToHost, c.stream));
  HIP_OK(hipStreamSynchronize(c.stream));
  const size_t nn = (size_t)neq * neq;
  std::vector<double> vals((size_t)L.n_blocks * nn, 0.);
  std::string err;
  auto put = [&](uint64_t r, uint64_t col, double v) {
    const uint32_t i = (uint32_t)(r / neq), j = (uint32_t)(col / neq);
    const uint32_t* lo = cl.data() + rp[i];
    const uint32_t* hi = cl.data() + rp[i + 1];
    const uint32_t* p = std::lower_bound(lo, hi, j);
    if(p == hi || *p != j) return false;
    vals[(size_t)(p - cl.data()) * nn + (r % neq) * neq + col % neq] = v;
    return true;
  };
  while(std::fgets(line, sizeof(line), f)) {
    if(!header) {
      std::string h(line);
      for(auto& ch : h) ch = (char)std::tolower(ch);
      if(h.rfind("%%matrixmarket", 0) != 0 || h.find("coordinate") == std::string::npos ||
         (h.find("real") == std::string::npos && h.find("integer") == std::string::npos)) {
        err = "not a Matrix Market real coordinate file";
        break;
      }
      symmetric = h.find("symmetric") != std::string::npos;
      header = true;
      continue;
    }
    if(line[0] == '%' || line[0] == '\n') continue;
    if(!sized) {
      if(std::sscanf(line, "%llu %llu %llu", (unsigned long long*)&M, (unsigned long long*)&Ncol,
                     (unsigned long long*)&nnz) != 3) { err = "bad size line"; break; }
      if(M != L.dofs || Ncol != L.dofs) { err = "matrix size differs from the system's"; break; }
      sized = true;
      continue;
    }
    unsigned long long r, col;
    double v;
    if(std::sscanf(line, "%llu %llu %lf", &r, &col, &v) != 3 || r < 1 || col < 1 || r > M || col > Ncol) {
      err = "bad entry line";
      break;
    }
    if(!put(r - 1, col - 1, v) || (symmetric && r != col && !put(col - 1, r - 1, v))) {
      std::fclose(f);
      CF3_THROW(CF3HIP_ERR_ILLEGAL_ENTRY, "Trying to access an illegal entry");
    }
    ++got;
  }
  std::fclose(f);
  if(err.empty() && (!sized || got != nnz)) err = "entry count differs from the size line";
  if(!err.empty()) CF3_THROW(CF3HIP_ERR_INVALID_ARG, err + " (" + path + ")");
  HIP_OK(hipMemcpyAsync(L.values.ptr, vals.data(), sizeof(double) * vals.size(), hipMemcpyHostToDevice, c.stream));
  HIP_OK(hipStreamSynchronize(c.stream));
  lss_clear_conditions(L);
  if(L.has_fp32) lss_round_fp32(L);
  HIP_OK(hipStreamSynchronize(c.stream));
}

void lss_dirichlet(Lss& L, uint32_t nbc, const uint32_t* nodes, const uint32_t* eqs, const double* values)
{
  if(!nbc) return;
  if(L.builder == CF3HIP_BUILDER_CRS) {
    crs_dirichlet(L, nbc, nodes, eqs, values);
    return;
  }
  Context& c = *L.ctx;
  const int neq = L.neq;
  // Host bookkeeping of the ordered list.  A dof constrained for the first time (in this call or an
  // earlier one since the last reset) gets its column eliminated with that first value; every
  // condition sets the row to the identity and the rhs / solution to its value, so the last value
  // wins there (TrilinosFEVbrMatrix::symmetric_dirichlet leaves an already zero column untouched).
  std::vector<uint64_t> fresh;              // dofs constrained for the first time by this call
  std::unordered_map<uint64_t, double> last_here;
  std::vector<uint64_t> here;               // dofs of this call in first-appearance order
  for(uint32_t b = 0; b < nbc; ++b) {
    if(nodes[b] >= L.n_nodes || eqs[b] >= (uint32_t)neq)
      CF3_THROW(CF3HIP_ERR_INVALID_ARG, "Dirichlet condition on a dof outside the system");
    const uint64_t d = (uint64_t)nodes[b] * neq + eqs[b];
    auto it = L.bc_host.find(d);
    if(it == L.bc_host.end()) {
      L.bc_host.emplace(d, std::make_pair(values[b], values[b]));
      L.bc_order.push_back(d);
      fresh.push_back(d);
    } else {
      it->second.second = values[b];
    }
    if(last_here.find(d) == last_here.end()) here.push_back(d);
    last_here[d] = values[b];
  }
  // this call: rows of its dofs, columns of its fresh dofs
  const uint32_t nh = (uint32_t)here.size(), nf = (uint32_t)fresh.size();
  std::vector<uint32_t> hdof(nh), fdof(nf);
  std::vector<double> hval(nh), fval(nf);
  for(uint32_t i = 0; i < nh; ++i) { hdof[i] = (uint32_t)here[i]; hval[i] = last_here[here[i]]; }
  std::unordered_map<uint32_t, uint8_t> fmask;
  for(uint32_t i = 0; i < nf; ++i) {
    fdof[i] = (uint32_t)fresh[i];
    fval[i] = L.bc_host[fresh[i]].first;
    fmask[fdof[i] / neq] |= (uint8_t)(1u << (fdof[i] % neq));
  }
  // persistent state of all conditions since the reset (used by the matrix-free residual)
  const uint32_t nall = (uint32_t)L.bc_order.size();
  std::vector<uint32_t> adof(nall);
  std::vector<double> alast(nall);
  std::unordered_map<uint32_t, uint8_t> amask;
  for(uint32_t i = 0; i < nall; ++i) {
    adof[i] = (uint32_t)L.bc_order[i];
    alast[i] = L.bc_host[L.bc_order[i]].second;
    amask[adof[i] / neq] |= (uint8_t)(1u << (adof[i] % neq));
  }
  if(!L.bc_mask.ptr) {
    L.bc_mask.alloc(&c, L.n_nodes);
    L.bc_newmask.alloc(&c, L.n_nodes);
    L.bc_vfirst.alloc(&c, sizeof(double) * L.dofs);
    hipLaunchKernelGGL(k_fill<uint8_t>, dim3(std::min<unsigned>(grid_for(L.n_nodes, 256), 4096)), dim3(256), 0,
                       c.stream, L.bc_mask.as<uint8_t>(), (uint64_t)L.n_nodes, (uint8_t)0);
  }
  if(L.bc_dof.bytes < sizeof(uint32_t) * nall) {
    L.bc_dof.alloc(&c, sizeof(uint32_t) * nall * 2);
    L.bc_vlast.alloc(&c, sizeof(double) * nall * 2);
  }
  L.bc_n = nall;
  std::vector<uint32_t> fnode, anode;
  std::vector<uint8_t> fmv, amv;
  for(auto& kv : fmask) { fnode.push_back(kv.first); fmv.push_back(kv.second); }
  for(auto& kv : amask) { anode.push_back(kv.first); amv.push_back(kv.second); }

  DevBuf d_hdof, d_hval, d_fdof, d_fval, d_fnode, d_fmv, d_anode, d_amv;
  d_hdof.alloc(&c, sizeof(uint32_t) * nh);
  d_hval.alloc(&c, sizeof(double) * nh);
  d_fdof.alloc(&c, sizeof(uint32_t) * nf);
  d_fval.alloc(&c, sizeof(double) * nf);
  d_fnode.alloc(&c, sizeof(uint32_t) * fnode.size());
  d_fmv.alloc(&c, fmv.size());
  d_anode.alloc(&c, sizeof(uint32_t) * anode.size());
  d_amv.alloc(&c, amv.size());
  auto up = [&](DevBuf& b, const void* src, size_t bytes) {
    if(bytes) HIP_OK(hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, c.stream));
  };
  up(d_hdof, hdof.data(), sizeof(uint32_t) * nh);
  up(d_hval, hval.data(), sizeof(double) * nh);
  up(d_fdof, fdof.data(), sizeof(uint32_t) * nf);
  up(d_fval, fval.data(), sizeof(double) * nf);
  up(d_fnode, fnode.data(), sizeof(uint32_t) * fnode.size());
  up(d_fmv, fmv.data(), fmv.size());
  up(d_anode, anode.data(), sizeof(uint32_t) * anode.size());
  up(d_amv, amv.data(), amv.size());
  up(L.bc_dof, adof.data(), sizeof(uint32_t) * nall);
  up(L.bc_vlast, alast.data(), sizeof(double) * nall);
  const unsigned gN = std::min<unsigned>(grid_for(L.n_nodes, 256), 4096);
  hipLaunchKernelGGL(k_fill<uint8_t>, dim3(gN), dim3(256), 0, c.stream, L.bc_newmask.as<uint8_t>(),
                     (uint64_t)L.n_nodes, (uint8_t)0);
  if(!fnode.empty())
    hipLaunchKernelGGL(k_scatter_mask, dim3(grid_for(fnode.size(), 256)), dim3(256), 0, c.stream,
                       (uint32_t)fnode.size(), d_fnode.as<uint32_t>(), d_fmv.as<uint8_t>(), L.bc_newmask.as<uint8_t>());
  if(!anode.empty())
    hipLaunchKernelGGL(k_scatter_mask, dim3(grid_for(anode.size(), 256)), dim3(256), 0, c.stream,
                       (uint32_t)anode.size(), d_anode.as<uint32_t>(), d_amv.as<uint8_t>(), L.bc_mask.as<uint8_t>());
  if(nf)
    hipLaunchKernelGGL(k_scatter_values, dim3(grid_for(nf, 256)), dim3(256), 0, c.stream, nf, d_fdof.as<uint32_t>(),
                       d_fval.as<double>(), L.bc_vfirst.as<double>());
  if(nf)
    hipLaunchKernelGGL(k_dirichlet_columns<double>, dim3(grid_for(L.n_nodes, 256)), dim3(256), 0, c.stream, L.n_nodes,
                       neq, L.row_ptr.as<uint32_t>(), L.cols.as<uint32_t>(), L.values.as<double>(),
                       L.bc_newmask.as<uint8_t>(), L.bc_vfirst.as<double>(), L.rhs.as<double>());
  if(nh)
    hipLaunchKernelGGL(k_dirichlet_rows<double>, dim3(grid_for(nh, 256)), dim3(256), 0, c.stream, nh, neq,
                       d_hdof.as<uint32_t>(), d_hval.as<double>(), L.row_ptr.as<uint32_t>(), L.cols.as<uint32_t>(),
                       L.values.as<double>(), L.rhs.as<double>(), L.sol.as<double>());
  HIP_OK(hipGetLastError());
  if(L.has_fp32) {
    // the FP32 values receive the same elimination (the right-hand side is the FP64 one)
    if(nf)
      hipLaunchKernelGGL(k_dirichlet_columns<float>, dim3(grid_for(L.n_nodes, 256)), dim3(256), 0, c.stream,
                         L.n_nodes, neq, L.row_ptr.as<uint32_t>(), L.cols.as<uint32_t>(), L.values32.as<float>(),
                         L.bc_newmask.as<uint8_t>(), L.bc_vfirst.as<double>(), (double*)nullptr);
    if(nh)
      hipLaunchKernelGGL(k_dirichlet_rows<float>, dim3(grid_for(nh, 256)), dim3(256), 0, c.stream, nh, neq,
                         d_hdof.as<uint32_t>(), d_hval.as<double>(), L.row_ptr.as<uint32_t>(), L.cols.as<uint32_t>(),
                         L.values32.as<float>(), (double*)nullptr, (double*)nullptr);
    HIP_OK(hipGetLastError());
  }
  HIP_OK(hipStreamSynchronize(c.stream));
}

// Matrix::tie_blockrow_pairs / System::periodicity for an ordered list of (to, from) block-row pairs
// (TrilinosFEVbrMatrix.cpp:558-590, System.cpp:259-282).  One workgroup walks the pairs in order, a
// barrier between pairs, so a row tied by one pair is seen tied by the next.  Per pair: the two rows
// must hold the same block columns (err 1: block counts differ, err 2: columns differ); row from is
// added into row to and becomes x_from - x_to = 0 (identity at column from, minus identity at column
// to); in row to, the block of column from is moved onto column to.  With rhs / sol (System): rhs_to
// += rhs_from, rhs_from = 0, and both solution entries take their mean.
template<typename T>
__global__ __launch_bounds__(256) void k_tie_pairs(uint32_t npairs, const uint32_t* __restrict__ pairs, int neq,
                                                   const uint32_t* __restrict__ row_ptr,
                                                   const uint32_t* __restrict__ cols, T* __restrict__ vals,
                                                   double* __restrict__ rhs, double* __restrict__ sol,
                                                   int* __restrict__ err)
{
  __shared__ int s_err, s_diag, s_pair;
  const int t = threadIdx.x;
  const int nn = neq * neq;
  for(uint32_t p = 0; p < npairs; ++p) {
    const uint32_t to = pairs[2 * p], from = pairs[2 * p + 1];
    const uint32_t bt = row_ptr[to], bf = row_ptr[from];
    const uint32_t nb = row_ptr[to + 1] - bt;
    if(t == 0) { s_err = 0; s_diag = -1; s_pair = -1; }
    __syncthreads();
    if(row_ptr[from + 1] - bf != nb) {
      if(t == 0) *err = 1;
      return;   // uniform: every thread takes this branch
    }
    for(uint32_t i = t; i < nb; i += blockDim.x) {
      const uint32_t c = cols[bt + i];
      if(c != cols[bf + i]) s_err = 2;
      if(c == from) s_diag = (int)i;
      if(c == to) s_pair = (int)i;
    }
    __syncthreads();
    if(s_err || s_diag < 0 || s_pair < 0) {
      if(t == 0) *err = s_err ? s_err : 2;
      return;
    }
    const int diag = s_diag, pr = s_pair;
    for(uint32_t o = t; o < nb * nn; o += blockDim.x) {
      T* vt = vals + (uint64_t)bt * nn + o;
      T* vf = vals + (uint64_t)bf * nn + o;
      *vt = *vt + *vf;
      *vf = T(0);
    }
    __syncthreads();
    if(t < neq) {
      vals[((uint64_t)bf + diag) * nn + t * neq + t] = T(1);
      vals[((uint64_t)bf + pr) * nn + t * neq + t] = T(-1);
    }
    if(t < nn) {
      T* vp = vals + ((uint64_t)bt + pr) * nn + t;
      T* vd = vals + ((uint64_t)bt + diag) * nn + t;
      *vp = *vp + *vd;
      *vd = T(0);
    }
    if(t < neq && rhs) {
      const uint64_t dt = (uint64_t)to * neq + t, df = (uint64_t)from * neq + t;
      rhs[dt] = rhs[dt] + rhs[df];
      rhs[df] = 0.;
      const double m = 0.5 * (sol[dt] + sol[df]);
      sol[df] = m;
      sol[dt] = m;
    }
    __syncthreads();
  }
}

void lss_tie(Lss& L, uint32_t npairs, const uint32_t* to, const uint32_t* from, bool system)
{
  if(!npairs) return;
  Context& c = *L.ctx;
  std::vector<uint32_t> pr(2 * (size_t)npairs);
  for(uint32_t i = 0; i < npairs; ++i) {
    if(to[i] >= L.n_nodes || from[i] >= L.n_nodes)
      CF3_THROW(CF3HIP_ERR_INVALID_ARG, "periodic pair outside the system");
    if(to[i] == from[i]) CF3_THROW(CF3HIP_ERR_INVALID_ARG, "a block row cannot be tied to itself");
    pr[2 * i] = to[i];
    pr[2 * i + 1] = from[i];
  }
  DevBuf d_pairs, d_err;
  d_pairs.alloc(&c, sizeof(uint32_t) * pr.size());
  d_err.alloc(&c, sizeof(int) * 2);
  HIP_OK(hipMemcpyAsync(d_pairs.ptr, pr.data(), sizeof(uint32_t) * pr.size(), hipMemcpyHostToDevice, c.stream));
  HIP_OK(hipMemsetAsync(d_err.ptr, 0, sizeof(int) * 2, c.stream));
  hipLaunchKernelGGL(k_tie_pairs<double>, dim3(1), dim3(256), 0, c.stream, npairs, d_pairs.as<uint32_t>(), L.neq,
                     L.row_ptr.as<uint32_t>(), L.cols.as<uint32_t>(), L.values.as<double>(),
                     system ? L.rhs.as<double>() : (double*)nullptr, system ? L.sol.as<double>() : (double*)nullptr,
                     d_err.as<int>());
  if(L.has_fp32)
    hipLaunchKernelGGL(k_tie_pairs<float>, dim3(1), dim3(256), 0, c.stream, npairs, d_pairs.as<uint32_t>(), L.neq,
                       L.row_ptr.as<uint32_t>(), L.cols.as<uint32_t>(), L.values32.as<float>(), (double*)nullptr,
                       (double*)nullptr, d_err.as<int>() + 1);
  HIP_OK(hipGetLastError());
  int err[2] = {0, 0};
  HIP_OK(hipMemcpyAsync(err, d_err.ptr, sizeof(err), hipMemcpyDeviceToHost, c.stream));
  HIP_OK(hipStreamSynchronize(c.stream));
  L.tied = true;
  if(err[0] == 1 || err[1] == 1)
    CF3_THROW(CF3HIP_ERR_INVALID_ARG, "Number of blocks do not match for the two block rows to be tied together.");
  if(err[0] || err[1])
    CF3_THROW(CF3HIP_ERR_INVALID_ARG, "Indices of the entries do not match for the two block rows to be tied together.");
}

void lss_clear_conditions(Lss& L)
{
  L.bc_host.clear();
  L.bc_order.clear();
  L.bc_n = 0;
  L.tied = false;
  L.crs_col.clear();
  L.crs_used = 0;
  if(L.bc_mask.ptr) lss_fill8(*L.ctx, L.bc_mask.as<uint8_t>(), L.n_nodes, 0);
}

void lss_fill8(Context& c, uint8_t* p, size_t n, uint8_t v)
{
  if(!n) return;
  hipLaunchKernelGGL(k_fill<uint8_t>, dim3(std::min<unsigned>(grid_for(n, 256), 4096)), dim3(256), 0, c.stream, p,
                     (uint64_t)n, v);
  HIP_OK(hipGetLastError());
}

void lss_point_csr(Lss& L, uint32_t* prow, uint32_t* pcol, double* pval)
{
  Context& c = *L.ctx;
  hipLaunchKernelGGL(k_point_csr, dim3(grid_for(L.n_nodes, 128)), dim3(128), 0, c.stream, L.n_nodes, L.neq,
                     L.row_ptr.as<uint32_t>(), L.cols.as<uint32_t>(), L.values.as<double>(), prow, pcol, pval);
  HIP_OK(hipGetLastError());
}

}  // namespace cf3hip

// ---------------------------------------------------------------------------------------------
// Matrix / Vector interface helpers (cf3/math/LSS/Matrix.hpp:98-120, Vector.hpp:146-154)
// ---------------------------------------------------------------------------------------------
namespace cf3hip {

namespace {

// diagonal of every node block: mode 0 set, 1 add (Matrix::set_diagonal / add_diagonal)
__global__ void k_diagonal(uint32_t N, int neq, const uint32_t* __restrict__ diag_blk, double* __restrict__ values,
                           const double* __restrict__ d, int mode)
{
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(t >= (uint64_t)N * neq) return;
  const uint32_t i = (uint32_t)(t / neq);
  const int a = (int)(t % neq);
  double* v = values + ((uint64_t)diag_blk[i] * neq + a) * neq + a;
  *v = mode == 0 ? d[t] : *v + d[t];
}

// column (iblockcol, ieq): out[k*neq + j] = A(k*neq + j, iblockcol*neq + ieq), then zeroed
// (TrilinosFEVbrMatrix::get_column_and_replace_to_zero, TrilinosFEVbrMatrix.cpp:452-478)
__global__ void k_column_zero(uint32_t N, int neq, const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ cols,
                              double* __restrict__ values, uint32_t col, int ieq, double* __restrict__ out)
{
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if(k >= N) return;
  uint32_t lo = row_ptr[k], hi = row_ptr[k + 1];
  while(lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if(cols[mid] < col) lo = mid + 1; else hi = mid;
  }
  const bool found = lo < row_ptr[k + 1] && cols[lo] == col;
  for(int j = 0; j < neq; ++j) {
    double v = 0.;
    if(found) {
      double* p = values + ((uint64_t)lo * neq + j) * neq + ieq;
      v = *p;
      *p = 0.;
    }
    out[(uint64_t)k * neq + j] = v;
  }
}

// y = alpha x + beta y (x may alias y)
__global__ void k_axpby(uint64_t n, double alpha, const double* x, double beta, double* y)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for(uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) y[j] = alpha * x[j] + beta * y[j];
}

}  // namespace

void lss_diagonal(Lss& L, const double* d_dev, int mode)
{
  Context& c = *L.ctx;
  hipLaunchKernelGGL(k_diagonal, dim3(grid_for(L.dofs, 256)), dim3(256), 0, c.stream, L.n_nodes, L.neq,
                     L.diag_blk.as<uint32_t>(), L.values.as<double>(), d_dev, mode);
  HIP_OK(hipGetLastError());
}

void lss_column_zero(Lss& L, uint32_t col, int ieq, double* out_dev)
{
  Context& c = *L.ctx;
  hipLaunchKernelGGL(k_column_zero, dim3(grid_for(L.n_nodes, 256)), dim3(256), 0, c.stream, L.n_nodes, L.neq,
                     L.row_ptr.as<uint32_t>(), L.cols.as<uint32_t>(), L.values.as<double>(), col, ieq, out_dev);
  HIP_OK(hipGetLastError());
}

void lss_axpby(Context& c, size_t n, double alpha, const double* x, double beta, double* y)
{
  if(!n) return;
  hipLaunchKernelGGL(k_axpby, dim3(std::min<unsigned>(grid_for(n, 256), 4096)), dim3(256), 0, c.stream, (uint64_t)n,
                     alpha, x, beta, y);
  HIP_OK(hipGetLastError());
}

}  // namespace cf3hip

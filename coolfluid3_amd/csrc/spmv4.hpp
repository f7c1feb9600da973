// Row kernels of the neq = 4 block-CSR product (16 lanes per node row), shared by the plain SpMV
// launches (lss.hip) and the Krylov kernels that fuse a dot product into the product (krylov.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cf3hip {
namespace spmv4 {

// FP64 values: lanes 0-7 take block 2k, lanes 8-15 block 2k+1; lane l streams 2 consecutive doubles
// (16 B) of block row (l & 7) >> 1, columns 2 (l & 1) .. +1.  After the call, lanes with
// l < 8 and even l hold y[row * 4 + (l >> 1)]; `a` receives that component.
__device__ __forceinline__ double row_f64(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ cols,
                                          const double* __restrict__ vals, const double* __restrict__ x, uint32_t row,
                                          int l, int& a, bool& out)
{
  const uint32_t q0 = row_ptr[row], q1 = row_ptr[row + 1];
  const int sub = l >> 3, li = l & 7;
  double acc = 0.;
  for(uint32_t q = q0 + sub; q < q1; q += 2) {
    const double2 v = *reinterpret_cast<const double2*>(vals + (uint64_t)q * 16 + 2 * li);
    const double2 xv = *reinterpret_cast<const double2*>(x + (uint64_t)cols[q] * 4 + 2 * (li & 1));
    acc += v.x * xv.x + v.y * xv.y;
  }
  acc += __shfl_xor(acc, 1, 16);
  acc += __shfl_xor(acc, 8, 16);
  a = li >> 1;
  out = l < 8 && (l & 1) == 0;
  return acc;
}

// FP32 values, FP64 vectors and accumulation: lane l streams one block row (4 floats, 16 B) of
// blocks 4k + (l >> 2), component l & 3; lanes l < 4 hold y[row * 4 + l] after the call.
__device__ __forceinline__ double row_mixed(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ cols,
                                            const float* __restrict__ vals, const double* __restrict__ x,
                                            uint32_t row, int l, int& a, bool& out)
{
  const uint32_t q0 = row_ptr[row], q1 = row_ptr[row + 1];
  const int sub = l >> 2;
  a = l & 3;
  double acc = 0.;
  for(uint32_t q = q0 + sub; q < q1; q += 4) {
    const float4 v = *reinterpret_cast<const float4*>(vals + (uint64_t)q * 16 + 4 * a);
    const double2 x01 = *reinterpret_cast<const double2*>(x + (uint64_t)cols[q] * 4);
    const double2 x23 = *reinterpret_cast<const double2*>(x + (uint64_t)cols[q] * 4 + 2);
    acc += (double)v.x * x01.x + (double)v.y * x01.y + (double)v.z * x23.x + (double)v.w * x23.y;
  }
  acc += __shfl_xor(acc, 4, 16);
  acc += __shfl_xor(acc, 8, 16);
  out = l < 4;
  return acc;
}

// FP32 values and vectors, FP32 accumulation (lane layout of row_mixed): the inner product of the
// precision form with FP32 vectors
__device__ __forceinline__ float row_f32(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ cols,
                                         const float* __restrict__ vals, const float* __restrict__ x, uint32_t row,
                                         int l, int& a, bool& out)
{
  const uint32_t q0 = row_ptr[row], q1 = row_ptr[row + 1];
  const int sub = l >> 2;
  a = l & 3;
  float acc = 0.f;
  for(uint32_t q = q0 + sub; q < q1; q += 4) {
    const float4 v = *reinterpret_cast<const float4*>(vals + (uint64_t)q * 16 + 4 * a);
    const float4 xv = *reinterpret_cast<const float4*>(x + (uint64_t)cols[q] * 4);
    acc += v.x * xv.x + v.y * xv.y + v.z * xv.z + v.w * xv.w;
  }
  acc += __shfl_xor(acc, 4, 16);
  acc += __shfl_xor(acc, 8, 16);
  out = l < 4;
  return acc;
}

template<typename VR>
__device__ __forceinline__ double row(const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ cols,
                                      const VR* __restrict__ vals, const double* __restrict__ x, uint32_t r, int l,
                                      int& a, bool& out)
{
  if constexpr(sizeof(VR) == sizeof(double)) return row_f64(row_ptr, cols, vals, x, r, l, a, out);
  else return row_mixed(row_ptr, cols, vals, x, r, l, a, out);
}

}  // namespace spmv4
}  // namespace cf3hip

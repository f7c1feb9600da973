// The relative residual ||rhs - A sol||_2 / ||rhs||_2 of a stored FP64 system without any device
// allocation: for callers that cannot allocate (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  It has no counterpart in the reference.
//
// The partial sums of the two norms and the two results live in `scratch`, a file-scope device array
// of fixed size (2 x kParts + 8 doubles, 32.8 KB per GPU, held until the process ends).  Host code
// never passes the array's name to a kernel: the library's initialisation (cf3hip_context_create) calls
// noalloc_init() with the device current, which asks hipGetSymbolAddress for that device's copy and
// keeps the address per device.  Every call overwrites all the partial sums and results it reads, so
// nothing needs clearing between calls; calls that overlap on one GPU share the array and race.
//
// One thread per scalar row accumulates its row of the block product in column order, subtracts it from
// the rhs and adds the squares to per-thread sums; workgroup sums are written to the partial arrays and a
// one-workgroup pass adds them in a fixed order.  Results are deterministic for a given system size.
#include <mutex>

#include "internal.hpp"

namespace cf3hip {

namespace {

constexpr unsigned kParts = 2048;      // workgroups of the norm pass (grid-stride over the rows)
constexpr unsigned kThreadsNA = 256;
constexpr int kMaxDevices = 64;

// partial sums: [0, kParts) ||r||^2, [kParts, 2 kParts) ||b||^2; results at 2 kParts, 2 kParts + 1
static __device__ __align__(256) double scratch[2 * kParts + 8];

std::mutex g_mutex;
double* g_scratch[kMaxDevices] = {};

template<int NEQ>
__global__ __launch_bounds__(kThreadsNA) void k_residual_parts(uint32_t n_nodes, const uint32_t* __restrict__ row_ptr,
                                                               const uint32_t* __restrict__ cols,
                                                               const double* __restrict__ vals,
                                                               const double* __restrict__ x,
                                                               const double* __restrict__ b, double* __restrict__ parts)
{
  constexpr int NN = NEQ * NEQ;
  const uint64_t n = (uint64_t)n_nodes * NEQ;
  double rr = 0., bb = 0.;
  for(uint64_t d = (uint64_t)blockIdx.x * kThreadsNA + threadIdx.x; d < n; d += (uint64_t)gridDim.x * kThreadsNA) {
    const uint32_t i = (uint32_t)(d / NEQ);
    const int al = (int)(d % NEQ);
    double s = 0.;
    for(uint32_t q = row_ptr[i]; q < row_ptr[i + 1]; ++q) {
      const double* v = vals + (uint64_t)q * NN + al * NEQ;
      const double* xc = x + (uint64_t)cols[q] * NEQ;
#pragma unroll
      for(int be = 0; be < NEQ; ++be) s += v[be] * xc[be];
    }
    const double bi = b[d];
    const double r = bi - s;
    rr += r * r;
    bb += bi * bi;
  }
  __shared__ double sh[2][kThreadsNA / 64];
  for(int off = 32; off > 0; off >>= 1) {
    rr += __shfl_down(rr, off, 64);
    bb += __shfl_down(bb, off, 64);
  }
  if((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = rr;
    sh[1][threadIdx.x >> 6] = bb;
  }
  __syncthreads();
  if(threadIdx.x == 0) {
    double a = 0., c = 0.;
    for(int w = 0; w < (int)(kThreadsNA / 64); ++w) { a += sh[0][w]; c += sh[1][w]; }
    parts[blockIdx.x] = a;
    parts[kParts + blockIdx.x] = c;
  }
}

__global__ __launch_bounds__(kThreadsNA) void k_residual_final(unsigned G, double* __restrict__ parts)
{
  __shared__ double sh[2][kThreadsNA / 64];
  double a = 0., c = 0.;
  for(unsigned i = threadIdx.x; i < G; i += kThreadsNA) { a += parts[i]; c += parts[kParts + i]; }
  for(int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    c += __shfl_down(c, off, 64);
  }
  if((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = a;
    sh[1][threadIdx.x >> 6] = c;
  }
  __syncthreads();
  if(threadIdx.x == 0) {
    double s = 0., t = 0.;
    for(int w = 0; w < (int)(kThreadsNA / 64); ++w) { s += sh[0][w]; t += sh[1][w]; }
    parts[2 * kParts] = s;
    parts[2 * kParts + 1] = t;
  }
}

}  // namespace

// Called by cf3hip_context_create with `device` current: the device address of this GPU's scratch.
void noalloc_init(int device)
{
  if(device < 0 || device >= kMaxDevices) CF3_THROW(CF3HIP_ERR_LIMIT, "device index beyond the no-allocation scratch table");
  std::lock_guard<std::mutex> lock(g_mutex);
  if(g_scratch[device]) return;
  void* p = nullptr;
  HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if(!p) CF3_THROW(CF3HIP_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
  g_scratch[device] = static_cast<double*>(p);
}

// ||rhs - A sol|| / ||rhs|| of the stored FP64 values (||rhs - A sol|| itself when the rhs is zero)
double lss_residual_noalloc(Lss& L)
{
  Context& c = *L.ctx;
  double* parts = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_mutex);
    parts = (c.device >= 0 && c.device < kMaxDevices) ? g_scratch[c.device] : nullptr;
  }
  if(!parts) CF3_THROW(CF3HIP_ERR_NOT_CREATED, "no-allocation scratch not initialised for this device");
  if(L.dofs >= 0xffffffffull) CF3_THROW(CF3HIP_ERR_LIMIT, "system has 2^32 or more dofs (32-bit row indices)");
  if(L.neq < 1 || L.neq > 4) CF3_THROW(CF3HIP_ERR_LIMIT, "no-allocation residual supports 1 to 4 equations per node");
  const unsigned G = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(kParts, (L.dofs + kThreadsNA - 1) / kThreadsNA));
  const uint32_t* rp = L.row_ptr.as<uint32_t>();
  const uint32_t* cl = L.cols.as<uint32_t>();
  const double* v = L.values.as<double>();
  const double* x = L.sol.as<double>();
  const double* b = L.rhs.as<double>();
  switch(L.neq) {
    case 1: hipLaunchKernelGGL(k_residual_parts<1>, dim3(G), dim3(kThreadsNA), 0, c.stream, L.n_nodes, rp, cl, v, x, b, parts); break;
    case 2: hipLaunchKernelGGL(k_residual_parts<2>, dim3(G), dim3(kThreadsNA), 0, c.stream, L.n_nodes, rp, cl, v, x, b, parts); break;
    case 3: hipLaunchKernelGGL(k_residual_parts<3>, dim3(G), dim3(kThreadsNA), 0, c.stream, L.n_nodes, rp, cl, v, x, b, parts); break;
    default: hipLaunchKernelGGL(k_residual_parts<4>, dim3(G), dim3(kThreadsNA), 0, c.stream, L.n_nodes, rp, cl, v, x, b, parts); break;
  }
  hipLaunchKernelGGL(k_residual_final, dim3(1), dim3(kThreadsNA), 0, c.stream, G, parts);
  HIP_OK(hipGetLastError());
  double* res = c.scalars_host;   // pinned host mirror of the context (allocated at its creation)
  HIP_OK(hipMemcpyAsync(res, parts + 2 * kParts, sizeof(double) * 2, hipMemcpyDeviceToHost, c.stream));
  HIP_OK(hipStreamSynchronize(c.stream));
  const double bn = res[1] > 0. ? std::sqrt(res[1]) : 1.;
  return std::sqrt(res[0]) / bn;
}

}  // namespace cf3hip

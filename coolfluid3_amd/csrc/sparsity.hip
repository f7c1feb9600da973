// Device construction of the LSS sparsity and of the element scatter structures.
//
// Node graph: plugins/UFEM/src/UFEM/SparsityBuilder.cpp:157-189 (per node the sorted set of all nodes
// sharing an element with it, self included).  Here: all element node pairs are encoded as
// row*N + col, radix-sorted and de-duplicated; row starts come from binary searches, so the result
// is independent of thread scheduling (bit-exact with the set-based construction).
//
// Element-to-slot map: block slot of (e, a, b) = position of conn[e][b] in row conn[e][a]
// (what TrilinosFEVbrMatrix::add_values finds by linear search, TrilinosFEVbrMatrix.cpp:330-352).
//
// Row-owner scatter structures used by the assembly kernel:
//   inc_ptr/inc     node -> incident (element, local node) pairs, sorted by element
//   contrib_ptr/... per block slot, the contributing (local incidence q, column local node b)
//                   pairs sorted by (b, q): the sorted segmented reduction order
//   wg_rows         node-row ranges handled by one assembly workgroup
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "internal.hpp"

namespace cf3hip {

namespace {


int bits_for(uint64_t v)
{
  int b = 0;
  while(b < 64 && (v >> b) != 0) ++b;
  return std::max(b, 1);
}

__global__ void k_pair_keys(const uint32_t* __restrict__ conn, uint32_t n_elems, int k, uint64_t N,
                            uint64_t* __restrict__ keys)
{
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t total = (uint64_t)n_elems * k * k;
  if(t >= total) return;
  const uint64_t e = t / (k * k);
  const int ab = (int)(t % (k * k));
  const uint32_t ra = conn[e * k + ab / k];
  const uint32_t cb = conn[e * k + ab % k];
  keys[t] = (uint64_t)ra * N + cb;
}

__device__ inline uint64_t lower_bound64(const uint64_t* a, uint64_t n, uint64_t v)
{
  uint64_t lo = 0, hi = n;
  while(lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if(a[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__device__ inline uint32_t lower_bound32(const uint32_t* a, uint32_t lo, uint32_t hi, uint32_t v)
{
  while(lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if(a[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__global__ void k_row_ptr_from_keys(const uint64_t* __restrict__ keys, uint64_t n, uint64_t N, uint64_t stride,
                                    uint32_t* __restrict__ row_ptr)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i > N) return;
  row_ptr[i] = (uint32_t)lower_bound64(keys, n, i * stride);
}

__global__ void k_cols_from_keys(const uint64_t* __restrict__ keys, uint64_t n, uint64_t N, uint32_t* __restrict__ cols)
{
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(j >= n) return;
  cols[j] = (uint32_t)(keys[j] % N);
}

__global__ void k_slots(const uint32_t* __restrict__ conn, uint32_t n_elems, int k, const uint32_t* __restrict__ row_ptr,
                        const uint32_t* __restrict__ cols, uint32_t* __restrict__ slots)
{
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t total = (uint64_t)n_elems * k * k;
  if(t >= total) return;
  const uint64_t e = t / (k * k);
  const int ab = (int)(t % (k * k));
  const uint32_t ra = conn[e * k + ab / k];
  const uint32_t cb = conn[e * k + ab % k];
  slots[t] = lower_bound32(cols, row_ptr[ra], row_ptr[ra + 1], cb);
}

__global__ void k_diag_blocks(uint32_t N, const uint32_t* __restrict__ row_ptr, const uint32_t* __restrict__ cols,
                              uint32_t* __restrict__ diag)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= N) return;
  const uint32_t q = lower_bound32(cols, row_ptr[i], row_ptr[i + 1], i);
  diag[i] = (q < row_ptr[i + 1] && cols[q] == i) ? q : 0xffffffffu;
}

__global__ void k_inc_keys(const uint32_t* __restrict__ conn, uint64_t nek, uint64_t* __restrict__ keys)
{
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(t >= nek) return;
  keys[t] = (uint64_t)conn[t] * nek + t;
}

__global__ void k_inc_from_keys(const uint64_t* __restrict__ keys, uint64_t nek, uint32_t* __restrict__ inc)
{
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(t >= nek) return;
  inc[t] = (uint32_t)(keys[t] % nek);
}

// contribution key: slot * 2^16 + b * 2^12 + q
__global__ void k_contrib_keys(const uint32_t* __restrict__ conn, const uint32_t* __restrict__ inc,
                               const uint32_t* __restrict__ inc_ptr, const uint32_t* __restrict__ slots,
                               uint64_t nek, int k, uint64_t* __restrict__ keys)
{
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(t >= nek) return;
  const uint32_t ea = inc[t];
  const uint32_t row = conn[ea];
  const uint32_t q = (uint32_t)(t - inc_ptr[row]);
  for(int b = 0; b < k; ++b) {
    const uint64_t s = slots[(uint64_t)ea * k + b];
    keys[t * k + b] = (s << 16) | ((uint64_t)b << 12) | q;
  }
}

__global__ void k_contrib_from_keys(const uint64_t* __restrict__ keys, uint64_t n, uint16_t* __restrict__ contrib)
{
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(j >= n) return;
  const uint64_t key = keys[j];
  const uint32_t b = (uint32_t)((key >> 12) & 15);
  const uint32_t q = (uint32_t)(key & 4095);
  contrib[j] = (uint16_t)((q << 4) | b);
}

template<typename KeyT>
void radix_sort(Context& c, DevBuf& in, DevBuf& out, uint64_t n, int end_bit)
{
  size_t temp_bytes = 0;
  HIP_OK(hipcub::DeviceRadixSort::SortKeys(nullptr, temp_bytes, in.as<KeyT>(), out.as<KeyT>(), (int64_t)n, 0, end_bit,
                                           c.stream));
  DevBuf temp;
  temp.alloc(&c, std::max<size_t>(temp_bytes, 16));
  HIP_OK(hipcub::DeviceRadixSort::SortKeys(temp.ptr, temp_bytes, in.as<KeyT>(), out.as<KeyT>(), (int64_t)n, 0, end_bit,
                                           c.stream));
}

}  // namespace

void build_graph_structures(Lss& L)
{
  Context& c = *L.ctx;
  L.diag_blk.alloc(&c, sizeof(uint32_t) * std::max<uint32_t>(L.n_nodes, 1));
  hipLaunchKernelGGL(k_diag_blocks, dim3(grid_for(L.n_nodes, 256)), dim3(256), 0, c.stream, L.n_nodes,
                     L.row_ptr.as<uint32_t>(), L.cols.as<uint32_t>(), L.diag_blk.as<uint32_t>());
  HIP_OK(hipGetLastError());
  std::vector<uint32_t> diag(L.n_nodes);
  HIP_OK(hipMemcpyAsync(diag.data(), L.diag_blk.ptr, sizeof(uint32_t) * L.n_nodes, hipMemcpyDeviceToHost, c.stream));
  HIP_OK(hipStreamSynchronize(c.stream));
  for(uint32_t i = 0; i < L.n_nodes; ++i)
    if(diag[i] == 0xffffffffu)
      CF3_THROW(CF3HIP_ERR_INVALID_ARG, "node graph row " + std::to_string(i) + " has no diagonal block");
}

void build_element_structures(Lss& L)
{
  Context& c = *L.ctx;
  const Mesh& M = *L.mesh;
  const uint64_t N = M.n_nodes, E = M.n_elems, k = M.k;
  const uint64_t npairs = E * k * k, nek = E * k;
  const uint32_t* conn = M.conn.as<uint32_t>();

  // ---- node graph ----
  {
    DevBuf keys, sorted;
    keys.alloc(&c, sizeof(uint64_t) * npairs);
    sorted.alloc(&c, sizeof(uint64_t) * npairs);
    hipLaunchKernelGGL(k_pair_keys, dim3(grid_for(npairs, 256)), dim3(256), 0, c.stream, conn, (uint32_t)E, (int)k, N,
                       keys.as<uint64_t>());
    HIP_OK(hipGetLastError());
    radix_sort<uint64_t>(c, keys, sorted, npairs, bits_for(N * N - 1));
    // unique
    DevBuf nsel;
    nsel.alloc(&c, sizeof(int64_t));
    size_t temp_bytes = 0;
    HIP_OK(hipcub::DeviceSelect::Unique(nullptr, temp_bytes, sorted.as<uint64_t>(), keys.as<uint64_t>(),
                                        nsel.as<int64_t>(), (int64_t)npairs, c.stream));
    DevBuf temp;
    temp.alloc(&c, std::max<size_t>(temp_bytes, 16));
    HIP_OK(hipcub::DeviceSelect::Unique(temp.ptr, temp_bytes, sorted.as<uint64_t>(), keys.as<uint64_t>(),
                                        nsel.as<int64_t>(), (int64_t)npairs, c.stream));
    int64_t nb = 0;
    HIP_OK(hipMemcpyAsync(&nb, nsel.ptr, sizeof(int64_t), hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    if(nb >= (int64_t)0xffffffffll)
      CF3_THROW(CF3HIP_ERR_LIMIT, "more than 2^32-1 node blocks");
    L.n_blocks = (uint64_t)nb;
    L.row_ptr.alloc(&c, sizeof(uint32_t) * (N + 1));
    L.cols.alloc(&c, sizeof(uint32_t) * std::max<uint64_t>(nb, 1));
    hipLaunchKernelGGL(k_row_ptr_from_keys, dim3(grid_for(N + 1, 256)), dim3(256), 0, c.stream, keys.as<uint64_t>(),
                       (uint64_t)nb, N, N, L.row_ptr.as<uint32_t>());
    hipLaunchKernelGGL(k_cols_from_keys, dim3(grid_for(nb, 256)), dim3(256), 0, c.stream, keys.as<uint64_t>(),
                       (uint64_t)nb, N, L.cols.as<uint32_t>());
    HIP_OK(hipGetLastError());
  }
  build_graph_structures(L);

  // ---- element-to-slot map ----
  L.slots.alloc(&c, sizeof(uint32_t) * npairs);
  hipLaunchKernelGGL(k_slots, dim3(grid_for(npairs, 256)), dim3(256), 0, c.stream, conn, (uint32_t)E, (int)k,
                     L.row_ptr.as<uint32_t>(), L.cols.as<uint32_t>(), L.slots.as<uint32_t>());
  HIP_OK(hipGetLastError());

  // ---- node -> (element, local node) incidences sorted by element ----
  L.inc_ptr.alloc(&c, sizeof(uint32_t) * (N + 1));
  L.inc.alloc(&c, sizeof(uint32_t) * nek);
  {
    DevBuf keys, sorted;
    keys.alloc(&c, sizeof(uint64_t) * nek);
    sorted.alloc(&c, sizeof(uint64_t) * nek);
    hipLaunchKernelGGL(k_inc_keys, dim3(grid_for(nek, 256)), dim3(256), 0, c.stream, conn, nek, keys.as<uint64_t>());
    HIP_OK(hipGetLastError());
    radix_sort<uint64_t>(c, keys, sorted, nek, bits_for(N * nek - 1));
    hipLaunchKernelGGL(k_row_ptr_from_keys, dim3(grid_for(N + 1, 256)), dim3(256), 0, c.stream,
                       sorted.as<uint64_t>(), nek, N, nek, L.inc_ptr.as<uint32_t>());
    hipLaunchKernelGGL(k_inc_from_keys, dim3(grid_for(nek, 256)), dim3(256), 0, c.stream, sorted.as<uint64_t>(), nek,
                       L.inc.as<uint32_t>());
    HIP_OK(hipGetLastError());
  }

  // ---- host-side schedule of rows into assembly workgroups (structure only) ----
  std::vector<uint32_t> inc_ptr(N + 1), row_ptr(N + 1);
  HIP_OK(hipMemcpyAsync(inc_ptr.data(), L.inc_ptr.ptr, sizeof(uint32_t) * (N + 1), hipMemcpyDeviceToHost, c.stream));
  HIP_OK(hipMemcpyAsync(row_ptr.data(), L.row_ptr.ptr, sizeof(uint32_t) * (N + 1), hipMemcpyDeviceToHost, c.stream));
  HIP_OK(hipStreamSynchronize(c.stream));
  // a workgroup holds at most kAsmThreads incidences (one per thread) and kAsmThreads CSR blocks
  // (one owner thread per block)
  std::vector<uint32_t> wg;
  wg.push_back(0);
  uint32_t r = 0;
  while(r < N) {
    uint32_t r1 = r;
    uint64_t ninc = 0, nblk = 0;
    while(r1 < N) {
      const uint64_t di = inc_ptr[r1 + 1] - inc_ptr[r1];
      const uint64_t db = row_ptr[r1 + 1] - row_ptr[r1];
      if(di > (uint64_t)kAsmThreads || db > (uint64_t)kAsmThreads)
        CF3_THROW(CF3HIP_ERR_LIMIT, "node " + std::to_string(r1) + " has " + std::to_string(di) +
                                        " incident elements and " + std::to_string(db) +
                                        " neighbours: above the assembly workgroup limit of " +
                                        std::to_string(kAsmThreads));
      if(ninc + di > (uint64_t)kAsmThreads || nblk + db > (uint64_t)kAsmThreads || r1 - r >= (uint32_t)kAsmThreads)
        break;
      ninc += di;
      nblk += db;
      ++r1;
    }
    wg.push_back(r1);
    r = r1;
  }
  L.n_wg = (uint32_t)(wg.size() - 1);
  L.wg_rows.alloc(&c, sizeof(uint32_t) * wg.size());
  HIP_OK(hipMemcpyAsync(L.wg_rows.ptr, wg.data(), sizeof(uint32_t) * wg.size(), hipMemcpyHostToDevice, c.stream));

  // ---- per-slot contribution lists sorted by (slot, b, q) ----
  L.contrib_ptr.alloc(&c, sizeof(uint32_t) * (L.n_blocks + 1));
  L.contrib.alloc(&c, sizeof(uint16_t) * npairs);
  {
    DevBuf keys, sorted;
    keys.alloc(&c, sizeof(uint64_t) * npairs);
    sorted.alloc(&c, sizeof(uint64_t) * npairs);
    hipLaunchKernelGGL(k_contrib_keys, dim3(grid_for(nek, 256)), dim3(256), 0, c.stream, conn, L.inc.as<uint32_t>(),
                       L.inc_ptr.as<uint32_t>(), L.slots.as<uint32_t>(), nek, (int)k, keys.as<uint64_t>());
    HIP_OK(hipGetLastError());
    radix_sort<uint64_t>(c, keys, sorted, npairs, bits_for(((uint64_t)L.n_blocks << 16) - 1));
    hipLaunchKernelGGL(k_row_ptr_from_keys, dim3(grid_for(L.n_blocks + 1, 256)), dim3(256), 0, c.stream,
                       sorted.as<uint64_t>(), npairs, L.n_blocks, (uint64_t)65536, L.contrib_ptr.as<uint32_t>());
    hipLaunchKernelGGL(k_contrib_from_keys, dim3(grid_for(npairs, 256)), dim3(256), 0, c.stream,
                       sorted.as<uint64_t>(), npairs, L.contrib.as<uint16_t>());
    HIP_OK(hipGetLastError());
  }
  HIP_OK(hipStreamSynchronize(c.stream));
}

}  // namespace cf3hip

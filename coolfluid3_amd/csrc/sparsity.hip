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
//   wg_rows         node-row ranges handled by one assembly workgroup
//   inc_pos/blk_beg staging maps.  Within a workgroup, the element blocks of one staging pass are
//                   written to LDS rows sorted by (CSR block slot, incidence): the block of incidence
//                   q and column local node b goes to row inc_pos[q][b]; the rows of block s in pass
//                   p are [blk_beg[s][p], blk_beg[s+1][p]).  Simplices stage one pass per column
//                   local node b (p = b); Q1 cells (hexahedra, quadrilaterals) stage all (q, b) at once.
//   blk_acc         accumulator threads of each block (first index within the workgroup, count):
//                   c threads share the rows of a block (thread h takes rows h, h+c, ...) so that no
//                   thread sums more than kAccRows rows per pass; their partial sums are added in
//                   thread order when the block is written.
//   The maps come from the per-slot contribution lists sorted by (slot, b, q): the summation order is
//   fixed, so results are deterministic and no floating-point atomics are used.
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "internal.hpp"

namespace cf3hip {

namespace {

constexpr uint32_t kAccRows = 2;   // staged rows one accumulator thread sums per pass (at most)

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

// Accumulator threads of a block: enough that none sums more than kAccRows staged rows per pass.
__global__ void k_block_acc(uint64_t n_blocks, int passes, const uint32_t* __restrict__ contrib_ptr,
                            const uint16_t* __restrict__ contrib, uint8_t* __restrict__ acc_cnt)
{
  const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(s >= n_blocks) return;
  uint32_t cnt[4] = {0, 0, 0, 0};
  for(uint32_t j = contrib_ptr[s]; j < contrib_ptr[s + 1]; ++j) ++cnt[passes == 1 ? 0 : (contrib[j] & 15u)];
  const uint32_t mx = max(max(cnt[0], cnt[1]), max(cnt[2], cnt[3]));
  acc_cnt[s] = (uint8_t)max(1u, (mx + kAccRows - 1) / kAccRows);
}

__global__ void k_row_acc(uint32_t n, const uint32_t* __restrict__ row_ptr, const uint8_t* __restrict__ acc_cnt,
                          uint32_t* __restrict__ row_acc)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  uint32_t a = 0;
  for(uint32_t s = row_ptr[i]; s < row_ptr[i + 1]; ++s) a += acc_cnt[s];
  row_acc[i] = a;
}

// One workgroup per assembly workgroup: derive the staging maps from the contribution lists of its
// blocks (contrib_ptr/contrib, sorted by (slot, b, q)) and number its accumulator threads.
__global__ __launch_bounds__(kAsmThreads) void k_stage_maps(int passes, int pos_stride,
                                                            const uint32_t* __restrict__ wg_rows,
                                                            const uint32_t* __restrict__ row_ptr,
                                                            const uint32_t* __restrict__ inc_ptr,
                                                            const uint32_t* __restrict__ contrib_ptr,
                                                            const uint16_t* __restrict__ contrib,
                                                            const uint8_t* __restrict__ acc_cnt,
                                                            uint8_t* __restrict__ inc_pos, uint8_t* __restrict__ blk_beg,
                                                            uint16_t* __restrict__ blk_acc)
{
  __shared__ uint32_t s_cnt[5][kAsmThreads];
  const uint32_t r0 = wg_rows[blockIdx.x], r1 = wg_rows[blockIdx.x + 1];
  const uint32_t blk0 = row_ptr[r0], nblk = row_ptr[r1] - blk0;
  const uint32_t t = threadIdx.x;
  uint32_t cnt[4] = {0, 0, 0, 0};
  if(t < nblk)
    for(uint32_t j = contrib_ptr[blk0 + t]; j < contrib_ptr[blk0 + t + 1]; ++j)
      ++cnt[passes == 1 ? 0 : (contrib[j] & 15u)];
  for(int p = 0; p < 4; ++p) s_cnt[p][t] = cnt[p];
  s_cnt[4][t] = t < nblk ? acc_cnt[blk0 + t] : 0;
  __syncthreads();
  if(t < 5) {   // exclusive scans over the blocks: one per pass, and the accumulator numbering
    uint32_t run = 0;
    for(uint32_t j = 0; j < nblk; ++j) {
      const uint32_t c = s_cnt[t][j];
      s_cnt[t][j] = run;
      run += c;
    }
  }
  __syncthreads();
  if(t < nblk) {
    // row of block t: row_ptr[row] <= blk0 + t < row_ptr[row + 1]; q counts the incidences of that row
    uint32_t lo = r0, hi = r1;
    while(hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if(row_ptr[mid] <= blk0 + t) lo = mid; else hi = mid;
    }
    const uint32_t ib = inc_ptr[lo];
    uint32_t beg[4];
    for(int p = 0; p < 4; ++p) {
      beg[p] = s_cnt[p][t];
      blk_beg[(uint64_t)(blk0 + t) * 4 + p] = (uint8_t)beg[p];
    }
    blk_acc[blk0 + t] = (uint16_t)(s_cnt[4][t] | ((uint32_t)acc_cnt[blk0 + t] << 8));
    for(uint32_t j = contrib_ptr[blk0 + t]; j < contrib_ptr[blk0 + t + 1]; ++j) {
      const uint32_t c = contrib[j], b = c & 15u, q = c >> 4;
      const uint32_t p = passes == 1 ? 0 : b;
      inc_pos[(uint64_t)(ib + q) * pos_stride + b] = (uint8_t)(beg[p]++);
    }
  }
}

// Node staging (simplices): the distinct nodes of the elements of one assembly workgroup, and per
// incidence the local indices of its element's nodes in that list.  One workgroup per assembly
// workgroup: the (at most kAsmThreads * 4) node ids are bitonic-sorted in LDS and de-duplicated.
// FILL = false writes the count of each workgroup; FILL = true writes the list at nodes_ptr[wg] and
// the local indices (byte m of inc_loc = local index of element node m).
template<bool FILL>
__global__ __launch_bounds__(kAsmThreads) void k_wg_nodes(int k, const uint32_t* __restrict__ wg_rows,
                                                         const uint32_t* __restrict__ inc_ptr,
                                                         const uint32_t* __restrict__ inc,
                                                         const uint32_t* __restrict__ conn,
                                                         const uint32_t* __restrict__ nodes_ptr,
                                                         uint32_t* __restrict__ out, uint32_t* __restrict__ inc_loc)
{
  constexpr uint32_t S = kAsmThreads * 4;
  __shared__ uint32_t s[S];
  __shared__ uint32_t s_u[S];
  __shared__ uint32_t s_cnt[kAsmThreads + 1];
  const uint32_t r0 = wg_rows[blockIdx.x], r1 = wg_rows[blockIdx.x + 1];
  const uint32_t i0 = inc_ptr[r0], ninc = inc_ptr[r1] - i0;
  const uint32_t t = threadIdx.x;
  uint32_t mine[4] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu};
  if(t < ninc) {
    const uint64_t e = inc[i0 + t] / (uint32_t)k;
    for(int m = 0; m < k; ++m) mine[m] = conn[e * k + m];
  }
  for(int m = 0; m < 4; ++m) s[m * kAsmThreads + t] = mine[m];
  __syncthreads();
  // bitonic sort, ascending (unused slots hold 0xffffffff and sort last)
  for(uint32_t size = 2; size <= S; size <<= 1)
    for(uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      for(uint32_t p = t; p < S / 2; p += kAsmThreads) {
        const uint32_t i = 2 * p - (p & (stride - 1)), j = i + stride;   // i has the stride bit clear
        const uint32_t a = s[i], b = s[j];
        if((a > b) == ((i & size) == 0)) { s[i] = b; s[j] = a; }
      }
      __syncthreads();
    }
  // distinct values: thread t takes positions 4t .. 4t+3
  uint32_t f = 0;
  for(uint32_t m = 0; m < 4; ++m) {
    const uint32_t i = 4 * t + m;
    if(s[i] != 0xffffffffu && (i == 0 || s[i] != s[i - 1])) f |= 1u << m;
  }
  s_cnt[t] = __popc(f);
  __syncthreads();
  if(t == 0) {
    uint32_t run = 0;
    for(uint32_t j = 0; j < kAsmThreads; ++j) {
      const uint32_t c = s_cnt[j];
      s_cnt[j] = run;
      run += c;
    }
    s_cnt[kAsmThreads] = run;
  }
  __syncthreads();
  const uint32_t total = s_cnt[kAsmThreads];
  if constexpr(!FILL) {
    if(t == 0) out[blockIdx.x] = total;
  } else {
    uint32_t o = s_cnt[t];
    const uint32_t base = nodes_ptr[blockIdx.x];
    for(uint32_t m = 0; m < 4; ++m)
      if(f & (1u << m)) {
        s_u[o] = s[4 * t + m];
        out[base + o] = s[4 * t + m];
        ++o;
      }
    __syncthreads();
    if(t < ninc) {
      uint32_t packed = 0;
      for(int m = 0; m < k; ++m) {
        uint32_t lo = 0, hi = total;   // s_u[lo] == mine[m]
        while(hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if(s_u[mid] <= mine[m]) lo = mid; else hi = mid;
        }
        packed |= lo << (8 * m);
      }
      inc_loc[i0 + t] = packed;
    }
  }
}

// out[i] = map[in[i]]
__global__ void k_map_nodes(uint64_t n, const uint32_t* __restrict__ in, const uint32_t* __restrict__ map,
                            uint32_t* __restrict__ out)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) out[i] = map[in[i]];
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
  const uint64_t N = L.n_nodes, E = M.n_elems, k = M.k;   // N block rows
  const uint64_t npairs = E * k * k, nek = E * k;
  // rows and columns of the element nodes (the mesh nodes, or their block rows under periodic links);
  // the gathers of the kernels (and the node staging lists) keep the mesh nodes of M.conn
  DevBuf dconn;
  const uint32_t* conn = M.conn.as<uint32_t>();
  if(L.periodic) {
    dconn.alloc(&c, sizeof(uint32_t) * nek);
    hipLaunchKernelGGL(k_map_nodes, dim3(grid_for(nek, 256)), dim3(256), 0, c.stream, nek, M.conn.as<uint32_t>(),
                       L.p2m.as<uint32_t>(), dconn.as<uint32_t>());
    HIP_OK(hipGetLastError());
    conn = dconn.as<uint32_t>();
  }

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

  const bool q1 = is_cell(M.etype);   // one lane per column node (and Gauss point)
  const int lanes = cell_lanes(M.etype, (int)k);
  L.passes = q1 ? 1 : (int)k;
  L.pos_stride = q1 ? lanes : 4;

  // ---- per-slot contribution lists sorted by (slot, b, q) ----
  DevBuf contrib_ptr, contrib, acc_cnt, row_acc;
  contrib_ptr.alloc(&c, sizeof(uint32_t) * (L.n_blocks + 1));
  contrib.alloc(&c, sizeof(uint16_t) * npairs);
  {
    DevBuf keys, sorted;
    keys.alloc(&c, sizeof(uint64_t) * npairs);
    sorted.alloc(&c, sizeof(uint64_t) * npairs);
    hipLaunchKernelGGL(k_contrib_keys, dim3(grid_for(nek, 256)), dim3(256), 0, c.stream, conn, L.inc.as<uint32_t>(),
                       L.inc_ptr.as<uint32_t>(), L.slots.as<uint32_t>(), nek, (int)k, keys.as<uint64_t>());
    HIP_OK(hipGetLastError());
    radix_sort<uint64_t>(c, keys, sorted, npairs, bits_for(((uint64_t)L.n_blocks << 16) - 1));
    hipLaunchKernelGGL(k_row_ptr_from_keys, dim3(grid_for(L.n_blocks + 1, 256)), dim3(256), 0, c.stream,
                       sorted.as<uint64_t>(), npairs, L.n_blocks, (uint64_t)65536, contrib_ptr.as<uint32_t>());
    hipLaunchKernelGGL(k_contrib_from_keys, dim3(grid_for(npairs, 256)), dim3(256), 0, c.stream,
                       sorted.as<uint64_t>(), npairs, contrib.as<uint16_t>());
    HIP_OK(hipGetLastError());
  }
  // accumulator threads per block and per row
  acc_cnt.alloc(&c, L.n_blocks);
  row_acc.alloc(&c, sizeof(uint32_t) * N);
  hipLaunchKernelGGL(k_block_acc, dim3(grid_for(L.n_blocks, 256)), dim3(256), 0, c.stream, L.n_blocks, L.passes,
                     contrib_ptr.as<uint32_t>(), contrib.as<uint16_t>(), acc_cnt.as<uint8_t>());
  hipLaunchKernelGGL(k_row_acc, dim3(grid_for(N, 256)), dim3(256), 0, c.stream, (uint32_t)N, L.row_ptr.as<uint32_t>(),
                     acc_cnt.as<uint8_t>(), row_acc.as<uint32_t>());
  HIP_OK(hipGetLastError());

  // ---- host-side schedule of rows into assembly workgroups (structure only) ----
  std::vector<uint32_t> inc_ptr(N + 1), racc(N);
  HIP_OK(hipMemcpyAsync(inc_ptr.data(), L.inc_ptr.ptr, sizeof(uint32_t) * (N + 1), hipMemcpyDeviceToHost, c.stream));
  HIP_OK(hipMemcpyAsync(racc.data(), row_acc.ptr, sizeof(uint32_t) * N, hipMemcpyDeviceToHost, c.stream));
  HIP_OK(hipStreamSynchronize(c.stream));
  // a workgroup holds at most inc_limit incidences (simplices: one incidence per thread; Q1 cells:
  // one (incidence, column node) pair per thread) and kAsmThreads accumulator threads
  const uint64_t inc_limit = q1 ? kAsmThreads / lanes : kAsmThreads;
  std::vector<uint32_t> wg;
  wg.push_back(0);
  uint32_t r = 0;
  while(r < N) {
    uint32_t r1 = r;
    uint64_t ninc = 0, nacc = 0;
    while(r1 < N) {
      const uint64_t di = inc_ptr[r1 + 1] - inc_ptr[r1];
      const uint64_t da = racc[r1];
      if(di > inc_limit || da > (uint64_t)kAsmThreads)
        CF3_THROW(CF3HIP_ERR_LIMIT, "node " + std::to_string(r1) + " has " + std::to_string(di) +
                                        " incident elements and needs " + std::to_string(da) +
                                        " accumulators: above the assembly workgroup limits of " +
                                        std::to_string(inc_limit) + " and " + std::to_string(kAsmThreads));
      if(ninc + di > inc_limit || nacc + da > (uint64_t)kAsmThreads || r1 - r >= (uint32_t)kAsmThreads)
        break;
      ninc += di;
      nacc += da;
      ++r1;
    }
    wg.push_back(r1);
    r = r1;
  }
  L.n_wg = (uint32_t)(wg.size() - 1);
  L.wg_rows.alloc(&c, sizeof(uint32_t) * wg.size());
  HIP_OK(hipMemcpyAsync(L.wg_rows.ptr, wg.data(), sizeof(uint32_t) * wg.size(), hipMemcpyHostToDevice, c.stream));

  // ---- staging maps ----
  L.inc_pos.alloc(&c, (size_t)nek * L.pos_stride);
  L.blk_beg.alloc(&c, (size_t)L.n_blocks * 4);
  L.blk_acc.alloc(&c, sizeof(uint16_t) * L.n_blocks);
  lss_fill8(c, L.inc_pos.as<uint8_t>(), (size_t)nek * L.pos_stride, 0);
  hipLaunchKernelGGL(k_stage_maps, dim3(L.n_wg), dim3(kAsmThreads), 0, c.stream, L.passes, L.pos_stride,
                     L.wg_rows.as<uint32_t>(), L.row_ptr.as<uint32_t>(), L.inc_ptr.as<uint32_t>(),
                     contrib_ptr.as<uint32_t>(), contrib.as<uint16_t>(), acc_cnt.as<uint8_t>(),
                     L.inc_pos.as<uint8_t>(), L.blk_beg.as<uint8_t>(), L.blk_acc.as<uint16_t>());
  HIP_OK(hipGetLastError());

  // ---- node staging (simplices) ----
  L.staged = false;
  if(!q1 && L.n_wg > 0) {
    DevBuf counts;
    counts.alloc(&c, sizeof(uint32_t) * L.n_wg);
    hipLaunchKernelGGL(k_wg_nodes<false>, dim3(L.n_wg), dim3(kAsmThreads), 0, c.stream, (int)k,
                       L.wg_rows.as<uint32_t>(), L.inc_ptr.as<uint32_t>(), L.inc.as<uint32_t>(), M.conn.as<uint32_t>(),
                       (const uint32_t*)nullptr, counts.as<uint32_t>(), (uint32_t*)nullptr);
    HIP_OK(hipGetLastError());
    std::vector<uint32_t> cnt(L.n_wg), row_ptr(N + 1);
    HIP_OK(hipMemcpyAsync(cnt.data(), counts.ptr, sizeof(uint32_t) * L.n_wg, hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipMemcpyAsync(row_ptr.data(), L.row_ptr.ptr, sizeof(uint32_t) * (N + 1), hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    const uint32_t mx = *std::max_element(cnt.begin(), cnt.end());
    if(mx <= kMaxWgNodes) {
      std::vector<uint32_t> nptr(L.n_wg + 1), head((size_t)L.n_wg * 8);
      nptr[0] = 0;
      for(uint32_t w = 0; w < L.n_wg; ++w) nptr[w + 1] = nptr[w] + cnt[w];
      for(uint32_t w = 0; w < L.n_wg; ++w) {
        const uint32_t r0 = wg[w], r1 = wg[w + 1];
        uint32_t* h = &head[(size_t)w * 8];
        h[0] = r0; h[1] = r1; h[2] = inc_ptr[r0]; h[3] = inc_ptr[r1];
        h[4] = row_ptr[r0]; h[5] = row_ptr[r1]; h[6] = nptr[w]; h[7] = nptr[w + 1];
      }
      DevBuf dptr;
      dptr.alloc(&c, sizeof(uint32_t) * (L.n_wg + 1));
      L.wg_head.alloc(&c, sizeof(uint32_t) * head.size());
      L.wg_nodes.alloc(&c, sizeof(uint32_t) * std::max<uint32_t>(nptr[L.n_wg], 1));
      L.inc_loc.alloc(&c, sizeof(uint32_t) * nek);
      HIP_OK(hipMemcpyAsync(dptr.ptr, nptr.data(), sizeof(uint32_t) * nptr.size(), hipMemcpyHostToDevice, c.stream));
      HIP_OK(hipMemcpyAsync(L.wg_head.ptr, head.data(), sizeof(uint32_t) * head.size(), hipMemcpyHostToDevice,
                            c.stream));
      hipLaunchKernelGGL(k_wg_nodes<true>, dim3(L.n_wg), dim3(kAsmThreads), 0, c.stream, (int)k,
                         L.wg_rows.as<uint32_t>(), L.inc_ptr.as<uint32_t>(), L.inc.as<uint32_t>(), M.conn.as<uint32_t>(),
                         dptr.as<uint32_t>(), L.wg_nodes.as<uint32_t>(), L.inc_loc.as<uint32_t>());
      HIP_OK(hipGetLastError());
      HIP_OK(hipStreamSynchronize(c.stream));   // dptr and the host vectors go out of scope
      L.wg_nodes_max = mx;
      L.staged = true;
    }
  }
  HIP_OK(hipStreamSynchronize(c.stream));
}

}  // namespace cf3hip

// C ABI of cf3hip (declared in include/cf3hip.h): argument checks, object lifetimes, host<->device
// copies; every entry point turns exceptions into status codes.
#include <algorithm>
#include <cstring>
#include <vector>

#include "internal.hpp"

struct cf3hip_context_s { cf3hip::Context c; };
struct cf3hip_mesh_s { cf3hip::Mesh m; };
struct cf3hip_lss_s { cf3hip::Lss l; };
struct cf3hip_fields_s { cf3hip::Fields f; };
struct cf3hip_batch_s { cf3hip::Batch b; };
struct cf3hip_si_s { cf3hip::SemiImplicit s; };

namespace cf3hip {

thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

void DevBuf::alloc(Context* c, size_t n)
{
  release();
  ctx = c;
  if(n == 0) n = 16;
  HIP_OK(hipSetDevice(c->device));
  hipError_t e = hipMalloc(&ptr, n);
  if(e != hipSuccess) {
    ptr = nullptr;
    (void)hipGetLastError();
    CF3_THROW(CF3HIP_ERR_OUT_OF_MEMORY, "device allocation of " + std::to_string(n) + " bytes failed");
  }
  bytes = n;
  c->bytes_in_use += n;
}

void timer_begin(Context& c, int k)
{
  if(!c.timing) return;
  KernelTimer& t = c.timers[k];
  if(t.used + 2 > t.ev.size()) {
    for(int i = 0; i < 64; ++i) {
      hipEvent_t e;
      HIP_OK(hipEventCreate(&e));
      t.ev.push_back(e);
    }
  }
  HIP_OK(hipEventRecord(t.ev[t.used], c.stream));
}

void timer_end(Context& c, int k)
{
  if(!c.timing) return;
  KernelTimer& t = c.timers[k];
  HIP_OK(hipEventRecord(t.ev[t.used + 1], c.stream));
  t.used += 2;
}

void DevBuf::release()
{
  if(ptr) {
    (void)hipFree(ptr);
    if(ctx) ctx->bytes_in_use -= bytes;
  }
  ptr = nullptr;
  bytes = 0;
}

}  // namespace cf3hip

using namespace cf3hip;

#define API_BEGIN try {
#define API_END                                                                                  \
  }                                                                                              \
  catch(const cf3hip::Failure& e) {                                                              \
    set_error(e.msg);                                                                            \
    return e.code;                                                                               \
  }                                                                                              \
  catch(const std::exception& e) {                                                               \
    set_error(e.what());                                                                         \
    return CF3HIP_ERR_INVALID_ARG;                                                               \
  }                                                                                              \
  return CF3HIP_OK;

#define REQUIRE(cond, code, msg) do { if(!(cond)) CF3_THROW(code, msg); } while(0)

namespace {

int etype_nodes(int et) { return et == CF3HIP_TRIAG2D ? 3 : (et == CF3HIP_HEXA3D ? 8 : (et == CF3HIP_PRISM3D ? 6 : 4)); }
int etype_dim(int et) { return et == CF3HIP_TRIAG2D || et == CF3HIP_QUAD2D ? 2 : 3; }

void h2d(Context& c, void* dst, const void* src, size_t bytes)
{
  HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c.stream));
}

void d2h(Context& c, void* dst, const void* src, size_t bytes)
{
  HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c.stream));
  HIP_OK(hipStreamSynchronize(c.stream));
}

int64_t find_block_host(Lss& L, uint32_t row, uint32_t col)
{
  // one node row of the graph (small) is copied to the host to locate a column
  Context& c = *L.ctx;
  uint32_t rp[2];
  d2h(c, rp, L.row_ptr.as<uint32_t>() + row, sizeof(rp));
  std::vector<uint32_t> cols(rp[1] - rp[0]);
  if(!cols.empty()) d2h(c, cols.data(), L.cols.as<uint32_t>() + rp[0], sizeof(uint32_t) * cols.size());
  auto it = std::lower_bound(cols.begin(), cols.end(), col);
  if(it == cols.end() || *it != col) return -1;
  return rp[0] + (it - cols.begin());
}

double* vec_ptr(Lss& L, int which)
{
  if(which == CF3HIP_RHS) return L.rhs.as<double>();
  if(which == CF3HIP_SOLUTION) return L.sol.as<double>();
  CF3_THROW(CF3HIP_ERR_INVALID_ARG, "vector selector must be CF3HIP_RHS or CF3HIP_SOLUTION");
}

void alloc_system_storage(Lss& L)
{
  Context& c = *L.ctx;
  const size_t nv = L.n_blocks * L.neq * L.neq;
  L.values.alloc(&c, sizeof(double) * nv);
  L.rhs.alloc(&c, sizeof(double) * L.dofs);
  L.sol.alloc(&c, sizeof(double) * L.dofs);
  lss_fill(c, L.values.as<double>(), nv, 0.);
  lss_fill(c, L.rhs.as<double>(), L.dofs, 0.);
  lss_fill(c, L.sol.as<double>(), L.dofs, 0.);
  HIP_OK(hipStreamSynchronize(c.stream));
}

DevBuf* field_buf(Fields& F, int which, int& ncomp)
{
  const int dim = F.mesh->dim;
  switch(which) {
    case CF3HIP_FIELD_U: ncomp = dim; return F.u;
    case CF3HIP_FIELD_P: ncomp = 1; return &F.p;
    case CF3HIP_FIELD_UADV: ncomp = dim; return F.uadv;
    case CF3HIP_FIELD_U1: ncomp = dim; return F.u1;
    case CF3HIP_FIELD_U2: ncomp = dim; return F.u2;
    case CF3HIP_FIELD_U3: ncomp = dim; return F.u3;
    case CF3HIP_FIELD_NU: ncomp = 1; return &F.nu;
    case CF3HIP_FIELD_G: ncomp = dim; return F.g;
    default: CF3_THROW(CF3HIP_ERR_INVALID_ARG, "unknown field selector");
  }
}

}  // namespace

extern "C" {

const char* cf3hip_last_error(void) { return g_last_error.c_str(); }
const char* cf3hip_version(void) { return "cf3hip 0.1 gfx950"; }

int cf3hip_context_create(int device, cf3hip_context* out)
{
  API_BEGIN
  REQUIRE(out, CF3HIP_ERR_INVALID_ARG, "null output");
  int n = 0;
  HIP_OK(hipGetDeviceCount(&n));
  REQUIRE(device >= 0 && device < n, CF3HIP_ERR_INVALID_ARG, "no such device");
  HIP_OK(hipSetDevice(device));
  noalloc_init(device);   // device address of this GPU's scratch of the no-allocation entry points
  auto* h = new cf3hip_context_s();
  h->c.device = device;
  h->c.stream = nullptr;
  try {
    h->c.red_partials.alloc(&h->c, sizeof(double) * kMaxRedBlocks * 8);
    h->c.scalars.alloc(&h->c, sizeof(double) * 64);
    HIP_OK(hipHostMalloc((void**)&h->c.scalars_host, sizeof(double) * 64, hipHostMallocDefault));
  } catch(...) {
    delete h;
    throw;
  }
  *out = h;
  API_END
}

int cf3hip_context_destroy(cf3hip_context ctx)
{
  API_BEGIN
  if(!ctx) return CF3HIP_OK;
  (void)hipStreamSynchronize(ctx->c.stream);
  for(auto& t : ctx->c.timers)
    for(hipEvent_t e : t.ev) (void)hipEventDestroy(e);
  if(ctx->c.scalars_host) (void)hipHostFree(ctx->c.scalars_host);
  hipStream_t own = ctx->c.own_stream;
  delete ctx;   // device buffers are released before the stream goes away
  if(own) (void)hipStreamDestroy(own);
  API_END
}

int cf3hip_context_set_stream(cf3hip_context ctx, void* stream)
{
  API_BEGIN
  REQUIRE(ctx, CF3HIP_ERR_INVALID_ARG, "null context");
  ctx->c.stream = static_cast<hipStream_t>(stream);
  API_END
}

int cf3hip_context_create_stream(cf3hip_context ctx)
{
  API_BEGIN
  REQUIRE(ctx, CF3HIP_ERR_INVALID_ARG, "null context");
  Context& c = ctx->c;
  if(!c.own_stream) {
    HIP_OK(hipSetDevice(c.device));
    HIP_OK(hipStreamCreateWithFlags(&c.own_stream, hipStreamNonBlocking));
  }
  c.stream = c.own_stream;
  API_END
}

int cf3hip_context_synchronize(cf3hip_context ctx)
{
  API_BEGIN
  REQUIRE(ctx, CF3HIP_ERR_INVALID_ARG, "null context");
  Context& c = ctx->c;
  HIP_OK(hipStreamSynchronize(c.stream));
  if(c.async_mode) {
    int st = 0;
    d2h(c, &st, c.status_flag(), sizeof(int));
    if(st) {
      const int zero = 0;
      h2d(c, c.status_flag(), &zero, sizeof(int));
      HIP_OK(hipStreamSynchronize(c.stream));
      CF3_THROW(CF3HIP_ERR_SINGULAR, "non-positive element Jacobian determinant");
    }
  }
  API_END
}

int cf3hip_context_set_async(cf3hip_context ctx, int enable)
{
  API_BEGIN
  REQUIRE(ctx, CF3HIP_ERR_INVALID_ARG, "null context");
  Context& c = ctx->c;
  HIP_OK(hipStreamSynchronize(c.stream));
  const int zero = 0;
  h2d(c, c.status_flag(), &zero, sizeof(int));
  HIP_OK(hipStreamSynchronize(c.stream));
  c.async_mode = enable != 0;
  API_END
}

int cf3hip_context_enable_timing(cf3hip_context ctx, int enable)
{
  API_BEGIN
  REQUIRE(ctx, CF3HIP_ERR_INVALID_ARG, "null context");
  Context& c = ctx->c;
  HIP_OK(hipStreamSynchronize(c.stream));
  for(auto& t : c.timers) t.used = 0;
  c.timing = enable != 0;
  API_END
}

int cf3hip_context_kernel_time(cf3hip_context ctx, int kernel, double* total_ms, uint64_t* launches)
{
  API_BEGIN
  REQUIRE(ctx && total_ms, CF3HIP_ERR_INVALID_ARG, "null argument");
  REQUIRE(kernel >= 0 && kernel < CF3HIP_KERNEL_COUNT, CF3HIP_ERR_INVALID_ARG, "unknown kernel id");
  Context& c = ctx->c;
  HIP_OK(hipStreamSynchronize(c.stream));
  KernelTimer& t = c.timers[kernel];
  double sum = 0.;
  for(size_t i = 0; i + 1 < t.used; i += 2) {
    float ms = 0.f;
    HIP_OK(hipEventElapsedTime(&ms, t.ev[i], t.ev[i + 1]));
    sum += ms;
  }
  *total_ms = sum;
  if(launches) *launches = t.used / 2;
  API_END
}

int cf3hip_context_memory(cf3hip_context ctx, uint64_t* bytes)
{
  API_BEGIN
  REQUIRE(ctx && bytes, CF3HIP_ERR_INVALID_ARG, "null argument");
  *bytes = ctx->c.bytes_in_use;
  API_END
}

// ---------------------------------------------------------------------------------------------
int cf3hip_mesh_create(cf3hip_context ctx, int etype, uint32_t n_nodes, uint32_t n_elems, const double* coords,
                       const uint32_t* conn, cf3hip_mesh* out)
{
  API_BEGIN
  REQUIRE(ctx && coords && conn && out, CF3HIP_ERR_INVALID_ARG, "null argument");
  REQUIRE(etype == CF3HIP_TRIAG2D || etype == CF3HIP_TETRA3D || etype == CF3HIP_HEXA3D || etype == CF3HIP_QUAD2D ||
              etype == CF3HIP_PRISM3D,
          CF3HIP_ERR_INVALID_ARG, "unknown element type");
  REQUIRE(n_nodes > 0 && n_elems > 0, CF3HIP_ERR_INVALID_ARG, "empty mesh");
  const int k = etype_nodes(etype), dim = etype_dim(etype);
  for(uint64_t i = 0; i < (uint64_t)n_elems * k; ++i)
    REQUIRE(conn[i] < n_nodes, CF3HIP_ERR_INVALID_ARG, "connectivity refers to a node outside the mesh");
  auto* h = new cf3hip_mesh_s();
  Mesh& M = h->m;
  try {
    Context& c = ctx->c;
    M.ctx = &c;
    M.etype = etype; M.dim = dim; M.k = k; M.n_nodes = n_nodes; M.n_elems = n_elems;
    std::vector<double> comp(n_nodes);
    for(int d = 0; d < dim; ++d) {
      for(uint32_t i = 0; i < n_nodes; ++i) comp[i] = coords[(uint64_t)i * dim + d];
      M.x[d].alloc(&c, sizeof(double) * n_nodes);
      h2d(c, M.x[d].ptr, comp.data(), sizeof(double) * n_nodes);
      HIP_OK(hipStreamSynchronize(c.stream));
    }
    M.conn.alloc(&c, sizeof(uint32_t) * (uint64_t)n_elems * k);
    h2d(c, M.conn.ptr, conn, sizeof(uint32_t) * (uint64_t)n_elems * k);
    mesh_pack_coords(M);
    HIP_OK(hipStreamSynchronize(c.stream));
  } catch(...) {
    delete h;
    throw;
  }
  *out = h;
  API_END
}

int cf3hip_mesh_destroy(cf3hip_mesh mesh)
{
  API_BEGIN
  delete mesh;
  API_END
}

int cf3hip_mesh_set_coordinates(cf3hip_mesh mesh, const double* coords)
{
  API_BEGIN
  REQUIRE(mesh && coords, CF3HIP_ERR_INVALID_ARG, "null argument");
  Mesh& M = mesh->m;
  std::vector<double> comp(M.n_nodes);
  for(int d = 0; d < M.dim; ++d) {
    for(uint32_t i = 0; i < M.n_nodes; ++i) comp[i] = coords[(uint64_t)i * M.dim + d];
    h2d(*M.ctx, M.x[d].ptr, comp.data(), sizeof(double) * M.n_nodes);
    HIP_OK(hipStreamSynchronize(M.ctx->stream));
  }
  mesh_pack_coords(M);
  HIP_OK(hipStreamSynchronize(M.ctx->stream));
  API_END
}

// ---------------------------------------------------------------------------------------------
int cf3hip_lss_create(cf3hip_context ctx, cf3hip_mesh mesh, int neq, cf3hip_lss* out)
{
  API_BEGIN
  REQUIRE(ctx && mesh && out, CF3HIP_ERR_INVALID_ARG, "null argument");
  REQUIRE(neq >= 1 && neq <= 4, CF3HIP_ERR_INVALID_ARG, "neq must be in [1, 4]");
  auto* h = new cf3hip_lss_s();
  Lss& L = h->l;
  try {
    L.ctx = &ctx->c;
    L.mesh = &mesh->m;
    L.neq = neq;
    L.n_nodes = mesh->m.n_nodes;
    L.dofs = (uint64_t)L.n_nodes * neq;
    build_element_structures(L);
    alloc_system_storage(L);
  } catch(...) {
    delete h;
    throw;
  }
  *out = h;
  API_END
}

int cf3hip_lss_create_periodic(cf3hip_context ctx, cf3hip_mesh mesh, int neq, const uint32_t* links,
                               const uint8_t* active, cf3hip_lss* out)
{
  API_BEGIN
  REQUIRE(ctx && mesh && links && active && out, CF3HIP_ERR_INVALID_ARG, "null argument");
  REQUIRE(neq >= 1 && neq <= 4, CF3HIP_ERR_INVALID_ARG, "neq must be in [1, 4]");
  const uint32_t N = mesh->m.n_nodes;
  // create_map_data (TrilinosDetail.cpp:147-227) on one rank: the unlinked nodes take the block rows
  // in node order; a linked node takes the row of the end of its chain of links
  std::vector<uint32_t> p2m(N);
  uint32_t rows = 0;
  for(uint32_t i = 0; i < N; ++i) {
    REQUIRE(!active[i] || links[i] < N, CF3HIP_ERR_INVALID_ARG, "periodic link to a node outside the mesh");
    if(!active[i]) p2m[i] = rows++;
  }
  REQUIRE(rows > 0, CF3HIP_ERR_INVALID_ARG, "every node is periodically linked");
  for(uint32_t i = 0; i < N; ++i) {
    if(!active[i]) continue;
    uint32_t f = links[i], hops = 0;
    while(active[f]) {
      REQUIRE(++hops <= N, CF3HIP_ERR_INVALID_ARG, "periodic links form a cycle");
      f = links[f];
    }
    p2m[i] = p2m[f];
  }
  auto* h = new cf3hip_lss_s();
  Lss& L = h->l;
  try {
    L.ctx = &ctx->c;
    L.mesh = &mesh->m;
    L.neq = neq;
    L.n_nodes = rows;
    L.dofs = (uint64_t)rows * neq;
    L.periodic = true;
    L.p2m.alloc(L.ctx, sizeof(uint32_t) * N);
    h2d(*L.ctx, L.p2m.ptr, p2m.data(), sizeof(uint32_t) * N);
    build_element_structures(L);
    alloc_system_storage(L);
  } catch(...) {
    delete h;
    throw;
  }
  *out = h;
  API_END
}

int cf3hip_lss_get_p2m(cf3hip_lss lss, uint32_t* p2m)
{
  API_BEGIN
  REQUIRE(lss && p2m, CF3HIP_ERR_INVALID_ARG, "null argument");
  const Lss& L = lss->l;
  REQUIRE(L.mesh, CF3HIP_ERR_NOT_CREATED, "system was created without a mesh");
  const uint32_t N = L.mesh->n_nodes;
  if(L.periodic) {
    HIP_OK(hipMemcpyAsync(p2m, L.p2m.ptr, sizeof(uint32_t) * N, hipMemcpyDeviceToHost, L.ctx->stream));
    HIP_OK(hipStreamSynchronize(L.ctx->stream));
  } else {
    for(uint32_t i = 0; i < N; ++i) p2m[i] = i;
  }
  API_END
}

int cf3hip_lss_create_from_graph(cf3hip_context ctx, uint32_t n_nodes, int neq, const uint32_t* start,
                                 const uint32_t* conn, cf3hip_lss* out)
{
  API_BEGIN
  REQUIRE(ctx && start && conn && out, CF3HIP_ERR_INVALID_ARG, "null argument");
  REQUIRE(neq >= 1 && neq <= 4, CF3HIP_ERR_INVALID_ARG, "neq must be in [1, 4]");
  REQUIRE(start[0] == 0, CF3HIP_ERR_INVALID_ARG, "starting_indices must start at 0");
  for(uint32_t i = 0; i < n_nodes; ++i) {
    REQUIRE(start[i + 1] >= start[i], CF3HIP_ERR_INVALID_ARG, "starting_indices must be non-decreasing");
    for(uint32_t q = start[i]; q < start[i + 1]; ++q) {
      REQUIRE(conn[q] < n_nodes, CF3HIP_ERR_INVALID_ARG, "node_connectivity refers to a node outside the graph");
      REQUIRE(q == start[i] || conn[q] > conn[q - 1], CF3HIP_ERR_INVALID_ARG, "rows must be sorted and unique");
    }
  }
  auto* h = new cf3hip_lss_s();
  Lss& L = h->l;
  try {
    Context& c = ctx->c;
    L.ctx = &c;
    L.neq = neq;
    L.n_nodes = n_nodes;
    L.dofs = (uint64_t)n_nodes * neq;
    L.n_blocks = start[n_nodes];
    L.row_ptr.alloc(&c, sizeof(uint32_t) * (n_nodes + 1));
    L.cols.alloc(&c, sizeof(uint32_t) * std::max<uint64_t>(L.n_blocks, 1));
    h2d(c, L.row_ptr.ptr, start, sizeof(uint32_t) * (n_nodes + 1));
    if(L.n_blocks) h2d(c, L.cols.ptr, conn, sizeof(uint32_t) * L.n_blocks);
    HIP_OK(hipStreamSynchronize(c.stream));
    build_graph_structures(L);
    alloc_system_storage(L);
  } catch(...) {
    delete h;
    throw;
  }
  *out = h;
  API_END
}

int cf3hip_lss_destroy(cf3hip_lss lss)
{
  API_BEGIN
  delete lss;
  API_END
}

int cf3hip_lss_info(cf3hip_lss lss, uint32_t* n_nodes, int* neq, uint64_t* n_blocks)
{
  API_BEGIN
  REQUIRE(lss, CF3HIP_ERR_NOT_CREATED, "system not created");
  if(n_nodes) *n_nodes = lss->l.n_nodes;
  if(neq) *neq = lss->l.neq;
  if(n_blocks) *n_blocks = lss->l.n_blocks;
  API_END
}

int cf3hip_lss_get_graph(cf3hip_lss lss, uint32_t* start, uint32_t* conn)
{
  API_BEGIN
  REQUIRE(lss && start && conn, CF3HIP_ERR_INVALID_ARG, "null argument");
  Lss& L = lss->l;
  d2h(*L.ctx, start, L.row_ptr.ptr, sizeof(uint32_t) * (L.n_nodes + 1));
  d2h(*L.ctx, conn, L.cols.ptr, sizeof(uint32_t) * L.n_blocks);
  API_END
}

int cf3hip_lss_get_slots(cf3hip_lss lss, uint32_t* slots)
{
  API_BEGIN
  REQUIRE(lss && slots, CF3HIP_ERR_INVALID_ARG, "null argument");
  Lss& L = lss->l;
  REQUIRE(L.mesh, CF3HIP_ERR_NOT_CREATED, "system was created from a graph: no element-to-slot map");
  d2h(*L.ctx, slots, L.slots.ptr, sizeof(uint32_t) * (uint64_t)L.mesh->n_elems * L.mesh->k * L.mesh->k);
  API_END
}

int cf3hip_lss_enable_fp32(cf3hip_lss lss)
{
  API_BEGIN
  REQUIRE(lss, CF3HIP_ERR_NOT_CREATED, "system not created");
  Lss& L = lss->l;
  if(!L.has_fp32) {
    const size_t nv = L.n_blocks * L.neq * L.neq;
    L.values32.alloc(L.ctx, sizeof(float) * nv);
    lss_fill32(*L.ctx, L.values32.as<float>(), nv, 0.f);
    HIP_OK(hipStreamSynchronize(L.ctx->stream));
    L.has_fp32 = true;
  }
  API_END
}

int cf3hip_lss_round_fp32(cf3hip_lss lss)
{
  API_BEGIN
  REQUIRE(lss, CF3HIP_ERR_NOT_CREATED, "system not created");
  lss_round_fp32(lss->l);
  HIP_OK(hipStreamSynchronize(lss->l.ctx->stream));
  API_END
}

// ---------------------------------------------------------------------------------------------
int cf3hip_matrix_reset(cf3hip_lss lss, double value)
{
  API_BEGIN
  REQUIRE(lss, CF3HIP_ERR_NOT_CREATED, "system not created");
  Lss& L = lss->l;
  lss_fill(*L.ctx, L.values.as<double>(), L.n_blocks * L.neq * L.neq, value);
  lss_clear_conditions(L);
  if(L.has_fp32) lss_fill32(*L.ctx, L.values32.as<float>(), L.n_blocks * L.neq * L.neq, (float)value);
  HIP_OK(hipStreamSynchronize(L.ctx->stream));
  API_END
}

static int entry_access(cf3hip_lss lss, uint32_t icol, uint32_t irow, double* v, int mode)
{
  API_BEGIN
  REQUIRE(lss, CF3HIP_ERR_NOT_CREATED, "system not created");
  Lss& L = lss->l;
  REQUIRE(irow < L.dofs && icol < L.dofs, CF3HIP_ERR_ILLEGAL_ENTRY, "Trying to access an illegal entry.");
  const int64_t q = find_block_host(L, irow / L.neq, icol / L.neq);
  REQUIRE(q >= 0, CF3HIP_ERR_ILLEGAL_ENTRY, "Trying to access an illegal entry.");
  double* p = L.values.as<double>() + (uint64_t)q * L.neq * L.neq + (irow % L.neq) * L.neq + (icol % L.neq);
  double cur = 0.;
  d2h(*L.ctx, &cur, p, sizeof(double));
  if(mode == 0) { *v = cur; return CF3HIP_OK; }
  const double nv = mode == 1 ? *v : cur + *v;
  h2d(*L.ctx, p, &nv, sizeof(double));
  HIP_OK(hipStreamSynchronize(L.ctx->stream));
  API_END
}

int cf3hip_matrix_set_value(cf3hip_lss lss, uint32_t icol, uint32_t irow, double v) { return entry_access(lss, icol, irow, &v, 1); }
int cf3hip_matrix_add_value(cf3hip_lss lss, uint32_t icol, uint32_t irow, double v) { return entry_access(lss, icol, irow, &v, 2); }
int cf3hip_matrix_get_value(cf3hip_lss lss, uint32_t icol, uint32_t irow, double* v)
{
  if(!v) { set_error("null argument"); return CF3HIP_ERR_INVALID_ARG; }
  return entry_access(lss, icol, irow, v, 0);
}

int cf3hip_matrix_set_row(cf3hip_lss lss, uint32_t blockrow, uint32_t ieq, double diag, double offdiag)
{
  API_BEGIN
  REQUIRE(lss, CF3HIP_ERR_NOT_CREATED, "system not created");
  Lss& L = lss->l;
  REQUIRE(blockrow < L.n_nodes && ieq < (uint32_t)L.neq, CF3HIP_ERR_ILLEGAL_ENTRY, "Trying to access an illegal entry.");
  uint32_t rp[2];
  d2h(*L.ctx, rp, L.row_ptr.as<uint32_t>() + blockrow, sizeof(rp));
  const uint32_t nb = rp[1] - rp[0];
  const int nn = L.neq * L.neq;
  std::vector<uint32_t> cols(nb);
  std::vector<double> blk((size_t)nb * nn);
  d2h(*L.ctx, cols.data(), L.cols.as<uint32_t>() + rp[0], sizeof(uint32_t) * nb);
  d2h(*L.ctx, blk.data(), L.values.as<double>() + (uint64_t)rp[0] * nn, sizeof(double) * blk.size());
  for(uint32_t s = 0; s < nb; ++s) {
    for(int j = 0; j < L.neq; ++j) blk[(size_t)s * nn + ieq * L.neq + j] = offdiag;
    if(cols[s] == blockrow) blk[(size_t)s * nn + ieq * L.neq + ieq] = diag;
  }
  h2d(*L.ctx, L.values.as<double>() + (uint64_t)rp[0] * nn, blk.data(), sizeof(double) * blk.size());
  HIP_OK(hipStreamSynchronize(L.ctx->stream));
  API_END
}

namespace {
// BlockAccumulator access to the matrix: mode 0 set, 1 add, 2 get; mat row-major (nb*neq)^2
void matrix_block_values(Lss& L, uint32_t nb, const uint32_t* indices, double* mat, int mode)
{
  const int neq = L.neq, nn = neq * neq;
  const uint32_t n = nb * neq;
  for(uint32_t a = 0; a < nb; ++a) {
    REQUIRE(indices[a] < L.n_nodes, CF3HIP_ERR_ILLEGAL_ENTRY, "Trying to access an illegal entry.");
    for(uint32_t b = 0; b < nb; ++b) {
      const int64_t q = find_block_host(L, indices[a], indices[b]);
      REQUIRE(q >= 0, CF3HIP_ERR_ILLEGAL_ENTRY, "Trying to access an illegal entry.");
      double blk[16];
      double* p = L.values.as<double>() + (uint64_t)q * nn;
      d2h(*L.ctx, blk, p, sizeof(double) * nn);
      for(int al = 0; al < neq; ++al)
        for(int be = 0; be < neq; ++be) {
          double& m = mat[(uint64_t)(a * neq + al) * n + b * neq + be];
          double& v = blk[al * neq + be];
          if(mode == 0) v = m; else if(mode == 1) v += m; else m = v;
        }
      if(mode != 2) {
        h2d(*L.ctx, p, blk, sizeof(double) * nn);
        HIP_OK(hipStreamSynchronize(L.ctx->stream));
      }
    }
  }
}

// BlockAccumulator access to a system vector: mode 0 set, 1 add, 2 get; vals node-major nb*neq
void vector_block_values(Lss& L, int which, uint32_t nb, const uint32_t* indices, double* vals, int mode)
{
  double* v = vec_ptr(L, which);
  const int neq = L.neq;
  for(uint32_t a = 0; a < nb; ++a) {
    REQUIRE(indices[a] < L.n_nodes, CF3HIP_ERR_ILLEGAL_ENTRY, "Trying to access an illegal entry.");
    double node[4];
    double* p = v + (uint64_t)indices[a] * neq;
    d2h(*L.ctx, node, p, sizeof(double) * neq);
    for(int e = 0; e < neq; ++e) {
      double& x = vals[(uint64_t)a * neq + e];
      if(mode == 0) node[e] = x; else if(mode == 1) node[e] += x; else x = node[e];
    }
    if(mode != 2) {
      h2d(*L.ctx, p, node, sizeof(double) * neq);
      HIP_OK(hipStreamSynchronize(L.ctx->stream));
    }
  }
}
}  // namespace

int cf3hip_matrix_add_values(cf3hip_lss lss, uint32_t nb, const uint32_t* indices, const double* mat)
{
  API_BEGIN
  REQUIRE(lss && indices && mat, CF3HIP_ERR_INVALID_ARG, "null argument");
  matrix_block_values(lss->l, nb, indices, const_cast<double*>(mat), 1);
  API_END
}

int cf3hip_matrix_set_values(cf3hip_lss lss, uint32_t nb, const uint32_t* indices, const double* mat)
{
  API_BEGIN
  REQUIRE(lss && indices && mat, CF3HIP_ERR_INVALID_ARG, "null argument");
  matrix_block_values(lss->l, nb, indices, const_cast<double*>(mat), 0);
  API_END
}

int cf3hip_matrix_get_block_values(cf3hip_lss lss, uint32_t nb, const uint32_t* indices, double* mat)
{
  API_BEGIN
  REQUIRE(lss && indices && mat, CF3HIP_ERR_INVALID_ARG, "null argument");
  matrix_block_values(lss->l, nb, indices, mat, 2);
  API_END
}

static int matrix_diagonal(cf3hip_lss lss, const double* diag, int mode)
{
  API_BEGIN
  REQUIRE(lss && diag, CF3HIP_ERR_INVALID_ARG, "null argument");
  Lss& L = lss->l;
  DevBuf d;
  d.alloc(L.ctx, sizeof(double) * L.dofs);
  h2d(*L.ctx, d.ptr, diag, sizeof(double) * L.dofs);
  lss_diagonal(L, d.as<double>(), mode);
  HIP_OK(hipStreamSynchronize(L.ctx->stream));
  API_END
}

int cf3hip_matrix_set_diagonal(cf3hip_lss lss, const double* diag) { return matrix_diagonal(lss, diag, 0); }
int cf3hip_matrix_add_diagonal(cf3hip_lss lss, const double* diag) { return matrix_diagonal(lss, diag, 1); }

int cf3hip_lss_set_builder(cf3hip_lss lss, int builder)
{
  API_BEGIN
  REQUIRE(lss, CF3HIP_ERR_NOT_CREATED, "system not created");
  REQUIRE(builder == CF3HIP_BUILDER_VBR || builder == CF3HIP_BUILDER_CRS, CF3HIP_ERR_INVALID_ARG, "unknown matrix builder");
  Lss& L = lss->l;
  REQUIRE(L.bc_order.empty() && L.crs_col.empty() && !L.tied, CF3HIP_ERR_INVALID_ARG,
          "choose the matrix builder before the first condition or tie");
  L.builder = builder;
  API_END
}

int cf3hip_lss_get_builder(cf3hip_lss lss, int* builder)
{
  API_BEGIN
  REQUIRE(lss && builder, CF3HIP_ERR_INVALID_ARG, "null argument");
  *builder = lss->l.builder;
  API_END
}

int cf3hip_matrix_clone_to(cf3hip_lss src, cf3hip_lss dst)
{
  API_BEGIN
  REQUIRE(src && dst, CF3HIP_ERR_NOT_CREATED, "system not created");
  REQUIRE(src != dst, CF3HIP_ERR_INVALID_ARG, "clone_to needs another system");
  lss_crs_clone(src->l, dst->l);
  API_END
}

int cf3hip_matrix_read_native(cf3hip_lss lss, const char* path)
{
  API_BEGIN
  REQUIRE(lss && path, CF3HIP_ERR_INVALID_ARG, "null argument");
  lss_read_native(lss->l, path);
  API_END
}

int cf3hip_matrix_get_column_and_replace_to_zero(cf3hip_lss lss, uint32_t iblockcol, uint32_t ieq, double* values)
{
  API_BEGIN
  REQUIRE(lss && values, CF3HIP_ERR_INVALID_ARG, "null argument");
  Lss& L = lss->l;
  REQUIRE(L.builder != CF3HIP_BUILDER_CRS, CF3HIP_ERR_NOT_IMPLEMENTED,
          "get_column_and_replace_to_zero is not implemented for the CRS matrix builder");
  REQUIRE(iblockcol < L.n_nodes && ieq < (uint32_t)L.neq, CF3HIP_ERR_ILLEGAL_ENTRY, "Trying to access an illegal entry.");
  DevBuf d;
  d.alloc(L.ctx, sizeof(double) * L.dofs);
  lss_column_zero(L, iblockcol, (int)ieq, d.as<double>());
  d2h(*L.ctx, values, d.ptr, sizeof(double) * L.dofs);
  API_END
}

int cf3hip_matrix_get_diagonal(cf3hip_lss lss, double* diag)
{
  API_BEGIN
  REQUIRE(lss && diag, CF3HIP_ERR_INVALID_ARG, "null argument");
  Lss& L = lss->l;
  std::vector<uint32_t> db(L.n_nodes);
  d2h(*L.ctx, db.data(), L.diag_blk.ptr, sizeof(uint32_t) * L.n_nodes);
  const int nn = L.neq * L.neq;
  std::vector<double> vals(L.n_blocks * nn);
  d2h(*L.ctx, vals.data(), L.values.ptr, sizeof(double) * vals.size());
  for(uint32_t i = 0; i < L.n_nodes; ++i)
    for(int a = 0; a < L.neq; ++a) diag[(uint64_t)i * L.neq + a] = vals[(uint64_t)db[i] * nn + a * L.neq + a];
  API_END
}

int cf3hip_matrix_apply(cf3hip_lss lss, double alpha, const double* x, double beta, double* y)
{
  API_BEGIN
  REQUIRE(lss && x && y, CF3HIP_ERR_INVALID_ARG, "null argument");
  lss_spmv(lss->l, alpha, x, beta, y);
  if(!lss->l.ctx->async_mode) HIP_OK(hipStreamSynchronize(lss->l.ctx->stream));
  API_END
}

int cf3hip_matrix_apply_fp32(cf3hip_lss lss, float alpha, const float* x, float beta, float* y)
{
  API_BEGIN
  REQUIRE(lss && x && y, CF3HIP_ERR_INVALID_ARG, "null argument");
  lss_spmv32(lss->l, alpha, x, beta, y);
  if(!lss->l.ctx->async_mode) HIP_OK(hipStreamSynchronize(lss->l.ctx->stream));
  API_END
}

int cf3hip_matrix_get_values(cf3hip_lss lss, double* values)
{
  API_BEGIN
  REQUIRE(lss && values, CF3HIP_ERR_INVALID_ARG, "null argument");
  Lss& L = lss->l;
  d2h(*L.ctx, values, L.values.ptr, sizeof(double) * L.n_blocks * L.neq * L.neq);
  API_END
}

int cf3hip_matrix_get_values_fp32(cf3hip_lss lss, float* values)
{
  API_BEGIN
  REQUIRE(lss && values, CF3HIP_ERR_INVALID_ARG, "null argument");
  Lss& L = lss->l;
  REQUIRE(L.has_fp32, CF3HIP_ERR_NOT_CREATED, "FP32 values not enabled");
  d2h(*L.ctx, values, L.values32.ptr, sizeof(float) * L.n_blocks * L.neq * L.neq);
  API_END
}

int cf3hip_matrix_debug_data(cf3hip_lss lss, uint32_t* row_ptr, uint32_t* col_idx, double* vals)
{
  API_BEGIN
  REQUIRE(lss && row_ptr && col_idx && vals, CF3HIP_ERR_INVALID_ARG, "null argument");
  Lss& L = lss->l;
  const uint64_t nnz = L.n_blocks * L.neq * L.neq;
  DevBuf pr, pc, pv;
  pr.alloc(L.ctx, sizeof(uint32_t) * (L.dofs + 1));
  pc.alloc(L.ctx, sizeof(uint32_t) * nnz);
  pv.alloc(L.ctx, sizeof(double) * nnz);
  lss_point_csr(L, pr.as<uint32_t>(), pc.as<uint32_t>(), pv.as<double>());
  d2h(*L.ctx, row_ptr, pr.ptr, sizeof(uint32_t) * (L.dofs + 1));
  d2h(*L.ctx, col_idx, pc.ptr, sizeof(uint32_t) * nnz);
  d2h(*L.ctx, vals, pv.ptr, sizeof(double) * nnz);
  API_END
}

// ---------------------------------------------------------------------------------------------
int cf3hip_vector_reset(cf3hip_lss lss, int which, double value)
{
  API_BEGIN
  REQUIRE(lss, CF3HIP_ERR_NOT_CREATED, "system not created");
  lss_fill(*lss->l.ctx, vec_ptr(lss->l, which), lss->l.dofs, value);
  HIP_OK(hipStreamSynchronize(lss->l.ctx->stream));
  API_END
}

int cf3hip_vector_get(cf3hip_lss lss, int which, double* host)
{
  API_BEGIN
  REQUIRE(lss && host, CF3HIP_ERR_INVALID_ARG, "null argument");
  d2h(*lss->l.ctx, host, vec_ptr(lss->l, which), sizeof(double) * lss->l.dofs);
  API_END
}

int cf3hip_vector_set(cf3hip_lss lss, int which, const double* host)
{
  API_BEGIN
  REQUIRE(lss && host, CF3HIP_ERR_INVALID_ARG, "null argument");
  h2d(*lss->l.ctx, vec_ptr(lss->l, which), host, sizeof(double) * lss->l.dofs);
  HIP_OK(hipStreamSynchronize(lss->l.ctx->stream));
  API_END
}

int cf3hip_vector_set_value(cf3hip_lss lss, int which, uint32_t blockrow, uint32_t ieq, double v)
{
  API_BEGIN
  REQUIRE(lss, CF3HIP_ERR_NOT_CREATED, "system not created");
  REQUIRE(blockrow < lss->l.n_nodes && ieq < (uint32_t)lss->l.neq, CF3HIP_ERR_ILLEGAL_ENTRY, "Trying to access an illegal entry.");
  h2d(*lss->l.ctx, vec_ptr(lss->l, which) + (uint64_t)blockrow * lss->l.neq + ieq, &v, sizeof(double));
  HIP_OK(hipStreamSynchronize(lss->l.ctx->stream));
  API_END
}

int cf3hip_vector_get_value(cf3hip_lss lss, int which, uint32_t blockrow, uint32_t ieq, double* v)
{
  API_BEGIN
  REQUIRE(lss && v, CF3HIP_ERR_INVALID_ARG, "null argument");
  REQUIRE(blockrow < lss->l.n_nodes && ieq < (uint32_t)lss->l.neq, CF3HIP_ERR_ILLEGAL_ENTRY, "Trying to access an illegal entry.");
  d2h(*lss->l.ctx, v, vec_ptr(lss->l, which) + (uint64_t)blockrow * lss->l.neq + ieq, sizeof(double));
  API_END
}

int cf3hip_vector_device_ptr(cf3hip_lss lss, int which, double** dev)
{
  API_BEGIN
  REQUIRE(lss && dev, CF3HIP_ERR_INVALID_ARG, "null argument");
  *dev = vec_ptr(lss->l, which);
  API_END
}

int cf3hip_vector_assign(cf3hip_lss lss, int to, int from)
{
  API_BEGIN
  REQUIRE(lss, CF3HIP_ERR_NOT_CREATED, "system not created");
  Lss& L = lss->l;
  lss_copy(*L.ctx, vec_ptr(L, to), vec_ptr(L, from), L.dofs);
  if(!L.ctx->async_mode) HIP_OK(hipStreamSynchronize(L.ctx->stream));
  API_END
}

int cf3hip_vector_add_value(cf3hip_lss lss, int which, uint32_t blockrow, uint32_t ieq, double v)
{
  API_BEGIN
  REQUIRE(lss, CF3HIP_ERR_NOT_CREATED, "system not created");
  REQUIRE(blockrow < lss->l.n_nodes && ieq < (uint32_t)lss->l.neq, CF3HIP_ERR_ILLEGAL_ENTRY, "Trying to access an illegal entry.");
  double* p = vec_ptr(lss->l, which) + (uint64_t)blockrow * lss->l.neq + ieq;
  double x = 0.;
  d2h(*lss->l.ctx, &x, p, sizeof(double));
  x += v;
  h2d(*lss->l.ctx, p, &x, sizeof(double));
  HIP_OK(hipStreamSynchronize(lss->l.ctx->stream));
  API_END
}

int cf3hip_vector_set_values(cf3hip_lss lss, int which, uint32_t nb, const uint32_t* indices, const double* vals)
{
  API_BEGIN
  REQUIRE(lss && indices && vals, CF3HIP_ERR_INVALID_ARG, "null argument");
  vector_block_values(lss->l, which, nb, indices, const_cast<double*>(vals), 0);
  API_END
}

int cf3hip_vector_add_values(cf3hip_lss lss, int which, uint32_t nb, const uint32_t* indices, const double* vals)
{
  API_BEGIN
  REQUIRE(lss && indices && vals, CF3HIP_ERR_INVALID_ARG, "null argument");
  vector_block_values(lss->l, which, nb, indices, const_cast<double*>(vals), 1);
  API_END
}

int cf3hip_vector_get_values(cf3hip_lss lss, int which, uint32_t nb, const uint32_t* indices, double* vals)
{
  API_BEGIN
  REQUIRE(lss && indices && vals, CF3HIP_ERR_INVALID_ARG, "null argument");
  vector_block_values(lss->l, which, nb, indices, vals, 2);
  API_END
}

int cf3hip_vector_update(cf3hip_lss lss, int to, int from, double alpha)
{
  API_BEGIN
  REQUIRE(lss, CF3HIP_ERR_NOT_CREATED, "system not created");
  Lss& L = lss->l;
  lss_axpby(*L.ctx, L.dofs, alpha, vec_ptr(L, from), 1., vec_ptr(L, to));
  if(!L.ctx->async_mode) HIP_OK(hipStreamSynchronize(L.ctx->stream));
  API_END
}

int cf3hip_vector_scale(cf3hip_lss lss, int which, double alpha)
{
  API_BEGIN
  REQUIRE(lss, CF3HIP_ERR_NOT_CREATED, "system not created");
  Lss& L = lss->l;
  double* v = vec_ptr(L, which);
  if(alpha != 1.) lss_axpby(*L.ctx, L.dofs, 0., v, alpha, v);
  if(!L.ctx->async_mode) HIP_OK(hipStreamSynchronize(L.ctx->stream));
  API_END
}

int cf3hip_lss_dirichlet(cf3hip_lss lss, uint32_t nbc, const uint32_t* nodes, const uint32_t* eqs, const double* values)
{
  API_BEGIN
  REQUIRE(lss, CF3HIP_ERR_NOT_CREATED, "system not created");
  REQUIRE(nbc == 0 || (nodes && eqs && values), CF3HIP_ERR_INVALID_ARG, "null argument");
  lss_dirichlet(lss->l, nbc, nodes, eqs, values);
  API_END
}

int cf3hip_matrix_tie_blockrow_pairs(cf3hip_lss lss, uint32_t npairs, const uint32_t* to, const uint32_t* from)
{
  API_BEGIN
  REQUIRE(lss, CF3HIP_ERR_NOT_CREATED, "system not created");
  REQUIRE(npairs == 0 || (to && from), CF3HIP_ERR_INVALID_ARG, "null argument");
  lss_tie(lss->l, npairs, to, from, false);
  API_END
}

int cf3hip_lss_periodicity(cf3hip_lss lss, uint32_t npairs, const uint32_t* to, const uint32_t* from)
{
  API_BEGIN
  REQUIRE(lss, CF3HIP_ERR_NOT_CREATED, "system not created");
  REQUIRE(npairs == 0 || (to && from), CF3HIP_ERR_INVALID_ARG, "null argument");
  lss_tie(lss->l, npairs, to, from, true);
  API_END
}

int cf3hip_lss_solve(cf3hip_lss lss, const cf3hip_solver_params* prm, cf3hip_solve_stats* stats)
{
  API_BEGIN
  REQUIRE(lss && prm, CF3HIP_ERR_INVALID_ARG, "null argument");
  REQUIRE(prm->tol > 0. && prm->max_iter > 0, CF3HIP_ERR_INVALID_ARG, "tolerance and iteration limit must be positive");
  cf3hip_solve_stats st;
  lss_solve(lss->l, *prm, st);
  if(stats) *stats = st;
  if(!st.converged) {
    set_error("Krylov solve did not reach the tolerance");
    return CF3HIP_ERR_NOT_CONVERGED;
  }
  API_END
}

int cf3hip_lss_residual(cf3hip_lss lss, double* rel)
{
  API_BEGIN
  REQUIRE(lss && rel, CF3HIP_ERR_INVALID_ARG, "null argument");
  *rel = lss_residual(lss->l);
  API_END
}

int cf3hip_lss_residual_noalloc(cf3hip_lss lss, double* rel)
{
  API_BEGIN
  REQUIRE(lss && rel, CF3HIP_ERR_INVALID_ARG, "null argument");
  *rel = lss_residual_noalloc(lss->l);
  API_END
}

// ---------------------------------------------------------------------------------------------
int cf3hip_fields_create(cf3hip_mesh mesh, cf3hip_fields* out)
{
  API_BEGIN
  REQUIRE(mesh && out, CF3HIP_ERR_INVALID_ARG, "null argument");
  auto* h = new cf3hip_fields_s();
  Fields& F = h->f;
  try {
    Mesh& M = mesh->m;
    Context& c = *M.ctx;
    F.mesh = &M;
    const size_t bytes = sizeof(double) * M.n_nodes;
    for(int d = 0; d < M.dim; ++d) {
      DevBuf* all[6] = {&F.u[d], &F.uadv[d], &F.u1[d], &F.u2[d], &F.u3[d], &F.g[d]};
      for(DevBuf* b : all) {
        b->alloc(&c, bytes);
        lss_fill(c, b->as<double>(), M.n_nodes, 0.);
      }
    }
    F.p.alloc(&c, bytes);
    F.nu.alloc(&c, bytes);
    F.st.alloc(&c, 4 * bytes);
    F.ad.alloc(&c, 4 * bytes);
    lss_fill(c, F.p.as<double>(), M.n_nodes, 0.);
    lss_fill(c, F.nu.as<double>(), M.n_nodes, 0.);
    HIP_OK(hipStreamSynchronize(c.stream));
  } catch(...) {
    delete h;
    throw;
  }
  *out = h;
  API_END
}

int cf3hip_fields_destroy(cf3hip_fields fields)
{
  API_BEGIN
  delete fields;
  API_END
}

int cf3hip_fields_set(cf3hip_fields fields, int which, const double* host)
{
  API_BEGIN
  REQUIRE(fields && host, CF3HIP_ERR_INVALID_ARG, "null argument");
  Fields& F = fields->f;
  int nc = 0;
  DevBuf* b = field_buf(F, which, nc);
  const uint32_t n = F.mesh->n_nodes;
  std::vector<double> comp(n);
  bool nonzero = false;
  for(int d = 0; d < nc; ++d) {
    for(uint32_t i = 0; i < n; ++i) {
      comp[i] = host[(uint64_t)i * nc + d];
      nonzero = nonzero || comp[i] != 0.;
    }
    h2d(*F.mesh->ctx, b[d].ptr, comp.data(), sizeof(double) * n);
    HIP_OK(hipStreamSynchronize(F.mesh->ctx->stream));
  }
  if(which == CF3HIP_FIELD_G) F.has_g = nonzero;
  API_END
}

int cf3hip_fields_get(cf3hip_fields fields, int which, double* host)
{
  API_BEGIN
  REQUIRE(fields && host, CF3HIP_ERR_INVALID_ARG, "null argument");
  Fields& F = fields->f;
  int nc = 0;
  DevBuf* b = field_buf(F, which, nc);
  const uint32_t n = F.mesh->n_nodes;
  std::vector<double> comp(n);
  for(int d = 0; d < nc; ++d) {
    d2h(*F.mesh->ctx, comp.data(), b[d].ptr, sizeof(double) * n);
    for(uint32_t i = 0; i < n; ++i) host[(uint64_t)i * nc + d] = comp[i];
  }
  API_END
}

int cf3hip_fields_initialize(cf3hip_fields fields, double nu)
{
  API_BEGIN
  REQUIRE(fields, CF3HIP_ERR_INVALID_ARG, "null argument");
  Fields& F = fields->f;
  Mesh& M = *F.mesh;
  Context& c = *M.ctx;
  const size_t bytes = sizeof(double) * M.n_nodes;
  for(int d = 0; d < M.dim; ++d) {
    HIP_OK(hipMemcpyAsync(F.uadv[d].ptr, F.u[d].ptr, bytes, hipMemcpyDeviceToDevice, c.stream));
    HIP_OK(hipMemcpyAsync(F.u1[d].ptr, F.u[d].ptr, bytes, hipMemcpyDeviceToDevice, c.stream));
    HIP_OK(hipMemcpyAsync(F.u2[d].ptr, F.u[d].ptr, bytes, hipMemcpyDeviceToDevice, c.stream));
    HIP_OK(hipMemcpyAsync(F.u3[d].ptr, F.u[d].ptr, bytes, hipMemcpyDeviceToDevice, c.stream));
  }
  lss_fill(c, F.nu.as<double>(), M.n_nodes, nu);
  HIP_OK(hipStreamSynchronize(c.stream));
  API_END
}

int cf3hip_batch_create(cf3hip_context ctx, uint32_t max_cases, cf3hip_batch* out)
{
  API_BEGIN
  REQUIRE(ctx && out, CF3HIP_ERR_INVALID_ARG, "null argument");
  REQUIRE(max_cases > 0 && max_cases <= 4096, CF3HIP_ERR_INVALID_ARG, "max_cases must be in [1, 4096]");
  auto* h = new cf3hip_batch_s();
  Batch& B = h->b;
  try {
    B.ctx = &ctx->c;
    B.max_cases = max_cases;
    HIP_OK(hipSetDevice(ctx->c.device));
    HIP_OK(hipStreamCreateWithFlags(&B.stream, hipStreamNonBlocking));
    const size_t bytes = batch_table_bytes(max_cases);
    B.table.alloc(B.ctx, bytes);
    HIP_OK(hipHostMalloc(&B.host_table, bytes, hipHostMallocDefault));
    std::memset(B.host_table, 0, bytes);
  } catch(...) {
    if(B.host_table) (void)hipHostFree(B.host_table);
    if(B.stream) (void)hipStreamDestroy(B.stream);
    delete h;
    throw;
  }
  *out = h;
  API_END
}

int cf3hip_batch_destroy(cf3hip_batch batch)
{
  API_BEGIN
  if(!batch) return CF3HIP_OK;
  Batch& B = batch->b;
  if(B.stream) (void)hipStreamSynchronize(B.stream);
  if(B.exec) (void)hipGraphExecDestroy(B.exec);
  if(B.graph) (void)hipGraphDestroy(B.graph);
  if(B.host_table) (void)hipHostFree(B.host_table);
  B.table.release();
  if(B.stream) (void)hipStreamDestroy(B.stream);
  delete batch;
  API_END
}

int cf3hip_batch_assemble(cf3hip_batch batch, uint32_t n, const cf3hip_lss* lss, const cf3hip_fields* fields,
                          const cf3hip_ns_params* params)
{
  API_BEGIN
  REQUIRE(batch && (n == 0 || (lss && fields && params)), CF3HIP_ERR_INVALID_ARG, "null argument");
  std::vector<Lss*> L(n);
  std::vector<Fields*> F(n);
  for(uint32_t i = 0; i < n; ++i) {
    REQUIRE(lss[i] && fields[i], CF3HIP_ERR_INVALID_ARG, "null system or fields");
    L[i] = &lss[i]->l;
    F[i] = &fields[i]->f;
  }
  batch_assemble(batch->b, n, L.data(), F.data(), params);
  API_END
}

int cf3hip_fields_copy(cf3hip_fields fields, int dst, int src)
{
  API_BEGIN
  REQUIRE(fields, CF3HIP_ERR_INVALID_ARG, "null argument");
  Fields& F = fields->f;
  Context& c = *F.mesh->ctx;
  int nd = 0, ns = 0;
  DevBuf* bd = field_buf(F, dst, nd);
  DevBuf* bs = field_buf(F, src, ns);
  REQUIRE(nd == ns, CF3HIP_ERR_INVALID_ARG, "fields of different shapes");
  for(int d = 0; d < nd; ++d)
    if(bd[d].ptr != bs[d].ptr) lss_copy(c, bd[d].as<double>(), bs[d].as<double>(), F.mesh->n_nodes);
  if(dst == CF3HIP_FIELD_G) F.has_g = F.has_g || src != CF3HIP_FIELD_G;
  if(!c.async_mode) HIP_OK(hipStreamSynchronize(c.stream));
  API_END
}

int cf3hip_ns_linearize(cf3hip_fields fields)
{
  API_BEGIN
  REQUIRE(fields, CF3HIP_ERR_INVALID_ARG, "null argument");
  ns_linearize(fields->f);
  HIP_OK(hipStreamSynchronize(fields->f.mesh->ctx->stream));
  API_END
}

int cf3hip_ns_assemble(cf3hip_lss lss, cf3hip_fields fields, const cf3hip_ns_params* prm)
{
  API_BEGIN
  REQUIRE(lss && fields && prm, CF3HIP_ERR_INVALID_ARG, "null argument");
  assemble(lss->l, fields->f, *prm);
  API_END
}

int cf3hip_ns_solve(cf3hip_lss lss, cf3hip_fields fields, const cf3hip_ns_params* nsp, const cf3hip_solver_params* prm,
                    cf3hip_solve_stats* stats)
{
  API_BEGIN
  REQUIRE(lss && fields && nsp && prm, CF3HIP_ERR_INVALID_ARG, "null argument");
  REQUIRE(prm->tol > 0. && prm->max_iter > 0, CF3HIP_ERR_INVALID_ARG, "tolerance and iteration limit must be positive");
  cf3hip_solve_stats st;
  ns_solve(lss->l, fields->f, *nsp, *prm, st);
  if(stats) *stats = st;
  if(!st.converged) {
    set_error("Krylov solve did not reach the tolerance");
    return CF3HIP_ERR_NOT_CONVERGED;
  }
  API_END
}

int cf3hip_ns_residual(cf3hip_lss lss, cf3hip_fields fields, const cf3hip_ns_params* nsp, const double* x, double* r)
{
  API_BEGIN
  REQUIRE(lss && fields && nsp && x && r, CF3HIP_ERR_INVALID_ARG, "null argument");
  ns_residual(lss->l, fields->f, *nsp, x, r);
  if(!lss->l.ctx->async_mode) HIP_OK(hipStreamSynchronize(lss->l.ctx->stream));
  API_END
}

int cf3hip_ns_update(cf3hip_lss lss, cf3hip_fields fields)
{
  API_BEGIN
  REQUIRE(lss && fields, CF3HIP_ERR_INVALID_ARG, "null argument");
  ns_update(lss->l, fields->f);
  HIP_OK(hipStreamSynchronize(fields->f.mesh->ctx->stream));
  API_END
}


// ---------------------------------------------------------------------------------------------
// NavierStokesSemiImplicit
// ---------------------------------------------------------------------------------------------
int cf3hip_si_create(cf3hip_lss u_lss, cf3hip_lss p_lss, cf3hip_si* out)
{
  API_BEGIN
  REQUIRE(u_lss && p_lss && out, CF3HIP_ERR_INVALID_ARG, "null argument");
  auto* h = new cf3hip_si_s();
  try {
    si_create(h->s, u_lss->l, p_lss->l);
  } catch(...) {
    delete h;
    throw;
  }
  *out = h;
  API_END
}

int cf3hip_si_destroy(cf3hip_si si)
{
  API_BEGIN
  if(si) {
    HIP_OK(hipStreamSynchronize(si->s.ctx->stream));
    delete si;
  }
  API_END
}

int cf3hip_si_initial(cf3hip_si si, cf3hip_fields fields, const cf3hip_ns_params* prm)
{
  API_BEGIN
  REQUIRE(si && fields && prm, CF3HIP_ERR_INVALID_ARG, "null argument");
  si_initial(si->s, fields->f, *prm);
  API_END
}

int cf3hip_si_velocity_assemble(cf3hip_si si, cf3hip_fields fields, const cf3hip_ns_params* prm)
{
  API_BEGIN
  REQUIRE(si && fields && prm, CF3HIP_ERR_INVALID_ARG, "null argument");
  si_velocity_assemble(si->s, fields->f, *prm);
  HIP_OK(hipStreamSynchronize(si->s.ctx->stream));
  API_END
}

int cf3hip_si_set_solution(cf3hip_si si, cf3hip_fields fields)
{
  API_BEGIN
  REQUIRE(si && fields, CF3HIP_ERR_INVALID_ARG, "null argument");
  si_set_solution(si->s, fields->f);
  HIP_OK(hipStreamSynchronize(si->s.ctx->stream));
  API_END
}

int cf3hip_si_action(cf3hip_si si, cf3hip_fields fields, const cf3hip_ns_params* prm, int which)
{
  API_BEGIN
  REQUIRE(si && fields && prm, CF3HIP_ERR_INVALID_ARG, "null argument");
  si_action(si->s, fields->f, *prm, which);
  HIP_OK(hipStreamSynchronize(si->s.ctx->stream));
  API_END
}

extern "C++" {
namespace {

DevBuf& si_vector(SemiImplicit& S, int which, size_t& n)
{
  const size_t rows = S.u->n_nodes, dim = S.u->neq;
  switch(which) {
    case CF3HIP_SI_VEC_A: n = rows * dim; return S.A;
    case CF3HIP_SI_VEC_U: n = rows * dim; return S.Uv;
    case CF3HIP_SI_VEC_P: n = rows; return S.Pv;
    case CF3HIP_SI_VEC_DPS: n = rows; return S.dps;
    case CF3HIP_SI_VEC_MINV: n = rows * dim; return S.minv;
  }
  CF3_THROW(CF3HIP_ERR_INVALID_ARG, "unknown semi-implicit vector");
}

// Dirichlet increments new - old against a field, at the rows of the given mesh nodes
void si_conditions(SemiImplicit& S, const std::vector<const double*>& field, uint32_t n, const uint32_t* nodes,
                   const double* values, std::vector<uint32_t>& rows, std::vector<uint32_t>& eqs,
                   std::vector<double>& inc)
{
  const int nc = (int)field.size();
  if(n == 0) return;
  const std::vector<uint32_t>& p2m = S.p2m_host;
  for(uint32_t b = 0; b < n; ++b)
    REQUIRE(nodes[b] < S.mesh->n_nodes, CF3HIP_ERR_INVALID_ARG, "boundary node out of range");
  // one bulk copy per field component, then the old values are read on the host (conditions in
  // node order, the components of a node consecutive)
  const size_t nn = S.mesh->n_nodes;
  std::vector<double> old(nn * nc);
  HIP_OK(hipStreamSynchronize(S.ctx->stream));
  for(int c = 0; c < nc; ++c)
    HIP_OK(hipMemcpy(old.data() + c * nn, field[c], sizeof(double) * nn, hipMemcpyDeviceToHost));
  for(uint32_t b = 0; b < n; ++b)
    for(int c = 0; c < nc; ++c) {
      rows.push_back(p2m[nodes[b]]);
      eqs.push_back((uint32_t)c);
      inc.push_back(values[(size_t)b * nc + c] - old[c * nn + nodes[b]]);
    }
}

}  // namespace
}  // extern "C++"

int cf3hip_si_set_options(cf3hip_si si, const cf3hip_si_options* options)
{
  API_BEGIN
  REQUIRE(si && options, CF3HIP_ERR_INVALID_ARG, "null argument");
  si->s.body_force = options->enable_body_force != 0;
  si->s.rcg = options->pressure_rcg_solve != 0;
  API_END
}

int cf3hip_si_get_options(cf3hip_si si, cf3hip_si_options* options)
{
  API_BEGIN
  REQUIRE(si && options, CF3HIP_ERR_INVALID_ARG, "null argument");
  options->enable_body_force = si->s.body_force ? 1 : 0;
  options->pressure_rcg_solve = si->s.rcg ? 1 : 0;
  API_END
}

int cf3hip_si_get(cf3hip_si si, int which, double* host)
{
  API_BEGIN
  REQUIRE(si && host, CF3HIP_ERR_INVALID_ARG, "null argument");
  size_t n = 0;
  DevBuf& b = si_vector(si->s, which, n);
  HIP_OK(hipStreamSynchronize(si->s.ctx->stream));
  HIP_OK(hipMemcpy(host, b.ptr, sizeof(double) * n, hipMemcpyDeviceToHost));
  API_END
}

int cf3hip_si_set(cf3hip_si si, int which, const double* host)
{
  API_BEGIN
  REQUIRE(si && host, CF3HIP_ERR_INVALID_ARG, "null argument");
  size_t n = 0;
  DevBuf& b = si_vector(si->s, which, n);
  HIP_OK(hipStreamSynchronize(si->s.ctx->stream));
  HIP_OK(hipMemcpy(b.ptr, host, sizeof(double) * n, hipMemcpyHostToDevice));
  API_END
}

int cf3hip_si_step(cf3hip_si si, cf3hip_fields fields, const cf3hip_ns_params* prm, int nb_iterations,
                   const cf3hip_solver_params* u_solver, const cf3hip_solver_params* p_solver, uint32_t n_ubc,
                   const uint32_t* ubc_nodes, const double* ubc_values, uint32_t n_pbc, const uint32_t* pbc_nodes,
                   const double* pbc_values, cf3hip_solve_stats* stats)
{
  API_BEGIN
  REQUIRE(si && fields && prm && u_solver && p_solver, CF3HIP_ERR_INVALID_ARG, "null argument");
  REQUIRE(nb_iterations > 0, CF3HIP_ERR_INVALID_ARG, "nb_iterations must be positive");
  REQUIRE((n_ubc == 0 || (ubc_nodes && ubc_values)) && (n_pbc == 0 || (pbc_nodes && pbc_values)),
          CF3HIP_ERR_INVALID_ARG, "null boundary condition arrays");
  REQUIRE(u_solver->precision == CF3HIP_FP64 && p_solver->precision == CF3HIP_FP64, CF3HIP_ERR_NOT_IMPLEMENTED,
          "the semi-implicit systems are solved in FP64");
  SemiImplicit& S = si->s;
  REQUIRE(S.initialized, CF3HIP_ERR_NOT_CREATED, "cf3hip_si_initial has not run");
  Fields& F = fields->f;
  Lss &U = *S.u, &P = *S.p;
  Context& c = *S.ctx;
  const double dt = prm->dt;
  // conditions against the fields before the step (the fields change only in Update)
  std::vector<uint32_t> urows, ueqs, prows, peqs;
  std::vector<double> uinc, pinc;
  std::vector<const double*> uf;
  for(int d = 0; d < S.mesh->dim; ++d) uf.push_back(F.u[d].as<double>());
  si_conditions(S, uf, n_ubc, ubc_nodes, ubc_values, urows, ueqs, uinc);
  si_conditions(S, {F.p.as<double>()}, n_pbc, pbc_nodes, pbc_values, prows, peqs, pinc);
  std::vector<uint32_t> udofs(urows.size()), pdofs(prows);
  for(size_t j = 0; j < urows.size(); ++j) udofs[j] = urows[j] * U.neq + ueqs[j];
  // one condition per dof: a second symmetric_dirichlet on a dof of a TrilinosCrsMatrix subtracts the
  // cached column once more (TrilinosCrsMatrix.cpp:389-396), which this path does not reproduce
  for(const std::vector<uint32_t>* d : {&udofs, &pdofs}) {
    std::vector<uint32_t> sorted(*d);
    std::sort(sorted.begin(), sorted.end());
    REQUIRE(std::adjacent_find(sorted.begin(), sorted.end()) == sorted.end(), CF3HIP_ERR_INVALID_ARG,
            "a degree of freedom appears in more than one Dirichlet condition");
  }
  ns_linearize(F);                                     // LinearizeU
  si_velocity_assemble(S, F, *prm);                    // VelocityLSS: ZeroLSS + assembly
  si_set_solution(S, F);                               // SetSolution
  // InnerLoop::execute
  lss_fill(c, S.A.as<double>(), U.dofs, 0.);
  lss_fill(c, S.dps.as<double>(), P.dofs, 0.);
  lss_copy(c, S.Uv.as<double>(), U.sol.as<double>(), U.dofs);
  lss_copy(c, S.Pv.as<double>(), P.sol.as<double>(), P.dofs);
  for(int i = 0; i < nb_iterations; ++i) {
    lss_fill(c, U.rhs.as<double>(), U.dofs, 0.);
    si_action(S, F, *prm, CF3HIP_SI_URHS);
    if(i == 0) {
      if(!urows.empty()) {   // velocity BC, written in terms of the acceleration
        lss_axpby(c, U.dofs, 0., U.rhs.as<double>(), dt, U.rhs.as<double>());
        lss_dirichlet(U, (uint32_t)urows.size(), urows.data(), ueqs.data(), uinc.data());
        lss_axpby(c, U.dofs, 0., U.rhs.as<double>(), 1. / dt, U.rhs.as<double>());
      }
    } else {
      si_zero_dofs(S, udofs, U.rhs.as<double>());
    }
    lss_fill(c, U.sol.as<double>(), U.dofs, 0.);
    cf3hip_solve_stats st{};
    lss_solve(U, *u_solver, st);
    if(stats) stats[2 * i] = st;
    REQUIRE(st.converged, CF3HIP_ERR_NOT_CONVERGED, "velocity solve did not reach the tolerance");
    lss_fill(c, P.rhs.as<double>(), P.dofs, 0.);
    si_action(S, F, *prm, CF3HIP_SI_PRHS);
    lss_fill(c, P.sol.as<double>(), P.dofs, 0.);
    if(i == 0) {
      // the matrix as assembled, then its conditions (TrilinosCrsMatrix::symmetric_dirichlet reaches the
      // same matrix and rhs through its cache of eliminated column values)
      lss_copy(c, P.values.as<double>(), S.pvals0.as<double>(), P.n_blocks);
      lss_clear_conditions(P);
      if(!prows.empty()) lss_dirichlet(P, (uint32_t)prows.size(), prows.data(), peqs.data(), pinc.data());
    } else {
      si_zero_dofs(S, pdofs, P.rhs.as<double>());
    }
    if(S.rcg) {   // FirstPressureStrategy / SecondPressureStrategy
      cf3hip_solver_params rp = *p_solver;
      if(rp.method != CF3HIP_RCG_JACOBI) {
        rp.method = CF3HIP_RCG_JACOBI;
        rp.restart = 40;
        rp.recycle = 16;
      }
      rp.rcg_space = i == 0 ? 0 : 1;
      lss_solve(P, rp, st);
    } else {
      lss_solve(P, *p_solver, st);
    }
    if(stats) stats[2 * i + 1] = st;
    REQUIRE(st.converged, CF3HIP_ERR_NOT_CONVERGED, "pressure solve did not reach the tolerance");
    lss_fill(c, U.rhs.as<double>(), U.dofs, 0.);
    si_action(S, F, *prm, CF3HIP_SI_AUP);
    si_zero_dofs(S, udofs, U.rhs.as<double>());
    si_correct(S, dt);
  }
  lss_copy(c, U.sol.as<double>(), S.Uv.as<double>(), U.dofs);
  lss_copy(c, P.sol.as<double>(), S.Pv.as<double>(), P.dofs);
  si_update(S, F);                                     // Update
  HIP_OK(hipStreamSynchronize(c.stream));
  API_END
}

}  // extern "C"

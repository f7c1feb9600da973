// Krylov solve on the device (gfx950): BiCGStab with Jacobi right preconditioning in FP64 or FP32,
// FP64 iterative refinement around the FP32 solve, residual.
//
// Reference: math::LSS::SolutionStrategy::solve / compute_residual
// (cf3/math/LSS/Trilinos/TrilinosStratimikosStrategy.cpp:135-207).  The Belos/Ifpack arithmetic is
// third-party; the method here is the textbook BiCGStab of van der Vorst (1992) with the stopping rule
// ||r|| <= tol ||b||, restated identically in the oracle.
//
// All scalars live in device memory; reductions are two-pass with a fixed grid, so every dot product
// has a fixed summation order (deterministic results, no floating-point atomics).  The convergence
// and breakdown tests run on the device: they raise a stop flag that turns every later kernel of the
// loop into a no-op, so the host enqueues batches of iterations and reads the scalars once per batch
// (the iterates are the same as with a test after every iteration).
#include <algorithm>
#include <cmath>

#include "arnoldi.hpp"
#include "internal.hpp"
#include "spmv4.hpp"

namespace cf3hip {

namespace {

constexpr int kRT = 256;   // reduction block size

inline unsigned red_grid(uint64_t n) { return std::min<unsigned>(kMaxRedBlocks, grid_for(n, kRT)); }

template<int NDOT>
__device__ inline void block_reduce_store(double (&v)[NDOT], double* partials, unsigned G)
{
  __shared__ double sh[NDOT][kRT / 64];
  for(int d = 0; d < NDOT; ++d) {
    double x = v[d];
    for(int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if((threadIdx.x & 63) == 0) sh[d][threadIdx.x >> 6] = x;
  }
  __syncthreads();
  if(threadIdx.x == 0)
    for(int d = 0; d < NDOT; ++d) {
      double s = 0.;
      for(int w = 0; w < kRT / 64; ++w) s += sh[d][w];
      partials[d * G + blockIdx.x] = s;
    }
}

// S layout (double scalars)
enum { S_RHO = 0, S_RHO_PREV, S_ALPHA, S_OMEGA, S_RHV, S_SN2, S_TS, S_TT, S_RN2, S_BN2, S_SCONV, S_THR2,
       S_TMP0, S_TMP1, S_STOP, S_ITER, S_MAXR, S_DELTA, S_MAXR_ODD, S_COUNT };

// finalize: sum partials of NDOT dots into S[dst + d], then apply a mode-specific scalar update.
// FIN_RNORM ends an iteration: it counts it and raises S_STOP on convergence or breakdown;
// FIN_RNORM_REL also raises it when the residual has dropped by S_DELTA since the last reliable
// update (running maximum S_MAXR).
enum { FIN_STORE = 0, FIN_ALPHA, FIN_SNORM, FIN_OMEGA, FIN_RNORM, FIN_RNORM_REL };

template<int NDOT>
__global__ void k_finalize(const double* __restrict__ partials, unsigned G, double* __restrict__ S, int dst, int mode)
{
  __shared__ double sh[NDOT][kRT / 64];
  double v[NDOT];
  for(int d = 0; d < NDOT; ++d) {
    double s = 0.;
    for(unsigned i = threadIdx.x; i < G; i += kRT) s += partials[d * G + i];
    for(int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if((threadIdx.x & 63) == 0) sh[d][threadIdx.x >> 6] = s;
  }
  __syncthreads();
  if(threadIdx.x != 0) return;
  if(mode != FIN_STORE && S[S_STOP] != 0.) return;
  for(int d = 0; d < NDOT; ++d) {
    double s = 0.;
    for(int w = 0; w < kRT / 64; ++w) s += sh[d][w];
    v[d] = s;
  }
  if(mode == FIN_STORE) {
    for(int d = 0; d < NDOT; ++d) S[dst + d] = v[d];
  } else if(mode == FIN_ALPHA) {          // v0 = rh . v
    S[S_RHV] = v[0];
    S[S_ALPHA] = S[S_RHO] / v[0];
  } else if(mode == FIN_SNORM) {          // v0 = s . s
    S[S_SN2] = v[0];
    S[S_SCONV] = v[0] <= S[S_THR2] ? 1. : 0.;
  } else if(mode == FIN_OMEGA) {          // v0 = t . s, v1 = t . t
    S[S_TS] = v[0];
    S[S_TT] = v[NDOT > 1 ? 1 : 0];
    S[S_OMEGA] = S[S_SCONV] != 0. ? 0. : v[0] / v[NDOT > 1 ? 1 : 0];
  } else if(mode == FIN_RNORM || mode == FIN_RNORM_REL) {   // v0 = r . r, v1 = rh . r
    const double rn2 = v[0], rho = v[NDOT > 1 ? 1 : 0];
    S[S_RN2] = rn2;
    S[S_RHO_PREV] = S[S_RHO];
    S[S_RHO] = rho;
    S[S_ITER] += 1.;
    const bool conv = S[S_SCONV] != 0. || rn2 <= S[S_THR2];
    if(mode == FIN_RNORM) {
      if(conv || !isfinite(rn2) || rho == 0. || S[S_OMEGA] == 0.) S[S_STOP] = 1.;
    } else {
      const double rn = sqrt(rn2);
      if(conv || !isfinite(rn2) || rho == 0. || rn < S[S_DELTA] * S[S_MAXR]) S[S_STOP] = 1.;
      else S[S_MAXR] = fmax(S[S_MAXR], rn);
    }
  }
}

// ---- fused iteration (FP64 vectors, neq = 1 to 4): five launches, no one-workgroup finalisations ----
// Each kernel that needs a scalar reduces the partial sums of the kernel that produced them itself, in
// every workgroup, in the fixed order of k_finalize (so all workgroups obtain the same bits); workgroup
// 0 records the scalars in S for the host and for later kernels.  A kernel never reads a slot of S
// that a workgroup of the same launch writes, except the stop flag, whose value every workgroup also
// derives itself.  rho and the running maximum of the residual norm alternate between two slots by
// iteration parity.
enum { S_RHO_EVEN = S_RHO, S_RHO_ODD = S_RHO_PREV, S_MAXR_EVEN = S_MAXR };

// sum of G partials per dot product (partials[d * G + i]), result in every thread of the workgroup
template<int NDOT>
__device__ inline void workgroup_sum(const double* __restrict__ partials, unsigned G, double (&out)[NDOT])
{
  __shared__ double sh[NDOT][kRT / 64];
  for(int d = 0; d < NDOT; ++d) {
    double s = 0.;
    for(unsigned i = threadIdx.x; i < G; i += kRT) s += partials[d * G + i];
    for(int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if((threadIdx.x & 63) == 0) sh[d][threadIdx.x >> 6] = s;
  }
  __syncthreads();
  for(int d = 0; d < NDOT; ++d) {
    double s = 0.;
    for(int w = 0; w < kRT / 64; ++w) s += sh[d][w];
    out[d] = s;
  }
}

__device__ inline bool leader() { return blockIdx.x == 0 && threadIdx.x == 0; }

// Start of iteration `it`: finalise iteration it-1 from the partials px = (||r||^2, rh . r) -- count,
// convergence / breakdown test (rel: also the residual-replacement test of the reliable phase) -- then,
// unless stopped or fin_only, p = r + beta (p - omega v), ph = D^-1 p.
template<typename R>
__global__ void k_fb_p(uint64_t n, double* __restrict__ S, int it, int rel, int fin_only, unsigned Gx,
                       const double* __restrict__ px, const R* __restrict__ r, R* __restrict__ p,
                       const R* __restrict__ v, R* __restrict__ ph, const R* __restrict__ dinv)
{
  // the leader of this launch may raise the stop flag while other workgroups start: read it once per
  // workgroup, so a workgroup returns or proceeds as a whole (workgroup_sum needs every wave)
  __shared__ int s_stop;
  if(threadIdx.x == 0) s_stop = S[S_STOP] != 0. ? 1 : 0;
  __syncthreads();
  if(s_stop) return;
  double red[2];
  workgroup_sum<2>(px, Gx, red);
  const double rn2 = red[0], rho = red[1];
  const bool conv = S[S_SCONV] != 0. || rn2 <= S[S_THR2];
  bool stop = conv || !isfinite(rn2) || rho == 0. || S[S_OMEGA] == 0.;
  const double maxr = S[(it & 1) ? S_MAXR_EVEN : S_MAXR_ODD];   // after iteration it - 1
  if(rel) stop = stop || sqrt(rn2) < S[S_DELTA] * maxr;
  if(leader()) {
    S[S_RN2] = rn2;
    S[(it & 1) ? S_RHO_ODD : S_RHO_EVEN] = rho;
    S[S_ITER] = (double)it;
    if(stop) S[S_STOP] = 1.;
    S[(it & 1) ? S_MAXR_ODD : S_MAXR_EVEN] = (rel && !stop) ? fmax(maxr, sqrt(rn2)) : maxr;
  }
  if(stop || fin_only) return;
  const double rho_old = S[(it & 1) ? S_RHO_EVEN : S_RHO_ODD];
  const R beta = (R)((rho / rho_old) * (S[S_ALPHA] / S[S_OMEGA]));
  const R omega = (R)S[S_OMEGA];
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for(uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const R pi = r[i] + beta * (p[i] - omega * v[i]);
    p[i] = pi;
    ph[i] = dinv[i] * pi;
  }
}

// y = A x (neq = 4, values VR, vectors XR: FP64, or FP32 with FP32 values) with the partial dot
// products of the BiCGStab step, summed in FP64: NDOT = 1: d0 . y (d0 = rh); NDOT = 2: d0 . y and
// y . y (d0 = s).  Fixed grid, rows in a grid-stride loop (16 lanes per node row).
template<typename VR, int NDOT, typename XR = double>
__global__ __launch_bounds__(kRT) void k_spmv4_dot(uint32_t N, const uint32_t* __restrict__ row_ptr,
                                                   const uint32_t* __restrict__ cols, const VR* __restrict__ vals,
                                                   const XR* __restrict__ x, XR* __restrict__ y,
                                                   const XR* __restrict__ d0, const double* __restrict__ S,
                                                   double* __restrict__ partials)
{
  if(S[S_STOP] != 0.) return;
  double acc[NDOT];
  for(int d = 0; d < NDOT; ++d) acc[d] = 0.;
  const int l = threadIdx.x & 15;
  const uint32_t step = gridDim.x * (kRT / 16);
  for(uint32_t row = blockIdx.x * (kRT / 16) + (threadIdx.x >> 4); row < N; row += step) {
    int a;
    bool out;
    XR yv;
    if constexpr(sizeof(XR) == sizeof(float)) yv = spmv4::row_f32(row_ptr, cols, vals, x, row, l, a, out);
    else yv = spmv4::row<VR>(row_ptr, cols, vals, x, row, l, a, out);
    if(out) {
      const uint64_t o = (uint64_t)row * 4 + a;
      y[o] = yv;
      acc[0] += (double)d0[o] * (double)yv;
      if constexpr(NDOT > 1) acc[NDOT - 1] += (double)yv * (double)yv;
    }
  }
  block_reduce_store<NDOT>(acc, partials, gridDim.x);
}

// The same for neq = 1, 2 or 3: LANES lanes per node row (8 for scalar rows, 16 otherwise) stride over
// its blocks, each lane keeps the NEQ row sums of its blocks, and the partial sums are combined in a
// fixed shuffle order.
template<typename VR, int NDOT, int NEQ>
__global__ __launch_bounds__(kRT) void k_spmvn_dot(uint32_t N, const uint32_t* __restrict__ row_ptr,
                                                   const uint32_t* __restrict__ cols, const VR* __restrict__ vals,
                                                   const double* __restrict__ x, double* __restrict__ y,
                                                   const double* __restrict__ d0, const double* __restrict__ S,
                                                   double* __restrict__ partials)
{
  constexpr int LANES = NEQ == 1 ? 8 : 16;
  if(S[S_STOP] != 0.) return;
  double acc[NDOT];
  for(int d = 0; d < NDOT; ++d) acc[d] = 0.;
  const int l = threadIdx.x & (LANES - 1);
  const uint32_t step = gridDim.x * (kRT / LANES);
  for(uint32_t row = blockIdx.x * (kRT / LANES) + (threadIdx.x / LANES); row < N; row += step) {
    double ys[NEQ];
    for(int i = 0; i < NEQ; ++i) ys[i] = 0.;
    for(uint32_t q = row_ptr[row] + l; q < row_ptr[row + 1]; q += LANES) {
      const VR* b = vals + (uint64_t)q * NEQ * NEQ;
      const double* xc = x + (uint64_t)cols[q] * NEQ;
      for(int i = 0; i < NEQ; ++i)
        for(int j = 0; j < NEQ; ++j) ys[i] += (double)b[i * NEQ + j] * xc[j];
    }
    for(int i = 0; i < NEQ; ++i)
      for(int off = LANES / 2; off > 0; off >>= 1) ys[i] += __shfl_xor(ys[i], off, LANES);
    if(l == 0)
      for(int i = 0; i < NEQ; ++i) {
        const uint64_t o = (uint64_t)row * NEQ + i;
        y[o] = ys[i];
        acc[0] += d0[o] * ys[i];
        if constexpr(NDOT > 1) acc[NDOT - 1] += ys[i] * ys[i];
      }
  }
  block_reduce_store<NDOT>(acc, partials, gridDim.x);
}

// alpha = rho / (rh . v) from the partials pv; s = r - alpha v, sh = D^-1 s, partials ||s||^2
template<typename R>
__global__ void k_fb_s(uint64_t n, double* __restrict__ S, int it, unsigned Gv, const double* __restrict__ pv,
                       const R* __restrict__ r, const R* __restrict__ v, R* __restrict__ s, R* __restrict__ sh,
                       const R* __restrict__ dinv, double* __restrict__ ps)
{
  if(S[S_STOP] != 0.) return;
  double red[1];
  workgroup_sum<1>(pv, Gv, red);
  const double alpha_d = S[(it & 1) ? S_RHO_ODD : S_RHO_EVEN] / red[0];
  if(leader()) {
    S[S_RHV] = red[0];
    S[S_ALPHA] = alpha_d;
  }
  const R alpha = (R)alpha_d;
  double acc[1] = {0.};
  const uint64_t stride = (uint64_t)gridDim.x * kRT;
  for(uint64_t i = (uint64_t)blockIdx.x * kRT + threadIdx.x; i < n; i += stride) {
    const R si = r[i] - alpha * v[i];
    s[i] = si;
    sh[i] = dinv[i] * si;
    acc[0] += (double)si * (double)si;
  }
  block_reduce_store<1>(acc, ps, gridDim.x);
}

// omega = (t . s) / (t . t) (0 once ||s|| meets the tolerance) from the partials ps, pt;
// x += alpha ph + omega sh, r = s - omega t, partials (||r||^2, rh . r)
template<typename R>
__global__ void k_fb_x(uint64_t n, double* __restrict__ S, unsigned Gs, const double* __restrict__ ps, unsigned Gt,
                       const double* __restrict__ pt, R* __restrict__ x, const R* __restrict__ ph,
                       const R* __restrict__ sh, R* __restrict__ r, const R* __restrict__ s, const R* __restrict__ t,
                       const R* __restrict__ rh, double* __restrict__ px)
{
  if(S[S_STOP] != 0.) return;
  double rs[1], rt[2];
  workgroup_sum<1>(ps, Gs, rs);
  __syncthreads();
  workgroup_sum<2>(pt, Gt, rt);
  const bool sconv = rs[0] <= S[S_THR2];
  const double omega_d = sconv ? 0. : rt[0] / rt[1];
  if(leader()) {
    S[S_SN2] = rs[0];
    S[S_SCONV] = sconv ? 1. : 0.;
    S[S_TS] = rt[0];
    S[S_TT] = rt[1];
    S[S_OMEGA] = omega_d;
  }
  const R alpha = (R)S[S_ALPHA], omega = (R)omega_d;
  double acc[2] = {0., 0.};
  const uint64_t stride = (uint64_t)gridDim.x * kRT;
  for(uint64_t i = (uint64_t)blockIdx.x * kRT + threadIdx.x; i < n; i += stride) {
    x[i] += alpha * ph[i] + omega * sh[i];
    const R ri = s[i] - omega * t[i];
    r[i] = ri;
    acc[0] += (double)ri * (double)ri;
    acc[1] += (double)rh[i] * (double)ri;
  }
  block_reduce_store<2>(acc, px, gridDim.x);
}

template<typename R>
__global__ void k_dot2(uint64_t n, const R* __restrict__ a0, const R* __restrict__ b0, const R* __restrict__ a1,
                       const R* __restrict__ b1, double* partials, const double* __restrict__ stop)
{
  if(stop && stop[0] != 0.) return;
  double v[2] = {0., 0.};
  const uint64_t stride = (uint64_t)gridDim.x * kRT;
  for(uint64_t i = (uint64_t)blockIdx.x * kRT + threadIdx.x; i < n; i += stride) {
    v[0] += (double)a0[i] * (double)b0[i];
    if(a1) v[1] += (double)a1[i] * (double)b1[i];
  }
  block_reduce_store<2>(v, partials, gridDim.x);
}

// dinv = 1 / diag(A) in the vector type R from values of type VR
template<typename R, typename VR>
__global__ void k_jacobi(uint32_t N, int neq, const uint32_t* __restrict__ diag_blk, const VR* __restrict__ vals,
                         R* __restrict__ dinv)
{
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(r >= (uint64_t)N * neq) return;
  const uint32_t i = (uint32_t)(r / neq);
  const int a = (int)(r % neq);
  dinv[r] = R(1) / (R)vals[(uint64_t)diag_blk[i] * neq * neq + a * neq + a];
}

// r = b - r (r holds A x on entry), rh = r; partial ||r||^2 and ||b||^2
template<typename R>
__global__ void k_bicg_init(uint64_t n, const R* __restrict__ b, R* __restrict__ r, R* __restrict__ rh,
                            R* __restrict__ p, R* __restrict__ v, double* partials)
{
  double acc[2] = {0., 0.};
  const uint64_t stride = (uint64_t)gridDim.x * kRT;
  for(uint64_t i = (uint64_t)blockIdx.x * kRT + threadIdx.x; i < n; i += stride) {
    const R ri = b[i] - r[i];
    r[i] = ri; rh[i] = ri; p[i] = R(0); v[i] = R(0);
    acc[0] += (double)ri * (double)ri;
    acc[1] += (double)b[i] * (double)b[i];
  }
  block_reduce_store<2>(acc, partials, gridDim.x);
}

template<typename R>
__global__ void k_bicg_p(uint64_t n, const double* __restrict__ S, const R* __restrict__ r, R* __restrict__ p,
                         const R* __restrict__ v, R* __restrict__ ph, const R* __restrict__ dinv)
{
  if(S[S_STOP] != 0.) return;
  const double beta_d = (S[S_RHO] / S[S_RHO_PREV]) * (S[S_ALPHA] / S[S_OMEGA]);
  const R beta = (R)beta_d, omega = (R)S[S_OMEGA];
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for(uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const R pi = r[i] + beta * (p[i] - omega * v[i]);
    p[i] = pi;
    ph[i] = dinv[i] * pi;
  }
}

template<typename R>
__global__ void k_bicg_s(uint64_t n, const double* __restrict__ S, const R* __restrict__ r, const R* __restrict__ v,
                         R* __restrict__ s, R* __restrict__ sh, const R* __restrict__ dinv, double* partials)
{
  if(S[S_STOP] != 0.) return;
  const R alpha = (R)S[S_ALPHA];
  double acc[1] = {0.};
  const uint64_t stride = (uint64_t)gridDim.x * kRT;
  for(uint64_t i = (uint64_t)blockIdx.x * kRT + threadIdx.x; i < n; i += stride) {
    const R si = r[i] - alpha * v[i];
    s[i] = si;
    sh[i] = dinv[i] * si;
    acc[0] += (double)si * (double)si;
  }
  block_reduce_store<1>(acc, partials, gridDim.x);
}

template<typename R>
__global__ void k_bicg_x(uint64_t n, const double* __restrict__ S, R* __restrict__ x, const R* __restrict__ ph,
                         const R* __restrict__ sh, R* __restrict__ r, const R* __restrict__ s, const R* __restrict__ t,
                         const R* __restrict__ rh, double* partials)
{
  if(S[S_STOP] != 0.) return;
  const R alpha = (R)S[S_ALPHA], omega = (R)S[S_OMEGA];
  double acc[2] = {0., 0.};
  const uint64_t stride = (uint64_t)gridDim.x * kRT;
  for(uint64_t i = (uint64_t)blockIdx.x * kRT + threadIdx.x; i < n; i += stride) {
    x[i] += alpha * ph[i] + omega * sh[i];
    const R ri = s[i] - omega * t[i];
    r[i] = ri;
    acc[0] += (double)ri * (double)ri;
    acc[1] += (double)rh[i] * (double)ri;
  }
  block_reduce_store<2>(acc, partials, gridDim.x);
}

__global__ void k_set_scalar(double* S, int idx, double v) { S[idx] = v; }

// fresh recurrence (p = v = 0) at iteration it0
__global__ void k_set_scalars(double* S, double thr2, int it0)
{
  S[S_RHO] = 1.; S[S_RHO_PREV] = 1.; S[S_ALPHA] = 1.; S[S_OMEGA] = 1.; S[S_THR2] = thr2; S[S_SCONV] = 0.;
  S[S_STOP] = 0.; S[S_ITER] = (double)it0;
}

constexpr int kMaxBatch = 16;   // iterations enqueued between two host reads of the scalars
constexpr unsigned kFusedSpmvBlocks = 2048;   // workgroups of the fused products (8 per CU)
// partials of the fused iteration: (||r||^2, rh . r), rh . v, ||s||^2, (t . s, t . t)
constexpr size_t kFusedPartials = 2 * kMaxRedBlocks + kFusedSpmvBlocks + kMaxRedBlocks + 2 * kFusedSpmvBlocks;

template<typename Src, typename Dst>
__global__ void k_convert_scaled(uint64_t n, const Src* __restrict__ a, Dst* __restrict__ b, double scale)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for(uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = (Dst)((double)a[i] * scale);
}

// y = A x with vectors of type R and values of type VR (FP64, FP32, or FP32 values with FP64 vectors)
template<typename R, typename VR>
void spmv_t(Lss& L, const R* x, R* y, const double* stop);
template<>
void spmv_t<double, double>(Lss& L, const double* x, double* y, const double* stop)
{
  lss_spmv(L, 1., x, 0., y, stop);
}
template<>
void spmv_t<float, float>(Lss& L, const float* x, float* y, const double* stop)
{
  lss_spmv32(L, 1.f, x, 0.f, y, stop);
}
template<>
void spmv_t<double, float>(Lss& L, const double* x, double* y, const double* stop)
{
  lss_spmv_mixed(L, x, y, stop);
}

struct Scalars {
  double v[S_COUNT];
};

void read_scalars(Context& c, Scalars& s)
{
  HIP_OK(hipMemcpyAsync(c.scalars_host, c.scalars.ptr, sizeof(double) * S_COUNT, hipMemcpyDeviceToHost, c.stream));
  HIP_OK(hipStreamSynchronize(c.stream));
  std::copy(c.scalars_host, c.scalars_host + S_COUNT, s.v);
}

// Device state of one BiCGStab recurrence (Jacobi right preconditioning, dinv = D^-1) and the launches
// that advance it; the scalars stay on the device and the host reads them once per batch.  With FP64
// vectors and neq = 1 to 4 an iteration is five launches (dot products fused into the products, each
// reduction completed by the kernel that consumes it); otherwise eleven (separate dot products and
// one-workgroup finalisations).  Both sum in a fixed order: the results are deterministic.  rel selects
// the residual-replacement stop test of the reliable phase.
template<typename R, typename VR>
struct Recurrence {
  Lss& L;
  Context& c;
  uint64_t n;
  double* S;
  R *x, *r, *rh, *p, *v, *ph, *s, *sh, *t;
  const R* dinv;
  int rel;
  bool fused;
  unsigned Gv, Gs;
  double *part, *px, *pv, *ps, *pt;
  int next = 0;   // index of the next iteration to enqueue

  Recurrence(Lss& L_, R* x_, R* work, int rel_)
      : L(L_), c(*L_.ctx), n(L_.dofs), S(L_.ctx->scalars.as<double>()), x(x_), rel(rel_)
  {
    r = work; rh = r + n; p = rh + n; v = p + n; ph = v + n; s = ph + n; sh = s + n; t = sh + n;
    dinv = t + n;
    // FP64 vectors with neq = 1 to 4; FP32 vectors (with FP32 values) with neq = 4
    fused = sizeof(R) == sizeof(double) ? (L.neq >= 1 && L.neq <= 4) : (sizeof(VR) == sizeof(float) && L.neq == 4);
    Gv = red_grid(n);
    // rows per workgroup of the fused products: 16 lanes per node row, 8 for scalar rows (k_spmvn_dot)
    Gs = std::min<unsigned>(kFusedSpmvBlocks, grid_for(L.n_nodes, kRT / (L.neq == 1 ? 8 : 16)));
    part = c.red_partials.as<double>();
    if(fused) {
      if(!c.fused_partials.ptr) c.fused_partials.alloc(&c, sizeof(double) * kFusedPartials);
      px = c.fused_partials.as<double>();
      pv = px + 2 * kMaxRedBlocks;
      ps = pv + kFusedSpmvBlocks;
      pt = ps + kMaxRedBlocks;
    }
  }
  const VR* vals() const { return sizeof(VR) == sizeof(double) ? (const VR*)L.values.ptr : (const VR*)L.values32.ptr; }
  double rho(const Scalars& sc, int it) const
  {
    return fused ? sc.v[(it & 1) ? S_RHO_ODD : S_RHO_EVEN] : sc.v[S_RHO];
  }

  // r holds a new residual.  restart: fresh recurrence from iteration it0 (rh = r, p = v = 0, scalars
  // reset); otherwise a residual replacement that keeps the search directions and scalars.
  void seed(double thr2, bool restart, int it0)
  {
    if(restart) {
      if(rh != r) HIP_OK(hipMemcpyAsync(rh, r, sizeof(R) * n, hipMemcpyDeviceToDevice, c.stream));
      fill(p, 0.);
      fill(v, 0.);
      hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(1), 0, c.stream, S, thr2, it0);
    } else {
      hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, c.stream, S, (int)S_SCONV, 0.);
      hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, c.stream, S, (int)S_STOP, 0.);
    }
    if(fused) {
      hipLaunchKernelGGL(k_dot2<R>, dim3(Gv), dim3(kRT), 0, c.stream, n, r, r, rh, r, px, (const double*)nullptr);
    } else {
      hipLaunchKernelGGL(k_dot2<R>, dim3(Gv), dim3(kRT), 0, c.stream, n, rh, r, (const R*)nullptr, (const R*)nullptr,
                         part, (const double*)nullptr);
      hipLaunchKernelGGL(k_finalize<1>, dim3(1), dim3(kRT), 0, c.stream, part, Gv, S, (int)S_RHO, (int)FIN_STORE);
    }
    HIP_OK(hipGetLastError());
    next = it0;
  }

  // fused product y = A x with its partial dot products (see k_spmv4_dot / k_spmvn_dot)
  template<int NDOT>
  void spmv_dot(const uint32_t* rp, const uint32_t* cl, const VR* vl, const R* xx, R* yy, const R* d0,
                double* partials)
  {
    if constexpr(sizeof(R) == sizeof(float)) {
      hipLaunchKernelGGL((k_spmv4_dot<VR, NDOT, float>), dim3(Gs), dim3(kRT), 0, c.stream, L.n_nodes, rp, cl, vl, xx, yy,
                         d0, S, partials);
      return;
    } else
    switch(L.neq) {
      case 4: hipLaunchKernelGGL((k_spmv4_dot<VR, NDOT>), dim3(Gs), dim3(kRT), 0, c.stream, L.n_nodes, rp, cl, vl, xx, yy,
                                 d0, S, partials); break;
      case 3: hipLaunchKernelGGL((k_spmvn_dot<VR, NDOT, 3>), dim3(Gs), dim3(kRT), 0, c.stream, L.n_nodes, rp, cl, vl, xx,
                                 yy, d0, S, partials); break;
      case 2: hipLaunchKernelGGL((k_spmvn_dot<VR, NDOT, 2>), dim3(Gs), dim3(kRT), 0, c.stream, L.n_nodes, rp, cl, vl, xx,
                                 yy, d0, S, partials); break;
      default: hipLaunchKernelGGL((k_spmvn_dot<VR, NDOT, 1>), dim3(Gs), dim3(kRT), 0, c.stream, L.n_nodes, rp, cl, vl,
                                  xx, yy, d0, S, partials); break;
    }
  }

  // nb iterations; afterwards S holds the iteration count, the stop flag and the last residual norm
  void enqueue(int nb)
  {
    if(fused) {
      const VR* vl = vals();
      const uint32_t* rp = L.row_ptr.as<uint32_t>();
      const uint32_t* cl = L.cols.as<uint32_t>();
      for(int k = 0; k < nb; ++k, ++next) {
        hipLaunchKernelGGL(k_fb_p<R>, dim3(Gv), dim3(kRT), 0, c.stream, n, S, next, rel, 0, Gv, px, r, p, v, ph, dinv);
        spmv_dot<1>(rp, cl, vl, ph, v, rh, pv);
        hipLaunchKernelGGL(k_fb_s<R>, dim3(Gv), dim3(kRT), 0, c.stream, n, S, next, Gs, pv, r, v, s, sh, dinv, ps);
        spmv_dot<2>(rp, cl, vl, sh, t, s, pt);
        hipLaunchKernelGGL(k_fb_x<R>, dim3(Gv), dim3(kRT), 0, c.stream, n, S, Gv, ps, Gs, pt, x, ph, sh, r, s, t, rh, px);
      }
      // close the batch: finalise its last iteration
      hipLaunchKernelGGL(k_fb_p<R>, dim3(1), dim3(kRT), 0, c.stream, n, S, next, rel, 1, Gv, px, r, p, v, ph, dinv);
      HIP_OK(hipGetLastError());
      return;
    }
    const int fin_last = rel ? (int)FIN_RNORM_REL : (int)FIN_RNORM;
    const unsigned G = Gv;
    for(int k = 0; k < nb; ++k, ++next) {
      hipLaunchKernelGGL(k_bicg_p<R>, dim3(G), dim3(kRT), 0, c.stream, n, S, r, p, v, ph, dinv);
      spmv_t<R, VR>(L, ph, v, S + S_STOP);
      hipLaunchKernelGGL(k_dot2<R>, dim3(G), dim3(kRT), 0, c.stream, n, rh, v, (const R*)nullptr, (const R*)nullptr,
                         part, S + S_STOP);
      hipLaunchKernelGGL(k_finalize<1>, dim3(1), dim3(kRT), 0, c.stream, part, G, S, 0, (int)FIN_ALPHA);
      hipLaunchKernelGGL(k_bicg_s<R>, dim3(G), dim3(kRT), 0, c.stream, n, S, r, v, s, sh, dinv, part);
      hipLaunchKernelGGL(k_finalize<1>, dim3(1), dim3(kRT), 0, c.stream, part, G, S, 0, (int)FIN_SNORM);
      spmv_t<R, VR>(L, sh, t, S + S_STOP);
      hipLaunchKernelGGL(k_dot2<R>, dim3(G), dim3(kRT), 0, c.stream, n, t, s, t, t, part, S + S_STOP);
      hipLaunchKernelGGL(k_finalize<2>, dim3(1), dim3(kRT), 0, c.stream, part, G, S, 0, (int)FIN_OMEGA);
      hipLaunchKernelGGL(k_bicg_x<R>, dim3(G), dim3(kRT), 0, c.stream, n, S, x, ph, sh, r, s, t, rh, part);
      hipLaunchKernelGGL(k_finalize<2>, dim3(1), dim3(kRT), 0, c.stream, part, G, S, 0, fin_last);
    }
    HIP_OK(hipGetLastError());
  }

 private:
  void fill(R* a, double val)
  {
    if constexpr(sizeof(R) == sizeof(double)) lss_fill(c, a, n, val);
    else lss_fill32(c, a, n, (float)val);
  }
};

// BiCGStab on A (values of type VR, vectors of type R) for A x = b; x is the initial guess and the
// result.  Returns the number of iterations; conv = recursive residual below tol*||b||.
template<typename R, typename VR = R>
int bicgstab(Lss& L, const R* b, R* x, R* work, double tol, int maxit, bool& conv, int& spmvs)
{
  Context& c = *L.ctx;
  const uint64_t n = L.dofs;
  Recurrence<R, VR> rc(L, x, work, 0);
  double* S = rc.S;
  double* part = rc.part;
  const unsigned G = rc.Gv;
  hipLaunchKernelGGL((k_jacobi<R, VR>), dim3(grid_for(n, 256)), dim3(256), 0, c.stream, L.n_nodes, L.neq,
                     L.diag_blk.as<uint32_t>(), rc.vals(), const_cast<R*>(rc.dinv));
  spmv_t<R, VR>(L, x, rc.r, nullptr);
  hipLaunchKernelGGL(k_bicg_init<R>, dim3(G), dim3(kRT), 0, c.stream, n, b, rc.r, rc.rh, rc.p, rc.v, part);
  hipLaunchKernelGGL(k_finalize<2>, dim3(1), dim3(kRT), 0, c.stream, part, G, S, (int)S_RN2, (int)FIN_STORE);
  spmvs = 1;
  Scalars sc;
  read_scalars(c, sc);
  double bn2 = sc.v[S_RN2 + 1];
  if(bn2 == 0.) bn2 = 1.;
  const double thr2 = tol * tol * bn2;
  rc.seed(thr2, true, 0);
  conv = sc.v[S_RN2] <= thr2;
  int it = 0, batch = 1;
  while(!conv && it < maxit) {
    rc.enqueue(std::min(batch, maxit - it));
    read_scalars(c, sc);
    it = (int)sc.v[S_ITER];
    if(sc.v[S_STOP] != 0.) {   // converged, or broke down
      conv = sc.v[S_SCONV] != 0. || sc.v[S_RN2] <= thr2;
      break;
    }
    batch = std::min(2 * batch, kMaxBatch);
  }
  spmvs = 1 + 2 * it;
  return it;
}

double norm2(Context& c, const double* a, uint64_t n)
{
  double* part = c.red_partials.as<double>();
  const unsigned G = red_grid(n);
  hipLaunchKernelGGL(k_dot2<double>, dim3(G), dim3(kRT), 0, c.stream, n, a, a, (const double*)nullptr,
                     (const double*)nullptr, part, (const double*)nullptr);
  hipLaunchKernelGGL(k_finalize<1>, dim3(1), dim3(kRT), 0, c.stream, part, G, c.scalars.as<double>(), (int)S_TMP0,
                     (int)FIN_STORE);
  HIP_OK(hipGetLastError());
  Scalars sc;
  read_scalars(c, sc);
  return std::sqrt(sc.v[S_TMP0]);
}

void ensure_work(Lss& L, size_t nvec, int real_bytes)
{
  const size_t need = nvec * L.dofs * real_bytes;
  if(L.work.bytes < need) L.work.alloc(L.ctx, need);
}

}  // namespace

void convert_f64_to_f32(Context& c, const double* a, float* b, size_t n)
{
  hipLaunchKernelGGL((k_convert_scaled<double, float>), dim3(std::min<unsigned>(grid_for(n, 256), 8192)), dim3(256), 0,
                     c.stream, (uint64_t)n, a, b, 1.0);
  HIP_OK(hipGetLastError());
}

double lss_residual(Lss& L)
{
  Context& c = *L.ctx;
  ensure_work(L, 1, sizeof(double));
  double* r = L.work.as<double>();
  HIP_OK(hipMemcpyAsync(r, L.rhs.ptr, sizeof(double) * L.dofs, hipMemcpyDeviceToDevice, c.stream));
  lss_spmv(L, -1., L.sol.as<double>(), 1., r);
  const double rn = norm2(c, r, L.dofs);
  double bn = norm2(c, L.rhs.as<double>(), L.dofs);
  if(bn == 0.) bn = 1.;
  return rn / bn;
}

void lss_solve(Lss& L, const cf3hip_solver_params& prm, cf3hip_solve_stats& st)
{
  Context& c = *L.ctx;
  st = cf3hip_solve_stats{0, 0, 0, 0, 0.};
  if(prm.method == CF3HIP_GMRES_ILU0) {
    extern void gmres_ilu0(Lss&, const cf3hip_solver_params&, cf3hip_solve_stats&);
    gmres_ilu0(L, prm, st);
    return;
  }
  if(prm.method == CF3HIP_RCG_JACOBI) {
    rcg_solve(L, prm, st);
    return;
  }
  if(prm.method != CF3HIP_BICGSTAB_JACOBI) CF3_THROW(CF3HIP_ERR_INVALID_ARG, "unknown Krylov method");
  if(prm.precision == CF3HIP_FP64) {
    ensure_work(L, 9, sizeof(double));
    bool conv = false;
    int spmvs = 0;
    st.iterations = bicgstab<double>(L, L.rhs.as<double>(), L.sol.as<double>(), L.work.as<double>(), prm.tol,
                                     prm.max_iter, conv, spmvs);
    st.outer = 1;
    st.spmvs = spmvs;
    st.rel_residual = lss_residual(L);
    st.converged = conv ? 1 : 0;
    return;
  }
  // FP32 values under FP64 refinement of the stored FP64 system
  (void)c;
  lss_solve_precision(L, prm, st);
}

namespace {

template<typename XL>
__global__ void k_flush_increment(uint64_t n, double* __restrict__ xhi, XL* __restrict__ xlo)
{
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for(uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    xhi[i] += (double)xlo[i];
    xlo[i] = XL(0);
  }
}

// ---- block-column scaling of the precision form: B = fl32(A32 D^-1), x = D^-1 y ----
// D^-1 of each node's diagonal block, by Gauss-Jordan elimination with partial pivoting in FP64 on the
// FP32 values; a block with a zero or non-finite pivot keeps D^-1 = I (its column block is not scaled)
template<int NEQ>
__global__ void k_block_inverse(uint32_t N, const uint32_t* __restrict__ diag_blk, const float* __restrict__ vals,
                                double* __restrict__ dinv)
{
  constexpr int NN = NEQ * NEQ;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= N) return;
  double a[NEQ][NEQ], b[NEQ][NEQ];
  const float* d = vals + (uint64_t)diag_blk[i] * NN;
#pragma unroll
  for(int r = 0; r < NEQ; ++r)
#pragma unroll
    for(int c = 0; c < NEQ; ++c) { a[r][c] = (double)d[r * NEQ + c]; b[r][c] = r == c ? 1. : 0.; }
  bool ok = true;
#pragma unroll
  for(int k = 0; k < NEQ; ++k) {
    int piv = k;
    double pv = fabs(a[k][k]);
#pragma unroll
    for(int r = k + 1; r < NEQ; ++r)
      if(fabs(a[r][k]) > pv) { pv = fabs(a[r][k]); piv = r; }
#pragma unroll
    for(int r = k + 1; r < NEQ; ++r)
      if(r == piv)
#pragma unroll
        for(int c = 0; c < NEQ; ++c) {
          const double ta = a[k][c]; a[k][c] = a[r][c]; a[r][c] = ta;
          const double tb = b[k][c]; b[k][c] = b[r][c]; b[r][c] = tb;
        }
    ok = ok && pv > 0. && isfinite(pv);
    const double inv = ok ? 1. / a[k][k] : 0.;
#pragma unroll
    for(int c = 0; c < NEQ; ++c) { a[k][c] *= inv; b[k][c] *= inv; }
#pragma unroll
    for(int r = 0; r < NEQ; ++r) {
      if(r == k) continue;
      const double f = a[r][k];
#pragma unroll
      for(int c = 0; c < NEQ; ++c) { a[r][c] -= f * a[k][c]; b[r][c] -= f * b[k][c]; }
    }
  }
  double* o = dinv + (uint64_t)i * NN;
#pragma unroll
  for(int r = 0; r < NEQ; ++r)
#pragma unroll
    for(int c = 0; c < NEQ; ++c) o[r * NEQ + c] = ok ? b[r][c] : (r == c ? 1. : 0.);
}

// out(block, row r) = fl32(vals(block, row r) D^-1 of the block's column node), sums in FP64
template<int NEQ>
__global__ void k_scale_cols(uint64_t n_rows, const uint32_t* __restrict__ cols, const float* __restrict__ vals,
                             const double* __restrict__ dinv, float* __restrict__ out)
{
  constexpr int NN = NEQ * NEQ;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(t >= n_rows) return;
  const uint64_t blk = t / NEQ;
  const int r = (int)(t % NEQ);
  const double* D = dinv + (uint64_t)cols[blk] * NN;
  double a[NEQ];
#pragma unroll
  for(int k = 0; k < NEQ; ++k) a[k] = (double)vals[blk * NN + r * NEQ + k];
#pragma unroll
  for(int c = 0; c < NEQ; ++c) {
    double s = 0.;
#pragma unroll
    for(int k = 0; k < NEQ; ++k) s += a[k] * D[k * NEQ + c];
    out[blk * NN + r * NEQ + c] = (float)s;
  }
}

// xhi += D^-1 xlo per node (xlo holds the increment in the scaled variable y), xlo = 0; with xhi ==
// nullptr the product replaces xlo in place
template<int NEQ, typename XL>
__global__ void k_apply_bdiag(uint32_t N, const double* __restrict__ dinv, double* xhi, XL* xlo)
{
  constexpr int NN = NEQ * NEQ;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= N) return;
  const double* D = dinv + (uint64_t)i * NN;
  double y[NEQ];
#pragma unroll
  for(int k = 0; k < NEQ; ++k) y[k] = (double)xlo[(uint64_t)i * NEQ + k];
#pragma unroll
  for(int r = 0; r < NEQ; ++r) {
    double s = 0.;
#pragma unroll
    for(int k = 0; k < NEQ; ++k) s += D[r * NEQ + k] * y[k];
    if(xhi) {
      xhi[(uint64_t)i * NEQ + r] += s;
      xlo[(uint64_t)i * NEQ + r] = XL(0);
    } else {
      xlo[(uint64_t)i * NEQ + r] = (XL)s;
    }
  }
}

template<int NEQ>
void block_scale_n(Lss& L)
{
  Context& c = *L.ctx;
  const size_t nv = (size_t)L.n_blocks * NEQ * NEQ;
  if(L.values32b.bytes < sizeof(float) * nv) L.values32b.alloc(&c, sizeof(float) * nv);
  if(L.bdinv.bytes < sizeof(double) * L.n_nodes * NEQ * NEQ) L.bdinv.alloc(&c, sizeof(double) * L.n_nodes * NEQ * NEQ);
  hipLaunchKernelGGL(k_block_inverse<NEQ>, dim3(grid_for(L.n_nodes, 256)), dim3(256), 0, c.stream, L.n_nodes,
                     L.diag_blk.as<uint32_t>(), L.values32.as<float>(), L.bdinv.as<double>());
  const uint64_t nr = (uint64_t)L.n_blocks * NEQ;
  hipLaunchKernelGGL(k_scale_cols<NEQ>, dim3(grid_for(nr, 256)), dim3(256), 0, c.stream, nr, L.cols.as<uint32_t>(),
                     L.values32.as<float>(), L.bdinv.as<double>(), L.values32b.as<float>());
  HIP_OK(hipGetLastError());
}

// L.values32b = fl32(A32 D^-1) and L.bdinv = D^-1 from the current FP32 values
void block_scale_fp32(Lss& L)
{
  switch(L.neq) {
    case 1: block_scale_n<1>(L); break;
    case 2: block_scale_n<2>(L); break;
    case 3: block_scale_n<3>(L); break;
    case 4: block_scale_n<4>(L); break;
    default: CF3_THROW(CF3HIP_ERR_NOT_IMPLEMENTED, "block scaling of the precision form supports 1 to 4 equations per node");
  }
}

// xhi += D^-1 xlo, xlo = 0 (xhi != nullptr), or xlo := D^-1 xlo
template<typename XL>
void apply_bdiag(Lss& L, double* xhi, XL* xlo)
{
  Context& c = *L.ctx;
  const dim3 g(grid_for(L.n_nodes, 256)), b(256);
  const double* D = L.bdinv.as<double>();
  switch(L.neq) {
    case 1: hipLaunchKernelGGL((k_apply_bdiag<1, XL>), g, b, 0, c.stream, L.n_nodes, D, xhi, xlo); break;
    case 2: hipLaunchKernelGGL((k_apply_bdiag<2, XL>), g, b, 0, c.stream, L.n_nodes, D, xhi, xlo); break;
    case 3: hipLaunchKernelGGL((k_apply_bdiag<3, XL>), g, b, 0, c.stream, L.n_nodes, D, xhi, xlo); break;
    default: hipLaunchKernelGGL((k_apply_bdiag<4, XL>), g, b, 0, c.stream, L.n_nodes, D, xhi, xlo); break;
  }
  HIP_OK(hipGetLastError());
}

// During a block-scaled solve the FP32 kernels read the scaled operator: the two buffers trade places
// and are restored when the solve ends (also on an exception).
struct ScaledFp32 {
  Lss& L;
  bool on;
  ScaledFp32(Lss& L_, bool on_) : L(L_), on(on_) { if(on) swap(); }
  ~ScaledFp32() { if(on) swap(); }
  void swap()
  {
    std::swap(L.values32.ptr, L.values32b.ptr);
    std::swap(L.values32.bytes, L.values32b.bytes);
    std::swap(L.values32.ctx, L.values32b.ctx);
  }
};

template<typename XL>
void flush_increment(Lss& L, bool bs, double* xhi, XL* xlo)
{
  const uint64_t n = L.dofs;
  if(bs) apply_bdiag(L, xhi, xlo);
  else
    hipLaunchKernelGGL(k_flush_increment<XL>, dim3(std::min<unsigned>(grid_for(n, 256), 8192)), dim3(256), 0,
                       L.ctx->stream, n, xhi, xlo);
  HIP_OK(hipGetLastError());
}

constexpr int kMaxOuterBasis = 16;   // FGMRES restart length cap of the precision form

// Precision form: flexible GMRES(m) in FP64 around an FP32 inner solve (GMRES-based iterative
// refinement).  The outer iteration applies the FP64 operator through `resid` (r = b' - A' x, from the
// stored FP64 values or matrix-free from the FP64 element matrices), so A' z_j is exact in FP64; each
// preconditioning step z_j = M^-1 v_j is a Jacobi-BiCGStab solve of A32 z = v_j with FP32 values and
// FP64 vectors from a zero guess to the relative tolerance inner_tol.  The outer Krylov space absorbs
// the difference between the FP32 and FP64 operators, which plain refinement (or residual replacement
// inside one BiCGStab) cannot when the solution is sensitive to the matrix.  Convergence is declared
// on the FP64 true residual ||b' - A' x|| <= tol ||b'||.
template<class Resid>
void fgmres_ir(Lss& L, Resid resid, const cf3hip_solver_params& prm, cf3hip_solve_stats& st, bool bs = false)
{
  using namespace arnoldi;
  if(prm.method != CF3HIP_BICGSTAB_JACOBI)
    CF3_THROW(CF3HIP_ERR_NOT_IMPLEMENTED, "the precision form's inner solver is BiCGStab + Jacobi");
  if(!L.has_fp32) CF3_THROW(CF3HIP_ERR_NOT_CREATED, "FP32 values not enabled");
  Context& c = *L.ctx;
  const uint64_t n = L.dofs;
  const int m = std::min(std::max(prm.restart, 1), kMaxOuterBasis);
  const double itol = (prm.inner_tol > 0. && prm.inner_tol < 1.) ? prm.inner_tol : 1e-2;
  // basis V (m+1), preconditioned directions Z (m), w, b', bookkeeping; the inner BiCGStab vectors in
  // the generic workspace
  const size_t need = sizeof(double) * ((size_t)(2 * m + 3) * n + GmresDev::size(m));
  if(L.ir.bytes < need) L.ir.alloc(&c, need);
  ensure_work(L, 9, sizeof(double));
  double* V = L.ir.as<double>();
  double* Z = V + (size_t)(m + 1) * n;
  double* w = Z + (size_t)m * n;
  double* b0 = w + n;
  double* base = b0 + n;
  const GmresDev gd{m, base};
  double* sc = c.scalars.as<double>() + 40;   // [0] squared norms of the Arnoldi step
  const unsigned gb = grid_for(n, 256);
  double* x = L.sol.as<double>();
  st = cf3hip_solve_stats{0, 0, 0, 0, 0.};
  auto read1 = [&](const double* dev) {
    double v = 0.;
    HIP_OK(hipMemcpyAsync(&v, dev, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    return v;
  };

  // b' = r(0)
  lss_fill(c, w, n, 0.);
  resid(w, b0);
  multidot(c, n, 1, b0, b0, sc);
  double bn = std::sqrt(read1(sc));
  if(bn == 0.) bn = 1.;
  const double thr = prm.tol * bn;
  int its = 0, outer = 0, spmvs = 0;
  bool conv = false;
  double rtrue = 0.;
  for(;;) {
    resid(x, w);
    multidot(c, n, 1, w, w, sc);
    hipLaunchKernelGGL(k_gmres_start, dim3(1), dim3(1), 0, c.stream, m, base, sc);
    HIP_OK(hipGetLastError());
    rtrue = read1(base);
    if(rtrue <= thr) { conv = true; break; }
    if(its >= prm.max_iter || (prm.max_outer > 0 && outer >= prm.max_outer) || !std::isfinite(rtrue)) break;
    hipLaunchKernelGGL(k_scale, dim3(gb), dim3(256), 0, c.stream, n, w, base, V);
    int j = 0;
    for(; j < m; ++j) {
      double* vj = V + (size_t)j * n;
      double* zj = Z + (size_t)j * n;
      lss_fill(c, zj, n, 0.);
      bool cv = false;
      int sp = 0;
      its += bicgstab<double, float>(L, vj, zj, L.work.as<double>(), itol, std::max(1, prm.max_iter - its), cv, sp);
      spmvs += sp;
      if(bs) apply_bdiag<double>(L, nullptr, zj);   // z_j = D^-1 y_j
      ++outer;
      resid(zj, w);   // A' z_j = b' - r(z_j)
      hipLaunchKernelGGL(k_sub, dim3(gb), dim3(256), 0, c.stream, n, b0, w);
      multidot(c, n, j + 1, V, w, gd.h1());
      hipLaunchKernelGGL(k_orth_update, dim3(gb), dim3(256), 0, c.stream, n, j + 1, V, gd.h1(), w);
      multidot(c, n, j + 1, V, w, gd.h2());
      hipLaunchKernelGGL(k_orth_update, dim3(gb), dim3(256), 0, c.stream, n, j + 1, V, gd.h2(), w);
      multidot(c, n, 1, w, w, sc);
      hipLaunchKernelGGL(k_arnoldi_close, dim3(1), dim3(1), 0, c.stream, m, base, j, gd.h1(), gd.h2(), sc);
      hipLaunchKernelGGL(k_scale, dim3(gb), dim3(256), 0, c.stream, n, w, base + 1, V + (size_t)(j + 1) * n);
      HIP_OK(hipGetLastError());
      const double est = read1(base);
      if(!std::isfinite(est)) break;   // breakdown: keep the j columns before this one
      if(est <= thr || its >= prm.max_iter || (prm.max_outer > 0 && outer >= prm.max_outer)) { ++j; break; }
    }
    if(j == 0) break;   // no usable direction
    hipLaunchKernelGGL(k_back_subst, dim3(1), dim3(1), 0, c.stream, m, base, j, gd.y());
    hipLaunchKernelGGL(k_add_lincomb, dim3(gb), dim3(256), 0, c.stream, n, j, Z, gd.y(), x);
    HIP_OK(hipGetLastError());
  }
  st.iterations = its;
  st.outer = outer;
  st.spmvs = spmvs;
  st.rel_residual = rtrue / bn;
  st.converged = conv ? 1 : 0;
}

}  // namespace

// Precision form, first phase: BiCGStab with Jacobi preconditioning on the FP32 values (FP32 stored;
// vectors R = FP64 with FP64 accumulation in the products, or R = FP32 with FP32 accumulation), with
// reliable updates: the FP64 solution accumulates the increments and, whenever the recursive residual has
// dropped by the factor delta (or meets the tolerance), the residual is replaced by the true FP64 residual
// b - A x given by `resid`.  The search directions are kept across replacements (no restart), so on most
// systems the iteration count stays close to that of the FP64 method.  With block scaling the recurrence
// runs on B = fl32(A32 D^-1) and its increment is y, x = D^-1 y.  When the FP32 operator is too far from
// the FP64 one for that (the true residual grows beyond kDiverge times its best value, or kSwitchStall
// replacements in a row fail to halve it: single rises are ordinary for BiCGStab), the solve continues
// from the current FP64 iterate with fgmres_ir.  Convergence is declared on the FP64 true residual only.
template<typename R, class Resid>
void reliable_core(Lss& L, Resid resid, const cf3hip_solver_params& prm, cf3hip_solve_stats& st)
{
  Context& c = *L.ctx;
  const uint64_t n = L.dofs;
  // the recurrence (9 vectors) and the increment xlo in type R; with R = FP32 also an FP64 true residual
  // and an FP64 zero vector behind them (7 n doubles in all)
  ensure_work(L, 10, sizeof(double));
  R* work = L.work.as<R>();
  R* xlo = work + 9 * n;
  double* xhi = L.sol.as<double>();
  Recurrence<R, float> rc(L, xlo, work, 1);
  R* r = rc.r;
  double* S = rc.S;
  constexpr bool f32 = sizeof(R) == sizeof(float);
  double* rt = f32 ? reinterpret_cast<double*>(work + 10 * n) : reinterpret_cast<double*>(r);   // FP64 true residual
  double* zero64 = f32 ? rt + n : reinterpret_cast<double*>(xlo);
  // replacement after every 10-fold drop of the recursive residual: on the 128^3 channel and the 160^3
  // cavity this gave fewer iterations than replacing after 100-fold drops
  const double delta = 0.1;
  st = cf3hip_solve_stats{0, 0, 0, 0, 0.};
  const bool bs = prm.block_scale != 0;
  if(bs) block_scale_fp32(L);
  ScaledFp32 scaled(L, bs);   // from here the FP32 values are B = fl32(A32 D^-1); xlo holds y
  auto to_r = [&]() {
    if constexpr(f32) convert_f64_to_f32(c, rt, reinterpret_cast<float*>(r), n);
  };

  // ||b'|| = ||r(0)||
  lss_fill(c, zero64, n, 0.);
  if constexpr(f32) lss_fill32(c, reinterpret_cast<float*>(xlo), n, 0.f);
  resid(zero64, rt);
  double bn = norm2(c, rt, n);
  if(bn == 0.) bn = 1.;
  const double thr2 = prm.tol * prm.tol * bn * bn;
  resid(xhi, rt);
  int updates = 1;
  double rtrue = norm2(c, rt, n);
  to_r();
  hipLaunchKernelGGL((k_jacobi<R, float>), dim3(grid_for(n, 256)), dim3(256), 0, c.stream, L.n_nodes, L.neq,
                     L.diag_blk.as<uint32_t>(), L.values32.as<float>(), const_cast<R*>(rc.dinv));
  rc.seed(thr2, true, 0);
  bool conv = rtrue * rtrue <= thr2;
  // the fallback is for operators the FP32 recurrence cannot follow: on the 128^3 channel (block-scaled
  // FP32 operator) segments that fail to halve the true residual come in long runs and still converge
  // sooner than the FGMRES phase would, so only a long stall or a true divergence switches
  constexpr int kSwitchStall = 40;
  constexpr double kDiverge = 1e3;
  // iterations enqueued after a replacement before the first host read (a segment is ~10-40 iterations;
  // launches past a stop return at once)
  constexpr int kBatchAfterReplace = 4;
  double rprog = rtrue, rbest = rtrue;
  int stall = 0;
  bool switched = false;
  hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, c.stream, S, (int)S_MAXR, rtrue);
  hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, c.stream, S, (int)S_MAXR_ODD, rtrue);
  hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, c.stream, S, (int)S_DELTA, delta);
  int it = 0, batch = 1;
  Scalars sc;
  while(!conv && it < prm.max_iter) {
    rc.enqueue(std::min(batch, prm.max_iter - it));
    read_scalars(c, sc);
    it = (int)sc.v[S_ITER];
    if(sc.v[S_STOP] == 0.) {
      batch = std::min(2 * batch, kMaxBatch);
      continue;
    }
    // reliable update: move the increment into the FP64 solution, replace the residual
    const bool breakdown = !std::isfinite(sc.v[S_RN2]) || rc.rho(sc, it) == 0.;
    flush_increment(L, bs, xhi, xlo);
    resid(xhi, rt);
    ++updates;
    rtrue = norm2(c, rt, n);
    if(rtrue * rtrue <= thr2) { conv = true; break; }
    if(rtrue < 0.5 * rprog) {
      rprog = rtrue;
      stall = 0;
    } else {
      ++stall;
    }
    rbest = std::min(rbest, rtrue);
    if(!std::isfinite(rtrue) || rtrue > kDiverge * rbest || stall >= kSwitchStall) { switched = true; break; }
    to_r();
    // breakdown: restart the recurrences from the true residual; otherwise keep the directions
    rc.seed(thr2, breakdown, it);
    hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, c.stream, S, (int)S_MAXR, rtrue);
    hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, c.stream, S, (int)S_MAXR_ODD, rtrue);
    HIP_OK(hipGetLastError());
    batch = kBatchAfterReplace;
  }
  const int spmvs = 2 * it;
  if(switched) {
    if(!std::isfinite(rtrue)) lss_fill(c, xhi, n, 0.);
    cf3hip_solver_params p2 = prm;
    p2.max_iter = std::max(1, prm.max_iter - it);
    cf3hip_solve_stats s2;
    fgmres_ir(L, resid, p2, s2, bs);
    st.iterations = it + s2.iterations;
    st.outer = updates + s2.outer;
    st.spmvs = spmvs + s2.spmvs;
    st.rel_residual = s2.rel_residual;
    st.converged = s2.converged;
    return;
  }
  if(!conv) {
    flush_increment(L, bs, xhi, xlo);
    resid(xhi, rt);
    rtrue = norm2(c, rt, n);
    conv = rtrue * rtrue <= thr2;
  }
  st.iterations = it;
  st.outer = updates;
  st.spmvs = spmvs;
  st.rel_residual = rtrue / bn;
  st.converged = conv ? 1 : 0;
}

template<class Resid>
void reliable_bicgstab(Lss& L, Resid resid, const cf3hip_solver_params& prm, cf3hip_solve_stats& st)
{
  if(prm.method != CF3HIP_BICGSTAB_JACOBI) CF3_THROW(CF3HIP_ERR_NOT_IMPLEMENTED, "precision form uses BiCGStab");
  if(!L.has_fp32) CF3_THROW(CF3HIP_ERR_NOT_CREATED, "FP32 values not enabled");
  if(prm.fp32_vectors && L.neq == 4) reliable_core<float>(L, resid, prm, st);
  else reliable_core<double>(L, resid, prm, st);
}

void ns_solve(Lss& L, Fields& F, const cf3hip_ns_params& nsp, const cf3hip_solver_params& prm,
              cf3hip_solve_stats& st)
{
  if(prm.precision == CF3HIP_FP64) {
    lss_solve(L, prm, st);
    return;
  }
  if(L.tied)
    CF3_THROW(CF3HIP_ERR_NOT_IMPLEMENTED, "tied block rows: the matrix-free residual of the element operator does not "
                                          "describe the system; solve the stored system instead");
  if(L.builder == CF3HIP_BUILDER_CRS)
    CF3_THROW(CF3HIP_ERR_NOT_IMPLEMENTED, "the matrix-free residual models the conditions of the block (VBR) builder; "
                                          "solve the stored CRS system instead");
  // the true residual is evaluated matrix-free from the FP64 element matrices of the fields
  reliable_bicgstab(L, [&](const double* x, double* r) { ns_residual(L, F, nsp, x, r); }, prm, st);
}

// precision form on a stored system: the true residual uses the stored FP64 values and rhs
void lss_solve_precision(Lss& L, const cf3hip_solver_params& prm, cf3hip_solve_stats& st)
{
  Context& c = *L.ctx;
  reliable_bicgstab(L, [&](const double* x, double* r) {
    lss_copy(c, r, L.rhs.as<double>(), L.dofs);
    lss_spmv(L, -1., x, 1., r);
  }, prm, st);
}

}  // namespace cf3hip

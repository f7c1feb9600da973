// Recycling conjugate gradients (the RCGStrategy of cf3/math/LSS/Trilinos/RCGStrategy.cpp:59-171, which
// drives Belos::RCGSolMgr from Trilinos; the reference pins no Trilinos version,
// cmake/DetectSoftware.cmake:88) for symmetric positive definite systems such as the simplified
// pressure matrix of the semi-implicit solver.
//
// Algorithm (the published RCG: Saad, Yeung, Erhel and Guyomarc'h, "A deflated version of the conjugate
// gradient algorithm", SIAM J. Sci. Comput. 21 (2000), for the deflated iteration; Wang, de Sturler and
// Paulino, "Large-scale topology optimization using preconditioned Krylov subspace methods with
// recycling", IJNME 69 (2007), for the recycle space that carries over between solves):
//   recycle space W (k columns) with AW = A W and E = W^T A W kept between solves;
//   x += W E^-1 W^T r0, r = b - A x;  z = M r;  p_0 = z - W E^-1 (AW)^T z;
//   for j: q = A p_j; alpha = r.z / p_j.q; x += alpha p_j; r -= alpha q; z = M r;
//          beta = r.z(new) / r.z(old); p_{j+1} = z + beta p_j - W E^-1 (AW)^T z
//   (every search direction is A-orthogonal to W, so CG runs on the complement of span W).
//   After each cycle of m directions and at the end of each solve, W is replaced by the k harmonic Ritz
//   vectors of A over span [W, P] with the smallest harmonic Ritz values,
//     (AV)^T (AV) y = theta V^T A V y,   V = [W P],
//   so W = V Y and AW = (AV) Y need no further products with A.
// M is point Jacobi (the reference preconditions with an ML smoothed-aggregation hierarchy, which is
// not restated; the preconditioner changes iteration counts, not the converged solution).
//
// On the device: the scalars of the recurrence stay in a small device array and the host reads them
// once per batch of iterations; every reduction sums per-block partials in a fixed order, so a solve is
// deterministic.  The Gram matrices of the recycle update are computed by a tiled kernel; the small
// generalized eigenproblem (at most kMaxRecycle + kMaxCycle columns) is solved on the host (Cholesky and
// cyclic Jacobi).  A space is reused while the matrix values have the same fingerprint (two weighted
// sums); otherwise AW and E are recomputed from W.  The solve ends with the true residual
// ||b - A x|| / ||b|| <= tol; a recursive residual that met tol while the true one did not restarts
// the recurrence from the true residual.
#include "arnoldi.hpp"
#include "internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace cf3hip {

namespace {

constexpr int kB = 256;            // block size of the vector kernels
constexpr int kTile = 32;          // Gram tiles of 32 x 32 columns
constexpr int kGramRows = 64;      // rows staged per step of the Gram kernel
constexpr unsigned kGramChunks = 256;

// scalar slots
enum { R_RZ = 0, R_PQ, R_RR, R_STOP, R_ITER, R_THR2, R_ALPHA, R_BETA, R_BREAK, R_BN2, R_COUNT };

__device__ inline void reduce_store(double v0, double v1, int ndot, double* part, unsigned G)
{
  __shared__ double sh[2][kB / 64];
  for(int off = 32; off > 0; off >>= 1) {
    v0 += __shfl_down(v0, off, 64);
    v1 += __shfl_down(v1, off, 64);
  }
  if((threadIdx.x & 63) == 0) { sh[0][threadIdx.x >> 6] = v0; sh[1][threadIdx.x >> 6] = v1; }
  __syncthreads();
  if(threadIdx.x == 0) {
    part[blockIdx.x] = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
    if(ndot > 1) part[G + blockIdx.x] = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
  }
}

// sum of part[0..G) in a fixed order by one block
__device__ inline double block_sum(const double* part, unsigned G)
{
  __shared__ double sh[kB / 64];
  double s = 0.;
  for(unsigned i = threadIdx.x; i < G; i += kB) s += part[i];
  for(int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  const double t = sh[0] + sh[1] + sh[2] + sh[3];
  __syncthreads();
  return t;
}

// dinv = 1 / diag(A) (point Jacobi; a zero diagonal entry leaves its row unscaled)
__global__ void k_rcg_jacobi(uint64_t dofs, int neq, const uint32_t* __restrict__ diag_blk,
                             const double* __restrict__ vals, double* __restrict__ dinv)
{
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(r >= dofs) return;
  const uint64_t i = r / neq;
  const int a = (int)(r % neq);
  const double d = vals[(uint64_t)diag_blk[i] * neq * neq + a * neq + a];
  dinv[r] = d != 0. ? 1. / d : 1.;
}

// partials of a0.b0 (and a1.b1)
__global__ __launch_bounds__(kB) void k_rcg_dot(uint64_t n, const double* __restrict__ a0, const double* __restrict__ b0,
                                                const double* __restrict__ a1, const double* __restrict__ b1,
                                                double* __restrict__ part, const double* __restrict__ S)
{
  if(S && S[R_STOP] != 0.) return;
  double v0 = 0., v1 = 0.;
  const uint64_t stride = (uint64_t)gridDim.x * kB;
  for(uint64_t i = (uint64_t)blockIdx.x * kB + threadIdx.x; i < n; i += stride) {
    v0 += a0[i] * b0[i];
    if(a1) v1 += a1[i] * b1[i];
  }
  reduce_store(v0, v1, a1 ? 2 : 1, part, gridDim.x);
}

// fingerprint of the matrix values: sum v_i and sum w_i v_i with pseudo-random weights w_i in [1, 2)
__global__ __launch_bounds__(kB) void k_rcg_print(uint64_t n, const double* __restrict__ v, double* __restrict__ part)
{
  double v0 = 0., v1 = 0.;
  const uint64_t stride = (uint64_t)gridDim.x * kB;
  for(uint64_t i = (uint64_t)blockIdx.x * kB + threadIdx.x; i < n; i += stride) {
    const uint32_t h = (uint32_t)(i * 2654435761ull) ^ (uint32_t)(i >> 32);
    v0 += v[i];
    v1 += (1. + h * 2.3283064365386963e-10) * v[i];
  }
  reduce_store(v0, v1, 2, part, gridDim.x);
}

// mode 0: out[0..ndot) = sums (no scalar logic)
// mode 1 (start of a recurrence): rz, rr from the partials of (r.z, r.r); stop when rr <= thr2
// mode 2 (alpha): pq from the partials of p.q; alpha = rz / pq; breakdown when pq <= 0
// mode 3 (beta): rz', rr; iteration counted; stop when rr <= thr2; beta = rz' / rz
__global__ __launch_bounds__(kB) void k_rcg_fin(const double* __restrict__ part, unsigned G, int ndot, int mode,
                                                double* __restrict__ S, double* __restrict__ out)
{
  if(mode >= 2 && S[R_STOP] != 0.) return;
  const double s0 = block_sum(part, G);
  const double s1 = ndot > 1 ? block_sum(part + G, G) : 0.;
  if(threadIdx.x != 0) return;
  if(mode == 0) {
    out[0] = s0;
    if(ndot > 1) out[1] = s1;
  } else if(mode == 1) {
    S[R_RZ] = s0;
    S[R_RR] = s1;
    S[R_STOP] = s1 <= S[R_THR2] ? 1. : 0.;
  } else if(mode == 2) {
    S[R_PQ] = s0;
    if(!(s0 > 0.)) {
      S[R_STOP] = 1.;
      S[R_BREAK] = 1.;
      return;
    }
    S[R_ALPHA] = S[R_RZ] / s0;
  } else {
    S[R_ITER] += 1.;
    S[R_RR] = s1;
    S[R_BETA] = s0 / S[R_RZ];
    S[R_RZ] = s0;
    if(s1 <= S[R_THR2]) S[R_STOP] = 1.;
  }
}

// x += alpha p; r -= alpha q; z = dinv r; partials of r.z and r.r
__global__ __launch_bounds__(kB) void k_rcg_update(uint64_t n, const double* __restrict__ S, const double* __restrict__ p,
                                                   const double* __restrict__ q, double* __restrict__ x,
                                                   double* __restrict__ r, double* __restrict__ z,
                                                   const double* __restrict__ dinv, double* __restrict__ part)
{
  if(S[R_STOP] != 0.) return;
  const double alpha = S[R_ALPHA];
  double v0 = 0., v1 = 0.;
  const uint64_t stride = (uint64_t)gridDim.x * kB;
  for(uint64_t i = (uint64_t)blockIdx.x * kB + threadIdx.x; i < n; i += stride) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * q[i];
    r[i] = ri;
    const double zi = dinv[i] * ri;
    z[i] = zi;
    v0 += ri * zi;
    v1 += ri * ri;
  }
  reduce_store(v0, v1, 2, part, gridDim.x);
}

// z = dinv r
__global__ void k_rcg_precond(uint64_t n, const double* __restrict__ r, const double* __restrict__ dinv,
                              double* __restrict__ z)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) z[i] = dinv[i] * r[i];
}

// mu = Einv raw (k x k, one block); skipped once the recurrence stopped when S is given
__global__ void k_rcg_mu(int k, const double* __restrict__ Einv, const double* __restrict__ raw, double* __restrict__ mu,
                         const double* __restrict__ S)
{
  if(S && S[R_STOP] != 0.) return;
  const int i = threadIdx.x;
  if(i >= k) return;
  double s = 0.;
  for(int j = 0; j < k; ++j) s += Einv[i * k + j] * raw[j];
  mu[i] = s;
}

// p_next = z + beta p - W mu  (beta taken as 0 when S is null: the first direction of a recurrence)
__global__ void k_rcg_dir(uint64_t n, const double* __restrict__ S, const double* __restrict__ z,
                          const double* __restrict__ p, double* __restrict__ pn, const double* __restrict__ W, int k,
                          const double* __restrict__ mu)
{
  if(S && S[R_STOP] != 0.) return;
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  double v = z[i];
  if(S) v += S[R_BETA] * p[i];
  for(int l = 0; l < k; ++l) v -= mu[l] * W[(uint64_t)l * n + i];
  pn[i] = v;
}

// x += W mu, r -= AW mu
__global__ void k_rcg_project(uint64_t n, const double* __restrict__ W, const double* __restrict__ AW, int k,
                              const double* __restrict__ mu, double* __restrict__ x, double* __restrict__ r)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  double v = x[i], w = r[i];
  for(int l = 0; l < k; ++l) {
    v += mu[l] * W[(uint64_t)l * n + i];
    w -= mu[l] * AW[(uint64_t)l * n + i];
  }
  x[i] = v;
  r[i] = w;
}

// part[c][i][j] = sum over the rows of chunk c of X_i Y_j (X: na columns, Y: nb columns, length n each)
__global__ __launch_bounds__(kB) void k_rcg_gram(uint64_t n, uint64_t rows_per_chunk, int na, int nb,
                                                 const double* __restrict__ X, const double* __restrict__ Y,
                                                 double* __restrict__ part)
{
  __shared__ double xs[kGramRows][kTile + 1];
  __shared__ double ys[kGramRows][kTile + 1];
  const int i0 = blockIdx.x * kTile, j0 = blockIdx.y * kTile;
  const uint64_t r0 = (uint64_t)blockIdx.z * rows_per_chunk;
  const uint64_t r1 = std::min<uint64_t>(n, r0 + rows_per_chunk);
  const int t = threadIdx.x;
  const int ti = t >> 3, tj = (t & 7) * 4;
  double acc[4] = {0., 0., 0., 0.};
  for(uint64_t rb = r0; rb < r1; rb += kGramRows) {
    const int lr = t & 63;
    const uint64_t row = rb + lr;
    for(int cq = t >> 6; cq < kTile; cq += kB / 64) {
      const int ci = i0 + cq, cj = j0 + cq;
      xs[lr][cq] = (row < r1 && ci < na) ? X[(uint64_t)ci * n + row] : 0.;
      ys[lr][cq] = (row < r1 && cj < nb) ? Y[(uint64_t)cj * n + row] : 0.;
    }
    __syncthreads();
    for(int l = 0; l < kGramRows; ++l) {
      const double xv = xs[l][ti];
#pragma unroll
      for(int q = 0; q < 4; ++q) acc[q] += xv * ys[l][tj + q];
    }
    __syncthreads();
  }
  const int i = i0 + ti;
  if(i >= na) return;
  double* out = part + (uint64_t)blockIdx.z * na * nb;
#pragma unroll
  for(int q = 0; q < 4; ++q)
    if(j0 + tj + q < nb) out[(uint64_t)i * nb + j0 + tj + q] = acc[q];
}

__global__ void k_rcg_gram_fin(int na, int nb, unsigned chunks, const double* __restrict__ part, double* __restrict__ out)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if(e >= na * nb) return;
  double s = 0.;
  for(unsigned c = 0; c < chunks; ++c) s += part[(uint64_t)c * na * nb + e];
  out[e] = s;
}

// out_c = sum_l Y[l][c] V_l for c < kn (Y row-major nv x kn), one row per thread
template<int KMAX>
__global__ void k_rcg_combine(uint64_t n, int nv, int kn, const double* __restrict__ V, const double* __restrict__ Y,
                              double* __restrict__ out)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(i >= n) return;
  double acc[KMAX];
#pragma unroll
  for(int c = 0; c < KMAX; ++c) acc[c] = 0.;
  for(int l = 0; l < nv; ++l) {
    const double v = V[(uint64_t)l * n + i];
#pragma unroll
    for(int c = 0; c < KMAX; ++c)
      if(c < kn) acc[c] += Y[l * kn + c] * v;
  }
#pragma unroll
  for(int c = 0; c < KMAX; ++c)
    if(c < kn) out[(uint64_t)c * n + i] = acc[c];
}

// ------------------------------------------------------------------------------------------------
// small dense linear algebra on the host (row-major)
// ------------------------------------------------------------------------------------------------

// lower Cholesky factor of the symmetric positive definite a (n x n); false when a pivot is not
// positive
bool cholesky(std::vector<double>& a, int n)
{
  for(int j = 0; j < n; ++j) {
    double d = a[j * n + j];
    for(int l = 0; l < j; ++l) d -= a[j * n + l] * a[j * n + l];
    if(!(d > 0.)) return false;
    d = std::sqrt(d);
    a[j * n + j] = d;
    for(int i = j + 1; i < n; ++i) {
      double s = a[i * n + j];
      for(int l = 0; l < j; ++l) s -= a[i * n + l] * a[j * n + l];
      a[i * n + j] = s / d;
    }
    for(int i = 0; i < j; ++i) a[i * n + j] = 0.;
  }
  return true;
}

// inverse of the symmetric positive definite a from its Cholesky factor; false when not definite
bool spd_inverse(const std::vector<double>& a, int n, std::vector<double>& inv)
{
  std::vector<double> L(a);
  if(!cholesky(L, n)) return false;
  inv.assign((size_t)n * n, 0.);
  std::vector<double> col(n);
  for(int c = 0; c < n; ++c) {
    // L y = e_c, then L^T x = y
    for(int i = 0; i < n; ++i) {
      double s = i == c ? 1. : 0.;
      for(int l = 0; l < i; ++l) s -= L[i * n + l] * col[l];
      col[i] = s / L[i * n + i];
    }
    for(int i = n - 1; i >= 0; --i) {
      double s = col[i];
      for(int l = i + 1; l < n; ++l) s -= L[l * n + i] * col[l];
      col[i] = s / L[i * n + i];
    }
    for(int i = 0; i < n; ++i) inv[i * n + c] = col[i];
  }
  for(int i = 0; i < n; ++i)
    for(int j = 0; j < i; ++j) inv[i * n + j] = inv[j * n + i] = 0.5 * (inv[i * n + j] + inv[j * n + i]);
  return true;
}

// eigen-decomposition of the symmetric a (n x n) by cyclic Jacobi rotations (Golub and Van Loan,
// Matrix Computations, section 8.5): eigenvalues in lam, eigenvectors in the columns of Z
void sym_eigen(std::vector<double> a, int n, std::vector<double>& lam, std::vector<double>& Z)
{
  Z.assign((size_t)n * n, 0.);
  for(int i = 0; i < n; ++i) Z[i * n + i] = 1.;
  double fro = 0.;
  for(double v : a) fro += v * v;
  for(int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.;
    for(int i = 0; i < n; ++i)
      for(int j = i + 1; j < n; ++j) off += a[i * n + j] * a[i * n + j];
    if(off <= 1e-30 * fro) break;
    for(int p = 0; p < n; ++p)
      for(int q = p + 1; q < n; ++q) {
        const double apq = a[p * n + q];
        if(apq == 0.) continue;
        const double theta = (a[q * n + q] - a[p * n + p]) / (2. * apq);
        const double t = (theta >= 0. ? 1. : -1.) / (std::fabs(theta) + std::sqrt(theta * theta + 1.));
        const double c = 1. / std::sqrt(t * t + 1.), s = t * c;
        for(int l = 0; l < n; ++l) {   // columns p, q
          const double alp = a[l * n + p], alq = a[l * n + q];
          a[l * n + p] = c * alp - s * alq;
          a[l * n + q] = s * alp + c * alq;
        }
        for(int l = 0; l < n; ++l) {   // rows p, q
          const double apl = a[p * n + l], aql = a[q * n + l];
          a[p * n + l] = c * apl - s * aql;
          a[q * n + l] = s * apl + c * aql;
        }
        for(int l = 0; l < n; ++l) {
          const double zlp = Z[l * n + p], zlq = Z[l * n + q];
          Z[l * n + p] = c * zlp - s * zlq;
          Z[l * n + q] = s * zlp + c * zlq;
        }
      }
  }
  lam.resize(n);
  for(int i = 0; i < n; ++i) lam[i] = a[i * n + i];
}

}  // namespace

// ------------------------------------------------------------------------------------------------
// the solver
// ------------------------------------------------------------------------------------------------

namespace {

struct RcgRun {
  Lss& L;
  RcgSpace& R;
  Context& c;
  uint64_t n;
  int m;          // cycle length
  int kmax;       // recycled vectors kept
  double* S;
  double *x, *b, *r, *z, *dinv, *raw, *mu, *part;
  unsigned G;

  RcgRun(Lss& L_, RcgSpace& R_, int m_, int kmax_) : L(L_), R(R_), c(*L_.ctx), n(L_.dofs), m(m_), kmax(kmax_)
  {
    S = R.sc.as<double>();
    x = L.sol.as<double>();
    b = L.rhs.as<double>();
    r = R.vec.as<double>();
    z = r + n;
    dinv = z + n;
    raw = R.small.as<double>();
    mu = raw + kMaxRecycle;
    part = c.red_partials.as<double>();
    G = std::min<unsigned>(kMaxRedBlocks, grid_for(n, kB));
  }
  double* V(int col) const { return R.V.as<double>() + (uint64_t)col * n; }
  // column of search direction j of the current cycle; with nothing to recycle there are no cycles and
  // two columns hold the current and the next direction
  int pcol(int j) const { return kmax > 0 ? R.k + j : (j & 1); }
  double* AV(int col) const { return R.AV.as<double>() + (uint64_t)col * n; }
  const double* Einv() const { return R.einv.as<double>(); }

  void read(double* host)
  {
    HIP_OK(hipMemcpyAsync(host, S, sizeof(double) * R_COUNT, hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
  }

  // G = X^T Y on the host (na x nb, row-major)
  void gram(int na, const double* X, int nb, const double* Y, std::vector<double>& out)
  {
    const unsigned chunks = std::min<unsigned>(kGramChunks, grid_for(n, 4096));
    uint64_t rows = (n + chunks - 1) / chunks;
    rows = (rows + kGramRows - 1) / kGramRows * kGramRows;
    const size_t need = sizeof(double) * ((size_t)chunks * na * nb + (size_t)na * nb);
    if(R.gram.bytes < need) R.gram.alloc(&c, need);
    double* gp = R.gram.as<double>();
    double* go = gp + (size_t)chunks * na * nb;
    const dim3 grid(grid_for(na, kTile), grid_for(nb, kTile), chunks);
    hipLaunchKernelGGL(k_rcg_gram, grid, dim3(kB), 0, c.stream, n, rows, na, nb, X, Y, gp);
    hipLaunchKernelGGL(k_rcg_gram_fin, dim3(grid_for((size_t)na * nb, 256)), dim3(256), 0, c.stream, na, nb, chunks,
                       (const double*)gp, go);
    HIP_OK(hipGetLastError());
    out.resize((size_t)na * nb);
    HIP_OK(hipMemcpyAsync(out.data(), go, sizeof(double) * na * nb, hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
  }

  void upload_einv(const std::vector<double>& Ei)
  {
    HIP_OK(hipMemcpyAsync(R.einv.ptr, Ei.data(), sizeof(double) * Ei.size(), hipMemcpyHostToDevice, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
  }

  // AW = A W and E = W^T A W for the current W (after a change of the matrix)
  void refresh_space()
  {
    for(int l = 0; l < R.k; ++l) lss_spmv(L, 1., V(l), 0., AV(l));
    std::vector<double> E;
    gram(R.k, V(0), R.k, AV(0), E);
    for(int i = 0; i < R.k; ++i)
      for(int j = 0; j < i; ++j) E[i * R.k + j] = E[j * R.k + i] = 0.5 * (E[i * R.k + j] + E[j * R.k + i]);
    std::vector<double> Ei;
    if(!spd_inverse(E, R.k, Ei)) { R.k = 0; return; }
    upload_einv(Ei);
  }

  // W <- the kmax harmonic Ritz vectors over [W, P_0 .. P_{s-1}] with the smallest values
  void update_space(int s)
  {
    const int nv = R.k + s;
    if(s <= 0 || kmax <= 0) return;
    std::vector<double> Gm, Bm;
    gram(nv, AV(0), nv, AV(0), Gm);   // (AV)^T AV
    gram(nv, V(0), nv, AV(0), Bm);    // V^T A V
    std::vector<double> d(nv);
    for(int i = 0; i < nv; ++i) {
      if(!(Bm[i * nv + i] > 0.)) return;   // a direction with no energy: keep the old space
      d[i] = 1. / std::sqrt(Bm[i * nv + i]);
    }
    std::vector<double> Bs((size_t)nv * nv), Gs((size_t)nv * nv);
    for(int i = 0; i < nv; ++i)
      for(int j = 0; j < nv; ++j) {
        Bs[i * nv + j] = 0.5 * (Bm[i * nv + j] + Bm[j * nv + i]) * d[i] * d[j];
        Gs[i * nv + j] = 0.5 * (Gm[i * nv + j] + Gm[j * nv + i]) * d[i] * d[j];
      }
    std::vector<double> Lc(Bs);
    if(!cholesky(Lc, nv)) return;
    // C = L^-1 Gs L^-T
    std::vector<double> T((size_t)nv * nv), C((size_t)nv * nv);
    for(int j = 0; j < nv; ++j)   // T = L^-1 Gs, column by column
      for(int i = 0; i < nv; ++i) {
        double v = Gs[i * nv + j];
        for(int l = 0; l < i; ++l) v -= Lc[i * nv + l] * T[l * nv + j];
        T[i * nv + j] = v / Lc[i * nv + i];
      }
    for(int i = 0; i < nv; ++i)   // C = T L^-T, row by row: C L^T = T
      for(int j = 0; j < nv; ++j) {
        double v = T[i * nv + j];
        for(int l = 0; l < j; ++l) v -= C[i * nv + l] * Lc[j * nv + l];
        C[i * nv + j] = v / Lc[j * nv + j];
      }
    for(int i = 0; i < nv; ++i)
      for(int j = 0; j < i; ++j) C[i * nv + j] = C[j * nv + i] = 0.5 * (C[i * nv + j] + C[j * nv + i]);
    std::vector<double> lam, Z;
    sym_eigen(C, nv, lam, Z);
    std::vector<int> order(nv);
    for(int i = 0; i < nv; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return lam[a] < lam[b]; });
    const int kn = std::min(kmax, nv);
    // Y = D L^-T Z_kn  (nv x kn, row-major)
    std::vector<double> Y((size_t)nv * kn);
    for(int c2 = 0; c2 < kn; ++c2) {
      const int e = order[c2];
      for(int i = nv - 1; i >= 0; --i) {   // L^T y = z
        double v = Z[i * nv + e];
        for(int l = i + 1; l < nv; ++l) v -= Lc[l * nv + i] * Y[l * kn + c2];
        Y[i * kn + c2] = v / Lc[i * nv + i];
      }
    }
    for(int i = 0; i < nv; ++i)
      for(int c2 = 0; c2 < kn; ++c2) Y[i * kn + c2] *= d[i];
    // E = Y^T B Y (the identity up to rounding: Y is B-orthonormal)
    std::vector<double> BY((size_t)nv * kn, 0.), E((size_t)kn * kn, 0.);
    for(int i = 0; i < nv; ++i)
      for(int c2 = 0; c2 < kn; ++c2) {
        double v = 0.;
        for(int l = 0; l < nv; ++l) v += 0.5 * (Bm[i * nv + l] + Bm[l * nv + i]) * Y[l * kn + c2];
        BY[i * kn + c2] = v;
      }
    for(int a = 0; a < kn; ++a)
      for(int c2 = 0; c2 < kn; ++c2) {
        double v = 0.;
        for(int i = 0; i < nv; ++i) v += Y[i * kn + a] * BY[i * kn + c2];
        E[a * kn + c2] = v;
      }
    for(int a = 0; a < kn; ++a)
      for(int c2 = 0; c2 < a; ++c2) E[a * kn + c2] = E[c2 * kn + a] = 0.5 * (E[a * kn + c2] + E[c2 * kn + a]);
    std::vector<double> Ei;
    if(!spd_inverse(E, kn, Ei)) return;
    // W = V Y, AW = AV Y through the scratch columns, then into the first kn columns
    double* Yd = R.small.as<double>() + 2 * kMaxRecycle;
    HIP_OK(hipMemcpyAsync(Yd, Y.data(), sizeof(double) * Y.size(), hipMemcpyHostToDevice, c.stream));
    double* tw = R.tmp.as<double>();
    double* taw = tw + (uint64_t)kMaxRecycle * n;
    const dim3 g(grid_for(n, 256));
    hipLaunchKernelGGL(k_rcg_combine<kMaxRecycle>, g, dim3(256), 0, c.stream, n, nv, kn, (const double*)V(0),
                       (const double*)Yd, tw);
    hipLaunchKernelGGL(k_rcg_combine<kMaxRecycle>, g, dim3(256), 0, c.stream, n, nv, kn, (const double*)AV(0),
                       (const double*)Yd, taw);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(V(0), tw, sizeof(double) * kn * n, hipMemcpyDeviceToDevice, c.stream));
    HIP_OK(hipMemcpyAsync(AV(0), taw, sizeof(double) * kn * n, hipMemcpyDeviceToDevice, c.stream));
    R.k = kn;
    upload_einv(Ei);
  }

  // p_0 = z - W E^-1 (AW)^T z into column k of V (z computed from r)
  void first_direction()
  {
    const dim3 g(grid_for(n, 256));
    hipLaunchKernelGGL(k_rcg_precond, g, dim3(256), 0, c.stream, n, (const double*)r, (const double*)dinv, z);
    if(R.k > 0) {
      arnoldi::multidot(c, n, R.k, AV(0), z, raw);
      hipLaunchKernelGGL(k_rcg_mu, dim3(1), dim3(64), 0, c.stream, R.k, Einv(), (const double*)raw, mu,
                         (const double*)nullptr);
    }
    hipLaunchKernelGGL(k_rcg_dot, dim3(G), dim3(kB), 0, c.stream, n, (const double*)r, (const double*)z,
                       (const double*)r, (const double*)r, part, (const double*)nullptr);
    hipLaunchKernelGGL(k_rcg_fin, dim3(1), dim3(kB), 0, c.stream, (const double*)part, G, 2, 1, S, (double*)nullptr);
    hipLaunchKernelGGL(k_rcg_dir, g, dim3(256), 0, c.stream, n, (const double*)nullptr, (const double*)z,
                       (const double*)nullptr, V(pcol(0)), (const double*)V(0), R.k, (const double*)mu);
    HIP_OK(hipGetLastError());
  }

  // iterations j0 .. j0 + nb - 1 of the current cycle
  void enqueue(int j0, int nb)
  {
    const dim3 g(grid_for(n, 256));
    for(int j = j0; j < j0 + nb; ++j) {
      double* p = V(pcol(j));
      double* q = AV(pcol(j));
      lss_spmv(L, 1., p, 0., q, S + R_STOP);
      hipLaunchKernelGGL(k_rcg_dot, dim3(G), dim3(kB), 0, c.stream, n, (const double*)p, (const double*)q,
                         (const double*)nullptr, (const double*)nullptr, part, (const double*)S);
      hipLaunchKernelGGL(k_rcg_fin, dim3(1), dim3(kB), 0, c.stream, (const double*)part, G, 1, 2, S, (double*)nullptr);
      hipLaunchKernelGGL(k_rcg_update, dim3(G), dim3(kB), 0, c.stream, n, (const double*)S, (const double*)p,
                         (const double*)q, x, r, z, (const double*)dinv, part);
      hipLaunchKernelGGL(k_rcg_fin, dim3(1), dim3(kB), 0, c.stream, (const double*)part, G, 2, 3, S, (double*)nullptr);
      if(j + 1 < m) {   // the next direction (the last of a cycle is rebuilt against the new space)
        if(R.k > 0) {
          arnoldi::multidot(c, n, R.k, AV(0), z, raw);
          hipLaunchKernelGGL(k_rcg_mu, dim3(1), dim3(64), 0, c.stream, R.k, Einv(), (const double*)raw, mu,
                             (const double*)S);
        }
        hipLaunchKernelGGL(k_rcg_dir, g, dim3(256), 0, c.stream, n, (const double*)S, (const double*)z, (const double*)p,
                           V(pcol(j + 1)), (const double*)V(0), R.k, (const double*)mu);
      }
    }
    HIP_OK(hipGetLastError());
  }

  // r = b - A x
  void true_residual()
  {
    HIP_OK(hipMemcpyAsync(r, b, sizeof(double) * n, hipMemcpyDeviceToDevice, c.stream));
    lss_spmv(L, -1., x, 1., r);
  }

  // the error component in span W: x += W E^-1 W^T r, r -= AW E^-1 W^T r (r must be b - A x)
  void project()
  {
    if(R.k == 0) return;
    arnoldi::multidot(c, n, R.k, V(0), r, raw);
    hipLaunchKernelGGL(k_rcg_mu, dim3(1), dim3(64), 0, c.stream, R.k, Einv(), (const double*)raw, mu,
                       (const double*)nullptr);
    hipLaunchKernelGGL(k_rcg_project, dim3(grid_for(n, 256)), dim3(256), 0, c.stream, n, (const double*)V(0),
                       (const double*)AV(0), R.k, (const double*)mu, x, r);
    HIP_OK(hipGetLastError());
  }

  // mode-0 reduction of one or two dot products (partials in part, G blocks), read back
  void sums(int ndot, unsigned Gp, double* h)
  {
    double* out = R.small.as<double>() + 2 * kMaxRecycle + (kMaxRecycle + kMaxCycle) * kMaxRecycle;
    hipLaunchKernelGGL(k_rcg_fin, dim3(1), dim3(kB), 0, c.stream, (const double*)part, Gp, ndot, 0, S, out);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(h, out, sizeof(double) * ndot, hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
  }
  double dot(const double* a0, const double* b0)
  {
    hipLaunchKernelGGL(k_rcg_dot, dim3(G), dim3(kB), 0, c.stream, n, a0, b0, (const double*)nullptr,
                       (const double*)nullptr, part, (const double*)nullptr);
    double h[2];
    sums(1, G, h);
    return h[0];
  }
  // fingerprint of the matrix values (a space is reused on the matrix it was built on)
  void fingerprint(double out[2])
  {
    const uint64_t nv = L.n_blocks * (uint64_t)L.neq * L.neq;
    const unsigned Gp = std::min<unsigned>(kMaxRedBlocks, grid_for(nv, kB));
    hipLaunchKernelGGL(k_rcg_print, dim3(Gp), dim3(kB), 0, c.stream, nv, L.values.as<double>(), part);
    sums(2, Gp, out);
  }
};

}  // namespace

void rcg_solve(Lss& L, const cf3hip_solver_params& prm, cf3hip_solve_stats& st)
{
  Context& c = *L.ctx;
  st = cf3hip_solve_stats{0, 0, 0, 0, 0.};
  if(prm.precision != CF3HIP_FP64) CF3_THROW(CF3HIP_ERR_NOT_IMPLEMENTED, "RCG runs in FP64");
  if(prm.rcg_space < 0 || prm.rcg_space >= kRcgSpaces)
    CF3_THROW(CF3HIP_ERR_INVALID_ARG, "rcg_space must lie in [0, 2)");
  const int kmax = prm.recycle;
  if(prm.restart > kMaxCycle) CF3_THROW(CF3HIP_ERR_LIMIT, "RCG cycle length (restart) above the maximum of 96");
  if(kmax < 0 || kmax > kMaxRecycle) CF3_THROW(CF3HIP_ERR_LIMIT, "RCG recycled vectors outside [0, 32]");
  // nothing to recycle: one uninterrupted Jacobi CG recurrence (no cycles)
  const int m = kmax == 0 ? prm.max_iter + 1 : (prm.restart > 0 ? prm.restart : kDefaultCycle);
  const uint64_t n = L.dofs;
  if(!L.rcg[prm.rcg_space]) L.rcg[prm.rcg_space].reset(new RcgSpace());
  RcgSpace& R = *L.rcg[prm.rcg_space];
  if(R.n != n || !R.V.ptr) {   // (re)allocate for this system
    R.V.alloc(&c, sizeof(double) * (kMaxRecycle + kMaxCycle + 1) * n);
    R.AV.alloc(&c, sizeof(double) * (kMaxRecycle + kMaxCycle + 1) * n);
    R.tmp.alloc(&c, sizeof(double) * 2 * kMaxRecycle * n);
    R.vec.alloc(&c, sizeof(double) * 3 * n);
    R.sc.alloc(&c, sizeof(double) * R_COUNT);
    R.small.alloc(&c, sizeof(double) * (2 * kMaxRecycle + (kMaxRecycle + kMaxCycle) * kMaxRecycle + 2));
    R.einv.alloc(&c, sizeof(double) * kMaxRecycle * kMaxRecycle);
    R.n = n;
    R.k = 0;
  }
  if(R.k > kmax) R.k = 0;   // fewer recycled vectors requested than held: start afresh
  RcgRun run(L, R, m, kmax);
  int spmvs = 0;
  double fp[2];
  run.fingerprint(fp);
  if(R.k > 0 && (fp[0] != R.print[0] || fp[1] != R.print[1])) {
    run.refresh_space();
    spmvs += R.k;
  }
  R.print[0] = fp[0];
  R.print[1] = fp[1];
  hipLaunchKernelGGL(k_rcg_jacobi, dim3(grid_for(n, 256)), dim3(256), 0, c.stream, n, L.neq, L.diag_blk.as<uint32_t>(),
                     L.values.as<double>(), run.dinv);
  HIP_OK(hipGetLastError());
  const double bn2 = run.dot(run.b, run.b);
  if(bn2 == 0.) {   // b = 0: x = 0
    lss_fill(c, run.x, n, 0.);
    HIP_OK(hipStreamSynchronize(c.stream));
    st.converged = 1;
    st.outer = 1;
    return;
  }
  const double thr2 = prm.tol * prm.tol * bn2;
  {
    double h[R_COUNT] = {};
    h[R_THR2] = thr2;
    h[R_BN2] = bn2;
    HIP_OK(hipMemcpyAsync(run.S, h, sizeof(double) * R_COUNT, hipMemcpyHostToDevice, c.stream));
  }
  run.true_residual();
  ++spmvs;
  int it = 0, cycles = 0, restarts = 0;
  bool conv = false;
  double h[R_COUNT];
  while(true) {
    run.project();
    run.first_direction();
    run.read(h);
    int base = (int)h[R_ITER], j = 0, batch = 1;
    bool stop = h[R_STOP] != 0.;
    while(!stop && it < prm.max_iter) {
      const int nb = std::min(std::min(batch, m - j), prm.max_iter - it);
      run.enqueue(j, nb);
      run.read(h);
      const int done = (int)h[R_ITER] - base;
      it += done - j;
      spmvs += done - j;
      j = done;
      stop = h[R_STOP] != 0.;
      if(stop) break;
      batch = std::min(2 * batch, 16);
      if(j == m) {   // end of a cycle: a new space, then a fresh recurrence from the true residual
        ++cycles;
        run.update_space(j);
        run.true_residual();
        ++spmvs;
        run.project();
        run.first_direction();
        run.read(h);
        base = (int)h[R_ITER];
        j = 0;
        batch = 1;
        stop = h[R_STOP] != 0.;
      }
    }
    ++cycles;
    run.update_space(j);   // what this solve learned about A, for the next one
    run.true_residual();   // the true residual decides
    ++spmvs;
    const double rr = run.dot(run.r, run.r);
    st.rel_residual = std::sqrt(rr / bn2);
    conv = rr <= thr2;
    if(conv || h[R_BREAK] != 0. || it >= prm.max_iter || restarts >= 3) break;
    ++restarts;   // the recursive residual drifted below tol: continue from the true one
  }
  st.iterations = it;
  st.outer = cycles;
  st.spmvs = spmvs;
  st.converged = conv ? 1 : 0;
}

}  // namespace cf3hip

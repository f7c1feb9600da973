// Restarted GMRES(m) with ILU(0) right preconditioning on the device (gfx950), FP64.
//
// Reference: the default Trilinos strategy is Belos GMRES with an Ifpack ILU preconditioner rebuilt at
// every solve (cf3/math/LSS/Trilinos/TrilinosStratimikosStrategy.cpp:135-181,
// BelosGMRESParameters.cpp:28-41).  Belos/Ifpack are third-party; this is the textbook method:
// ILU(0) (IKJ variant) on the point matrix, classical Gram-Schmidt applied twice, Givens rotations,
// stop on ||b - A x|| <= tol ||b|| -- restated identically in the oracle.
//
// ILU(0) factorisation and the triangular solves use level schedules (rows of one level only read rows
// of earlier levels); one workgroup walks the levels with barriers.  These kernels serve the small
// two-dimensional system of the lid-driven cavity case; the large systems use BiCGStab.
#include <algorithm>
#include <cmath>

#include "arnoldi.hpp"
#include "internal.hpp"

namespace cf3hip {

namespace {

constexpr int kWG = 1024;

__device__ inline uint32_t find_col(const uint32_t* col, uint32_t lo, uint32_t hi, uint32_t c)
{
  while(lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if(col[mid] < c) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// levels (single thread: a dynamic programme in row order) and rows sorted by level
__global__ void k_levels(uint32_t n, const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                         uint32_t* __restrict__ diag, uint32_t* __restrict__ lev, uint32_t* __restrict__ ulev,
                         uint32_t* __restrict__ lorder, uint32_t* __restrict__ lptr, uint32_t* __restrict__ uorder,
                         uint32_t* __restrict__ uptr, uint32_t* __restrict__ counts)
{
  if(threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t lmax = 0, umax = 0;
  for(uint32_t i = 0; i < n; ++i) {
    const uint32_t d = find_col(col, rp[i], rp[i + 1], i);
    diag[i] = d;
    uint32_t l = 0;
    for(uint32_t q = rp[i]; q < d; ++q) l = max(l, lev[col[q]] + 1);
    lev[i] = l;
    lmax = max(lmax, l);
  }
  for(uint32_t ii = n; ii-- > 0;) {
    uint32_t l = 0;
    for(uint32_t q = diag[ii] + 1; q < rp[ii + 1]; ++q) l = max(l, ulev[col[q]] + 1);
    ulev[ii] = l;
    umax = max(umax, l);
  }
  for(uint32_t l = 0; l <= lmax + 1; ++l) lptr[l] = 0;
  for(uint32_t i = 0; i < n; ++i) ++lptr[lev[i] + 1];
  for(uint32_t l = 0; l <= lmax; ++l) lptr[l + 1] += lptr[l];
  for(uint32_t l = 0; l <= lmax; ++l) counts[l] = lptr[l];
  for(uint32_t i = 0; i < n; ++i) lorder[counts[lev[i]]++] = i;
  for(uint32_t l = 0; l <= umax + 1; ++l) uptr[l] = 0;
  for(uint32_t i = 0; i < n; ++i) ++uptr[ulev[i] + 1];
  for(uint32_t l = 0; l <= umax; ++l) uptr[l + 1] += uptr[l];
  for(uint32_t l = 0; l <= umax; ++l) counts[l] = uptr[l];
  for(uint32_t i = 0; i < n; ++i) uorder[counts[ulev[i]]++] = i;
  counts[0] = lmax + 1;
  counts[1] = umax + 1;
}

__global__ __launch_bounds__(kWG) void k_ilu0_factor(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                                                    double* __restrict__ val, const uint32_t* __restrict__ diag,
                                                    const uint32_t* __restrict__ lorder, const uint32_t* __restrict__ lptr,
                                                    uint32_t nlev)
{
  for(uint32_t l = 0; l < nlev; ++l) {
    for(uint32_t t = lptr[l] + threadIdx.x; t < lptr[l + 1]; t += kWG) {
      const uint32_t i = lorder[t];
      for(uint32_t q = rp[i]; q < diag[i]; ++q) {
        const uint32_t k = col[q];
        const double lik = val[q] / val[diag[k]];
        val[q] = lik;
        for(uint32_t r = diag[k] + 1; r < rp[k + 1]; ++r) {
          const uint32_t j = col[r];
          const uint32_t pos = find_col(col, q + 1, rp[i + 1], j);
          if(pos < rp[i + 1] && col[pos] == j) val[pos] -= lik * val[r];
        }
      }
    }
    __syncthreads();
  }
}

// x = U^-1 L^-1 b
__global__ __launch_bounds__(kWG) void k_ilu0_solve(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                                                   const double* __restrict__ val, const uint32_t* __restrict__ diag,
                                                   const uint32_t* __restrict__ lorder, const uint32_t* __restrict__ lptr,
                                                   uint32_t nlev, const uint32_t* __restrict__ uorder,
                                                   const uint32_t* __restrict__ uptr, uint32_t nulev,
                                                   const double* __restrict__ b, double* __restrict__ x)
{
  for(uint32_t l = 0; l < nlev; ++l) {
    for(uint32_t t = lptr[l] + threadIdx.x; t < lptr[l + 1]; t += kWG) {
      const uint32_t i = lorder[t];
      double s = b[i];
      for(uint32_t q = rp[i]; q < diag[i]; ++q) s -= val[q] * x[col[q]];
      x[i] = s;
    }
    __syncthreads();
  }
  for(uint32_t l = 0; l < nulev; ++l) {
    for(uint32_t t = uptr[l] + threadIdx.x; t < uptr[l + 1]; t += kWG) {
      const uint32_t i = uorder[t];
      double s = x[i];
      for(uint32_t q = diag[i] + 1; q < rp[i + 1]; ++q) s -= val[q] * x[col[q]];
      x[i] = s / val[diag[i]];
    }
    __syncthreads();
  }
}

}  // namespace

void gmres_ilu0(Lss& L, const cf3hip_solver_params& prm, cf3hip_solve_stats& st)
{
  using namespace arnoldi;
  Context& c = *L.ctx;
  if(prm.precision != CF3HIP_FP64) CF3_THROW(CF3HIP_ERR_NOT_IMPLEMENTED, "GMRES+ILU0 runs in FP64");
  const int m = prm.restart;
  if(m < 1 || m > arnoldi::kMaxV - 1) CF3_THROW(CF3HIP_ERR_INVALID_ARG, "GMRES restart must be in [1, 32]");
  const uint64_t n = L.dofs;
  const uint64_t nnz = L.n_blocks * L.neq * L.neq;
  if(n >= 0xffffffffull || nnz >= 0xffffffffull) CF3_THROW(CF3HIP_ERR_LIMIT, "system too large for GMRES+ILU0");
  // point CSR copy of the matrix, factorised in place (rebuilt at every solve like the reference)
  DevBuf prow, pcol, pval, diag, lev, ulev, lorder, lptr, uorder, uptr, counts;
  prow.alloc(&c, sizeof(uint32_t) * (n + 1));
  pcol.alloc(&c, sizeof(uint32_t) * nnz);
  pval.alloc(&c, sizeof(double) * nnz);
  diag.alloc(&c, sizeof(uint32_t) * n);
  lev.alloc(&c, sizeof(uint32_t) * n);
  ulev.alloc(&c, sizeof(uint32_t) * n);
  lorder.alloc(&c, sizeof(uint32_t) * n);
  uorder.alloc(&c, sizeof(uint32_t) * n);
  lptr.alloc(&c, sizeof(uint32_t) * (n + 2));
  uptr.alloc(&c, sizeof(uint32_t) * (n + 2));
  counts.alloc(&c, sizeof(uint32_t) * (n + 2));
  lss_point_csr(L, prow.as<uint32_t>(), pcol.as<uint32_t>(), pval.as<double>());
  hipLaunchKernelGGL(k_levels, dim3(1), dim3(1), 0, c.stream, (uint32_t)n, prow.as<uint32_t>(), pcol.as<uint32_t>(),
                     diag.as<uint32_t>(), lev.as<uint32_t>(), ulev.as<uint32_t>(), lorder.as<uint32_t>(),
                     lptr.as<uint32_t>(), uorder.as<uint32_t>(), uptr.as<uint32_t>(), counts.as<uint32_t>());
  HIP_OK(hipGetLastError());
  uint32_t nl[2];
  HIP_OK(hipMemcpyAsync(nl, counts.ptr, sizeof(nl), hipMemcpyDeviceToHost, c.stream));
  HIP_OK(hipStreamSynchronize(c.stream));
  hipLaunchKernelGGL(k_ilu0_factor, dim3(1), dim3(kWG), 0, c.stream, prow.as<uint32_t>(), pcol.as<uint32_t>(),
                     pval.as<double>(), diag.as<uint32_t>(), lorder.as<uint32_t>(), lptr.as<uint32_t>(), nl[0]);
  HIP_OK(hipGetLastError());
  auto precond = [&](const double* b, double* x) {
    hipLaunchKernelGGL(k_ilu0_solve, dim3(1), dim3(kWG), 0, c.stream, prow.as<uint32_t>(), pcol.as<uint32_t>(),
                       pval.as<double>(), diag.as<uint32_t>(), lorder.as<uint32_t>(), lptr.as<uint32_t>(), nl[0],
                       uorder.as<uint32_t>(), uptr.as<uint32_t>(), nl[1], b, x);
  };

  // workspace: V (m+1) n, w, z, u ; small: bookkeeping, h1, h2, y, scalars
  DevBuf V, w, z, small;
  V.alloc(&c, sizeof(double) * (m + 1) * n);
  w.alloc(&c, sizeof(double) * n);
  z.alloc(&c, sizeof(double) * n);
  small.alloc(&c, sizeof(double) * GmresDev::size(m));
  double* base = small.as<double>();
  const GmresDev gd{m, base};
  double* h1 = gd.h1();
  double* h2 = gd.h2();
  double* y = gd.y();
  double* sc = c.scalars.as<double>();   // sc[0] |w|^2, sc[1] ||b||^2
  const unsigned gb = grid_for(n, 256);
  double* x = L.sol.as<double>();
  const double* b = L.rhs.as<double>();
  double* Vd = V.as<double>();

  multidot(c, n, 1, b, b, sc + 1);
  double bn2 = 0.;
  HIP_OK(hipMemcpyAsync(&bn2, sc + 1, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIP_OK(hipStreamSynchronize(c.stream));
  double bnorm = std::sqrt(bn2);
  if(bnorm == 0.) bnorm = 1.;
  int it = 0, spmvs = 0;
  // a cycle ends when the recurrence residual meets the tolerance (or at the restart length); the
  // true residual b - A x computed at the start of the next cycle decides convergence
  while(it < prm.max_iter) {
    // r = b - A x ; V_0 = r / ||r||
    lss_spmv(L, 1., x, 0., w.as<double>());
    ++spmvs;
    hipLaunchKernelGGL(k_sub, dim3(gb), dim3(256), 0, c.stream, n, b, w.as<double>());
    multidot(c, n, 1, w.as<double>(), w.as<double>(), sc);
    hipLaunchKernelGGL(k_gmres_start, dim3(1), dim3(1), 0, c.stream, m, base, sc);
    double beta = 0.;
    HIP_OK(hipMemcpyAsync(&beta, base, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    if(beta <= prm.tol * bnorm) break;
    hipLaunchKernelGGL(k_scale, dim3(gb), dim3(256), 0, c.stream, n, w.as<double>(), base, Vd);
    int j = 0;
    for(; j < m && it < prm.max_iter; ++j) {
      ++it;
      precond(Vd + (uint64_t)j * n, z.as<double>());
      lss_spmv(L, 1., z.as<double>(), 0., w.as<double>());
      ++spmvs;
      multidot(c, n, j + 1, Vd, w.as<double>(), h1);
      hipLaunchKernelGGL(k_orth_update, dim3(gb), dim3(256), 0, c.stream, n, j + 1, Vd, h1, w.as<double>());
      multidot(c, n, j + 1, Vd, w.as<double>(), h2);
      hipLaunchKernelGGL(k_orth_update, dim3(gb), dim3(256), 0, c.stream, n, j + 1, Vd, h2, w.as<double>());
      multidot(c, n, 1, w.as<double>(), w.as<double>(), sc);
      hipLaunchKernelGGL(k_arnoldi_close, dim3(1), dim3(1), 0, c.stream, m, base, j, h1, h2, sc);
      // V_{j+1} = w / ||w||  (||w|| kept on the device in base[1])
      hipLaunchKernelGGL(k_scale, dim3(gb), dim3(256), 0, c.stream, n, w.as<double>(), base + 1,
                         Vd + (uint64_t)(j + 1) * n);
      HIP_OK(hipGetLastError());
      double res = 0.;
      HIP_OK(hipMemcpyAsync(&res, base, sizeof(double), hipMemcpyDeviceToHost, c.stream));
      HIP_OK(hipStreamSynchronize(c.stream));
      if(res <= prm.tol * bnorm) { ++j; break; }
    }
    hipLaunchKernelGGL(k_back_subst, dim3(1), dim3(1), 0, c.stream, m, base, j, y);
    hipLaunchKernelGGL(k_lincomb, dim3(gb), dim3(256), 0, c.stream, n, j, Vd, y, w.as<double>());
    precond(w.as<double>(), z.as<double>());
    hipLaunchKernelGGL(k_add, dim3(gb), dim3(256), 0, c.stream, n, z.as<double>(), x);
    HIP_OK(hipGetLastError());
  }
  st.iterations = it;
  st.outer = 1;
  st.spmvs = spmvs;
  st.rel_residual = lss_residual(L);
  st.converged = st.rel_residual <= prm.tol ? 1 : 0;
}

}  // namespace cf3hip

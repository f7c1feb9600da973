// Restarted GMRES(m) with ILU(0) right preconditioning on the device (gfx950), FP64.
//
// Reference: the default Trilinos strategy is Belos GMRES with an Ifpack ILU preconditioner rebuilt at
// every solve (cf3/math/LSS/Trilinos/TrilinosStratimikosStrategy.cpp:135-181,
// BelosGMRESParameters.cpp:28-41).  Belos/Ifpack are third-party; this is the textbook method:
// ILU(0) (IKJ variant) on the point matrix, classical Gram-Schmidt applied twice, Givens rotations,
// stop on ||b - A x|| <= tol ||b|| -- restated identically in the oracle.
//
// ILU(0) factorisation and the triangular solves use level schedules (rows of one level only read rows
// of earlier levels); one workgroup walks the levels with barriers.  The schedules depend on the pattern
// only and are built once per system on the host.  Systems of up to kSmallMaxN dofs (the two-dimensional
// cavity and Taylor-Green cases) run the whole restarted solve in one launch of one workgroup
// (k_gmres_small): the preconditioned vector lives in LDS, so the triangular solves and the product
// read their operand from LDS, every reduction is a workgroup reduction in a fixed order, and the host
// waits once per solve; its triangular solves walk node levels (the point rows of a node together).
// Larger systems take the multi-launch loop (device-wide multi-dots) and the point-level solves.
#include <algorithm>
#include <cmath>
#include <vector>

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

// Rows are taken by groups of kGL lanes (kWG / kGL rows of a level at a time).  A row of at most kSeg
// entries is staged in LDS; its lower entries are eliminated in column order, the lanes of the group
// sharing the row k of each step: l_ik = a_ik / u_kk, then a_ij -= l_ik u_kj over the upper entries
// u_kj of row k (lane-strided; each j appears once, so the updates never collide), then the row is
// written back.  Longer rows are eliminated by the group's first lane in global memory.
constexpr int kGL = 16;                 // lanes per row
constexpr int kGroups = kWG / kGL;
constexpr uint32_t kSeg = 48;           // staged entries per row (48 doubles + 48 columns per group)

__device__ inline void group_sync() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); __builtin_amdgcn_wave_barrier(); }

__device__ inline uint32_t find_lds(const uint32_t* c, uint32_t lo, uint32_t hi, uint32_t j)
{
  while(lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if(c[mid] < j) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// The update lists of the elimination (pattern only, built with the level schedules): lower entry q of
// row i (column k) takes the pivot u_kk at fpiv[q] and updates a_ij -= l_ik u_kj for the pairs p in
// [fptr[q], fptr[q+1]): u_kj at fsrc[p], a_ij at offset fdst[p] of row i.  A row whose lists fit
// kPairs is staged whole (its entries, pivots and the u_kj) by one batch of loads, then eliminated in
// LDS; other rows search row k's columns.
constexpr uint32_t kPairs = 128;
struct FactorPlan {
  const uint32_t *fptr, *fpiv, *fsrc;
  const uint16_t* fdst;
};

__global__ __launch_bounds__(kWG) void k_ilu0_factor(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                                                    double* __restrict__ val, const uint32_t* __restrict__ diag,
                                                    const uint32_t* __restrict__ lorder, const uint32_t* __restrict__ lptr,
                                                    uint32_t nlev, FactorPlan fp)
{
#pragma clang fp contract(off)   // the oracle's arithmetic: l_ik = a_ik / u_kk, then a_ij - (l_ik u_kj)
  __shared__ double sv[kGroups][kSeg];
  __shared__ uint32_t sc[kGroups][kSeg];
  __shared__ double spv[kGroups][kPairs];   // staged rows: pivots (first nlow), then the u_kj
  __shared__ uint16_t spd[kGroups][kPairs];
  const uint32_t g = threadIdx.x / kGL, lane = threadIdx.x % kGL;
  double* rv = sv[g];
  uint32_t* rc = sc[g];
  for(uint32_t l = 0; l < nlev; ++l) {
    for(uint32_t t = lptr[l] + g; t < lptr[l + 1]; t += kGroups) {
      const uint32_t i = lorder[t];
      const uint32_t q0 = rp[i], q1 = rp[i + 1], d = diag[i];
      const uint32_t len = q1 - q0, nlow = d - q0;
      const uint32_t pb = fp.fptr[q0], np = fp.fptr[d] - pb;
      if(len <= kSeg && nlow + np <= kPairs) {
        double* pv = spv[g];
        uint16_t* pd = spd[g];
        uint32_t* pp = rc;   // pair offsets of each lower entry (nlow + 1 <= kSeg)
        // every load of the row at once: indices past a count read its last entry (valid), and only
        // the entries within the counts are stored
        constexpr int kRowIt = (int)(kSeg + kGL - 1) / kGL, kPairIt = (int)kPairs / kGL;
        double ra[kRowIt], pa[kRowIt], qa[kPairIt];
        uint32_t oa[kRowIt + 1], sa[kPairIt];
        uint16_t da[kPairIt];
        const uint32_t lowm = nlow ? nlow - 1 : 0;
#pragma unroll
        for(int m = 0; m < kRowIt; ++m) {
          const uint32_t p = lane + m * kGL;
          ra[m] = val[q0 + (p < len ? p : len - 1)];
          pa[m] = val[fp.fpiv[q0 + (p < lowm ? p : lowm)]];
        }
#pragma unroll
        for(int m = 0; m <= kRowIt; ++m) {
          const uint32_t p = lane + m * kGL;
          oa[m] = fp.fptr[q0 + (p < nlow ? p : nlow)];
        }
        if(np) {
#pragma unroll
          for(int m = 0; m < kPairIt; ++m) {
            const uint32_t p = lane + m * kGL, pc = pb + (p < np ? p : np - 1);
            sa[m] = fp.fsrc[pc];
            da[m] = fp.fdst[pc];
          }
#pragma unroll
          for(int m = 0; m < kPairIt; ++m) qa[m] = val[sa[m]];
        }
#pragma unroll
        for(int m = 0; m < kRowIt; ++m) {
          const uint32_t p = lane + m * kGL;
          if(p < len) rv[p] = ra[m];
          if(p < nlow) pv[p] = pa[m];
        }
#pragma unroll
        for(int m = 0; m <= kRowIt; ++m) {
          const uint32_t p = lane + m * kGL;
          if(p <= nlow) pp[p] = oa[m] - pb;
        }
        if(np) {
#pragma unroll
          for(int m = 0; m < kPairIt; ++m) {
            const uint32_t p = lane + m * kGL;
            if(p < np) {
              pv[nlow + p] = qa[m];
              pd[p] = da[m];
            }
          }
        }
        group_sync();
        for(uint32_t qq = 0; qq < nlow; ++qq) {
          const double lik = rv[qq] / pv[qq];
          group_sync();
          if(lane == 0) rv[qq] = lik;
          for(uint32_t p = pp[qq] + lane; p < pp[qq + 1]; p += kGL) rv[pd[p]] -= lik * pv[nlow + p];
          group_sync();
        }
        for(uint32_t p = lane; p < len; p += kGL) val[q0 + p] = rv[p];
        group_sync();   // rc (pp) and the staging are reused by the group's next row
      } else if(len <= kSeg) {
        for(uint32_t p = lane; p < len; p += kGL) { rv[p] = val[q0 + p]; rc[p] = col[q0 + p]; }
        group_sync();
        for(uint32_t qq = 0; qq < nlow; ++qq) {
          const uint32_t k = rc[qq];
          const double lik = rv[qq] / val[diag[k]];
          group_sync();
          if(lane == 0) rv[qq] = lik;
          for(uint32_t r = diag[k] + 1 + lane; r < rp[k + 1]; r += kGL) {
            const uint32_t j = col[r];
            const uint32_t pos = find_lds(rc, qq + 1, len, j);
            if(pos < len && rc[pos] == j) rv[pos] -= lik * val[r];
          }
          group_sync();
        }
        for(uint32_t p = lane; p < len; p += kGL) val[q0 + p] = rv[p];
        group_sync();
      } else if(lane == 0) {
        for(uint32_t q = q0; q < d; ++q) {
          const uint32_t k = col[q];
          const double lik = val[q] / val[diag[k]];
          val[q] = lik;
          for(uint32_t r = diag[k] + 1; r < rp[k + 1]; ++r) {
            const uint32_t j = col[r];
            const uint32_t pos = find_col(col, q + 1, q1, j);
            if(pos < q1 && col[pos] == j) val[pos] -= lik * val[r];
          }
        }
      }
    }
    __syncthreads();
  }
}



// ---- the whole solve in one workgroup (systems of at most kSmallMaxN dofs) ----
constexpr int kSW = 512;                        // threads of the single workgroup (256 VGPRs per lane)
constexpr uint32_t kSmallMaxN = 4096;           // dofs of one LDS vector (32 KB)
constexpr int kSR = (int)(kSmallMaxN / kSW);    // rows per thread (row i = t + kSW r)
constexpr int kSWaves = kSW / 64;
using arnoldi::kMaxV;                           // basis vectors (restart length <= 32)

struct SmallArgs {
  uint32_t n;
  int m, max_iter;
  double tol;
  const uint32_t* prow;    // point row offsets (row lengths of the ELL product)
  const double* b;
  double *x, *V;
  double* out;   // [0] iterations, [1] relative true residual at exit
  uint32_t nsteps, nlow;   // node-level sweep steps (the first nlow are the lower sweep)
  const uint4* smeta;
  const uint32_t* sint;
  const double* sval;
  const uint2* emeta;      // ELL slices of the point matrix
  const uint32_t* ecol;
  const double* eval;
};

// sum of v over the workgroup, the same value in every thread (waves in a fixed order)
__device__ inline double wg_sum(double v, double* red)
{
  for(int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  __syncthreads();   // earlier readers of red are done
  if((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.;
  for(int k = 0; k < kSWaves; ++k) s += red[k];
  return s;
}

// Rows of a thread past n hold zeros in its row registers (wr): their loads read row n - 1 (a valid
// address) and add 0 * v, so no load sits behind a branch and its wait (p + 0 is p, up to the sign of
// a zero p).
__device__ inline uint32_t row_at(uint32_t i, uint32_t n) { return i < n ? i : n - 1; }

// c[k] = V_k . w for k < nv, w in registers (rows t + kSW r), into LDS c[] for every thread
__device__ inline void wg_mdot(int nv, uint32_t n, const double* V, const double wr[kSR], double* red, double* c)
{
  const int t = threadIdx.x;
  for(int k = 0; k < nv; ++k) {
    double p = 0.;
#pragma unroll
    for(int r = 0; r < kSR; ++r) p += V[(uint64_t)k * n + row_at(t + kSW * r, n)] * wr[r];
    for(int off = 32; off > 0; off >>= 1) p += __shfl_down(p, off, 64);
    if((t & 63) == 0) red[(t >> 6) * kMaxV + k] = p;
  }
  __syncthreads();
  if(t < nv) {
    double s = 0.;
    for(int w = 0; w < kSWaves; ++w) s += red[w * kMaxV + t];
    c[t] = s;
  }
  __syncthreads();
}

// ---- triangular sweeps by node levels ----
// A node's level is one more than the largest level of the nodes it couples to below it (lower sweep)
// or above it (upper sweep); the point rows of one node are solved together, in row order.  Step s of
// the sweeps (the lower levels, then the upper ones) holds R nodes and the record (R, W, ib, vb); slot k
// (k < R) holds a node and the rows a of that node, each in one contiguous record:
//   sint[ib + k (2 + W) + ...]   the node, its number c_k of off-node blocks (W = max c_k), then the node
//                                columns of those blocks, ascending (the node itself past c_k)
//   sval[vb + (k neq + a)(W + 1) neq + ...]   row a: entry (a, b) of the j-th off-node block at j neq + b
//                                (zeros past c_k), then entry (a, b) of the node's own diagonal block at
//                                W neq + b
// so that a lane finds every entry of its row at fixed offsets from one address, and every address of
// a step is known before the step before it ends: each lane loads its next rows while waiting at the
// barriers.  A level of more than kSlots nodes takes several steps.
// Four lanes per node (lane a takes point row a; kSlots nodes per step).  Every row is reduced in the
// oracle's order and without contraction: s = b (lower) or x (upper); s -= l_q x_q (u_q x_q) over the
// columns of the other nodes in ascending order, then over those of the node's own rows; upper:
// x = s / u_ii.  The sweeps give the oracle's results bit for bit (up to the sign of a zero).
constexpr uint32_t kSlots = kSW / 4;
constexpr int kAhead = 2;                // steps whose rows are in flight
constexpr uint32_t kMaxSteps = 4096;     // sweep steps of the single-workgroup solver (records in LDS)

template<int NEQ>
struct SlotRow {
  static constexpr int KB = NEQ == 1 ? 8 : NEQ == 2 ? 6 : 4;   // off-node blocks held in registers
  uint32_t node, cnt;
  uint32_t col[KB];
  double v[KB][NEQ];
  double d[NEQ];
};

// row a of slot k of step M (nothing when the slot or lane is idle)
template<int NEQ>
__device__ inline void slot_load(const SmallArgs& g, const uint4 M, uint32_t k, uint32_t a, SlotRow<NEQ>& P)
{
  constexpr int KB = SlotRow<NEQ>::KB;
  const uint32_t R = M.x, W = M.y;
  if(k >= R || a >= (uint32_t)NEQ) return;
  const uint32_t* ri = g.sint + M.z + k * (2 + W);
  const double* rv = g.sval + M.w + (k * NEQ + a) * ((W + 1) * NEQ);
  P.node = ri[0];
  P.cnt = ri[1];
#pragma unroll
  for(int j = 0; j < KB; ++j)
    if((uint32_t)j < W) {
      P.col[j] = ri[2 + j];
#pragma unroll
      for(int b = 0; b < NEQ; ++b) P.v[j][b] = rv[j * NEQ + b];
    } else {   // past the step's width: a valid column and a zero entry
      P.col[j] = P.node;
#pragma unroll
      for(int b = 0; b < NEQ; ++b) P.v[j][b] = 0.;
    }
#pragma unroll
  for(int b = 0; b < NEQ; ++b) P.d[b] = rv[W * NEQ + b];
}

// the off-node products l_q x_q of row a, in place of the entries (the mul of each term of the oracle's
// s = s - l_q x_q, which does not depend on s)
// every register block, without branches (the entries past cnt are zeros at valid columns), so that
// the LDS reads issue together
template<int NEQ>
__device__ inline void off_products(SlotRow<NEQ>& P, const double* zs)
{
#pragma clang fp contract(off)
  constexpr int KB = SlotRow<NEQ>::KB;
#pragma unroll
  for(int j = 0; j < KB; ++j) {
#pragma unroll
    for(int b = 0; b < NEQ; ++b) P.v[j][b] = P.v[j][b] * zs[P.col[j] * NEQ + b];
  }
}

// s -= (off-node terms of row a), ascending: the products in P, then the blocks past the registers.
// The register blocks past c_k hold zero products: s - 0 is s (up to the sign of a zero s).
template<int NEQ>
__device__ inline double off_subtract(const SmallArgs& g, const uint4 M, uint32_t k, uint32_t a,
                                      const SlotRow<NEQ>& P, const double* zs, double s)
{
#pragma clang fp contract(off)
  constexpr int KB = SlotRow<NEQ>::KB;
#pragma unroll
  for(int j = 0; j < KB; ++j) {
#pragma unroll
    for(int b = 0; b < NEQ; ++b) s = s - P.v[j][b];
  }
  if(P.cnt > (uint32_t)KB) {
    const uint32_t W = M.y;
    const uint32_t* ri = g.sint + M.z + k * (2 + W);
    const double* rv = g.sval + M.w + (k * NEQ + a) * ((W + 1) * NEQ);
    for(uint32_t j = KB; j < P.cnt; ++j) {
      const uint32_t c = ri[2 + j];
#pragma unroll
      for(int b = 0; b < NEQ; ++b) s = s - rv[j * NEQ + b] * zs[c * NEQ + b];
    }
  }
  return s;
}

// v of lane P of each group of four lanes (the lanes of one node), in all four (DPP quad_perm)
template<int P>
__device__ inline double quad_bcast(double v)
{
  constexpr int ctrl = P * 0x55;   // quad_perm [P, P, P, P]
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, ctrl, 0xf, 0xf, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), ctrl, 0xf, 0xf, false);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// the rows of slot k of step M (row a in this lane); every lane of the wave calls it (the broadcasts
// between the rows of a node run in all lanes)
template<int NEQ>
__device__ inline void slot_rows(const SmallArgs& g, const uint4 M, bool upper, uint32_t k, uint32_t a,
                                 SlotRow<NEQ>& P, double* zs)
{
#pragma clang fp contract(off)
  const bool act = k < M.x && a < (uint32_t)NEQ;
  const uint32_t r0 = act ? P.node * NEQ : 0;
  double s = 0.;
  if(act) {
    off_products<NEQ>(P, zs);
    s = zs[r0 + a];
  }
  if(!upper) {
    // s = b - off-node terms, then the node's earlier rows in row order: after round p, row p is final
    if(act) s = off_subtract<NEQ>(g, M, k, a, P, zs, s);
    static_for<0, NEQ - 1>([&](auto pc) {
      constexpr int p = decltype(pc)::value;
      const double xp = quad_bcast<p>(s);
      if(a > (uint32_t)p) s = s - P.d[p] * xp;
    });
    if(act) zs[r0 + a] = s;
  } else {
    // s = y - off-node terms, then from the last row: s -= the node's later rows (ascending), x = s / u_aa
    if(act) s = off_subtract<NEQ>(g, M, k, a, P, zs, s);
    double xr[NEQ];
    double x = 0.;
    static_for<0, NEQ>([&](auto pc) {
      constexpr int p = NEQ - 1 - decltype(pc)::value;
      if(a == (uint32_t)p) {
#pragma unroll
        for(int b = p + 1; b < NEQ; ++b) s = s - P.d[b] * xr[b];
        x = s / P.d[p];
      }
      xr[p] = quad_bcast<p>(x);
    });
    if(act) zs[r0 + a] = x;
  }
}

// zs := U^-1 L^-1 zs in place (the factors of A), one barrier per node level.  hdr(st): step record
// st (zero records follow the last step, kAhead of them).  zs: the LDS copy of the vector in the
// single-workgroup solver, the vector itself in global memory otherwise.  The rows of the next kAhead
// steps are in flight in a ring of registers.
template<int NEQ, typename Hdr>
__device__ inline void ilu_sweeps(const SmallArgs& g, Hdr hdr, double* zs)
{
  const uint32_t k0 = threadIdx.x >> 2, a = threadIdx.x & 3;
  SlotRow<NEQ> P[kAhead];
#pragma unroll
  for(int i = 0; i < kAhead; ++i) slot_load<NEQ>(g, hdr(i), k0, a, P[i]);
  for(uint32_t st = 0; st < g.nsteps; st += kAhead) {
#pragma unroll
    for(int i = 0; i < kAhead; ++i)
      if(st + i < g.nsteps) {
        const uint4 M = hdr(st + i), Mf = hdr(st + i + kAhead);
        slot_rows<NEQ>(g, M, st + i >= g.nlow, k0, a, P[i], zs);
        slot_load<NEQ>(g, Mf, k0, a, P[i]);
        __syncthreads();
      }
  }
}

// ml: the step records copied to LDS
template<int NEQ>
__device__ inline void ilu_solve_lds(const SmallArgs& g, const uint4* ml, double* zs)
{
  ilu_sweeps<NEQ>(g, [ml](uint32_t st) { return ml[st]; }, zs);
}

// x = U^-1 L^-1 b in one workgroup, the vector in global memory (the multi-launch solver).  The step
// records are read through the constant address space: the solver never writes them, so every lane
// gets them by scalar loads.
typedef const __attribute__((address_space(4))) uint32_t* ConstWords;
template<int NEQ>
__global__ __launch_bounds__(kSW) void k_ilu0_sweep(SmallArgs g, const double* __restrict__ b, double* x)
{
  for(uint32_t i = threadIdx.x; i < g.n; i += kSW) x[i] = b[i];
  __syncthreads();
  const uint4* m = g.smeta;
  ilu_sweeps<NEQ>(g, [m](uint32_t st) {
    const ConstWords p = (ConstWords)(uintptr_t)m + 4ull * st;
    return make_uint4(p[0], p[1], p[2], p[3]);
  }, x);
}

// wr = A zs for the thread's rows: 64-row ELL slices (per slice (base, width); entry q of row i at
// base + q 64 + i % 64), each row summed in ascending column order as in the point matrix; the rows of
// a thread advance kSpmvRows at a time, so each entry index has their loads in flight together
constexpr int kSpmvRows = 4;
__device__ inline void spmv_lds(const SmallArgs& g, const double* zs, double wr[kSR])
{
  static_assert(kSR % kSpmvRows == 0, "rows per thread");
  const uint32_t lane = threadIdx.x & 63;
#pragma unroll
  for(int r0 = 0; r0 < kSR; r0 += kSpmvRows) {
    uint32_t base[kSpmvRows], len[kSpmvRows], wid[kSpmvRows], wmax = 0;
#pragma unroll
    for(int r = 0; r < kSpmvRows; ++r) {   // rows past n: the slice of row n - 1, length 0
      const uint32_t i = threadIdx.x + kSW * (r0 + r);
      const uint32_t ic = row_at(i, g.n);
      const uint2 E = g.emeta[ic >> 6];
      base[r] = E.x + lane;
      wid[r] = E.y;
      len[r] = i < g.n ? g.prow[ic + 1] - g.prow[ic] : 0;
      wmax = wmax > E.y ? wmax : E.y;
      wr[r0 + r] = 0.;
    }
#pragma unroll 4
    for(uint32_t q = 0; q < wmax; ++q) {
#pragma unroll
      for(int r = 0; r < kSpmvRows; ++r) {   // past the row's length: a valid entry of its slice, times 0
        const uint32_t e = base[r] + (q < wid[r] ? q : wid[r] - 1) * 64;
        const double v = g.eval[e];
        wr[r0 + r] += (q < len[r] ? v : 0.) * zs[g.ecol[e]];
      }
    }
  }
}

// Restarted GMRES(m) with ILU(0) right preconditioning, classical Gram-Schmidt twice and Givens
// rotations (the method of the multi-launch loop below and of the oracle's orc_gmres_ilu0).  A cycle
// ends when the recurrence residual meets the tolerance or at the restart length; the true residual
// b - A x at the start of the next cycle decides convergence.
template<int NEQ>
__global__ __launch_bounds__(kSW) void k_gmres_small(SmallArgs a)
{
  __shared__ double zs[kSmallMaxN];
  __shared__ uint4 ml[kMaxSteps + kAhead];
  __shared__ double red[kSWaves * kMaxV];
  __shared__ double H[kMaxV * (kMaxV - 1)];
  __shared__ double cs[kMaxV], sn[kMaxV], g[kMaxV], hc[kMaxV], y[kMaxV], cc[kMaxV];
  __shared__ double s_res;
  const uint32_t t = threadIdx.x;
  const uint32_t n = a.n;
  const int m = a.m;
  double wr[kSR];
  for(uint32_t i = t; i < a.nsteps + kAhead; i += kSW) ml[i] = a.smeta[i];   // visible after wg_sum
  double p = 0.;
#pragma unroll
  for(int r = 0; r < kSR; ++r) {
    const uint32_t i = t + kSW * r;
    if(i < n) p += a.b[i] * a.b[i];
  }
  double bn = sqrt(wg_sum(p, red));
  if(bn == 0.) bn = 1.;
  const double thr = a.tol * bn;
  int it = 0;
  double rel = 0.;
  for(;;) {
    // w = b - A x (x staged in LDS; each thread stages and updates its own rows)
#pragma unroll
    for(int r = 0; r < kSR; ++r) {
      const uint32_t i = t + kSW * r;
      if(i < n) zs[i] = a.x[i];
    }
    __syncthreads();
    spmv_lds(a, zs, wr);
    p = 0.;
#pragma unroll
    for(int r = 0; r < kSR; ++r) {
      const uint32_t i = t + kSW * r;
      wr[r] = i < n ? a.b[i] - wr[r] : 0.;
      p += wr[r] * wr[r];
    }
    const double beta = sqrt(wg_sum(p, red));
    rel = beta / bn;
    if(beta <= thr || it >= a.max_iter) break;
#pragma unroll
    for(int r = 0; r < kSR; ++r) {
      const uint32_t i = t + kSW * r;
      if(i < n) {   // V_0, and the first input of the preconditioner (no reader of zs is left)
        const double v = wr[r] / beta;
        a.V[i] = v;
        zs[i] = v;
      }
    }
    if(t == 0) {
      g[0] = beta;
      for(int k = 1; k <= m; ++k) g[k] = 0.;
    }
    int j = 0;
    for(; j < m && it < a.max_iter; ++j) {
      ++it;
      // z = M^-1 V_j (zs holds V_j: each thread wrote its rows with V_j)
      __syncthreads();
      ilu_solve_lds<NEQ>(a, ml, zs);
      spmv_lds(a, zs, wr);
      // two passes of classical Gram-Schmidt against V_0 .. V_j
      for(int pass = 0; pass < 2; ++pass) {
        wg_mdot(j + 1, n, a.V, wr, red, cc);
        for(int k = 0; k <= j; ++k) {   // every row: w -= c_k V_k in the order of k
          const double ck = cc[k];
#pragma unroll
          for(int r = 0; r < kSR; ++r) wr[r] -= ck * a.V[(uint64_t)k * n + row_at(t + kSW * r, n)];
        }
#pragma unroll
        for(int r = 0; r < kSR; ++r) wr[r] = t + kSW * r < n ? wr[r] : 0.;
        if((int)t <= j) hc[t] = pass == 0 ? cc[t] : hc[t] + cc[t];
      }
      p = 0.;
#pragma unroll
      for(int r = 0; r < kSR; ++r) p += wr[r] * wr[r];
      const double hn = sqrt(wg_sum(p, red));
#pragma unroll
      for(int r = 0; r < kSR; ++r) {
        const uint32_t i = t + kSW * r;
        if(i < n) {   // V_{j+1}, and the next input of the preconditioner
          const double v = wr[r] / hn;
          a.V[(uint64_t)(j + 1) * n + i] = v;
          zs[i] = v;
        }
      }
      if(t == 0) {
        hc[j + 1] = hn;
        for(int k = 0; k < j; ++k) {
          const double t0 = cs[k] * hc[k] + sn[k] * hc[k + 1];
          hc[k + 1] = -sn[k] * hc[k] + cs[k] * hc[k + 1];
          hc[k] = t0;
        }
        const double den = sqrt(hc[j] * hc[j] + hc[j + 1] * hc[j + 1]);
        cs[j] = hc[j] / den;
        sn[j] = hc[j + 1] / den;
        hc[j] = den;
        hc[j + 1] = 0.;
        g[j + 1] = -sn[j] * g[j];
        g[j] = cs[j] * g[j];
        for(int k = 0; k <= j; ++k) H[k * m + j] = hc[k];
        s_res = fabs(g[j + 1]);
      }
      __syncthreads();
      if(!(s_res > thr)) { ++j; break; }
    }
    // x += M^-1 V y, H y = g (upper triangular, j x j)
    if(t == 0)
      for(int i = j - 1; i >= 0; --i) {
        double s = g[i];
        for(int l = i + 1; l < j; ++l) s -= H[i * m + l] * y[l];
        y[i] = s / H[i * m + i];
      }
    __syncthreads();
    double sy[kSR];
#pragma unroll
    for(int r = 0; r < kSR; ++r) sy[r] = 0.;
    for(int k = 0; k < j; ++k) {
      const double yk = y[k];
#pragma unroll
      for(int r = 0; r < kSR; ++r) sy[r] += yk * a.V[(uint64_t)k * n + row_at(t + kSW * r, n)];
    }
#pragma unroll
    for(int r = 0; r < kSR; ++r) {
      const uint32_t i = t + kSW * r;
      if(i < n) zs[i] = sy[r];
    }
    __syncthreads();
    ilu_solve_lds<NEQ>(a, ml, zs);
#pragma unroll
    for(int r = 0; r < kSR; ++r) {
      const uint32_t i = t + kSW * r;
      if(i < n) a.x[i] += zs[i];
    }
    __syncthreads();   // zs is restaged from x next
  }
  if(t == 0) {
    a.out[0] = (double)it;
    a.out[1] = rel;
  }
}

constexpr uint32_t kNoSrc = 0xffffffffu;

// dst[e] = from[src[e]] (0 where src[e] is kNoSrc) over two tables in one launch
__global__ void k_gather2(uint64_t n1, const uint32_t* __restrict__ src1, const double* __restrict__ from1,
                          double* __restrict__ dst1, uint64_t n2, const uint32_t* __restrict__ src2,
                          const double* __restrict__ from2, double* __restrict__ dst2)
{
  const uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if(e < n1) {
    const uint32_t q = src1[e];
    dst1[e] = q == kNoSrc ? 0. : from1[q];
  } else if(e < n1 + n2) {
    const uint32_t q = src2[e - n1];
    dst2[e - n1] = q == kNoSrc ? 0. : from2[q];
  }
}

// The single-workgroup solver's tables (layouts at slot_load and spmv_lds), on the host.  rp/cl: the
// block pattern (columns ascending in each row); prow: point row offsets (point row (i, a) holds, for
// each block column position s of node row i, the columns cl[s] neq + b at prow + s neq + b).
void build_slot_plan(Lss& L, const std::vector<uint32_t>& rp, const std::vector<uint32_t>& cl,
                     const std::vector<uint32_t>& prow)
{
  Context& c = *L.ctx;
  const uint32_t N = L.n_nodes;
  const uint32_t neq = (uint32_t)L.neq;
  const uint64_t n = (uint64_t)N * neq;
  std::vector<uint32_t> dpos(N);   // position of the diagonal block in each node row
  for(uint32_t i = 0; i < N; ++i)
    for(uint32_t s = rp[i]; s < rp[i + 1]; ++s)
      if(cl[s] == i) dpos[i] = s - rp[i];
  std::vector<uint32_t> blev(N, 0), bulev(N, 0);
  uint32_t nl = 0, nu = 0;
  for(uint32_t i = 0; i < N; ++i) {
    for(uint32_t s = rp[i]; s < rp[i + 1]; ++s)
      if(cl[s] < i) blev[i] = std::max(blev[i], blev[cl[s]] + 1);
    nl = std::max(nl, blev[i] + 1);
  }
  for(uint32_t i = N; i-- > 0;) {
    for(uint32_t s = rp[i]; s < rp[i + 1]; ++s)
      if(cl[s] > i) bulev[i] = std::max(bulev[i], bulev[cl[s]] + 1);
    nu = std::max(nu, bulev[i] + 1);
  }
  std::vector<std::vector<uint32_t>> levels(nl + nu);
  for(uint32_t i = 0; i < N; ++i) {
    levels[blev[i]].push_back(i);
    levels[nl + bulev[i]].push_back(i);
  }
  std::vector<std::vector<uint32_t>> steps;   // levels cut into steps of at most kSlots nodes
  uint32_t nlow = 0;
  for(uint32_t l = 0; l < nl + nu; ++l) {
    for(size_t k = 0; k < levels[l].size(); k += kSlots)
      steps.emplace_back(levels[l].begin() + k, levels[l].begin() + std::min<size_t>(levels[l].size(), k + kSlots));
    if(l + 1 == nl) nlow = (uint32_t)steps.size();
  }
  std::vector<uint32_t> meta, sint, ssrc;
  for(uint32_t st = 0; st < (uint32_t)steps.size(); ++st) {
    const bool upper = st >= nlow;
    const std::vector<uint32_t>& nodes = steps[st];
    const uint32_t R = (uint32_t)nodes.size();
    std::vector<std::vector<uint32_t>> pos(R);   // off-node block positions of each slot, ascending
    uint32_t W = 0;
    for(uint32_t k = 0; k < R; ++k) {
      const uint32_t i = nodes[k];
      for(uint32_t s = 0; s < rp[i + 1] - rp[i]; ++s)
        if(upper ? cl[rp[i] + s] > i : cl[rp[i] + s] < i) pos[k].push_back(s);
      W = std::max(W, (uint32_t)pos[k].size());
    }
    if(sint.size() > 0xffffffffull - (2ull + W) * R || ssrc.size() > 0xffffffffull - (W + 1ull) * neq * neq * R)
      CF3_THROW(CF3HIP_ERR_LIMIT, "ILU sweep tables too large");
    meta.insert(meta.end(), {R, W, (uint32_t)sint.size(), (uint32_t)ssrc.size()});
    for(uint32_t k = 0; k < R; ++k) {
      const uint32_t i = nodes[k];
      sint.push_back(i);
      sint.push_back((uint32_t)pos[k].size());
      for(uint32_t j = 0; j < W; ++j) sint.push_back(j < pos[k].size() ? cl[rp[i] + pos[k][j]] : i);
      for(uint32_t a = 0; a < neq; ++a) {
        const uint32_t r0 = prow[(uint64_t)i * neq + a];
        for(uint32_t j = 0; j < W; ++j)
          for(uint32_t b = 0; b < neq; ++b) ssrc.push_back(j < pos[k].size() ? r0 + pos[k][j] * neq + b : kNoSrc);
        for(uint32_t b = 0; b < neq; ++b) ssrc.push_back(r0 + dpos[i] * neq + b);
      }
    }
  }
  meta.insert(meta.end(), 4 * kAhead, 0u);   // zero records read ahead past the last step
  // 64-row ELL slices of the point matrix (the product of the single-workgroup solver only)
  const uint32_t nsl = n <= kSmallMaxN ? (uint32_t)((n + 63) / 64) : 0;
  std::vector<uint32_t> emeta, ecol, esrc;
  for(uint32_t sl = 0; sl < nsl; ++sl) {
    uint32_t W = 0;
    for(uint64_t r = (uint64_t)sl * 64; r < std::min<uint64_t>(n, (uint64_t)sl * 64 + 64); ++r)
      W = std::max(W, prow[r + 1] - prow[r]);
    emeta.insert(emeta.end(), {(uint32_t)ecol.size(), W});
    for(uint32_t q = 0; q < W; ++q)
      for(uint32_t l = 0; l < 64; ++l) {
        const uint64_t r = (uint64_t)sl * 64 + l;
        const bool in = r < n && q < prow[r + 1] - prow[r];
        const uint32_t i = (uint32_t)(r / neq);
        ecol.push_back(in ? cl[rp[i] + q / neq] * neq + q % neq : 0);
        esrc.push_back(in ? prow[r] + q : kNoSrc);
      }
  }
  auto up = [&](DevBuf& d, const std::vector<uint32_t>& h) {   // at least one uint4 (smeta[0] is read)
    d.alloc(&c, sizeof(uint32_t) * std::max<size_t>(h.size(), 4));
    HIP_OK(hipMemcpyAsync(d.ptr, h.data(), sizeof(uint32_t) * h.size(), hipMemcpyHostToDevice, c.stream));
  };
  auto& I = L.ilu;
  I.nsv = ssrc.size();
  I.nev = esrc.size();
  up(I.smeta, meta);
  up(I.sint, sint);
  up(I.ssrc, ssrc);
  up(I.emeta, emeta);
  up(I.ecol, ecol);
  up(I.esrc, esrc);
  I.sval.alloc(&c, sizeof(double) * std::max<size_t>(ssrc.size(), 1));
  I.eval.alloc(&c, sizeof(double) * std::max<size_t>(esrc.size(), 1));
  HIP_OK(hipStreamSynchronize(c.stream));   // the host vectors are released on return
  I.nsteps = (uint32_t)steps.size();
  I.nlow = nlow;
}

// the factorisation's level schedule of the point pattern (natural order), on the host: a row's level is
// one more than the largest of its lower columns'; its update lists; and the solvers' sweep tables
void build_ilu_plan(Lss& L)
{
  Context& c = *L.ctx;
  const uint32_t N = L.n_nodes;
  const int neq = L.neq;
  std::vector<uint32_t> rp(N + 1), cl(L.n_blocks);
  HIP_OK(hipMemcpyAsync(rp.data(), L.row_ptr.ptr, sizeof(uint32_t) * (N + 1), hipMemcpyDeviceToHost, c.stream));
  HIP_OK(hipMemcpyAsync(cl.data(), L.cols.ptr, sizeof(uint32_t) * L.n_blocks, hipMemcpyDeviceToHost, c.stream));
  HIP_OK(hipStreamSynchronize(c.stream));
  const uint64_t n = (uint64_t)N * neq;
  // point row r = i neq + a holds, for each block column s of row i, columns cl[s] neq + b (ascending)
  std::vector<uint32_t> prow(n + 1), diag(n), lev(n);
  const uint64_t nn = (uint64_t)neq * neq;
  for(uint32_t i = 0; i < N; ++i) {
    const uint32_t deg = rp[i + 1] - rp[i];
    uint32_t si = 0;
    for(uint32_t s2 = 0; s2 < deg; ++s2)
      if(cl[rp[i] + s2] == i) si = s2;
    for(int a2 = 0; a2 < neq; ++a2) {
      const uint64_t r = (uint64_t)i * neq + a2;
      prow[r] = (uint32_t)(nn * rp[i] + (uint64_t)a2 * neq * deg);
      diag[r] = prow[r] + si * neq + a2;
    }
  }
  prow[n] = (uint32_t)(nn * rp[N]);
  auto col_of = [&](uint32_t i, uint32_t k) {   // k-th entry of point row (i, .): column
    const uint32_t deg = rp[i + 1] - rp[i];
    (void)deg;
    return cl[rp[i] + k / neq] * neq + k % neq;
  };
  uint32_t lmax = 0;
  for(uint64_t r = 0; r < n; ++r) {
    const uint32_t i = (uint32_t)(r / neq);
    uint32_t l = 0;
    for(uint32_t q = prow[r]; q < diag[r]; ++q) l = std::max(l, lev[col_of(i, q - prow[r])] + 1);
    lev[r] = l;
    lmax = std::max(lmax, l);
  }
  auto order = [&](const std::vector<uint32_t>& lv, uint32_t nl, std::vector<uint32_t>& ord, std::vector<uint32_t>& ptr) {
    ptr.assign(nl + 1, 0);
    for(uint64_t r = 0; r < n; ++r) ++ptr[lv[r] + 1];
    for(uint32_t l = 0; l < nl; ++l) ptr[l + 1] += ptr[l];
    std::vector<uint32_t> fill(ptr.begin(), ptr.end() - 1);
    ord.resize(n);
    for(uint64_t r = 0; r < n; ++r) ord[fill[lv[r]]++] = (uint32_t)r;
  };
  std::vector<uint32_t> lorder, lptr;
  order(lev, lmax + 1, lorder, lptr);
  // update lists of the elimination (at k_ilu0_factor)
  const uint64_t nnz = prow[n];
  std::vector<uint32_t> fptr(nnz + 1, 0), fpiv(nnz, 0), fsrc;
  std::vector<uint16_t> fdst;
  std::vector<int64_t> at(n, -1);   // column -> offset in the current row
  for(uint64_t r = 0; r < n; ++r) {
    const uint32_t i = (uint32_t)(r / neq);
    const uint32_t len = prow[r + 1] - prow[r];
    if(len > 0xffffu) CF3_THROW(CF3HIP_ERR_LIMIT, "ILU row too long");
    for(uint32_t o = 0; o < len; ++o) at[col_of(i, o)] = o;
    for(uint32_t q = prow[r]; q < prow[r + 1]; ++q) {
      fptr[q] = (uint32_t)fsrc.size();
      if(q < diag[r]) {
        const uint32_t k = col_of(i, q - prow[r]);
        fpiv[q] = diag[k];
        const uint32_t ki = k / neq;
        for(uint32_t rr = diag[k] + 1; rr < prow[k + 1]; ++rr) {
          const int64_t o = at[col_of(ki, rr - prow[k])];
          if(o >= 0) {
            fsrc.push_back(rr);
            fdst.push_back((uint16_t)o);
          }
        }
        if(fsrc.size() >= 0xffffffffull) CF3_THROW(CF3HIP_ERR_LIMIT, "ILU update lists too large");
      }
    }
    for(uint32_t o = 0; o < len; ++o) at[col_of(i, o)] = -1;
  }
  fptr[nnz] = (uint32_t)fsrc.size();
  auto up = [&](DevBuf& d, const std::vector<uint32_t>& h) {
    d.alloc(&c, sizeof(uint32_t) * h.size());
    HIP_OK(hipMemcpyAsync(d.ptr, h.data(), sizeof(uint32_t) * h.size(), hipMemcpyHostToDevice, c.stream));
  };
  up(L.ilu.diag, diag);
  up(L.ilu.fptr, fptr);
  up(L.ilu.fpiv, fpiv);
  up(L.ilu.fsrc, fsrc);
  L.ilu.fdst.alloc(&c, sizeof(uint16_t) * std::max<size_t>(fdst.size(), 1));
  HIP_OK(hipMemcpyAsync(L.ilu.fdst.ptr, fdst.data(), sizeof(uint16_t) * fdst.size(), hipMemcpyHostToDevice, c.stream));
  up(L.ilu.lorder, lorder);
  up(L.ilu.lptr, lptr);
  L.ilu.prow.alloc(&c, sizeof(uint32_t) * (n + 1));
  L.ilu.pcol.alloc(&c, sizeof(uint32_t) * prow[n]);
  L.ilu.aval.alloc(&c, sizeof(double) * prow[n]);
  L.ilu.fval.alloc(&c, sizeof(double) * prow[n]);
  L.ilu.out.alloc(&c, sizeof(double) * 2);
  HIP_OK(hipStreamSynchronize(c.stream));   // the host vectors are released on return
  L.ilu.nlev = lmax + 1;
  build_slot_plan(L, rp, cl, prow);
  L.ilu.ready = true;
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
  // point CSR copy of the matrix, factorised in place at every solve like the reference; the level
  // schedules of its pattern are kept with the system
  if(!L.ilu.ready) build_ilu_plan(L);
  auto& I = L.ilu;
  lss_point_csr(L, I.prow.as<uint32_t>(), I.pcol.as<uint32_t>(), I.aval.as<double>());
  HIP_OK(hipMemcpyAsync(I.fval.ptr, I.aval.ptr, I.aval.bytes, hipMemcpyDeviceToDevice, c.stream));
  hipLaunchKernelGGL(k_ilu0_factor, dim3(1), dim3(kWG), 0, c.stream, I.prow.as<uint32_t>(), I.pcol.as<uint32_t>(),
                     I.fval.as<double>(), I.diag.as<uint32_t>(), I.lorder.as<uint32_t>(), I.lptr.as<uint32_t>(), I.nlev,
                     FactorPlan{I.fptr.as<uint32_t>(), I.fpiv.as<uint32_t>(), I.fsrc.as<uint32_t>(), I.fdst.as<uint16_t>()});
  HIP_OK(hipGetLastError());
  if(I.V.bytes < sizeof(double) * (m + 1) * n) I.V.alloc(&c, sizeof(double) * (m + 1) * n);
  if(n <= kSmallMaxN && I.nsteps <= kMaxSteps) {
    SmallArgs a{(uint32_t)n, m, prm.max_iter, prm.tol, I.prow.as<uint32_t>(), L.rhs.as<double>(), L.sol.as<double>(),
                I.V.as<double>(), I.out.as<double>(),
                I.nsteps, I.nlow, I.smeta.as<uint4>(), I.sint.as<uint32_t>(), I.sval.as<double>(),
                I.emeta.as<uint2>(), I.ecol.as<uint32_t>(), I.eval.as<double>()};
    hipLaunchKernelGGL(k_gather2, dim3(grid_for(I.nsv + I.nev, 256)), dim3(256), 0, c.stream, I.nsv,
                       I.ssrc.as<uint32_t>(), I.fval.as<double>(), I.sval.as<double>(), I.nev, I.esrc.as<uint32_t>(),
                       I.aval.as<double>(), I.eval.as<double>());
    switch(L.neq) {
      case 1: hipLaunchKernelGGL(k_gmres_small<1>, dim3(1), dim3(kSW), 0, c.stream, a); break;
      case 2: hipLaunchKernelGGL(k_gmres_small<2>, dim3(1), dim3(kSW), 0, c.stream, a); break;
      case 3: hipLaunchKernelGGL(k_gmres_small<3>, dim3(1), dim3(kSW), 0, c.stream, a); break;
      default: hipLaunchKernelGGL(k_gmres_small<4>, dim3(1), dim3(kSW), 0, c.stream, a); break;
    }
    HIP_OK(hipGetLastError());
    double o[2];
    HIP_OK(hipMemcpyAsync(o, I.out.ptr, sizeof(o), hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    st.iterations = (int)o[0];
    st.outer = 1;
    st.spmvs = 0;
    st.rel_residual = o[1];
    st.converged = o[1] <= prm.tol ? 1 : 0;
    return;
  }
  // node-level sweeps of the factors, as in the single-workgroup solver, on vectors in global memory
  hipLaunchKernelGGL(k_gather2, dim3(grid_for(I.nsv, 256)), dim3(256), 0, c.stream, I.nsv, I.ssrc.as<uint32_t>(),
                     I.fval.as<double>(), I.sval.as<double>(), (uint64_t)0, I.esrc.as<uint32_t>(), I.aval.as<double>(),
                     I.eval.as<double>());
  SmallArgs sw{};
  sw.n = (uint32_t)n;
  sw.nsteps = I.nsteps;
  sw.nlow = I.nlow;
  sw.smeta = I.smeta.as<uint4>();
  sw.sint = I.sint.as<uint32_t>();
  sw.sval = I.sval.as<double>();
  auto precond = [&](const double* b, double* x) {
    switch(L.neq) {
      case 1: hipLaunchKernelGGL(k_ilu0_sweep<1>, dim3(1), dim3(kSW), 0, c.stream, sw, b, x); break;
      case 2: hipLaunchKernelGGL(k_ilu0_sweep<2>, dim3(1), dim3(kSW), 0, c.stream, sw, b, x); break;
      case 3: hipLaunchKernelGGL(k_ilu0_sweep<3>, dim3(1), dim3(kSW), 0, c.stream, sw, b, x); break;
      default: hipLaunchKernelGGL(k_ilu0_sweep<4>, dim3(1), dim3(kSW), 0, c.stream, sw, b, x); break;
    }
  };

  // workspace: V (m+1) n, w, z, u ; small: bookkeeping, h1, h2, y, scalars
  if(I.w.bytes < sizeof(double) * n) I.w.alloc(&c, sizeof(double) * n);
  if(I.z.bytes < sizeof(double) * n) I.z.alloc(&c, sizeof(double) * n);
  if(I.small.bytes < sizeof(double) * GmresDev::size(m)) I.small.alloc(&c, sizeof(double) * GmresDev::size(m));
  DevBuf& w = I.w;
  DevBuf& z = I.z;
  double* base = I.small.as<double>();
  const GmresDev gd{m, base};
  double* h1 = gd.h1();
  double* h2 = gd.h2();
  double* y = gd.y();
  double* sc = c.scalars.as<double>();   // sc[0] |w|^2, sc[1] ||b||^2
  const unsigned gb = grid_for(n, 256);
  double* x = L.sol.as<double>();
  const double* b = L.rhs.as<double>();
  double* Vd = I.V.as<double>();

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

// precon.hip -- point-Jacobi and block-Jacobi ILU(0) / D-ILU preconditioners.
//
// Setup (host, per block of contiguous rows, OpenMP over blocks) follows the
// reference factorisations:
//   point Jacobi   get_DInv_crs          src/matrix.f90:3104-3145
//   ILU(0)         get_ILU_LU_crs        src/matrix.f90:8249-8392
//                  ILU_LU_crs            src/matrix.f90:7970-8072
//                  divide_ILU_crs        src/matrix.f90:7821-7897
//   D-ILU          DILU_DInv_crs         src/matrix.f90:3482-3650
//                  init_solve_ILU_crs    src/matrix.f90:5170-5350
//   blocks         divide_region         src/common.f90:125-182
// Apply replaces the reference's sequential per-block sweeps (solve_LU_crs
// src/matrix.f90:8548-8746, solve_ILU_crs :4987-5163; on the GPU cuSPARSE
// csrsv2, src/CUDA/src/cusparse.cu:712-784) by level-scheduled sweeps: one
// workgroup per block walks the block's levels; every row sums its entries
// in the reference's order, so each row's value is bit-identical to the
// sequential sweep whatever the schedule.
#include <omp.h>

#include <algorithm>
#include <type_traits>
#include <chrono>
#include <cstdlib>
#include <cstring>

#include "dd.hpp"
#include "internal.hpp"

namespace parcel {

namespace {

constexpr int kRingBudget = 6144;  // doubles of LDS for the ring (48 KiB)

// Entry count per row the apply kernel is instantiated for (1..4, 6, 8),
// or the row length itself above 8 (generic kernel).
int apply_k(int k) {
  if (k <= 4) return std::max(k, 1);
  if (k <= 6) return 6;
  if (k <= 8) return 8;
  return k;
}

// Block-Jacobi block b is rows [blo[b], bhi[b]): the plan's explicit block
// starts when it has them (the box blocks of the stencil entry point), else
// divide_region's split.
void block_ranges(const Plan& P, int nblk, int n, std::vector<int>& blo, std::vector<int>& bhi);

}  // namespace

// divide_region (src/common.f90:125-182), 0-based half-open result.
void divide_region0(int rank, int nproc, int n, int* lo, int* hi) {
  if (n == 1) { *lo = 0; *hi = 1; return; }
  int iblock = n / nproc, over = n % nproc;
  int l, h;  // 1-based inclusive as in the reference
  if (rank < over) { l = 1 + iblock * rank + rank; h = l + iblock; }
  else { l = 1 + iblock * rank + over; h = l + iblock - 1; }
  *lo = l - 1; *hi = h;
}

namespace {

void block_ranges(const Plan& P, int nblk, int n, std::vector<int>& blo, std::vector<int>& bhi) {
  if (!P.block_starts.empty()) {
    for (int b = 0; b < nblk; ++b) { blo[b] = P.block_starts[b]; bhi[b] = P.block_starts[b + 1]; }
    return;
  }
  for (int b = 0; b < nblk; ++b) divide_region0(b, nblk, n, &blo[b], &bhi[b]);
}

struct BlockSweep {
  // rows of the block in level order, with per-level boundaries
  std::vector<int> lev_start;   // offsets into order (size nlev+1)
  std::vector<int> order;       // local rows (0-based within the block)
  std::vector<int> ptr;         // entry ranges per local row (size nrow+1)
  std::vector<int> dep;         // local dependency rows
  std::vector<double> val;
  std::vector<double> diag;     // per local row
  // double-double factors: lo parts of val and diag (empty otherwise)
  std::vector<double> val_lo;
  std::vector<double> diag_lo;
};

// Level analysis.  forward: deps of row i are < i; backward: > i.
void build_levels(BlockSweep& S, int nrow, bool forward) {
  std::vector<int> lev(nrow, 0);
  int maxl = -1;
  if (forward) {
    for (int i = 0; i < nrow; ++i) {
      int l = 0;
      for (int q = S.ptr[i]; q < S.ptr[i + 1]; ++q) l = std::max(l, lev[S.dep[q]] + 1);
      lev[i] = l; maxl = std::max(maxl, l);
    }
  } else {
    for (int i = nrow - 1; i >= 0; --i) {
      int l = 0;
      for (int q = S.ptr[i]; q < S.ptr[i + 1]; ++q) l = std::max(l, lev[S.dep[q]] + 1);
      lev[i] = l; maxl = std::max(maxl, l);
    }
  }
  int nlev = maxl + 1;
  if (nrow == 0) nlev = 0;
  S.lev_start.assign(nlev + 1, 0);
  for (int i = 0; i < nrow; ++i) S.lev_start[lev[i] + 1]++;
  for (int l = 0; l < nlev; ++l) S.lev_start[l + 1] += S.lev_start[l];
  S.order.assign(nrow, 0);
  std::vector<int> fill(S.lev_start.begin(), S.lev_start.end() - (nlev ? 1 : 0));
  for (int i = 0; i < nrow; ++i) S.order[fill[lev[i]]++] = i;
}

// ILU(0) of one block, reference semantics (see file header).
// Returns false if a row of the block has no diagonal.
// ddm: factorise in double-double (ddILU_LU_crs, src/ddmatrix.f90:5686-5830,
// with A's lo parts from A.val_lo when present) and fill val_lo / diag_lo.
bool ilu0_block(const HostCSR& A, int lo, int hi, int nvec, BlockSweep& F, BlockSweep& B, bool ddm = false) {
  const int nb = hi - lo;
  std::vector<int> ptr(nb + 1), c2;
  std::vector<double> v2, w2;  // w2: lo parts in dd mode
  const bool alo = ddm && !A.val_lo.empty();
  c2.reserve((size_t)(A.rowptr[hi] - A.rowptr[lo]));
  v2.reserve(c2.capacity());
  for (int ii = lo; ii < hi; ii += nvec) {
    int iimax = std::min(ii + nvec, hi);  // chunk [ii, iimax)
    for (int i = ii; i < iimax; ++i) {
      ptr[i - lo] = (int)c2.size();
      for (int j = A.rowptr[i]; j < A.rowptr[i + 1]; ++j) {
        int c = A.col[j];
        if (c < lo || c >= hi) continue;
        if (c < ii || c >= iimax || c == i) {
          c2.push_back(c - lo);
          v2.push_back(A.val[j]);
          if (ddm) w2.push_back(alo ? A.val_lo[j] : 0.0);
        }
      }
    }
  }
  ptr[nb] = (int)c2.size();
  std::vector<int> dptr(nb, -1);
  for (int i = 0; i < nb; ++i)
    for (int q = ptr[i]; q < ptr[i + 1]; ++q)
      if (c2[q] == i) { dptr[i] = q; break; }
  for (int i = 0; i < nb; ++i) if (dptr[i] < 0) return false;
  // IKJ elimination (ILU_LU_crs): row i, for entries k < i in stored order
  for (int i = 1; i < nb; ++i) {
    for (int kc = ptr[i]; kc < ptr[i + 1]; ++kc) {
      int k = c2[kc];
      if (k > i - 1) break;
      int kd = dptr[k];
      if (ddm) {
        const dd l = dd_div(mkdd(v2[kc], w2[kc]), mkdd(v2[kd], w2[kd]));
        v2[kc] = l.hi;
        w2[kc] = l.lo;
      } else {
        v2[kc] = v2[kc] / v2[kd];
      }
      for (int jc = ptr[i]; jc < ptr[i + 1]; ++jc) {
        int j = c2[jc];
        if (j < k + 1) continue;
        for (int q = kd; q < ptr[k + 1]; ++q)
          if (c2[q] == j) {
            if (ddm) {
              const dd u = dd_sub(mkdd(v2[jc], w2[jc]), dd_mul(mkdd(v2[kc], w2[kc]), mkdd(v2[q], w2[q])));
              v2[jc] = u.hi;
              w2[jc] = u.lo;
            } else {
              v2[jc] = std::fma(-v2[kc], v2[q], v2[jc]);
            }
            break;
          }
      }
    }
  }
  // forward: strictly-lower entries in stored order, unit diagonal
  F.ptr.assign(nb + 1, 0); F.dep.clear(); F.val.clear(); F.diag.assign(nb, 1.0);
  B.ptr.assign(nb + 1, 0); B.dep.clear(); B.val.clear(); B.diag.assign(nb, 1.0);
  F.val_lo.clear(); B.val_lo.clear(); F.diag_lo.clear(); B.diag_lo.clear();
  if (ddm) { F.diag_lo.assign(nb, 0.0); B.diag_lo.assign(nb, 0.0); }
  for (int i = 0; i < nb; ++i) {
    F.ptr[i] = (int)F.dep.size();
    for (int q = ptr[i]; q < ptr[i + 1]; ++q)
      if (c2[q] < i) {
        F.dep.push_back(c2[q]);
        F.val.push_back(v2[q]);
        if (ddm) F.val_lo.push_back(w2[q]);
      }
  }
  F.ptr[nb] = (int)F.dep.size();
  // backward: strictly-upper entries in stored order, divide by u_ii
  for (int i = 0; i < nb; ++i) {
    B.ptr[i] = (int)B.dep.size();
    for (int q = ptr[i]; q < ptr[i + 1]; ++q)
      if (c2[q] > i) {
        B.dep.push_back(c2[q]);
        B.val.push_back(v2[q]);
        if (ddm) B.val_lo.push_back(w2[q]);
      }
    B.diag[i] = v2[dptr[i]];
    if (ddm) B.diag_lo[i] = w2[dptr[i]];
  }
  B.ptr[nb] = (int)B.dep.size();
  return true;
}

// D-ILU of one block (DILU_DInv_crs + init_solve_ILU_crs), ilu_method 1
// (diagonal match, ILUD) or 2 (row-sum match, MLUD).
// ddm: D and Dinv in double-double (ddDILU_DInv_crs, src/ddmatrix.f90:1451-1662,
// zero pivots replaced by 1e-100), val_lo / diag_lo filled.
bool dilu_block(const HostCSR& A, int lo, int hi, int nvec, int method, BlockSweep& F,
                BlockSweep& B, bool ddm = false) {
  const int nb = hi - lo;
  const int* rp = A.rowptr.data(); const int* cl = A.col.data(); const double* a = A.val.data();
  const double* al = (ddm && !A.val_lo.empty()) ? A.val_lo.data() : nullptr;
  auto ad = [&](int j) { return mkdd(a[j], al ? al[j] : 0.0); };
  for (int k = lo; k < hi; ++k) {
    bool found = false;
    for (int j = rp[k]; j < rp[k + 1]; ++j) {
      if (cl[j] < lo || cl[j] >= hi) continue;
      if (cl[j] == k) { found = true; break; }
    }
    if (!found) return false;
  }
  std::vector<double> D(nb);
  std::vector<dd> Ddd(ddm ? nb : 0);
  for (int k = lo; k < hi; ++k) {
    int ist = lo + ((k - lo) / nvec) * nvec;
    int ien = std::min(ist + nvec - 1, hi - 1);
    int jstart = -1;
    for (int j = rp[k]; j < rp[k + 1]; ++j) {
      int c = cl[j];
      if (c < lo || c >= hi) continue;
      if (jstart == -1) jstart = j;
      if (c != k) continue;
      double d = a[j];
      dd d2 = ad(j);
      for (int jr = jstart; jr <= j - 1; ++jr) {
        int kr = cl[jr];
        if (ist <= kr && kr <= ien && kr != k) continue;
        for (int i = rp[kr]; i < rp[kr + 1]; ++i) {
          int rc = cl[i];
          if (rc < lo || rc >= hi) continue;
          const bool upd = (method == PARCEL_ILUD) ? (rc == k) : (kr < rc);
          if (!upd) continue;
          if (ddm) {
            const dd dk = Ddd[kr - lo];
            dd t = dd_mul(ad(jr), ad(i));
            t = (dk.hi == 0.0 && dk.lo == 0.0) ? dd_div(t, mkdd(1.0e-100, 0.0)) : dd_div(t, dk);
            d2 = dd_sub(d2, t);
          } else {
            d = d - a[jr] * a[i] / D[kr - lo];
          }
          if (method == PARCEL_ILUD) break;
        }
      }
      D[k - lo] = d;
      if (ddm) Ddd[k - lo] = d2;
      break;
    }
  }
  F.ptr.assign(nb + 1, 0); F.dep.clear(); F.val.clear(); F.diag.assign(nb, 0.0);
  B.ptr.assign(nb + 1, 0); B.dep.clear(); B.val.clear(); B.diag.assign(nb, 0.0);
  F.val_lo.clear(); B.val_lo.clear(); F.diag_lo.clear(); B.diag_lo.clear();
  if (ddm) { F.diag_lo.assign(nb, 0.0); B.diag_lo.assign(nb, 0.0); }
  for (int k = lo; k < hi; ++k) {
    double dinv = 1.0 / D[k - lo];
    if (ddm) {
      const dd dk = Ddd[k - lo];
      const dd v = (dk.hi == 0.0 && dk.lo == 0.0) ? dd_div(mkdd(1.0, 0.0), mkdd(1.0e-100, 0.0))
                                                  : dd_div(mkdd(1.0, 0.0), dk);
      dinv = v.hi;
      F.diag_lo[k - lo] = v.lo;
      B.diag_lo[k - lo] = v.lo;
    }
    F.diag[k - lo] = dinv;
    B.diag[k - lo] = dinv;
    int ist = lo + ((k - lo) / nvec) * nvec;
    int ien = std::min(ist + nvec - 1, hi - 1);
    // L range: first in-block entry .. (first in-block entry with col >= ist) - 1
    int ls = rp[k + 1] - 1, le = rp[k];
    for (int j = rp[k]; j < rp[k + 1]; ++j) { if (cl[j] < lo) continue; ls = j; break; }
    for (int j = rp[k]; j < rp[k + 1]; ++j) {
      if (cl[j] < lo) continue;
      if (ist <= cl[j]) { le = j - 1; break; }
    }
    F.ptr[k - lo] = (int)F.dep.size();
    for (int j = ls; j <= le; ++j) {
      F.dep.push_back(cl[j] - lo);
      F.val.push_back(a[j]);
      if (ddm) F.val_lo.push_back(al ? al[j] : 0.0);
    }
    // U range walked downward: last in-block entry .. first entry right of the chunk
    int us = rp[k], ue = rp[k + 1] - 1;
    for (int j = rp[k + 1] - 1; j >= rp[k]; --j) { if (hi - 1 < cl[j]) continue; us = j; break; }
    for (int j = rp[k + 1] - 1; j >= rp[k]; --j) {
      if (ien < cl[j]) continue;
      ue = j + 1; break;
    }
    B.ptr[k - lo] = (int)B.dep.size();
    for (int j = us; j >= ue; --j) {
      B.dep.push_back(cl[j] - lo);
      B.val.push_back(a[j]);
      if (ddm) B.val_lo.push_back(al ? al[j] : 0.0);
    }
  }
  F.ptr[nb] = (int)F.dep.size();
  B.ptr[nb] = (int)B.dep.size();
  return true;
}

// Pack per-block level-ordered sweeps into the device layout of `Sweep`.
// ring_budget: doubles of LDS available for the ring of recent levels.
// Every level is stored with `padk` entries per row (at least the widest
// row of the sweep).
// `nth`: workgroup size of the apply kernel; the LDS ring stores a level's
// values at stride nth (every thread writes its slot).
void pack_sweep(const std::vector<BlockSweep>& blocks, const std::vector<int>& blo, Sweep& S,
                int ring_budget, int padk, int nth, hipStream_t st) {
  const int nbk = (int)blocks.size();
  std::vector<int> blk_lev(nbk + 1, 0);
  int maxw = 0, maxk = 0;
  for (int b = 0; b < nbk; ++b) {
    const BlockSweep& B = blocks[b];
    int nl = B.lev_start.empty() ? 0 : (int)B.lev_start.size() - 1;
    blk_lev[b + 1] = blk_lev[b] + nl;
    for (int l = 0; l < nl; ++l) {
      int m = B.lev_start[l + 1] - B.lev_start[l];
      maxw = std::max(maxw, m);
      for (int t = 0; t < m; ++t) {
        int i = B.order[B.lev_start[l] + t];
        maxk = std::max(maxk, B.ptr[i + 1] - B.ptr[i]);
      }
    }
  }
  const int nlev = blk_lev[nbk];
  maxk = std::max(maxk, padk);
  // ring of the last W levels in LDS (W >= 2 to be useful: level l reads l-1)
  int W = (maxw > 0 && maxw <= nth) ? ring_budget / nth : 0;
  if (W > 4) W = 4;
  if (W < 2) W = 0;
  std::vector<int4> meta(nlev);
  std::vector<long> eo(nlev + 1, 0);
  std::vector<int> ro(nlev + 1, 0);
  for (int b = 0; b < nbk; ++b) {
    const BlockSweep& B = blocks[b];
    int nl = blk_lev[b + 1] - blk_lev[b];
    for (int l = 0; l < nl; ++l) {
      int g = blk_lev[b] + l;
      int m = B.lev_start[l + 1] - B.lev_start[l];
      int k = 0;
      for (int t = 0; t < m; ++t) {
        int i = B.order[B.lev_start[l] + t];
        k = std::max(k, B.ptr[i + 1] - B.ptr[i]);
      }
      (void)k;
      // every level of the sweep gets the sweep's widest row count, so the
      // apply kernel can load a fixed number of entries per row
      meta[g] = make_int4(m, maxk, 0, 0);
      eo[g + 1] = (long)maxk * m;
      ro[g + 1] = m;
    }
  }
  for (int g = 0; g < nlev; ++g) { eo[g + 1] += eo[g]; ro[g + 1] += ro[g]; }
  const long nent = eo[nlev];
  if (nent > 0x7fffffffL) throw hip_error("block-Jacobi factor too large for int32 offsets", PARCEL_HIP_ERR_UNSUPPORTED);
  for (int g = 0; g < nlev; ++g) { meta[g].z = (int)eo[g]; meta[g].w = ro[g]; }
  const int nrows = ro[nlev];
  // ring mode needs every dependency within W-1 levels; check first
  bool ring_ok = W > 0;
  if (ring_ok) {
    int bad = 0;
#pragma omp parallel for schedule(dynamic) reduction(| : bad)
    for (int b = 0; b < nbk; ++b) {
      const BlockSweep& B = blocks[b];
      const int nrow = (int)B.ptr.size() - 1;
      const int nl = blk_lev[b + 1] - blk_lev[b];
      std::vector<int> lev(nrow);
      for (int l = 0; l < nl; ++l)
        for (int t = B.lev_start[l]; t < B.lev_start[l + 1]; ++t) lev[B.order[t]] = l;
      for (int i = 0; i < nrow && !bad; ++i)
        for (int q = B.ptr[i]; q < B.ptr[i + 1]; ++q) {
          int d = lev[i] - lev[B.dep[q]];
          if (d < 1 || d > W - 1) { bad = 1; break; }
        }
    }
    ring_ok = !bad;
  }
  if (!ring_ok) W = 0;
  const int zero_slot = W * nth;  // ring slot that always holds 0.0
  const int pad = W ? -(1 + zero_slot) : kPad;
  std::vector<int> rows(nrows), dep((size_t)nent, pad);
  std::vector<double> eval((size_t)nent, 0.0), diag(nrows);
  const bool lo_parts = !blocks.empty() && !blocks[0].diag_lo.empty();
  std::vector<double> eval_lo(lo_parts ? (size_t)nent : 0, 0.0), diag_lo(lo_parts ? nrows : 0);
  long real = 0, gdeps = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : real, gdeps)
  for (int b = 0; b < nbk; ++b) {
    const BlockSweep& B = blocks[b];
    const int nrow = (int)B.ptr.size() - 1;
    int nl = blk_lev[b + 1] - blk_lev[b];
    // level and position of every local row
    std::vector<int> lev(nrow), pos(nrow);
    for (int l = 0; l < nl; ++l)
      for (int t = B.lev_start[l]; t < B.lev_start[l + 1]; ++t) {
        lev[B.order[t]] = l;
        pos[B.order[t]] = t - B.lev_start[l];
      }
    for (int l = 0; l < nl; ++l) {
      int g = blk_lev[b] + l;
      int m = meta[g].x;
      for (int t = 0; t < m; ++t) {
        int i = B.order[B.lev_start[l] + t];
        rows[ro[g] + t] = blo[b] + i;
        diag[ro[g] + t] = B.diag[i];
        if (lo_parts) diag_lo[ro[g] + t] = B.diag_lo[i];
        int e = 0;
        for (int q = B.ptr[i]; q < B.ptr[i + 1]; ++q, ++e) {
          size_t at = (size_t)eo[g] + (size_t)e * m + t;
          int d = B.dep[q];
          if (W) {
            dep[at] = -(1 + (lev[d] % W) * nth + pos[d]);
          } else {
            dep[at] = blo[b] + d;
            gdeps++;
          }
          eval[at] = B.val[q];
          if (lo_parts) eval_lo[at] = B.val_lo[q];
          real++;
        }
      }
    }
  }
  S.nlev = nlev;
  S.max_width = maxw;
  S.max_k = maxk;
  S.ring = W;
  S.ring_stride = nth;
  S.entries = nent;
  S.real_entries = real;
  S.global_deps = gdeps;
  S.blk_lev.upload(blk_lev, st);
  S.meta.upload(meta, st);
  S.rows.upload(rows, st);
  S.dep.upload(dep, st);
  S.eval.upload(eval, st);
  S.diag.upload(diag, st);
  if (lo_parts) {
    S.eval_lo.upload(eval_lo, st);
    S.diag_lo.upload(diag_lo, st);
  }
  PARCEL_HIP_CHECK(hipStreamSynchronize(st));
}

// --- kernels -----------------------------------------------------------------

template <class T>
__global__ void k_scale_pj(int n, const T* __restrict__ r, const T* __restrict__ dinv, T* __restrict__ s,
                           const int* skip) {
  if (skip && *skip) return;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    s[i] = dinv ? r[i] * dinv[i] : r[i];
}

// Sweep arrays as the kernels see them; T is the arithmetic type of the
// factor values and vectors (double, or float for the FP32 apply).
template <class T>
struct SweepArgs {
  const int* blk_lev;
  const int4* meta;
  const int* rows;
  const int* dep;
  const T* eval;
  const T* diag;
  int ring;
  int stride;
  int ldsbar;  // every dependency is in the LDS ring: level barriers need no HBM wait
};

// Barrier that orders LDS accesses only.  __syncthreads() also waits for
// every outstanding global load and store of the wave (CDNA counts both in
// vmcnt), which would drain the prefetch pipeline at every level.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

constexpr int kKR = 8;  // widest rows handled by the register-pipelined path

// Level data of one thread, filled in two pipeline stages: structure
// (row id, dependencies, values, scale) two levels ahead, vector operands
// (r and q, or the forward value s) one level ahead.
template <int K, class T>
struct LevelRegs {
  int row;
  int m;     // rows in the level (threads >= m only shadow the last row)
  int dep[K];
  T val[K];
  T d;
  T a;  // forward: r[row]; backward: s[row] (the forward value)
  T b;  // forward with the fused update: q[row]
};

constexpr int kMetaLds = 2048;  // levels of metadata held in LDS at a time

// The pipelined path keeps its loads branch-free: threads beyond the
// level's width load and compute the level's last row as shadows, and
// look-ahead levels past the block's end re-read its last level.  Only the
// owning thread (t < m) stores to HBM: a shadow's inputs may be read after
// the owner has already updated them (r in the fused update, s in the
// backward sweep), so its result must never be written.
template <bool FWD, int MODE, int K, class T>
__device__ __forceinline__ void stage_struct(const SweepArgs<T>& A, const int4* sm, int base, int l, int l1,
                                             LevelRegs<K, T>& R) {
  const int4 mt = sm[min(l, l1 - 1) - base];
  const int t = min((int)threadIdx.x, mt.x - 1);
  R.m = mt.x;
  R.row = A.rows[mt.w + t];
#pragma unroll
  for (int e = 0; e < K; ++e) {
    R.dep[e] = A.dep[mt.z + e * mt.x + t];
    R.val[e] = A.eval[mt.z + e * mt.x + t];
  }
  if (!(FWD && MODE == BJ_ILU0)) R.d = A.diag[mt.w + t];
}

template <bool FWD, bool FUSED, int K, class T>
__device__ __forceinline__ void stage_vec(LevelRegs<K, T>& R, const T* rin, const T* q, const T* s) {
  if (FWD) {
    R.a = rin[R.row];
    if (FUSED) R.b = q[R.row];
  } else {
    R.a = s[R.row];
  }
}

template <bool FWD, int MODE, int K, bool RING, bool FUSED, class T>
__device__ __forceinline__ void compute_level(const SweepArgs<T>& A, const int4* sm, int base, int l0, int L,
                                              const LevelRegs<K, T>& C, T* s, T* rout, double alpha, T* ring) {
  const int t = threadIdx.x;
  const bool own = t < C.m;
  T z0 = C.a;
  if (FWD && FUSED) {
    z0 = fma(-(T)alpha, C.b, C.a);
    if (own) rout[C.row] = z0;
  }
  // FWD:      z = z0 - sum dep*val ; ILU0: z ; DILU: d*z
  // BWD ILU0: z = z0 - sum dep*val ; z / d
  // BWD DILU: z = sum dep*val ; z0 - d*z
  // Padding entries only follow a row's real entries, so skipping them
  // keeps the stored order.
  T z = (!FWD && MODE == BJ_DILU) ? (T)0 : z0;
#pragma unroll
  for (int e = 0; e < K; ++e) {
    const int c = C.dep[e];
    T v;
    if (RING) {
      // padding points at a slot holding 0.0 with value 0.0:
      // fma(+-0, 0, z) == z exactly, so no branch is needed
      v = ring[-c - 1];
    } else {
      if (c == kPad) continue;
      v = s[c];
    }
    if (!FWD && MODE == BJ_DILU) z = fma(v, C.val[e], z);
    else z = fma(-v, C.val[e], z);
  }
  T v;
  if (FWD) v = MODE == BJ_ILU0 ? z : C.d * z;
  else v = MODE == BJ_ILU0 ? z / C.d : fma(-C.d, z, z0);
  if (own) s[C.row] = v;
  if (RING) ring[((L - l0) % A.ring) * A.stride + t] = v;
  (void)sm; (void)base;
}

// Generic level (wide levels, long rows): direct loads, no pipelining.
template <bool FWD, int MODE, class T>
__device__ void compute_level_generic(const SweepArgs<T>& A, const int4 mt, int l0, int L, const T* rin, T* s,
                                      const T* q, T* rout, double alpha, T* ring) {
  const int m = mt.x, k = mt.y;
  const int W = A.ring;
  for (int tt = threadIdx.x; tt < m; tt += blockDim.x) {
    const int row = A.rows[mt.w + tt];
    T z0;
    if (FWD) {
      z0 = rin[row];
      if (q) { z0 = fma(-(T)alpha, q[row], z0); rout[row] = z0; }
    } else {
      z0 = s[row];
    }
    T z = (!FWD && MODE == BJ_DILU) ? (T)0 : z0;
    for (int e = 0; e < k; ++e) {
      const int c = A.dep[mt.z + e * m + tt];
      const T a = A.eval[mt.z + e * m + tt];
      T v;
      if (W) {
        v = ring[-c - 1];
      } else {
        if (c == kPad) break;
        v = s[c];
      }
      if (!FWD && MODE == BJ_DILU) z = fma(v, a, z);
      else z = fma(-v, a, z);
    }
    T v;
    if (FWD) v = MODE == BJ_ILU0 ? z : A.diag[mt.w + tt] * z;
    else v = MODE == BJ_ILU0 ? z / A.diag[mt.w + tt] : fma(-A.diag[mt.w + tt], z, z0);
    s[row] = v;
    if (W) ring[((L - l0) % W) * A.stride + tt] = v;
  }
}

template <class T>
__device__ __forceinline__ void level_barrier(const SweepArgs<T>& A) {
  if (A.ldsbar) lds_barrier();
  else __syncthreads();
}

// Load the metadata of levels [c0, min(c0 + kMetaLds, l1)) into LDS.
template <class T>
__device__ __forceinline__ void load_meta(const SweepArgs<T>& A, int4* sm, int c0, int l1) {
  __syncthreads();
  const int cnt = min(kMetaLds, l1 - c0);
  for (int j = threadIdx.x; j < cnt; j += blockDim.x) sm[j] = A.meta[c0 + j];
  __syncthreads();
}

// Register-pipelined sweep: every row of the sweep has K entries (padded)
// and every level fits the workgroup.  Level l computes with its structure
// loaded kPS levels earlier and its vector operands kPV levels earlier
// (kPS + 1 register sets in rotation, the rotation unrolled so every set is
// a fixed set of registers): a level waits on a global round trip only when
// kPV levels of work are shorter than one.
constexpr int kPS = 4, kPV = 2, kPD = kPS + 1;

template <bool FWD, int MODE, int K, bool RING, bool FUSED, class T>
__device__ void run_sweep_fast(const SweepArgs<T>& A, const T* rin, T* s, const T* q, T* rout, double alpha,
                               T* ring, int4* sm) {
  const int l0 = A.blk_lev[blockIdx.x], l1 = A.blk_lev[blockIdx.x + 1];
  if (l1 <= l0) return;
  // metadata window of kMetaLds levels starting at `base`; it moves forward
  // before the staging of a round would leave it
  int base = l0;
  if (RING && threadIdx.x == 0) ring[A.ring * A.stride] = (T)0;  // zero slot for padding
  load_meta(A, sm, base, l1);  // its barriers also publish the zero slot
  LevelRegs<K, T> R[kPD];
#pragma unroll
  for (int d = 0; d < kPS; ++d) stage_struct<FWD, MODE, K, T>(A, sm, base, l0 + d, l1, R[d]);
#pragma unroll
  for (int d = 0; d < kPV; ++d) stage_vec<FWD, FUSED, K, T>(R[d], rin, q, s);
  for (int l = l0; l < l1; l += kPD) {
    if (l + kPD - 1 + kPS >= base + kMetaLds) {
      base = min(l + kPS, l1 - 1);
      load_meta(A, sm, base, l1);
    }
#pragma unroll
    for (int j = 0; j < kPD; ++j) {
      // R[j] holds level l + j; the set after it in the rotation takes the
      // structure of level l + j + kPS (it held level l + j - 1, done)
      stage_struct<FWD, MODE, K, T>(A, sm, base, l + j + kPS, l1, R[(j + kPS) % kPD]);
      stage_vec<FWD, FUSED, K, T>(R[(j + kPV) % kPD], rin, q, s);
      if (l + j < l1) {
        compute_level<FWD, MODE, K, RING, FUSED, T>(A, sm, base, l0, l + j, R[j], s, rout, alpha, ring);
        level_barrier(A);
      }
    }
  }
}

template <bool FWD, int MODE, class T>
__device__ void run_sweep_generic(const SweepArgs<T>& A, const T* rin, T* s, const T* q, T* rout, double alpha,
                                  T* ring) {
  const int l0 = A.blk_lev[blockIdx.x], l1 = A.blk_lev[blockIdx.x + 1];
  if (A.ring && threadIdx.x == 0) ring[A.ring * A.stride] = (T)0;  // zero slot for padding
  __syncthreads();
  for (int l = l0; l < l1; ++l) {
    compute_level_generic<FWD, MODE, T>(A, A.meta[l], l0, l, rin, s, q, rout, alpha, ring);
    __syncthreads();
  }
}

// One workgroup per block-Jacobi block: forward then backward sweep of the
// block (blocks are independent, so no grid-wide synchronisation).  With
// `fuse`, the forward sweep first applies r -= alpha*q of CG iteration
// fuse.it (alpha = slot.rho0 / gamma, gamma reduced from the SpMV partials,
// the same value every workgroup computes) and returns at once if the
// solve has already ended.  K > 0: every sweep row has at most K entries and
// every level fits the workgroup (register-pipelined path); K = 0: generic.
template <int MODE, int NTH, int K, class T>
__global__ void __launch_bounds__(NTH) k_bj_apply(SweepArgs<T> F, SweepArgs<T> B, const T* rin, T* s,
                                                  CGFuseT<T> fuse, const int* skip) {
  if (skip && *skip) return;
  extern __shared__ double ring_lds[];
  T* ring = reinterpret_cast<T*>(ring_lds);
  __shared__ int4 smeta[kMetaLds];
  __shared__ double sh[NTH / 64];
  double alpha = 0.0;
  const T* q = nullptr;
  T* rout = nullptr;
  if (fuse.st) {
    const int d = fuse.st->done_at;
    if (d != 0 && d <= fuse.it) return;
    CGSlot* sl = &fuse.st->slot[fuse.it & 1];
    // gamma in the fixed order: threads 0..kNT-1 sum partials t, t+kNT, ...
    double a = 0.0;
    if (threadIdx.x < kNT)
      for (int j = threadIdx.x; j < fuse.nb_prod; j += kNT) a = a + fuse.part_gamma[j];
    a = wave_sum(a);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
    __syncthreads();
    double gamma = sh[0];
    for (int k = 1; k < kWaves; ++k) gamma = gamma + sh[k];
    alpha = sl->rho0 / gamma;
    if (blockIdx.x == 0 && threadIdx.x == 0) sl->alpha = alpha;
    q = fuse.q;
    rout = fuse.r;
    rin = fuse.r;
  }
  if (K > 0) {
    constexpr int KK = K > 0 ? K : 1;
    if (q) {
      if (F.ring) run_sweep_fast<true, MODE, KK, true, true, T>(F, rin, s, q, rout, alpha, ring, smeta);
      else run_sweep_fast<true, MODE, KK, false, true, T>(F, rin, s, q, rout, alpha, ring, smeta);
    } else {
      if (F.ring) run_sweep_fast<true, MODE, KK, true, false, T>(F, rin, s, q, rout, alpha, ring, smeta);
      else run_sweep_fast<true, MODE, KK, false, false, T>(F, rin, s, q, rout, alpha, ring, smeta);
    }
    __syncthreads();  // forward values in HBM before the backward sweep reads them
    if (B.ring) run_sweep_fast<false, MODE, KK, true, false, T>(B, nullptr, s, nullptr, nullptr, 0.0, ring, smeta);
    else run_sweep_fast<false, MODE, KK, false, false, T>(B, nullptr, s, nullptr, nullptr, 0.0, ring, smeta);
  } else {
    run_sweep_generic<true, MODE, T>(F, rin, s, q, rout, alpha, ring);
    __syncthreads();
    run_sweep_generic<false, MODE, T>(B, nullptr, s, nullptr, nullptr, 0.0, ring);
  }
}

template <class T>
SweepArgs<T> sweep_args(const Sweep& S);

template <>
SweepArgs<double> sweep_args<double>(const Sweep& S) {
  SweepArgs<double> a;
  a.eval = S.eval.p;
  a.diag = S.diag.p;
  a.blk_lev = S.blk_lev.p;
  a.meta = S.meta.p;
  a.rows = S.rows.p;
  a.dep = S.dep.p;
  a.ring = S.ring;
  a.stride = S.ring_stride;
  a.ldsbar = (S.ring > 0 && S.global_deps == 0) ? 1 : 0;
  return a;
}

template <>
SweepArgs<float> sweep_args<float>(const Sweep& S) {
  SweepArgs<float> a;
  a.eval = S.eval32.p;
  a.diag = S.diag32.p;
  a.blk_lev = S.blk_lev.p;
  a.meta = S.meta.p;
  a.rows = S.rows.p;
  a.dep = S.dep.p;
  a.ring = S.ring;
  a.stride = S.ring_stride;
  a.ldsbar = (S.ring > 0 && S.global_deps == 0) ? 1 : 0;
  return a;
}

// --- line-wavefront form of the FP64 apply ------------------------------------
//
// A block of P grid planes of ny lines of nx cells (P = 1: a strip of lines
// of one plane, ny its line count) whose factor rows couple only to the cell
// neighbour (offset 1), the line neighbour (offset nx) and the plane
// neighbour (offset nx ny) is swept as a wavefront.  In sweep coordinates
// (forward: cell i, line j, plane k; backward: nx-1-i, ny-1-j, P-1-k)
// wavefront w = k WP + jw (WP = ceil(ny / 64) wavefronts per plane) owns
// lines j = 64 jw + lane of plane k, and a thread computes cell
// i = t - j - lag(w) at step t, lag(w) = 8 jw + 9 k.  Then the cell neighbour
// was computed at step t-1 by the thread itself, the line neighbour at t-1 by
// the lane below (a DPP wave shift) or, for lane 0, at t-9 by lane 63 of the
// wavefront before (through an LDS line indexed by cell), and the plane
// neighbour at t-9 by the same lane of wavefront w - WP (through an LDS ring
// of the last kLR steps of every line).  One LDS barrier per kLG = 8 steps
// publishes both: a value written at step t-9 or earlier was written in an
// earlier group.  Every row sums its couplings in the stored order with the
// same fma sequence as the level-scheduled kernels, so the values are
// bit-identical to theirs.
//
// Memory: the rows of one step of a wavefront are one anti-diagonal of its
// lines, nx - 1 apart in the natural order of the vectors.  The factor
// arrays are stored per wavefront in anti-diagonal order (one coalesced load
// per step); the vectors go through LDS in groups of kLG steps: each line of
// a group is kLG consecutive cells, which the wavefront loads and stores as
// 64-byte segments (lane 8j + c: line 8k + j, cell c, k = 0..7) and the
// owning lane reads and writes in its slot row.  The fused update r -= alpha q
// is elementwise and is applied (and stored) while a group's inputs are staged.
constexpr int kLG = 8;
static_assert(kLG == 8, "the staging moves 8 lines x 8 cells per instruction");
constexpr int kLS = kLG + 1;  // doubles per line of a wavefront's slots (padding against bank conflicts)
constexpr int kLPlag = kLG + 1;  // steps between a line and the same line of the next plane
// Ring depth of the plane hand-over: the value of step s is read at s + 9, in
// group floor((s + 9) / 8); it may be overwritten from the next group on,
// at s + kLR >= 8 (floor((s + 9) / 8) + 1), which holds for every s at 17.
constexpr int kLR = 17;

// Entries on the anti-diagonals 0..d-1 of an nl x nx sub-block (anti-diagonal
// d' holds min(d', nl-1) - max(0, d'-nx+1) + 1 cells); the cell of line l on
// anti-diagonal d sits at line_doff(d) + l - max(0, d - nx + 1).
__host__ __device__ inline long line_doff(int d, int nl, int nx) {
  const long a = d < nl ? d : nl;
  const long over = d > nx ? d - nx : 0;
  return a * (a + 1) / 2 + (long)(d > nl ? d - nl : 0) * nl - over * (over + 1) / 2;
}

struct LineArgs {
  const int* bstart;
  const double* fc;  // per row: line and cell couplings, in the sweep's stored order
  const double* fp;  // per row: plane coupling (P > 1)
  const double* fd;
  const double* bc;
  const double* bp;
  const double* bd;
  double* sink;  // 64 doubles per block: the stores of lanes off the block's cells
  int nx;
  int ny;        // lines per plane (0: every block is one plane of its own line count)
};

// v of the lane below (DPP wave_shr:1); lane 0 receives `first`.
__device__ __forceinline__ double from_lane_below(double first, double v) {
  const long long o = __double_as_longlong(first), x = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp((int)o, (int)x, 0x138, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(x >> 32), 0x138, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Orders the LDS accesses of the lanes of one wavefront (no s_barrier: the
// slots are private to the wavefront).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// Factor operands of one step, loaded kLG steps ahead.
struct LineOps {
  double c0, c1;  // the row's line and cell couplings in stored order
  double cp;      // plane coupling
  double d;       // scale
};

// The couplings' stored order is fixed by the factor: the forward rows of
// both factors and the backward rows of D-ILU hold the plane neighbour
// first, then the line neighbour, then the cell neighbour (ILU(0) keeps the
// CRS order of the block, D-ILU walks its upper part downward); ILU(0)'s
// backward rows hold them the other way round.
template <bool FWD, int MODE>
constexpr bool line_cell_first() { return !FWD && MODE == BJ_ILU0; }

// Blocks of one plane with at most four wavefronts stage through a 16-cell
// ring per line instead (kLA doubles with padding): each group loads one
// aligned 8-cell chunk of every line and stores its newest complete one, so
// the vector traffic moves whole 64-byte sectors when lines start on one.
// Row stride 20 doubles (16 ring cells, padding slot 16): a step's lanes read
// cell (a - lane) mod 16 of their rows, which at this stride fall in distinct
// banks (at 17 they would all share a bank pair).
constexpr int kLA = 20;
__host__ __device__ constexpr bool line_aligned(bool planes, int W) { return !planes && W <= 4; }

// Dynamic LDS of k_bj_lines, in doubles: slots, line hand-over, plane ring.
__host__ __device__ inline size_t line_lds_doubles(int W, int nx, int planes) {
  const size_t slots = (size_t)64 * W * (line_aligned(planes > 1, W) ? kLA : kLS);
  return slots + (size_t)W * nx + (planes > 1 ? (size_t)64 * W * kLR : 0);
}

template <bool FWD, int MODE, bool FUSED, bool PL, int W>
__device__ __forceinline__ void line_sweep(const LineArgs& A, int lo, int ny, int np, const double* rin,
                                           double* s, const double* q, double* rout, double alpha,
                                           double* lds) {
  constexpr bool kScale = !(FWD && MODE == BJ_ILU0);
  constexpr bool kCellFirst = line_cell_first<FWD, MODE>();
  const int nx = A.nx;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int WP = (ny + 63) >> 6;             // wavefronts per plane
  const int kp = w / WP, jw = w - kp * WP;   // this wavefront's plane and line group
  const bool wact = kp < np;                 // the others idle through the barriers
  const int j = 64 * jw + lane;
  const bool lact = wact && j < ny;
  const int nlw = wact ? min(64, ny - 64 * jw) : 1;  // lines of this wavefront
  const int lag = 8 * jw + kLPlag * kp;
  const int nsteps = nx + ny - 1 + 8 * (WP - 1) + kLPlag * (np - 1);
  const double* cc = FWD ? A.fc : A.bc;
  const double* cpl = FWD ? A.fp : A.bp;
  const double* dv = FWD ? A.fd : A.bd;
  double* sink = A.sink + (long)blockIdx.x * 64 + lane;
  // sweep line (j, kp) -> first row of the line; the wavefront's sub-block of entries
  auto line_row = [&](int jj) -> int {
    return lo + ((FWD ? kp : np - 1 - kp) * ny + (FWD ? jj : ny - 1 - jj)) * nx;
  };
  const long sub = (long)lo + (long)(kp * ny + 64 * jw) * nx;
  constexpr bool ALN = line_aligned(PL, W);
  constexpr int kSL = ALN ? kLA : kLS;
  double* ws = lds + w * 64 * kSL;  // this wavefront's slots: [line][step of the group] / [line][cell mod 16]
  double* my = ws + lane * kSL;
  double* hand = lds + 64 * W * kSL;  // [wavefront][cell]: lane 63's values
  double* ring = hand + W * nx;       // [thread][step mod kLR]: every line's values (P > 1)
  // staging lanes: lane 8j + c moves cell c of the group for local line 8k + j;
  // rows of the current group (stores) and of the next (loads), -1 off the block
  const int sj = lane >> 3, sc = lane & 7;
  int rcur[8], rnext[8], cnext[8];
  // cells t1 - j - lag + 0..7 of the group starting at step t1 (ALN = false)
  auto stage_rows = [&](int t1, int off) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int jj = 64 * jw + 8 * k + sj;
      const int ip = t1 + off + sc - jj - lag;
      const bool v = wact && jj < ny && (unsigned)ip < (unsigned)nx;
      rnext[k] = v ? line_row(jj) + (FWD ? ip : nx - 1 - ip) : -1;
      cnext[k] = ip;
    }
  };
  // ALN: the chunk cells advance by exactly 8 per group, so each staging lane
  // carries the cell of its input chunk (loaded during a group for the next)
  // and of its output chunk (stored after a group) and the first row of its
  // line (backward: the row of cell 0, rows descending)
  int lrow[8], cin[8], cout[8];
  bool lv[8];
  if (ALN) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int jj = 64 * jw + 8 * k + sj;
      const int a0 = -jj - lag;  // the line's cell at step 0
      lv[k] = wact && jj < ny;
      lrow[k] = line_row(min(jj, ny - 1)) + (FWD ? 0 : nx - 1);
      cout[k] = ((a0 >> 3) << 3) + sc;
      cin[k] = (((a0 + 15) >> 3) << 3) + sc;
    }
  }
  auto chunk_rows = [&](const int (&cell)[8], bool first) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int jj = 64 * jw + 8 * k + sj;
      const bool v = lv[k] && (unsigned)cell[k] < (unsigned)nx && !(first && ((-jj - lag) & 7) == 0);
      rnext[k] = v ? (FWD ? lrow[k] + cell[k] : lrow[k] - cell[k]) : -1;
      cnext[k] = cell[k];
    }
  };
  double va[8], vb[8];
  auto stage_load = [&]() {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const long row = rnext[k] >= 0 ? rnext[k] : lo;
      va[k] = FWD ? rin[row] : s[row];
      if (FUSED) vb[k] = q[row];
    }
  };
  auto stage_fill = [&]() {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      double z0 = va[k];
      if (FUSED) {
        z0 = fma(-alpha, vb[k], va[k]);
        *(rnext[k] >= 0 ? rout + rnext[k] : sink) = z0;
      }
      // aligned ring: cells off the block go to the padding slot 16
      ws[(8 * k + sj) * kSL + (ALN ? (rnext[k] >= 0 ? (cnext[k] & 15) : 16) : sc)] = z0;
      rcur[k] = rnext[k];
    }
  };
  auto stage_store = [&]() {
#pragma unroll
    for (int k = 0; k < 8; ++k)
      *(rcur[k] >= 0 ? s + rcur[k] : sink) = ws[(8 * k + sj) * kSL + (ALN ? (cnext[k] & 15) : sc)];
  };
  // factor loads: anti-diagonal dl = t - lag - 64 jw of the wavefront's own
  // lines for the next load and the entry offset of its line 0 on it
  // (wave-uniform, advanced one step per load)
  int dl = -lag - 64 * jw;
  long offl = 0;
  auto load = [&](int t, LineOps& o) {
    const int ip = t - j - lag;
    const bool on = lact && (unsigned)ip < (unsigned)nx;
    const long e = on ? sub + offl + lane : (long)lo;
    const double2 cv = reinterpret_cast<const double2*>(cc)[e];
    o.c0 = cv.x;
    o.c1 = cv.y;
    if (PL) o.cp = cpl[e];
    if (kScale) o.d = dv[e];
    if (dl >= 0) offl += min(dl, nlw - 1) - max(0, dl - nx + 1) + 1 - (dl >= nx - 1 ? 1 : 0);
    ++dl;
  };
  // An absent coupling has coefficient +0 and neighbour value kZero, whose
  // product term is -0 in the row's fma (fma(-(+0), +0, z) forward and in
  // ILU(0)'s backward rows, fma(-0, +0, z) in D-ILU's), and z + (-0) == z for
  // every z, the signed zeros included: the same bits as skipping the term.
  // The values start as kZero (the cell before a line's first, the line below
  // a plane's first, the plane below the block's first).
  constexpr double kZero = (!FWD && MODE == BJ_DILU) ? -0.0 : 0.0;
  double yprev = kZero;  // the thread's value of the previous step (its cell neighbour)
  auto step = [&](int t, double a, double hv, double pv, const LineOps& o) -> double {
    const int ip = t - j - lag;
    const bool on = lact && (unsigned)ip < (unsigned)nx;
    const double yd = from_lane_below(hv, yprev);  // line neighbour
    double y;
    if (!FWD && MODE == BJ_DILU) {
      // plane, line, cell; y = a - d * sum
      double z = PL ? fma(pv, o.cp, 0.0) : fma(yd, o.c0, 0.0);
      z = PL ? fma(yd, o.c0, z) : z;
      z = fma(yprev, o.c1, z);
      y = fma(-o.d, z, a);
    } else if (kCellFirst) {
      double z = fma(-yprev, o.c0, a);
      z = fma(-yd, o.c1, z);
      if (PL) z = fma(-pv, o.cp, z);
      y = z / o.d;
    } else {
      double z = PL ? fma(-pv, o.cp, a) : a;
      z = fma(-yd, o.c0, z);
      z = fma(-yprev, o.c1, z);
      y = MODE == BJ_ILU0 ? z : o.d * z;
    }
    if (W > 1 && lane == 63 && jw < WP - 1 && on) hand[w * nx + ip] = y;
    if (PL) ring[tid * kLR + t % kLR] = y;
    yprev = on ? y : yprev;
    return y;
  };
  LineOps R[kLG];
  // one group of kLG steps: the next group's vectors are in flight during it
  auto group = [&](int t0) {
    if (ALN) chunk_rows(cin, false);
    else stage_rows(t0 + kLG, 0);
    stage_load();
    wave_lds_sync();
    // the group's inputs and hand-over values, read before its first step
    // (the hand-over values were written in earlier groups)
    double av[kLG], hv[kLG], pv[kLG];
#pragma unroll
    for (int u = 0; u < kLG; ++u) {
      av[u] = my[ALN ? ((t0 + u - j - lag) & 15) : u];
      const int ip = min(max(t0 + u - j - lag, 0), nx - 1);
      hv[u] = (W > 1 && jw > 0) ? hand[(w - 1) * nx + ip] : kZero;
      if (PL) pv[u] = (kp > 0) ? ring[(tid - 64 * WP) * kLR + (t0 + u + kLR - kLPlag) % kLR] : kZero;
    }
#pragma unroll
    for (int u = 0; u < kLG; ++u) {
      const int ip = t0 + u - j - lag;
      const bool on = lact && (unsigned)ip < (unsigned)nx;
      my[ALN ? (on ? (ip & 15) : 16) : u] = step(t0 + u, av[u], hv[u], PL ? pv[u] : kZero, R[u]);
      load(t0 + u + kLG, R[u]);
    }
    wave_lds_sync();
    if (ALN) {
      // the newest complete chunk of every line: its cells were the ring's
      // since the group before (in-place: input, then the row's value)
      const int rl[8] = {rnext[0], rnext[1], rnext[2], rnext[3], rnext[4], rnext[5], rnext[6], rnext[7]};
      const int cl[8] = {cnext[0], cnext[1], cnext[2], cnext[3], cnext[4], cnext[5], cnext[6], cnext[7]};
      chunk_rows(cout, false);
#pragma unroll
      for (int k = 0; k < 8; ++k) rcur[k] = rnext[k];
      stage_store();
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        rnext[k] = rl[k];
        cnext[k] = cl[k];
        cin[k] += kLG;
        cout[k] += kLG;
      }
    } else {
      stage_store();
    }
    wave_lds_sync();
    stage_fill();
    if (W > 1) lds_barrier();
  };
  if (ALN) {
    // the chunks group 0 reads: cells a .. a + 7 of every line lie in chunks
    // a >> 3 (cout's) and (a + 7) >> 3 (cin's minus one)
    chunk_rows(cout, true);
    stage_load();
    stage_fill();
    wave_lds_sync();
    int c1[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) c1[k] = cin[k] - kLG;
    chunk_rows(c1, false);
  } else {
    stage_rows(0, 0);
  }
  stage_load();
  stage_fill();
#pragma unroll
  for (int u = 0; u < kLG; ++u) load(u, R[u]);
  // the first group runs outside the loop, so that the loop is entered with
  // the memory operations of a whole group in flight on both of its paths
  // (the compiler's wait counts then match the steady state); with aligned
  // chunks one more group stores the lines' last chunks
  group(0);
  for (int t0 = kLG; t0 < nsteps + (ALN ? kLG - 1 : 0); t0 += kLG) group(t0);
}

// One workgroup of 64 W threads per block: forward then backward sweep.  The
// fused residual update computes alpha as k_bj_apply does, with gamma summed
// in the same order (partials j, j + kNT, ... per virtual thread j < kNT,
// wavefront sums, then the kWaves sums in order).
template <int MODE, int W, bool PL>
__global__ void __launch_bounds__(64 * W) k_bj_lines(LineArgs A, const double* rin, double* s, CGFuse fuse,
                                                     const int* skip) {
  if (skip && *skip) return;
  extern __shared__ double lds_lines[];
  __shared__ double sh[kWaves];
  double alpha = 0.0;
  const double* q = nullptr;
  double* rout = nullptr;
  if (fuse.st) {
    const int d = fuse.st->done_at;
    if (d != 0 && d <= fuse.it) return;
    CGSlot* sl = &fuse.st->slot[fuse.it & 1];
    for (int vw = threadIdx.x >> 6; vw < kWaves; vw += W) {
      double a = 0.0;
      for (int j = vw * 64 + (threadIdx.x & 63); j < fuse.nb_prod; j += kNT) a = a + fuse.part_gamma[j];
      a = wave_sum(a);
      if ((threadIdx.x & 63) == 0) sh[vw] = a;
    }
    __syncthreads();
    double gamma = sh[0];
    for (int k = 1; k < kWaves; ++k) gamma = gamma + sh[k];
    alpha = sl->rho0 / gamma;
    if (blockIdx.x == 0 && threadIdx.x == 0) sl->alpha = alpha;
    q = fuse.q;
    rout = fuse.r;
    rin = fuse.r;
  }
  const int lo = A.bstart[blockIdx.x];
  const int lines = (A.bstart[blockIdx.x + 1] - lo) / A.nx;
  const int ny = A.ny > 0 ? A.ny : lines;
  const int np = lines / ny;
  if (q) line_sweep<true, MODE, true, PL, W>(A, lo, ny, np, rin, s, q, rout, alpha, lds_lines);
  else line_sweep<true, MODE, false, PL, W>(A, lo, ny, np, rin, s, nullptr, nullptr, 0.0, lds_lines);
  __syncthreads();  // forward values in memory before other threads' backward rows read them
  line_sweep<false, MODE, false, PL, W>(A, lo, ny, np, nullptr, s, nullptr, nullptr, 0.0, lds_lines);
}

constexpr size_t kLineLdsMax = 64 * 1024 - 64;  // dynamic LDS of k_bj_lines

// Set up the line-wavefront form when the block factors have its pattern
// (FP64 applies only; the level-scheduled sweeps stay for the other paths).
// PARCEL_HIP_BJ_LEVELS=1 in the environment keeps the level-scheduled apply.
bool line_setup(Plan& P, const std::vector<BlockSweep>& F, const std::vector<BlockSweep>& B,
                const std::vector<int>& blo, const std::vector<int>& bhi) {
  Precon& M = P.M;
  LineBJ& G = M.line;
  G = LineBJ();
  if (const char* e = std::getenv("PARCEL_HIP_BJ_LEVELS"))
    if (e[0] && e[0] != '0') return false;
  const int n = P.hA.n, nblk = (int)F.size();
  // the forward dependency offsets above 1: nx, and nx ny when blocks hold several planes
  long nx = 0, nxy = 0;
  for (int b = 0; b < nblk; ++b)
    for (size_t i = 0; i + 1 < F[b].ptr.size(); ++i)
      for (int q = F[b].ptr[i]; q < F[b].ptr[i + 1]; ++q) {
        const long off = (long)i - F[b].dep[q];
        if (off <= 1 || off == nx || off == nxy) continue;
        if (!nx) nx = off;
        else if (!nxy) nxy = off;
        else return false;
      }
  if (nx && nxy && nxy < nx) std::swap(nx, nxy);
  if (nx < 2 || (nxy && nxy % nx)) return false;
  const int X = (int)nx;
  const int NY = nxy ? (int)(nxy / nx) : 0;  // lines per plane (0: one plane per block)
  for (int b = 0; b + 1 < nblk; ++b)
    if (bhi[b] != blo[b + 1]) return false;
  int W = 1, np_max = 1;
  for (int b = 0; b < nblk; ++b) {
    const long len = (long)bhi[b] - blo[b];
    const long unit = nxy ? nxy : nx;
    if (len <= 0 || blo[b] % unit || len % unit) return false;
    const int lines = (int)(len / nx), ny = NY ? NY : lines, np = lines / ny;
    const int waves = ((ny + 63) / 64) * np;
    if (waves > 8) return false;
    while (W < waves) W *= 2;
    np_max = std::max(np_max, np);
  }
  const bool PL = np_max > 1;
  if (sizeof(double) * line_lds_doubles(W, X, np_max) > kLineLdsMax || (PL && W > 4)) return false;
  const bool ilu0 = M.mode == BJ_ILU0;
  std::vector<double> fc(2 * (size_t)n, 0.0), bc(2 * (size_t)n, 0.0), fd(n, 0.0), bd(n, 0.0);
  std::vector<double> fp(PL ? n : 0, 0.0), bp(PL ? n : 0, 0.0);
  int bad = 0;
#pragma omp parallel for schedule(dynamic) reduction(| : bad)
  for (int b = 0; b < nblk; ++b) {
    const int lo = blo[b], lines = (bhi[b] - lo) / X, ny = NY ? NY : lines, np = lines / ny;
    for (int k = 0; k < np && !bad; ++k)
      for (int jj = 0; jj < ny && !bad; ++jj)
        for (int x = 0; x < X; ++x) {
          const int i = (k * ny + jj) * X + x;  // local row
          for (int dir = 0; dir < 2; ++dir) {
            const BlockSweep& S = dir == 0 ? F[b] : B[b];
            const bool fwd = dir == 0;
            // the neighbours the sweep reads, in the kernel's order
            const int cn = fwd ? i - 1 : i + 1, ln = fwd ? i - X : i + X, pn = fwd ? i - X * ny : i + X * ny;
            const bool hc = fwd ? x > 0 : x < X - 1, hl = fwd ? jj > 0 : jj < ny - 1, hp = fwd ? k > 0 : k < np - 1;
            const bool cell_first = !fwd && ilu0;
            int want[3], nw = 0;
            if (cell_first) {
              if (hc) want[nw++] = cn;
              if (hl) want[nw++] = ln;
              if (hp) want[nw++] = pn;
            } else {
              if (hp) want[nw++] = pn;
              if (hl) want[nw++] = ln;
              if (hc) want[nw++] = cn;
            }
            if (S.ptr[i + 1] - S.ptr[i] != nw) { bad = 1; break; }
            double vc = 0.0, vl = 0.0, vp = 0.0;
            for (int q = S.ptr[i], m = 0; q < S.ptr[i + 1]; ++q, ++m) {
              const int dp = S.dep[q];
              if (dp != want[m]) { bad = 1; break; }
              (dp == cn ? vc : dp == ln ? vl : vp) = S.val[q];
            }
            if (bad) break;
            // sweep coordinates, the wavefront's sub-block and its anti-diagonal entry
            const int ip = fwd ? x : X - 1 - x, jp = fwd ? jj : ny - 1 - jj, kq = fwd ? k : np - 1 - k;
            const int jw = jp / 64, l = jp % 64, nlw = std::min(64, ny - 64 * jw), dg = ip + l;
            const long e = (long)lo + (long)(kq * ny + 64 * jw) * X + line_doff(dg, nlw, X) + l -
                           std::max(0, dg - X + 1);
            std::vector<double>& cv = fwd ? fc : bc;
            cv[2 * (size_t)e] = cell_first ? vc : vl;
            cv[2 * (size_t)e + 1] = cell_first ? vl : vc;
            if (PL) (fwd ? fp : bp)[(size_t)e] = vp;
            (fwd ? fd : bd)[(size_t)e] = S.diag[i];
          }
          if (bad) break;
        }
  }
  if (bad) return false;
  std::vector<int> bs(nblk + 1);
  for (int b = 0; b < nblk; ++b) bs[b] = blo[b];
  bs[nblk] = bhi[nblk - 1];
  G.bstart.upload(bs, P.stream);
  G.fc.upload(fc, P.stream);
  G.bc.upload(bc, P.stream);
  G.fd.upload(fd, P.stream);
  G.bd.upload(bd, P.stream);
  if (PL) {
    G.fp.upload(fp, P.stream);
    G.bp.upload(bp, P.stream);
  }
  G.sink.alloc((size_t)64 * nblk);
  PARCEL_HIP_CHECK(hipStreamSynchronize(P.stream));
  G.nx = X;
  G.ny = NY;
  G.planes = np_max;
  G.waves = W;
  G.on = true;
  return true;
}

void line_apply(Plan& P, const double* r, double* s, const CGFuse* fuse, const int* skip) {
  const Precon& M = P.M;
  const LineBJ& G = M.line;
  LineArgs A{G.bstart.p, G.fc.p, G.fp.p, G.fd.p, G.bc.p, G.bp.p, G.bd.p, G.sink.p, G.nx, G.ny};
  CGFuse f;
  if (fuse) f = *fuse;
  const size_t shm = sizeof(double) * line_lds_doubles(G.waves, G.nx, G.planes);
  const dim3 grid(M.nblocks);
#define PARCEL_LINE_K(W, PL)                                                                                    \
  do {                                                                                                          \
    if (M.mode == BJ_ILU0)                                                                                      \
      hipLaunchKernelGGL((k_bj_lines<BJ_ILU0, W, PL>), grid, dim3(64 * W), shm, P.stream, A, r, s, f, skip);   \
    else                                                                                                        \
      hipLaunchKernelGGL((k_bj_lines<BJ_DILU, W, PL>), grid, dim3(64 * W), shm, P.stream, A, r, s, f, skip);   \
  } while (0)
  if (G.planes > 1) {
    // eight wavefronts with the plane ring exceed the LDS budget (line_setup)
    if (G.waves == 2) PARCEL_LINE_K(2, true);
    else PARCEL_LINE_K(4, true);
  } else {
    switch (G.waves) {
      case 1: PARCEL_LINE_K(1, false); break;
      case 2: PARCEL_LINE_K(2, false); break;
      case 4: PARCEL_LINE_K(4, false); break;
      default: PARCEL_LINE_K(8, false); break;
    }
  }
#undef PARCEL_LINE_K
}

// Line-scan form of the FP32 inner apply (mixed-precision solver): when every
// block-Jacobi block is a strip of whole grid lines of one nx-cell line
// length, the strips are the equal split [ny k / nblk, ny (k+1) / nblk) of
// the ny = n / nx lines, the blocks couple only to the previous cell (offset
// 1, same line) and the line below (offset nx), and those couplings are
// symmetric bit for bit, the block ILU(0) factor L D L^T equals the block
// IC(0) factor (L D^1/2)(L D^1/2)^T in exact arithmetic, and the FP32 apply
// runs it as line scans (batched.hip, k_lscan: one workgroup per strip):
//   Ls_i = a(i,i-nx) / Ld_{i-nx} (0 on a strip's first line),
//   Lw_i = a(i,i-1) / Ld_{i-1} (0 on a line's first cell),
//   Ld_i = sqrt(fma(-Lw, Lw, fma(-Ls, Ls, a_ii))),
// computed in FP64 from the block factors' pattern and rounded to FP32 in
// the apply arrays ls, 1/Ld, -Lw/Ld, -Lw(i+1)/Ld.  Returns false (the
// level-scheduled FP32 kernel runs) for anything else.
bool lscan_setup(Plan& P, const std::vector<BlockSweep>& F, const std::vector<BlockSweep>& B,
                 const std::vector<int>& blo, const std::vector<int>& bhi) {
  Precon& M = P.M;
  LscanBJ& G = M.lscan;
  G = LscanBJ();
  if (P.prm.solver != PARCEL_SOLVER_CG_MIXED || M.mode != BJ_ILU0) return false;
  if (P.prm.independ_nvec > 1) return false;
  const HostCSR& A = P.hA;
  const int n = A.n, nblk = (int)F.size();
  // nx: the one dependency offset above 1
  long nx = 0;
  for (int b = 0; b < nblk; ++b)
    for (size_t i = 0; i + 1 < F[b].ptr.size(); ++i)
      for (int q = F[b].ptr[i]; q < F[b].ptr[i + 1]; ++q) {
        const long off = (long)i - F[b].dep[q];
        if (off <= 1 || off == nx) continue;
        if (nx) return false;  // a second offset (couplings between planes)
        nx = off;
      }
  if (nx < 1 || nx > 1024 || n % nx) return false;
  for (int b = 0; b < nblk; ++b)  // the upper couplings mirror them
    for (size_t i = 0; i + 1 < B[b].ptr.size(); ++i)
      for (int q = B[b].ptr[i]; q < B[b].ptr[i + 1]; ++q) {
        const long off = (long)B[b].dep[q] - (long)i;
        if (off != 1 && off != nx) return false;
      }
  const int ny = (int)(n / nx);
  if (nblk > ny) return false;
  for (int b = 0; b < nblk; ++b)
    if (blo[b] != (int)((long)ny * b / nblk) * nx || bhi[b] != (int)((long)ny * (b + 1) / nblk) * nx) return false;
  // a(i, j) of the block entry (i, j), 0.0 when the pattern has none
  auto entry = [&](int i, int j) -> double {
    for (int e = A.rowptr[i]; e < A.rowptr[i + 1]; ++e)
      if (A.col[e] == j) return A.val[e];
    return 0.0;
  };
  std::vector<double> L(3 * (size_t)n, 0.0);
  int bad = 0;
#pragma omp parallel for schedule(dynamic) reduction(| : bad)
  for (int b = 0; b < nblk; ++b) {
    double* Ls = L.data();
    double* Lw = Ls + n;
    double* Ld = Lw + n;
    for (int i = blo[b]; i < bhi[b] && !bad; ++i) {
      const int x = i % nx;
      double ls = 0.0, lw = 0.0;
      if (i - nx >= blo[b]) {
        const double a = entry(i, i - (int)nx);
        if (a != entry(i - (int)nx, i) || std::signbit(a) != std::signbit(entry(i - (int)nx, i))) { bad = 1; break; }
        ls = a / Ld[i - nx];
      }
      if (x > 0) {
        const double a = entry(i, i - 1);
        if (a != entry(i - 1, i) || std::signbit(a) != std::signbit(entry(i - 1, i))) { bad = 1; break; }
        lw = a / Ld[i - 1];
      }
      const double t = std::fma(-lw, lw, std::fma(-ls, ls, entry(i, i)));
      if (!(t > 0.0)) { bad = 1; break; }
      Ls[i] = ls;
      Lw[i] = lw;
      Ld[i] = std::sqrt(t);
    }
  }
  if (bad) return false;
  std::vector<float> Fa(4 * (size_t)n);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const float inv = 1.0f / (float)L[2 * (size_t)n + i];
    const float lwn = (i % nx) + 1 < nx ? (float)L[(size_t)n + i + 1] : 0.0f;
    Fa[i] = (float)L[i];
    Fa[(size_t)n + i] = inv;
    Fa[2 * (size_t)n + i] = -(float)L[(size_t)n + i] * inv;
    Fa[3 * (size_t)n + i] = -lwn * inv;
  }
  G.F.upload(Fa, P.stream);
  G.w.alloc((size_t)n);
  G.sink.alloc(1024);
  PARCEL_HIP_CHECK(hipStreamSynchronize(P.stream));
  G.nx = (int)nx;
  G.ny = ny;
  G.nstrip = nblk;
  G.on = true;
  return true;
}

}  // namespace

void precon_setup(Plan& P) {
  auto t0 = std::chrono::steady_clock::now();
  Precon& M = P.M;
  const HostCSR& A = P.hA;
  const int n = A.n;
  M.kind = P.prm.ipreflag;
  if (M.kind == PRECON_PJ) {
    std::vector<double> dinv(n, 0.0);
    for (int k = 0; k < n; ++k)
      for (int j = A.rowptr[k]; j < A.rowptr[k + 1]; ++j)
        if (A.col[j] == k) dinv[k] = (A.val[j] == 0.0) ? 1.0 / 1.0e-300 : 1.0 / A.val[j];
    M.dinv.alloc(n);
    M.dinv.upload(dinv, P.stream);
    PARCEL_HIP_CHECK(hipStreamSynchronize(P.stream));
  } else if (M.kind == PRECON_BJ) {
    int nblk = P.prm.precon_thblock;
    if (nblk < 1) {
      hipDeviceProp_t prop;
      PARCEL_HIP_CHECK(hipGetDeviceProperties(&prop, P.device));
      nblk = prop.multiProcessorCount;
    }
    if (nblk > n) nblk = n;
    if (nblk < 1) nblk = 1;
    if (!P.block_starts.empty()) nblk = (int)P.block_starts.size() - 1;
    M.nblocks = nblk;
    M.mode = (P.prm.ILU_method == PARCEL_ILU0) ? BJ_ILU0 : BJ_DILU;
    const int nvec = P.prm.independ_nvec < 1 ? 1 : P.prm.independ_nvec;
    std::vector<BlockSweep> F(nblk), B(nblk);
    std::vector<int> blo(nblk), bhi(nblk);
    block_ranges(P, nblk, n, blo, bhi);
    int bad = 0;
#pragma omp parallel for schedule(dynamic) reduction(| : bad)
    for (int b = 0; b < nblk; ++b) {
      bool ok = (M.mode == BJ_ILU0)
                    ? ilu0_block(A, blo[b], bhi[b], nvec, F[b], B[b])
                    : dilu_block(A, blo[b], bhi[b], nvec, P.prm.ILU_method, F[b], B[b]);
      if (!ok) { bad |= 1; continue; }
      build_levels(F[b], bhi[b] - blo[b], true);
      build_levels(B[b], bhi[b] - blo[b], false);
    }
    if (bad) throw hip_error("zero diagonal in a block-Jacobi block", PARCEL_HIP_ERR_ZERO_DIAG);
    // both sweeps padded to the entry count the apply kernel is built for
    int kmax = 1;
    for (int b = 0; b < nblk; ++b)
      for (const BlockSweep* S : {&F[b], &B[b]})
        for (size_t i = 0; i + 1 < S->ptr.size(); ++i) kmax = std::max(kmax, S->ptr[i + 1] - S->ptr[i]);
    int maxw = 1;
    for (int b = 0; b < nblk; ++b)
      for (const BlockSweep* S : {&F[b], &B[b]})
        for (size_t l = 0; l + 1 < S->lev_start.size(); ++l)
          maxw = std::max(maxw, S->lev_start[l + 1] - S->lev_start[l]);
    M.nth = maxw <= 256 ? 256 : maxw <= 512 ? 512 : 1024;
    pack_sweep(F, blo, M.fwd, kRingBudget, apply_k(kmax), M.nth, P.stream);
    pack_sweep(B, blo, M.bwd, kRingBudget, apply_k(kmax), M.nth, P.stream);
    lscan_setup(P, F, B, blo, bhi);
    if (P.prm.solver != PARCEL_SOLVER_CG_MIXED) line_setup(P, F, B, blo, bhi);
  } else if (M.kind != PRECON_NONE) {
    throw hip_error("preconditioner not supported on this path", PARCEL_HIP_ERR_UNSUPPORTED);
  }
  M.setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

namespace {

template <class T>
const T* dinv_of(const Precon& M);
template <>
const double* dinv_of<double>(const Precon& M) { return M.dinv.p; }
template <>
const float* dinv_of<float>(const Precon& M) { return M.dinv32.p; }

// restricted additive Schwarz: expanded row p takes r[src[p]]; after the
// block solve, s[dst[p]] = z[p] for the owned rows (dst >= 0)
__global__ void __launch_bounds__(kNT) k_ras_gather(int ne, const int* __restrict__ src, const double* __restrict__ r,
                                                    double* __restrict__ re, const int* skip) {
  if (skip && *skip) return;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < ne; p += (long)gridDim.x * blockDim.x) {
    const int q = src[p];
    re[p] = q >= 0 ? r[q] : 0.0;
  }
}
// additive output: s[k] = (0 + ze[oa[k]]) + ze[ob[k]] (second term where ob >= 0)
__global__ void __launch_bounds__(kNT) k_as_add(int n, const int* __restrict__ oa, const int* __restrict__ ob,
                                                const double* __restrict__ ze, double* __restrict__ s,
                                                const int* skip) {
  if (skip && *skip) return;
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
    double v = 0.0 + ze[oa[k]];
    const int b = ob[k];
    if (b >= 0) v = v + ze[b];
    s[k] = v;
  }
}
__global__ void __launch_bounds__(kNT) k_ras_scatter(int ne, const int* __restrict__ dst, const double* __restrict__ ze,
                                                     double* __restrict__ s, const int* skip) {
  if (skip && *skip) return;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < ne; p += (long)gridDim.x * blockDim.x) {
    const int d = dst[p];
    if (d >= 0) s[d] = ze[p];
  }
}

template <class T>
void precon_apply_t(Plan& P, const T* r, T* s, const CGFuseT<T>* fuse, const int* skip);

template <class T>
void ras_apply(Plan& P, const T* r, T* s, const int* skip) {
  if constexpr (!std::is_same<T, double>::value) {
    (void)P; (void)r; (void)s; (void)skip;
    throw hip_error("additive Schwarz runs in FP64 only", PARCEL_HIP_ERR_UNSUPPORTED);
  } else {
    const int ne = P.ras_n;
    const dim3 g((unsigned)std::min<long>(((long)ne + kNT - 1) / kNT, 4096)), b(kNT);
    hipLaunchKernelGGL(k_ras_gather, g, b, 0, P.stream, ne, (const int*)P.ras_src.p, r, P.ras_r.p, skip);
    precon_apply_t<double>(*P.ras_plan, P.ras_r.p, P.ras_s.p, nullptr, skip);
    if (P.ras_add) {
      const int n = P.A.n;
      const dim3 gn((unsigned)std::min<long>(((long)n + kNT - 1) / kNT, 4096));
      hipLaunchKernelGGL(k_as_add, gn, b, 0, P.stream, n, (const int*)P.ras_oa.p, (const int*)P.ras_ob.p,
                         (const double*)P.ras_s.p, s, skip);
    } else {
      hipLaunchKernelGGL(k_ras_scatter, g, b, 0, P.stream, ne, (const int*)P.ras_dst.p, (const double*)P.ras_s.p, s,
                         skip);
    }
  }
}

template <class T>
void precon_apply_t(Plan& P, const T* r, T* s, const CGFuseT<T>* fuse, const int* skip) {
  Precon& M = P.M;
  const int n = P.A.n;
  if (M.kind == PRECON_RAS) {
    if (fuse) throw hip_error("additive Schwarz has no fused residual update", PARCEL_HIP_ERR_ARG);
    ras_apply<T>(P, r, s, skip);
    return;
  }
  if (M.kind == PRECON_NONE || M.kind == PRECON_PJ) {
    if (fuse) throw hip_error("fused update needs the block-Jacobi preconditioner", PARCEL_HIP_ERR_ARG);
    int grid = tiled_nb(n);
    hipLaunchKernelGGL(k_scale_pj<T>, dim3(grid), dim3(kNT), 0, P.stream, n, r,
                       M.kind == PRECON_PJ ? dinv_of<T>(M) : nullptr, s, skip);
  } else if (std::is_same<T, float>::value && M.lscan.on) {
    if (fuse) throw hip_error("the line-scan apply has no fused residual update", PARCEL_HIP_ERR_ARG);
    if (skip) throw hip_error("the line-scan apply skips through the CG state", PARCEL_HIP_ERR_ARG);
    const LscanBJ& G = M.lscan;
    lscan_apply32(G.F.p, (const float*)r, G.w.p, (float*)s, G.nx, G.ny, G.nstrip, (long)n, nullptr, 0, G.sink.p,
                  P.stream);
  } else if (std::is_same<T, double>::value && M.line.on) {
    if constexpr (std::is_same<T, double>::value) line_apply(P, r, s, fuse, skip);
  } else {
    const int nb = M.nblocks;
    const int nth = M.nth;  // >= kNT: the fused gamma reduction uses kNT threads
    size_t shm = sizeof(double) * (size_t)(std::max(M.fwd.ring * M.fwd.ring_stride, M.bwd.ring * M.bwd.ring_stride) + 1);
    CGFuseT<T> f;
    if (fuse) f = *fuse;
    SweepArgs<T> F = sweep_args<T>(M.fwd), B = sweep_args<T>(M.bwd);
    // register-pipelined path when every level fits the workgroup and rows
    // are at most kKR long; the generic kernel otherwise
    int kk = std::max(M.fwd.max_k, M.bwd.max_k);
    if (kk < 1) kk = 1;
    const bool fast = kk <= kKR && nth <= 512;
    if (!fast) kk = 0;
#define PARCEL_BJ_K(NTH, KK)                                                                            \
  do {                                                                                                  \
    if (M.mode == BJ_ILU0)                                                                              \
      hipLaunchKernelGGL((k_bj_apply<BJ_ILU0, NTH, KK, T>), dim3(nb), dim3(NTH), shm, P.stream, F, B, r, s, f, skip); \
    else                                                                                                \
      hipLaunchKernelGGL((k_bj_apply<BJ_DILU, NTH, KK, T>), dim3(nb), dim3(NTH), shm, P.stream, F, B, r, s, f, skip); \
  } while (0)
#define PARCEL_BJ_LAUNCH(NTH)                                                     \
  do {                                                                            \
    switch (kk) {                                                                 \
      case 1: PARCEL_BJ_K(NTH, 1); break;                                         \
      case 2: PARCEL_BJ_K(NTH, 2); break;                                         \
      case 3: PARCEL_BJ_K(NTH, 3); break;                                         \
      case 4: PARCEL_BJ_K(NTH, 4); break;                                         \
      case 6: PARCEL_BJ_K(NTH, 6); break;                                         \
      case 8: PARCEL_BJ_K(NTH, 8); break;                                         \
      default: PARCEL_BJ_K(NTH, 0); break;                                        \
    }                                                                             \
  } while (0)
    if (nth <= 256) PARCEL_BJ_LAUNCH(256);
    else if (nth <= 512) PARCEL_BJ_LAUNCH(512);
    else PARCEL_BJ_K(1024, 0);
#undef PARCEL_BJ_LAUNCH
#undef PARCEL_BJ_K
  }
  PARCEL_HIP_CHECK(hipGetLastError());
}

__global__ void k_to_float(long n, const double* __restrict__ a, float* __restrict__ b) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

void to_float(const DevBuf<double>& a, DevBuf<float>& b, hipStream_t st) {
  if (!a.n) return;
  if (b.n != a.n) b.alloc(a.n);
  const long blocks = std::min<long>(((long)a.n + kNT - 1) / kNT, 4096);
  hipLaunchKernelGGL(k_to_float, dim3((unsigned)blocks), dim3(kNT), 0, st, (long)a.n, (const double*)a.p, b.p);
}

}  // namespace

// ---- double-double preconditioners ------------------------------------------
namespace {

struct DDSweep {
  const int* blk_lev;
  const int4* meta;
  const int* rows;
  const int* dep;
  const double* eh;
  const double* el;
  const double* dh;
  const double* dl;
};

DDSweep dd_sweep(const Sweep& S) {
  return DDSweep{S.blk_lev.p, S.meta.p, S.rows.p, S.dep.p, S.eval.p, S.eval_lo.p, S.diag.p, S.diag_lo.p};
}

// One workgroup per block; level by level, rows of a level in parallel.
// Every row accumulates its entries in the reference's order (ddsolve_LU_crs,
// src/ddmatrix.f90:6580-7132; ddsolve_ILU_crs, :3737-4230).  two == 0
// applies the factors' hi parts only (PRECISION_PRECONDITION = 1).
template <bool FWD, int MODE>
__device__ void dd_sweep_run(const DDSweep& A, const double* rh, const double* rl, double* sh, double* sl,
                             int two) {
  const int l0 = A.blk_lev[blockIdx.x], l1 = A.blk_lev[blockIdx.x + 1];
  for (int l = l0; l < l1; ++l) {
    const int4 mt = A.meta[l];
    const int m = mt.x, k = mt.y;
    for (int t = threadIdx.x; t < m; t += blockDim.x) {
      const int row = A.rows[mt.w + t];
      dd z;
      if (FWD) z = mkdd(rh[row], rl[row]);
      else z = (MODE == BJ_DILU) ? mkdd(0.0, 0.0) : mkdd(sh[row], sl[row]);
      for (int e = 0; e < k; ++e) {
        const int c = A.dep[mt.z + e * m + t];
        if (c == kPad) break;
        const dd a = mkdd(A.eh[mt.z + e * m + t], two ? A.el[mt.z + e * m + t] : 0.0);
        const dd prod = dd_mul(mkdd(sh[c], sl[c]), a);
        z = (!FWD && MODE == BJ_DILU) ? dd_add(z, prod) : dd_sub(z, prod);
      }
      const dd d = mkdd(A.dh[mt.w + t], two ? A.dl[mt.w + t] : 0.0);
      dd v;
      if (MODE == BJ_ILU0) v = dd_div(z, d);
      else if (FWD) v = dd_mul(z, d);
      else v = dd_sub(mkdd(sh[row], sl[row]), dd_mul(d, z));
      sh[row] = v.hi;
      sl[row] = v.lo;
    }
    __syncthreads();
  }
}

template <int MODE>
__global__ void __launch_bounds__(256) k_dd_bj(DDSweep F, DDSweep B, const double* rh, const double* rl, double* sh,
                                               double* sl, int two, const int* skip) {
  if (skip && *skip) return;
  dd_sweep_run<true, MODE>(F, rh, rl, sh, sl, two);
  dd_sweep_run<false, MODE>(B, nullptr, nullptr, sh, sl, two);
}

// Register-pipelined double-double sweep for sweeps whose dependencies all
// sit in the LDS ring (the last W levels, hi and lo rings side by side): the
// structure of level l+2 and the operand of level l+1 are loaded while level
// l computes, and a level reads its dependencies from LDS, so the level
// chain waits on no global round trip.  Every row computes exactly as in
// dd_sweep_run (entries in stored order up to the first padding entry, the
// same dd operations), so the results are the same bits.
template <int K>
struct DDLevel {
  int row, m;
  int dep[K];
  double vh[K], vl[K];
  double dh, dl;
  double ah, al;  // forward: r[row]; backward: the forward value s[row]
};

template <int K>
__device__ __forceinline__ void dd_stage_struct(const DDSweep& A, const int4* sm, int base, int l, int l1, int two,
                                                DDLevel<K>& R) {
  const int4 mt = sm[min(l, l1 - 1) - base];
  const int t = min((int)threadIdx.x, mt.x - 1);
  R.m = mt.x;
  R.row = A.rows[mt.w + t];
#pragma unroll
  for (int e = 0; e < K; ++e) {
    R.dep[e] = A.dep[mt.z + e * mt.x + t];
    R.vh[e] = A.eh[mt.z + e * mt.x + t];
    R.vl[e] = two ? A.el[mt.z + e * mt.x + t] : 0.0;
  }
  R.dh = A.dh[mt.w + t];
  R.dl = two ? A.dl[mt.w + t] : 0.0;
}

template <int K>
__device__ __forceinline__ void dd_stage_vec(DDLevel<K>& R, const double* h, const double* l) {
  R.ah = h[R.row];
  R.al = l[R.row];
}

template <bool FWD, int MODE, int K>
__device__ __forceinline__ void dd_compute(const DDLevel<K>& C, int l0, int L, int W, int pad, double* rgh,
                                           double* rgl, double* sh, double* sl) {
  dd z = (!FWD && MODE == BJ_DILU) ? mkdd(0.0, 0.0) : mkdd(C.ah, C.al);
#pragma unroll
  for (int e = 0; e < K; ++e) {
    const int c = C.dep[e];
    if (c == pad) break;
    const dd prod = dd_mul(mkdd(rgh[-c - 1], rgl[-c - 1]), mkdd(C.vh[e], C.vl[e]));
    z = (!FWD && MODE == BJ_DILU) ? dd_add(z, prod) : dd_sub(z, prod);
  }
  const dd d = mkdd(C.dh, C.dl);
  dd v;
  if (MODE == BJ_ILU0) v = dd_div(z, d);
  else if (FWD) v = dd_mul(z, d);
  else v = dd_sub(mkdd(C.ah, C.al), dd_mul(d, z));
  if ((int)threadIdx.x < C.m) {
    sh[C.row] = v.hi;
    sl[C.row] = v.lo;
  }
  const int slot = ((L - l0) % W) * (int)blockDim.x + (int)threadIdx.x;
  rgh[slot] = v.hi;
  rgl[slot] = v.lo;
}

__device__ __forceinline__ void dd_load_meta(const int4* meta, int4* sm, int c0, int l1) {
  __syncthreads();
  const int cnt = min(kMetaLds, l1 - c0);
  for (int j = threadIdx.x; j < cnt; j += blockDim.x) sm[j] = meta[c0 + j];
  __syncthreads();
}

template <bool FWD, int MODE, int K>
__device__ void dd_sweep_fast(const DDSweep& A, int W, const double* ah, const double* al, double* sh, double* sl,
                              int two, double* rgh, double* rgl, int4* sm) {
  const int l0 = A.blk_lev[blockIdx.x], l1 = A.blk_lev[blockIdx.x + 1];
  if (l1 <= l0) return;
  const int zero = W * (int)blockDim.x;  // the ring slot that holds 0 (padding)
  const int pad = -(1 + zero);
  int base = l0;
  if (threadIdx.x == 0) { rgh[zero] = 0.0; rgl[zero] = 0.0; }
  dd_load_meta(A.meta, sm, base, l1);
  DDLevel<K> R0, R1, R2;
  dd_stage_struct<K>(A, sm, base, l0, l1, two, R0);
  dd_stage_struct<K>(A, sm, base, l0 + 1, l1, two, R1);
  dd_stage_vec<K>(R0, ah, al);
  int l = l0;
  for (; l + 2 < l1; l += 3) {
    if (l + 4 >= base + kMetaLds) {
      base = l;
      dd_load_meta(A.meta, sm, base, l1);
    }
    dd_stage_struct<K>(A, sm, base, l + 2, l1, two, R2);
    dd_stage_vec<K>(R1, ah, al);
    dd_compute<FWD, MODE, K>(R0, l0, l, W, pad, rgh, rgl, sh, sl);
    lds_barrier();
    dd_stage_struct<K>(A, sm, base, l + 3, l1, two, R0);
    dd_stage_vec<K>(R2, ah, al);
    dd_compute<FWD, MODE, K>(R1, l0, l + 1, W, pad, rgh, rgl, sh, sl);
    lds_barrier();
    dd_stage_struct<K>(A, sm, base, l + 4, l1, two, R1);
    dd_stage_vec<K>(R0, ah, al);
    dd_compute<FWD, MODE, K>(R2, l0, l + 2, W, pad, rgh, rgl, sh, sl);
    lds_barrier();
  }
  if (l < l1) {
    if (l + 1 < l1) dd_stage_vec<K>(R1, ah, al);
    dd_compute<FWD, MODE, K>(R0, l0, l, W, pad, rgh, rgl, sh, sl);
    lds_barrier();
    if (l + 1 < l1) {
      dd_compute<FWD, MODE, K>(R1, l0, l + 1, W, pad, rgh, rgl, sh, sl);
      lds_barrier();
    }
  }
}

// One workgroup per block, both sweeps in ring mode; NTH >= the widest level.
template <int MODE, int NTH, int K>
__global__ void __launch_bounds__(NTH) k_dd_bj_fast(DDSweep F, DDSweep B, int WF, int WB, const double* rh,
                                                    const double* rl, double* sh, double* sl, int two,
                                                    const int* skip) {
  if (skip && *skip) return;
  extern __shared__ double dd_ring[];
  __shared__ int4 smeta[kMetaLds];
  const int W = max(WF, WB);
  double* rgh = dd_ring;
  double* rgl = dd_ring + W * NTH + 1;
  dd_sweep_fast<true, MODE, K>(F, WF, rh, rl, sh, sl, two, rgh, rgl, smeta);
  __syncthreads();  // forward values in memory before the backward sweep reads them
  dd_sweep_fast<false, MODE, K>(B, WB, sh, sl, sh, sl, two, rgh, rgl, smeta);
}

__global__ void k_dd_pj(int n, const double* __restrict__ rh, const double* __restrict__ rl,
                        const double* __restrict__ dh, const double* __restrict__ dl, double* __restrict__ sh,
                        double* __restrict__ sl, const int* skip) {
  if (skip && *skip) return;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    dd v;
    if (dh) v = dd_mul(mkdd(dh[i], dl ? dl[i] : 0.0), mkdd(rh[i], rl[i]));
    else v = mkdd(rh[i], rl[i]);
    sh[i] = v.hi;
    sl[i] = v.lo;
  }
}

}  // namespace

void precon_setup_dd(Plan& P, int prec) {
  auto t0 = std::chrono::steady_clock::now();
  Precon& M = P.M;
  const HostCSR& A = P.hA;
  const int n = A.n;
  M.kind = P.prm.ipreflag;
  M.dd_prec = prec;
  if (M.kind == PRECON_PJ) {
    // ddget_DInv_crs (src/ddmatrix.f90:290-330): Dinv = 1/a_kk in dd, 1/1e-100 for a zero diagonal
    std::vector<double> dh(n, 0.0), dl(n, 0.0);
    for (int k = 0; k < n; ++k)
      for (int j = A.rowptr[k]; j < A.rowptr[k + 1]; ++j)
        if (A.col[j] == k) {
          const dd a = mkdd(A.val[j], A.val_lo.empty() ? 0.0 : A.val_lo[j]);
          const dd v = (a.hi == 0.0 && a.lo == 0.0) ? dd_div(mkdd(1.0, 0.0), mkdd(1.0e-100, 0.0))
                                                    : dd_div(mkdd(1.0, 0.0), a);
          dh[k] = v.hi;
          dl[k] = v.lo;
        }
    M.dinv.alloc(n);
    M.dinv.upload(dh, P.stream);
    M.dinv_lo.alloc(n);
    M.dinv_lo.upload(dl, P.stream);
    PARCEL_HIP_CHECK(hipStreamSynchronize(P.stream));
  } else if (M.kind == PRECON_BJ) {
    int nblk = P.prm.precon_thblock;
    if (nblk < 1) {
      hipDeviceProp_t prop;
      PARCEL_HIP_CHECK(hipGetDeviceProperties(&prop, P.device));
      nblk = prop.multiProcessorCount;
    }
    nblk = std::max(1, std::min(nblk, n));
    if (!P.block_starts.empty()) nblk = (int)P.block_starts.size() - 1;
    M.nblocks = nblk;
    M.mode = (P.prm.ILU_method == PARCEL_ILU0) ? BJ_ILU0 : BJ_DILU;
    const int nvec = P.prm.independ_nvec < 1 ? 1 : P.prm.independ_nvec;
    std::vector<BlockSweep> F(nblk), B(nblk);
    std::vector<int> blo(nblk), bhi(nblk);
    block_ranges(P, nblk, n, blo, bhi);
    int bad = 0;
#pragma omp parallel for schedule(dynamic) reduction(| : bad)
    for (int b = 0; b < nblk; ++b) {
      bool ok = (M.mode == BJ_ILU0) ? ilu0_block(A, blo[b], bhi[b], nvec, F[b], B[b], true)
                                    : dilu_block(A, blo[b], bhi[b], nvec, P.prm.ILU_method, F[b], B[b], true);
      if (!ok) { bad |= 1; continue; }
      build_levels(F[b], bhi[b] - blo[b], true);
      build_levels(B[b], bhi[b] - blo[b], false);
    }
    if (bad) throw hip_error("zero diagonal in a block-Jacobi block", PARCEL_HIP_ERR_ZERO_DIAG);
    int kmax = 1;
    for (int b = 0; b < nblk; ++b)
      for (const BlockSweep* S : {&F[b], &B[b]})
        for (size_t i = 0; i + 1 < S->ptr.size(); ++i) kmax = std::max(kmax, S->ptr[i + 1] - S->ptr[i]);
    int maxw = 1;
    for (int b = 0; b < nblk; ++b)
      for (const BlockSweep* S : {&F[b], &B[b]})
        for (size_t l = 0; l + 1 < S->lev_start.size(); ++l)
          maxw = std::max(maxw, S->lev_start[l + 1] - S->lev_start[l]);
    // the pipelined kernel (k_dd_bj_fast) when every level fits 256 or 512
    // threads, rows have at most kKR entries and both sweeps keep their
    // dependencies in the LDS ring; otherwise global row ids for k_dd_bj
    const int kk = apply_k(kmax);
    M.nth = maxw <= 256 ? 256 : 512;
    bool fast = maxw <= 512 && kk <= kKR;
    if (fast) {
      pack_sweep(F, blo, M.fwd, kRingBudget, kk, M.nth, P.stream);
      pack_sweep(B, blo, M.bwd, kRingBudget, kk, M.nth, P.stream);
      fast = M.fwd.ring > 0 && M.bwd.ring > 0;
    }
    if (!fast) {
      M.nth = 256;
      pack_sweep(F, blo, M.fwd, 0, kmax, M.nth, P.stream);
      pack_sweep(B, blo, M.bwd, 0, kmax, M.nth, P.stream);
    }
  } else if (M.kind != PRECON_NONE) {
    throw hip_error("preconditioner not supported on the double-double path", PARCEL_HIP_ERR_UNSUPPORTED);
  }
  M.setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void precon_apply_dd(Plan& P, const double* rh, const double* rl, double* sh, double* sl, const int* skip) {
  Precon& M = P.M;
  const int n = P.A.n;
  const int two = M.dd_prec == 2 ? 1 : 0;
  if (M.kind == PRECON_NONE || M.kind == PRECON_PJ) {
    const bool pj = M.kind == PRECON_PJ;
    hipLaunchKernelGGL(k_dd_pj, dim3(tiled_nb(n)), dim3(kNT), 0, P.stream, n, rh, rl, pj ? (const double*)M.dinv.p : nullptr,
                       (pj && two) ? (const double*)M.dinv_lo.p : nullptr, sh, sl, skip);
  } else {
    const DDSweep F = dd_sweep(M.fwd), B = dd_sweep(M.bwd);
    if (M.fwd.ring > 0 && M.bwd.ring > 0) {
      const int W = std::max(M.fwd.ring, M.bwd.ring), nth = M.nth;
      const size_t shm = sizeof(double) * 2 * ((size_t)W * nth + 1);
      const int kk = std::max(M.fwd.max_k, M.bwd.max_k);
#define PARCEL_DD_K(NTH, KK)                                                                                     \
  do {                                                                                                           \
    if (M.mode == BJ_ILU0)                                                                                       \
      hipLaunchKernelGGL((k_dd_bj_fast<BJ_ILU0, NTH, KK>), dim3(M.nblocks), dim3(NTH), shm, P.stream, F, B,     \
                         M.fwd.ring, M.bwd.ring, rh, rl, sh, sl, two, skip);                                     \
    else                                                                                                         \
      hipLaunchKernelGGL((k_dd_bj_fast<BJ_DILU, NTH, KK>), dim3(M.nblocks), dim3(NTH), shm, P.stream, F, B,     \
                         M.fwd.ring, M.bwd.ring, rh, rl, sh, sl, two, skip);                                     \
  } while (0)
#define PARCEL_DD_NTH(NTH)                                                                                       \
  do {                                                                                                           \
    switch (kk) {                                                                                                \
      case 1: PARCEL_DD_K(NTH, 1); break;                                                                        \
      case 2: PARCEL_DD_K(NTH, 2); break;                                                                        \
      case 3: PARCEL_DD_K(NTH, 3); break;                                                                        \
      case 4: PARCEL_DD_K(NTH, 4); break;                                                                        \
      case 6: PARCEL_DD_K(NTH, 6); break;                                                                        \
      default: PARCEL_DD_K(NTH, 8); break;                                                                       \
    }                                                                                                            \
  } while (0)
      if (nth <= 256) PARCEL_DD_NTH(256);
      else PARCEL_DD_NTH(512);
#undef PARCEL_DD_NTH
#undef PARCEL_DD_K
    } else if (M.mode == BJ_ILU0) {
      hipLaunchKernelGGL(k_dd_bj<BJ_ILU0>, dim3(M.nblocks), dim3(256), 0, P.stream, F, B, rh, rl, sh, sl, two, skip);
    } else {
      hipLaunchKernelGGL(k_dd_bj<BJ_DILU>, dim3(M.nblocks), dim3(256), 0, P.stream, F, B, rh, rl, sh, sl, two, skip);
    }
  }
  PARCEL_HIP_CHECK(hipGetLastError());
}

void precon_apply(Plan& P, const double* r, double* s, const CGFuse* fuse, const int* skip) {
  precon_apply_t<double>(P, r, s, fuse, skip);
}

void precon_apply32(Plan& P, const float* r, float* s, const CGFuse32* fuse, const int* skip) {
  if (!P.M.has32) throw hip_error("FP32 preconditioner not set up", PARCEL_HIP_ERR_ARG);
  precon_apply_t<float>(P, r, s, fuse, skip);
}

void precon_setup32(Plan& P) {
  Precon& M = P.M;
  if (M.has32) return;
  // FP32 copies of the FP64 factors (rounded to nearest)
  to_float(M.dinv, M.dinv32, P.stream);
  to_float(M.fwd.eval, M.fwd.eval32, P.stream);
  to_float(M.fwd.diag, M.fwd.diag32, P.stream);
  to_float(M.bwd.eval, M.bwd.eval32, P.stream);
  to_float(M.bwd.diag, M.bwd.diag32, P.stream);
  PARCEL_HIP_CHECK(hipGetLastError());
  PARCEL_HIP_CHECK(hipStreamSynchronize(P.stream));
  M.has32 = true;
}

}  // namespace parcel

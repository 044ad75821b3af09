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
#include <chrono>
#include <cstring>

#include "internal.hpp"

namespace parcel {

namespace {

// divide_region (src/common.f90:125-182), 0-based half-open result.
void divide_region0(int rank, int nproc, int n, int* lo, int* hi) {
  if (n == 1) { *lo = 0; *hi = 1; return; }
  int iblock = n / nproc, over = n % nproc;
  int l, h;  // 1-based inclusive as in the reference
  if (rank < over) { l = 1 + iblock * rank + rank; h = l + iblock; }
  else { l = 1 + iblock * rank + over; h = l + iblock - 1; }
  *lo = l - 1; *hi = h;
}

// One row of a sweep: global row id, dependencies and values in
// accumulation order, and the row's scale.
struct RowEntries {
  std::vector<int> dep;
  std::vector<double> val;
};

struct BlockSweep {
  // rows of the block in level order, with per-level boundaries
  std::vector<int> lev_start;   // offsets into order (size nlev+1)
  std::vector<int> order;       // local rows (0-based within the block)
  std::vector<int> ptr;         // entry ranges per local row (size nrow+1)
  std::vector<int> dep;         // local dependency rows
  std::vector<double> val;
  std::vector<double> diag;     // per local row
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
  if (forward) {
    for (int i = 0; i < nrow; ++i) S.order[fill[lev[i]]++] = i;
  } else {
    for (int i = 0; i < nrow; ++i) S.order[fill[lev[i]]++] = i;
  }
}

// ILU(0) of one block, reference semantics (see file header).
// Returns false if a row of the block has no diagonal.
bool ilu0_block(const HostCSR& A, int lo, int hi, int nvec, BlockSweep& F, BlockSweep& B) {
  const int nb = hi - lo;
  std::vector<int> ptr(nb + 1), c2;
  std::vector<double> v2;
  c2.reserve((size_t)(A.rowptr[hi] - A.rowptr[lo]));
  v2.reserve(c2.capacity());
  for (int ii = lo; ii < hi; ii += nvec) {
    int iimax = std::min(ii + nvec, hi);  // chunk [ii, iimax)
    for (int i = ii; i < iimax; ++i) {
      ptr[i - lo] = (int)c2.size();
      for (int j = A.rowptr[i]; j < A.rowptr[i + 1]; ++j) {
        int c = A.col[j];
        if (c < lo || c >= hi) continue;
        if (c < ii || c >= iimax || c == i) { c2.push_back(c - lo); v2.push_back(A.val[j]); }
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
      v2[kc] = v2[kc] / v2[kd];
      for (int jc = ptr[i]; jc < ptr[i + 1]; ++jc) {
        int j = c2[jc];
        if (j < k + 1) continue;
        for (int q = kd; q < ptr[k + 1]; ++q)
          if (c2[q] == j) { v2[jc] = std::fma(-v2[kc], v2[q], v2[jc]); break; }
      }
    }
  }
  // forward: strictly-lower entries in stored order, unit diagonal
  F.ptr.assign(nb + 1, 0); F.dep.clear(); F.val.clear(); F.diag.assign(nb, 1.0);
  B.ptr.assign(nb + 1, 0); B.dep.clear(); B.val.clear(); B.diag.assign(nb, 1.0);
  for (int i = 0; i < nb; ++i) {
    F.ptr[i] = (int)F.dep.size();
    for (int q = ptr[i]; q < ptr[i + 1]; ++q)
      if (c2[q] < i) { F.dep.push_back(c2[q]); F.val.push_back(v2[q]); }
  }
  F.ptr[nb] = (int)F.dep.size();
  // backward: strictly-upper entries in stored order, divide by u_ii
  for (int i = 0; i < nb; ++i) {
    B.ptr[i] = (int)B.dep.size();
    for (int q = ptr[i]; q < ptr[i + 1]; ++q)
      if (c2[q] > i) { B.dep.push_back(c2[q]); B.val.push_back(v2[q]); }
    B.diag[i] = v2[dptr[i]];
  }
  B.ptr[nb] = (int)B.dep.size();
  return true;
}

// D-ILU of one block (DILU_DInv_crs + init_solve_ILU_crs), ilu_method 1
// (diagonal match, ILUD) or 2 (row-sum match, MLUD).
bool dilu_block(const HostCSR& A, int lo, int hi, int nvec, int method, BlockSweep& F,
                BlockSweep& B) {
  const int nb = hi - lo;
  const int* rp = A.rowptr.data(); const int* cl = A.col.data(); const double* a = A.val.data();
  for (int k = lo; k < hi; ++k) {
    bool found = false;
    for (int j = rp[k]; j < rp[k + 1]; ++j) {
      if (cl[j] < lo || cl[j] >= hi) continue;
      if (cl[j] == k) { found = true; break; }
    }
    if (!found) return false;
  }
  std::vector<double> D(nb);
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
      for (int jr = jstart; jr <= j - 1; ++jr) {
        int kr = cl[jr];
        if (ist <= kr && kr <= ien && kr != k) continue;
        for (int i = rp[kr]; i < rp[kr + 1]; ++i) {
          int rc = cl[i];
          if (rc < lo || rc >= hi) continue;
          if (method == PARCEL_ILUD) {
            if (rc == k) { d = d - a[jr] * a[i] / D[kr - lo]; break; }
          } else {
            if (kr < rc) d = d - a[jr] * a[i] / D[kr - lo];
          }
        }
      }
      D[k - lo] = d;
      break;
    }
  }
  F.ptr.assign(nb + 1, 0); F.dep.clear(); F.val.clear(); F.diag.assign(nb, 0.0);
  B.ptr.assign(nb + 1, 0); B.dep.clear(); B.val.clear(); B.diag.assign(nb, 0.0);
  for (int k = lo; k < hi; ++k) {
    double dinv = 1.0 / D[k - lo];
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
    for (int j = ls; j <= le; ++j) { F.dep.push_back(cl[j] - lo); F.val.push_back(a[j]); }
    // U range walked downward: last in-block entry .. first entry right of the chunk
    int us = rp[k], ue = rp[k + 1] - 1;
    for (int j = rp[k + 1] - 1; j >= rp[k]; --j) { if (hi - 1 < cl[j]) continue; us = j; break; }
    for (int j = rp[k + 1] - 1; j >= rp[k]; --j) {
      if (ien < cl[j]) continue;
      ue = j + 1; break;
    }
    B.ptr[k - lo] = (int)B.dep.size();
    for (int j = us; j >= ue; --j) { B.dep.push_back(cl[j] - lo); B.val.push_back(a[j]); }
  }
  F.ptr[nb] = (int)F.dep.size();
  B.ptr[nb] = (int)B.dep.size();
  return true;
}

// Pack per-block level-ordered sweeps into the device layout of `Sweep`.
void pack_sweep(const std::vector<BlockSweep>& blocks, const std::vector<int>& blo, Sweep& S,
                hipStream_t st) {
  const int nbk = (int)blocks.size();
  std::vector<int> blk_lev(nbk + 1, 0);
  for (int b = 0; b < nbk; ++b)
    blk_lev[b + 1] = blk_lev[b] + (int)blocks[b].lev_start.size() - (blocks[b].lev_start.empty() ? 0 : 1);
  const int nlev = blk_lev[nbk];
  std::vector<int> row_off(nlev + 1, 0), ent_off(nlev + 1, 0), kmax(nlev, 0);
  // sizes
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
      kmax[g] = k;
      row_off[g + 1] = m;
      ent_off[g + 1] = k * m;
    }
  }
  std::vector<long> ent_off_l(nlev + 1, 0);
  for (int g = 0; g < nlev; ++g) {
    row_off[g + 1] += row_off[g];
    ent_off_l[g + 1] = ent_off_l[g] + ent_off[g + 1];
  }
  const long nent = ent_off_l[nlev];
  if (nent > 0x7fffffffL) throw hip_error("block-Jacobi factor too large for int32 offsets", -1);
  for (int g = 0; g <= nlev; ++g) ent_off[g] = (int)ent_off_l[g];
  const int nrows = row_off[nlev];
  std::vector<int> rows(nrows), ecol((size_t)nent, -1);
  std::vector<double> eval((size_t)nent, 0.0), diag(nrows);
  long real = 0;
  int maxw = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : real) reduction(max : maxw)
  for (int b = 0; b < nbk; ++b) {
    const BlockSweep& B = blocks[b];
    int nl = blk_lev[b + 1] - blk_lev[b];
    for (int l = 0; l < nl; ++l) {
      int g = blk_lev[b] + l;
      int m = row_off[g + 1] - row_off[g];
      maxw = std::max(maxw, m);
      for (int t = 0; t < m; ++t) {
        int i = B.order[B.lev_start[l] + t];
        rows[row_off[g] + t] = blo[b] + i;
        diag[row_off[g] + t] = B.diag[i];
        int e = 0;
        for (int q = B.ptr[i]; q < B.ptr[i + 1]; ++q, ++e) {
          size_t at = (size_t)ent_off[g] + (size_t)e * m + t;
          ecol[at] = blo[b] + B.dep[q];
          eval[at] = B.val[q];
          real++;
        }
      }
    }
  }
  S.nlev = nlev;
  S.max_width = maxw;
  S.entries = nent;
  S.real_entries = real;
  S.blk_lev.upload(blk_lev, st);
  S.row_off.upload(row_off, st);
  S.ent_off.upload(ent_off, st);
  S.kmax.upload(kmax, st);
  S.rows.upload(rows, st);
  S.ecol.upload(ecol, st);
  S.eval.upload(eval, st);
  S.diag.upload(diag, st);
  PARCEL_HIP_CHECK(hipStreamSynchronize(st));
}

// --- kernels -----------------------------------------------------------------

__global__ void k_scale_pj(int n, const double* __restrict__ r, const double* __restrict__ dinv,
                           double* __restrict__ s, const int* skip) {
  if (skip && *skip) return;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    s[i] = dinv ? r[i] * dinv[i] : r[i];
}

// MODE 0: forward, unit diagonal (ILU(0) L):   s = r - sum s*l
// MODE 1: forward, D-ILU:                        s = dinv * (r - sum s*l)
// MODE 2: backward, ILU(0) U:                    s = (s - sum s*u) / u_ii
// MODE 3: backward, D-ILU (entries descending):  s = s - dinv * (sum s*u)
template <int MODE>
__global__ void __launch_bounds__(256) k_sweep(const int* __restrict__ blk_lev, const int* __restrict__ row_off,
                                               const int* __restrict__ ent_off, const int* __restrict__ kmax,
                                               const int* __restrict__ rows, const int* __restrict__ ecol,
                                               const double* __restrict__ eval, const double* __restrict__ diag,
                                               const double* __restrict__ rin, double* s, const int* skip) {
  if (skip && *skip) return;
  const int b = blockIdx.x;
  const int l0 = blk_lev[b], l1 = blk_lev[b + 1];
  for (int l = l0; l < l1; ++l) {
    const int ro = row_off[l], m = row_off[l + 1] - ro, eo = ent_off[l], k = kmax[l];
    for (int t = threadIdx.x; t < m; t += blockDim.x) {
      const int row = rows[ro + t];
      double z;
      if (MODE == 0 || MODE == 1) z = rin[row];
      else if (MODE == 2) z = s[row];
      else z = 0.0;
      for (int e = 0; e < k; ++e) {
        const int c = ecol[eo + e * m + t];
        if (c < 0) break;
        const double v = eval[eo + e * m + t];
        if (MODE == 3) z = fma(s[c], v, z);
        else z = fma(-s[c], v, z);
      }
      if (MODE == 0) s[row] = z;
      else if (MODE == 1) s[row] = diag[ro + t] * z;
      else if (MODE == 2) s[row] = z / diag[ro + t];
      else s[row] = fma(-diag[ro + t], z, s[row]);
    }
    __syncthreads();
  }
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
    M.nblocks = nblk;
    M.mode = (P.prm.ILU_method == PARCEL_ILU0) ? BJ_ILU0 : BJ_DILU;
    const int nvec = P.prm.independ_nvec < 1 ? 1 : P.prm.independ_nvec;
    std::vector<BlockSweep> F(nblk), B(nblk);
    std::vector<int> blo(nblk), bhi(nblk);
    for (int b = 0; b < nblk; ++b) divide_region0(b, nblk, n, &blo[b], &bhi[b]);
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
    pack_sweep(F, blo, M.fwd, P.stream);
    pack_sweep(B, blo, M.bwd, P.stream);
  } else if (M.kind != PRECON_NONE) {
    throw hip_error("preconditioner not supported on this path", PARCEL_HIP_ERR_UNSUPPORTED);
  }
  M.setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void precon_apply(Plan& P, const double* r, double* s, const int* skip) {
  Precon& M = P.M;
  const int n = P.A.n;
  if (M.kind == PRECON_NONE || M.kind == PRECON_PJ) {
    int grid = tiled_nb(n);
    hipLaunchKernelGGL(k_scale_pj, dim3(grid), dim3(kNT), 0, P.stream, n, r,
                       M.kind == PRECON_PJ ? (const double*)M.dinv.p : nullptr, s, skip);
  } else {
    const int nb = M.nblocks;
    if (M.mode == BJ_ILU0) {
      hipLaunchKernelGGL(k_sweep<0>, dim3(nb), dim3(256), 0, P.stream, M.fwd.blk_lev.p, M.fwd.row_off.p,
                         M.fwd.ent_off.p, M.fwd.kmax.p, M.fwd.rows.p, M.fwd.ecol.p, M.fwd.eval.p,
                         M.fwd.diag.p, r, s, skip);
      hipLaunchKernelGGL(k_sweep<2>, dim3(nb), dim3(256), 0, P.stream, M.bwd.blk_lev.p, M.bwd.row_off.p,
                         M.bwd.ent_off.p, M.bwd.kmax.p, M.bwd.rows.p, M.bwd.ecol.p, M.bwd.eval.p,
                         M.bwd.diag.p, (const double*)nullptr, s, skip);
    } else {
      hipLaunchKernelGGL(k_sweep<1>, dim3(nb), dim3(256), 0, P.stream, M.fwd.blk_lev.p, M.fwd.row_off.p,
                         M.fwd.ent_off.p, M.fwd.kmax.p, M.fwd.rows.p, M.fwd.ecol.p, M.fwd.eval.p,
                         M.fwd.diag.p, r, s, skip);
      hipLaunchKernelGGL(k_sweep<3>, dim3(nb), dim3(256), 0, P.stream, M.bwd.blk_lev.p, M.bwd.row_off.p,
                         M.bwd.ent_off.p, M.bwd.kmax.p, M.bwd.rows.p, M.bwd.ecol.p, M.bwd.eval.p,
                         M.bwd.diag.p, (const double*)nullptr, s, skip);
    }
  }
  PARCEL_HIP_CHECK(hipGetLastError());
}

}  // namespace parcel

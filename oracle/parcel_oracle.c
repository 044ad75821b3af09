/*
 * parcel_oracle.c -- CPU restatement of PARCEL's Krylov hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline); the product
 * library never links or calls it.  See parcel_oracle.h for conventions.
 *
 * Parity status of each function (how it is pinned, not just retyped):
 *   oracle_spmv            pinned: scipy.sparse product, closed-form stencils,
 *                          and the golden CG histories below.
 *   oracle_dot (both)      pinned: math.fsum / exact rational sums (tests).
 *   oracle_divide_region   pinned: brute-force partition invariants (tests).
 *   oracle_precon (PJ)     pinned: golden-free closed form (1/diag).
 *   oracle_precon (D-ILU)  pinned: src/example/template_cpu/STDOUT_run (291
 *                          residuals, 80 blocks) and the 1-D continued
 *                          fraction D_k = a - b*c/D_{k-1}.
 *   oracle_precon (ILU(0)) pinned: src/example/template_gpu/STDOUT_run (227
 *                          residuals, 4 blocks) and ILU(0) == exact LU on a
 *                          tridiagonal block (scipy.linalg.lu).
 *   oracle_pcg             pinned: both golden histories above.
 */
#include "parcel_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Reductions.                                                               */
/* ------------------------------------------------------------------------ */

/* Tree of one 64-lane wavefront: xor butterfly, offsets 32..1.  After every
 * step lane l holds v[l] + v[l^o]; addition is commutative, so lanes l and
 * l^o hold the same bits and lane 0's value is the wave total. */
static double wave_butterfly(double* v) {
  double w[64];
  for (int o = 32; o >= 1; o >>= 1) {
    for (int l = 0; l < 64; ++l) w[l] = v[l] + v[l ^ o];
    memcpy(v, w, sizeof(w));
  }
  return v[0];
}

/* Block of nt threads (nt a multiple of 64): butterfly per wave, then the
 * wave totals summed left to right. */
static double block_tree(double* acc, int nt) {
  int nw = nt / 64;
  double s = wave_butterfly(acc);
  for (int w = 1; w < nw; ++w) s = s + wave_butterfly(acc + 64 * w);
  return s;
}

#define ORACLE_NT_MAX 1024

double oracle_reduce_partials(const double* part, int nb) {
  /* One block of 256 threads: thread t adds part[t], part[t+256], ...
   * left to right starting from 0, then the block tree. */
  const int nt = 256;
  double acc[256];
  for (int t = 0; t < nt; ++t) {
    double a = 0.0;
    for (int j = t; j < nb; j += nt) a = a + part[j];
    acc[t] = a;
  }
  return block_tree(acc, nt);
}

/* Tiled order (the HIP kernels' order): the vector is cut into tiles of nt
 * consecutive entries; tile k belongs to block k mod nb with
 * nb = min(ceil(n/nt), nbmax); thread t of a block accumulates
 * fma(x[i], y[i], acc) over i = k*nt + t of its tiles in increasing k;
 * each block reduces its threads with block_tree(); the nb block partials
 * are reduced by oracle_reduce_partials().
 * Sequential order: fma accumulation over i = 0..n-1, as a one-thread run of
 * src/vec.f90:1008-1024 (calc_dot_local_FP64_CPU). */
double oracle_dot(const oracle_red* red, int n, const double* x, const double* y) {
  if (!red || red->mode == ORACLE_RED_SEQ) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s = fma(x[i], y[i], s);
    return s;
  }
  const int nt = red->nt;
  int ntiles = (n + nt - 1) / nt;
  int nb = ntiles < red->nbmax ? ntiles : red->nbmax;
  if (nb < 1) nb = 1;
  double* part = (double*)malloc(sizeof(double) * (size_t)nb);
#pragma omp parallel for schedule(static)
  for (int b = 0; b < nb; ++b) {
    double acc[ORACLE_NT_MAX];
    for (int t = 0; t < nt; ++t) acc[t] = 0.0;
    for (long k = b; k < ntiles; k += nb) {
      for (int t = 0; t < nt; ++t) {
        long i = k * nt + t;
        if (i < n) acc[t] = fma(x[i], y[i], acc[t]);
      }
    }
    part[b] = block_tree(acc, nt);
  }
  double s = oracle_reduce_partials(part, nb);
  free(part);
  return s;
}

/* ------------------------------------------------------------------------ */
/* divide_region (src/common.f90:125-182): contiguous split of [idown, iup]  */
/* into nproc parts; the first mod(n, nproc) parts get one extra row.        */
/* ------------------------------------------------------------------------ */
void oracle_divide_region(int rank, int nproc, int n, int* lo, int* hi) {
  int idown = 1, iup = n;
  if (iup == idown) { *lo = idown; *hi = iup; return; }
  int iregion = iup - idown + 1;
  int iblock = iregion / nproc;
  int over = iregion % nproc;
  if (rank < over) {
    *lo = idown + iblock * rank + rank;
    *hi = *lo + iblock;
  } else {
    *lo = idown + iblock * rank + over;
    *hi = *lo + iblock - 1;
  }
}

/* ------------------------------------------------------------------------ */
/* CRS SpMV, MatVec_crs_FP64_CPU (src/matrix.f90:9051-9110): row k sums      */
/* A(j)*x(col) over its stored entries in stored order, starting from 0.     */
/* ------------------------------------------------------------------------ */
void oracle_spmv(int n, const int* rowptr, const int* col, const double* val,
                 const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (int k = 0; k < n; ++k) {
    double tmp = 0.0;
    for (int j = rowptr[k] - 1; j < rowptr[k + 1] - 1; ++j)
      tmp = fma(val[j], x[col[j] - 1], tmp);
    y[k] = tmp;
  }
}

/* ------------------------------------------------------------------------ */
/* Preconditioners.                                                          */
/* ------------------------------------------------------------------------ */
#include "oracle_internal.h"

/* get_DInv_crs (src/matrix.f90:3104-3145): Dinv = 1/a_kk, 1/1e-300 for an
 * explicit zero diagonal. */
static void get_dinv(oracle_precon* p) {
  for (int k = 0; k < p->n; ++k) {
    for (int j = p->rowptr[k] - 1; j < p->rowptr[k + 1] - 1; ++j) {
      if (p->col[j] - 1 == k) {
        p->dinv[k] = (p->val[j] == 0.0) ? 1.0 / 1.0e-300 : 1.0 / p->val[j];
      }
    }
  }
}

/* DILU_DInv_crs (src/matrix.f90:3482-3650) for one block [lo, hi] (1-based
 * rows).  D(k) = a_kk - sum_{j<k} a_kj * a_jk / D(j) (ILUD) or the row-sum
 * variant (MLUD), over in-block entries, skipping couplings inside the
 * independ_nvec chunk of k.  Returns -1 if a row has no diagonal. */
static int dilu_block(oracle_precon* p, int lo, int hi) {
  const int* rp = p->rowptr; const int* cl = p->col; const double* a = p->val;
  int nvec = p->independ_nvec;
  int cnt = 0;
  for (int k = lo; k <= hi; ++k) {
    for (int j = rp[k - 1] - 1; j < rp[k] - 1; ++j) {
      if (cl[j] < lo || hi < cl[j]) continue;
      if (cl[j] == k) { cnt++; break; }
    }
  }
  if (cnt != hi - lo + 1) return -1;
  double* D = (double*)malloc(sizeof(double) * (size_t)(hi - lo + 1));
  for (int k = lo; k <= hi; ++k) {
    int ist = lo + ((k - lo) / nvec) * nvec;
    int ien = ist + nvec - 1 < hi ? ist + nvec - 1 : hi;
    int jstart = -1;
    for (int j = rp[k - 1] - 1; j < rp[k] - 1; ++j) {
      int c = cl[j];
      if (c < lo || hi < c) continue;
      if (jstart == -1) jstart = j;
      if (c != k) continue;
      double d = a[j];
      for (int jr = jstart; jr <= j - 1; ++jr) {
        int kr = cl[jr];
        if (ist <= kr && kr <= ien && kr != k) continue;
        for (int i = rp[kr - 1] - 1; i < rp[kr] - 1; ++i) {
          int rc = cl[i];
          if (rc < lo || hi < rc) continue;
          if (p->ilu_method == 1) {
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
  for (int k = lo; k <= hi; ++k) p->dinv[k - 1 - p->base] = 1.0 / D[k - lo];
  free(D);
  return 0;
}

/* init_solve_ILU_crs (src/matrix.f90:5170-5350): for row k, L = in-block
 * entries left of its independ_nvec chunk, U = in-block entries right of it.
 * lstart..lend ascending; ustart..uend is walked DOWNWARD (ustart is the last
 * in-block entry, uend the first entry right of the chunk). */
void oracle_dilu_ranges(oracle_precon* p, int lo, int hi) {
  const int* rp = p->rowptr; const int* cl = p->col;
  int nvec = p->independ_nvec;
  for (int k = lo; k <= hi; ++k) {
    int ist = lo + ((k - lo) / nvec) * nvec;
    int ls = rp[k] - 2, le = rp[k - 1] - 1; /* defaults: empty */
    for (int j = rp[k - 1] - 1; j < rp[k] - 1; ++j) {
      if (cl[j] < lo) continue;
      ls = j; break;
    }
    for (int j = rp[k - 1] - 1; j < rp[k] - 1; ++j) {
      if (cl[j] < lo) continue;
      if (ist <= cl[j]) { le = j - 1; break; }
    }
    p->lstart[k - 1 - p->base] = ls; p->lend[k - 1 - p->base] = le;
  }
  for (int k = hi; k >= lo; --k) {
    int ist = lo + ((k - lo) / nvec) * nvec;
    int ien = ist + nvec - 1 < hi ? ist + nvec - 1 : hi;
    int us = rp[k - 1] - 1, ue = rp[k] - 2; /* defaults: empty */
    for (int j = rp[k] - 2; j >= rp[k - 1] - 1; --j) {
      if (hi < cl[j]) continue;
      us = j; break;
    }
    for (int j = rp[k] - 2; j >= rp[k - 1] - 1; --j) {
      if (ien < cl[j]) continue;
      ue = j + 1; break;
    }
    p->ustart[k - 1 - p->base] = us; p->uend[k - 1 - p->base] = ue;
  }
}

/* solve_ILU_crs (src/matrix.f90:4987-5163), one block:
 *   forward  z = r(k) - sum_L s(c)*a ;  s(k) = Dinv(k)*z
 *   backward z = sum_U s(c)*a walked from ustart down to uend;
 *            s(k) = s(k) - Dinv(k)*z */
static void dilu_apply_block(const oracle_precon* p, int lo, int hi,
                             const double* r, double* s) {
  const int* cl = p->col; const double* a = p->val;
  const int o = 1 + p->base; /* row / column k lives at index k - o */
  for (int k = lo; k <= hi; ++k) {
    double zz = r[k - o];
    for (int j = p->lstart[k - o]; j <= p->lend[k - o]; ++j)
      zz = fma(-s[cl[j] - o], a[j], zz);
    s[k - o] = p->dinv[k - o] * zz;
  }
  for (int k = hi; k >= lo; --k) {
    double zz = 0.0;
    for (int j = p->ustart[k - o]; j >= p->uend[k - o]; --j)
      zz = fma(s[cl[j] - o], a[j], zz);
    s[k - o] = fma(-p->dinv[k - o], zz, s[k - o]);
  }
}

int oracle_dilu_block(oracle_precon* p, int lo, int hi) { return dilu_block(p, lo, hi); }
void oracle_dilu_apply_block(const oracle_precon* p, int lo, int hi, const double* r, double* s) {
  dilu_apply_block(p, lo, hi, r, s);
}

/* get_ILU_LU_crs (src/matrix.f90:8249-8392) + ILU_LU_crs (:7970-8072) +
 * divide_ILU_crs (:7821-7897) for one block b with rows [lo, hi]:
 * keep in-block entries (dropping couplings inside an independ_nvec chunk),
 * IKJ ILU(0) in place, split into unit-lower L (diagonal 1.0 stored last)
 * and upper U (diagonal stored first). */
static int ilu0_block(oracle_precon* p, int b, int lo, int hi) {
  const int* rp = p->rowptr; const int* cl = p->col; const double* a = p->val;
  int nvec = p->independ_nvec;
  int nb = hi - lo + 1;
  int* ptr = (int*)malloc(sizeof(int) * (size_t)(nb + 1));
  int cap = rp[hi] - rp[lo - 1];
  int* c2 = (int*)malloc(sizeof(int) * (size_t)(cap > 0 ? cap : 1));
  double* v2 = (double*)malloc(sizeof(double) * (size_t)(cap > 0 ? cap : 1));
  int nz = 0;
  for (int ii = lo; ii <= hi; ii += nvec) {
    int iimin = ii + nvec - 1 < hi ? ii + nvec - 1 : hi;
    for (int i = ii; i <= iimin; ++i) {
      ptr[i - lo] = nz;
      for (int j = rp[i - 1] - 1; j < rp[i] - 1; ++j) {
        int c = cl[j];
        if (lo <= c && c <= hi) {
          if (c < ii || iimin < c || c == i) {
            c2[nz] = c - lo; /* 0-based local */
            v2[nz] = a[j];
            nz++;
          }
        }
      }
    }
  }
  ptr[nb] = nz;
  /* diagonal positions */
  int* dptr = (int*)malloc(sizeof(int) * (size_t)nb);
  int kk = 0;
  for (int i = 0; i < nb; ++i) {
    for (int q = ptr[i]; q < ptr[i + 1]; ++q)
      if (c2[q] == i) { dptr[i] = q; kk++; break; }
  }
  if (kk != nb) { free(ptr); free(c2); free(v2); free(dptr); return -1; }
  /* IKJ: for i = 2..n (1-based), entries k < i in stored order */
  for (int i = 1; i < nb; ++i) {
    for (int kc = ptr[i]; kc < ptr[i + 1]; ++kc) {
      int k = c2[kc];
      if (k > i - 1) break;
      int kd = dptr[k];
      for (int q = ptr[k]; q < ptr[k + 1]; ++q)
        if (c2[q] == k) { kd = q; break; }
      v2[kc] = v2[kc] / v2[kd];
      for (int jc = ptr[i]; jc < ptr[i + 1]; ++jc) {
        int j = c2[jc];
        if (j < k + 1) continue;
        for (int q = kd; q < ptr[k + 1]; ++q) {
          if (c2[q] == j) { v2[jc] = fma(-v2[kc], v2[q], v2[jc]); break; }
        }
      }
    }
  }
  /* split */
  int nl = 0, nu = 0;
  for (int i = 0; i < nb; ++i)
    for (int q = ptr[i]; q < ptr[i + 1]; ++q) {
      if (c2[q] < i) nl++;
      if (c2[q] > i) nu++;
      if (c2[q] == i) { nl++; nu++; }
    }
  int* lp = (int*)malloc(sizeof(int) * (size_t)(nb + 1));
  int* up = (int*)malloc(sizeof(int) * (size_t)(nb + 1));
  int* lc = (int*)malloc(sizeof(int) * (size_t)(nl > 0 ? nl : 1));
  int* uc = (int*)malloc(sizeof(int) * (size_t)(nu > 0 ? nu : 1));
  double* lv = (double*)malloc(sizeof(double) * (size_t)(nl > 0 ? nl : 1));
  double* uv = (double*)malloc(sizeof(double) * (size_t)(nu > 0 ? nu : 1));
  nl = 0; nu = 0;
  for (int i = 0; i < nb; ++i) {
    lp[i] = nl; up[i] = nu;
    for (int q = ptr[i]; q < ptr[i + 1]; ++q) {
      int c = c2[q];
      if (c < i) { lv[nl] = v2[q]; lc[nl] = c; nl++; }
      if (i < c) { uv[nu] = v2[q]; uc[nu] = c; nu++; }
      if (c == i) {
        lv[nl] = 1.0; lc[nl] = c; nl++;
        uv[nu] = v2[q]; uc[nu] = c; nu++;
      }
    }
  }
  lp[nb] = nl; up[nb] = nu;
  p->lptr[b] = lp; p->lcol[b] = lc; p->lval[b] = lv;
  p->uptr[b] = up; p->ucol[b] = uc; p->uval[b] = uv;
  free(ptr); free(c2); free(v2); free(dptr);
  return 0;
}

/* solve_LU_crs (src/matrix.f90:8548-8746), one block:
 *   forward  z = r(i) - sum_{L except last} s(c)*l ;  s(i) = z / l_diag
 *   backward z = s(i) - sum_{U except first} s(c)*u ; s(i) = z / u_diag */
static void ilu0_apply_block(const oracle_precon* p, int b, int lo, int hi,
                             const double* r, double* s) {
  int nb = hi - lo + 1;
  const int* lp = p->lptr[b]; const int* lc = p->lcol[b]; const double* lv = p->lval[b];
  const int* up = p->uptr[b]; const int* uc = p->ucol[b]; const double* uv = p->uval[b];
  const double* rr = r + (lo - 1);
  double* ss = s + (lo - 1);
  for (int i = 0; i < nb; ++i) {
    double z = rr[i];
    for (int j = lp[i]; j < lp[i + 1] - 1; ++j) z = fma(-ss[lc[j]], lv[j], z);
    ss[i] = z / lv[lp[i + 1] - 1];
  }
  for (int i = nb - 1; i >= 0; --i) {
    double z = ss[i];
    for (int j = up[i] + 1; j < up[i + 1]; ++j) z = fma(-ss[uc[j]], uv[j], z);
    ss[i] = z / uv[up[i]];
  }
}

static oracle_precon* precon_create(int ipreflag, int ilu_method, int nblocks, const int* starts,
                                    int independ_nvec, int n, const int* rowptr,
                                    const int* col, const double* val, int* err);

oracle_precon* oracle_precon_create(int ipreflag, int ilu_method, int nblocks,
                                    int independ_nvec, int n, const int* rowptr,
                                    const int* col, const double* val, int* err) {
  return precon_create(ipreflag, ilu_method, nblocks, NULL, independ_nvec, n, rowptr, col, val, err);
}

/* Block Jacobi with explicit blocks: block b is 0-based rows
 * [starts[b], starts[b+1]) (the box blocks of the stencil entry point). */
oracle_precon* oracle_precon_create_blocks(int ipreflag, int ilu_method, int nblocks, const int* starts,
                                           int independ_nvec, int n, const int* rowptr,
                                           const int* col, const double* val, int* err) {
  return precon_create(ipreflag, ilu_method, nblocks, starts, independ_nvec, n, rowptr, col, val, err);
}

static oracle_precon* precon_create(int ipreflag, int ilu_method, int nblocks, const int* starts,
                                    int independ_nvec, int n, const int* rowptr,
                                    const int* col, const double* val, int* err) {
  oracle_precon* p = (oracle_precon*)calloc(1, sizeof(oracle_precon));
  p->ipreflag = ipreflag;
  p->ilu_method = ilu_method;
  p->independ_nvec = independ_nvec < 1 ? 1 : independ_nvec; /* type.f90:1377 */
  p->n = n;
  p->rowptr = rowptr; p->col = col; p->val = val;
  if (!starts) {
    if (nblocks < 1) nblocks = 1;
    if (nblocks > n) nblocks = n;
  }
  p->nblocks = nblocks;
  p->blo = (int*)malloc(sizeof(int) * (size_t)nblocks);
  p->bhi = (int*)malloc(sizeof(int) * (size_t)nblocks);
  for (int b = 0; b < nblocks; ++b) {
    if (starts) { p->blo[b] = starts[b] + 1; p->bhi[b] = starts[b + 1]; }
    else oracle_divide_region(b, nblocks, n, &p->blo[b], &p->bhi[b]);
  }
  *err = 0;
  if (ipreflag == 1) {
    p->dinv = (double*)calloc((size_t)n, sizeof(double));
    get_dinv(p);
  } else if (ipreflag == 2) {
    if (ilu_method == 0) {
      p->lptr = (int**)calloc((size_t)nblocks, sizeof(int*));
      p->lcol = (int**)calloc((size_t)nblocks, sizeof(int*));
      p->lval = (double**)calloc((size_t)nblocks, sizeof(double*));
      p->uptr = (int**)calloc((size_t)nblocks, sizeof(int*));
      p->ucol = (int**)calloc((size_t)nblocks, sizeof(int*));
      p->uval = (double**)calloc((size_t)nblocks, sizeof(double*));
      int bad = 0;
#pragma omp parallel for schedule(dynamic) reduction(|:bad)
      for (int b = 0; b < nblocks; ++b)
        if (ilu0_block(p, b, p->blo[b], p->bhi[b]) != 0) bad |= 1;
      if (bad) { *err = -2; oracle_precon_free(p); return NULL; }
    } else {
      p->dinv = (double*)calloc((size_t)n, sizeof(double));
      p->lstart = (int*)malloc(sizeof(int) * (size_t)n);
      p->lend = (int*)malloc(sizeof(int) * (size_t)n);
      p->ustart = (int*)malloc(sizeof(int) * (size_t)n);
      p->uend = (int*)malloc(sizeof(int) * (size_t)n);
      int bad = 0;
#pragma omp parallel for schedule(dynamic) reduction(|:bad)
      for (int b = 0; b < nblocks; ++b) {
        if (dilu_block(p, p->blo[b], p->bhi[b]) != 0) bad |= 1;
        else oracle_dilu_ranges(p, p->blo[b], p->bhi[b]);
      }
      if (bad) { *err = -2; oracle_precon_free(p); return NULL; }
    }
  } else if (ipreflag != 0) {
    *err = -1;
    oracle_precon_free(p);
    return NULL;
  }
  return p;
}

void oracle_precon_apply(const oracle_precon* p, const double* r, double* s) {
  int n = p->n;
  if (p->ipreflag == ORACLE_PRECON_AS) {
    oracle_as_apply(p, r, s);
    return;
  }
  if (p->ipreflag == ORACLE_PRECON_RAS) {
    for (int q = 0; q < p->ras_n; ++q) p->ras_re[q] = r[p->ras_src[q]];
    oracle_precon_apply(p->inner, p->ras_re, p->ras_ze);
    for (int q = 0; q < p->ras_n; ++q)
      if (p->ras_dst[q] >= 0) s[p->ras_dst[q]] = p->ras_ze[q];
    return;
  }
  if (p->ipreflag == ORACLE_PRECON_FACTOR) {
    if (p->apply32in64)
      oracle_lscan_apply32in64(p->nx, p->ny, p->nstrip, p->F32, r, s);
    else
      oracle_lscan_apply_strips(p->nx, p->ny, p->nstrip, p->F64, r, s);
    return;
  }
  if (p->ipreflag == 0) {
    memcpy(s, r, sizeof(double) * (size_t)n); /* xcpy, krylov_h.f90:1708 */
  } else if (p->ipreflag == 1) {
    /* vec_scaling (src/vec.f90:1050): s = r * Dinv */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) s[i] = r[i] * p->dinv[i];
  } else if (p->ilu_method == 0) {
#pragma omp parallel for schedule(dynamic)
    for (int b = 0; b < p->nblocks; ++b) ilu0_apply_block(p, b, p->blo[b], p->bhi[b], r, s);
  } else {
#pragma omp parallel for schedule(dynamic)
    for (int b = 0; b < p->nblocks; ++b) dilu_apply_block(p, p->blo[b], p->bhi[b], r, s);
  }
}

/* FP32 apply: the same sweeps in float, with the FP64 factors rounded to
 * float where they are read (the mixed-precision inner solve). */
static void dilu_apply_block32(const oracle_precon* p, int lo, int hi, const float* r, float* s) {
  const int* cl = p->col; const double* a = p->val;
  for (int k = lo; k <= hi; ++k) {
    float zz = r[k - 1];
    for (int j = p->lstart[k - 1]; j <= p->lend[k - 1]; ++j) zz = fmaf(-s[cl[j] - 1], (float)a[j], zz);
    s[k - 1] = (float)p->dinv[k - 1] * zz;
  }
  for (int k = hi; k >= lo; --k) {
    float zz = 0.0f;
    for (int j = p->ustart[k - 1]; j >= p->uend[k - 1]; --j) zz = fmaf(s[cl[j] - 1], (float)a[j], zz);
    s[k - 1] = fmaf(-(float)p->dinv[k - 1], zz, s[k - 1]);
  }
}

static void ilu0_apply_block32(const oracle_precon* p, int b, int lo, int hi, const float* r, float* s) {
  int nb = hi - lo + 1;
  const int* lp = p->lptr[b]; const int* lc = p->lcol[b]; const double* lv = p->lval[b];
  const int* up = p->uptr[b]; const int* uc = p->ucol[b]; const double* uv = p->uval[b];
  const float* rr = r + (lo - 1);
  float* ss = s + (lo - 1);
  for (int i = 0; i < nb; ++i) {
    float z = rr[i];
    for (int j = lp[i]; j < lp[i + 1] - 1; ++j) z = fmaf(-ss[lc[j]], (float)lv[j], z);
    ss[i] = z / (float)lv[lp[i + 1] - 1];
  }
  for (int i = nb - 1; i >= 0; --i) {
    float z = ss[i];
    for (int j = up[i] + 1; j < up[i + 1]; ++j) z = fmaf(-ss[uc[j]], (float)uv[j], z);
    ss[i] = z / (float)uv[up[i]];
  }
}

void oracle_precon_apply32(const oracle_precon* p, const float* r, float* s) {
  if (p->ipreflag == ORACLE_PRECON_FACTOR) {
    oracle_lscan_apply32_strips(p->nx, p->ny, p->nstrip, p->F32, r, s);
    return;
  }
  int n = p->n;
  if (p->ipreflag == 0) {
    memcpy(s, r, sizeof(float) * (size_t)n);
  } else if (p->ipreflag == 1) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) s[i] = r[i] * (float)p->dinv[i];
  } else if (p->ilu_method == 0) {
#pragma omp parallel for schedule(dynamic)
    for (int b = 0; b < p->nblocks; ++b) ilu0_apply_block32(p, b, p->blo[b], p->bhi[b], r, s);
  } else {
#pragma omp parallel for schedule(dynamic)
    for (int b = 0; b < p->nblocks; ++b) dilu_apply_block32(p, p->blo[b], p->bhi[b], r, s);
  }
}

int oracle_precon_nblocks(const oracle_precon* p) { return p->nblocks; }
const double* oracle_precon_dinv(const oracle_precon* p) { return p->dinv; }

/* Restricted additive Schwarz (solve_DILU_AS_DDM_GPU at one rank,
 * src/src_DDM/precon_ddm_GPU.f90:1573-1868): expanded row q takes r[src[q]];
 * the expanded block-diagonal matrix (one block per extended box, 1-based
 * CSR, block b = 0-based rows [starts[b], starts[b+1])) is D-ILU (ILUD)
 * factorised per block; s[dst[q]] = z[q] where dst[q] >= 0. */
oracle_precon* oracle_precon_create_ras(int n, int ne, const int* src, const int* dst, const int* rp_e,
                                        const int* col_e, const double* val_e, int nblocks, const int* starts,
                                        int* err) {
  oracle_precon* in = oracle_precon_create_blocks(2, 1, nblocks, starts, 1, ne, rp_e, col_e, val_e, err);
  if (!in) return NULL;
  oracle_precon* p = (oracle_precon*)calloc(1, sizeof(oracle_precon));
  p->ipreflag = ORACLE_PRECON_RAS;
  p->n = n;
  p->inner = in;
  p->ras_n = ne;
  p->ras_src = (int*)malloc(sizeof(int) * (size_t)ne);
  p->ras_dst = (int*)malloc(sizeof(int) * (size_t)ne);
  memcpy(p->ras_src, src, sizeof(int) * (size_t)ne);
  memcpy(p->ras_dst, dst, sizeof(int) * (size_t)ne);
  p->ras_re = (double*)malloc(sizeof(double) * (size_t)ne);
  p->ras_ze = (double*)malloc(sizeof(double) * (size_t)ne);
  return p;
}

void oracle_precon_free(oracle_precon* p) {
  if (!p) return;
  if (p->ipreflag == ORACLE_PRECON_AS) {
    oracle_as_free(p);
    return;
  }
  if (p->ipreflag == ORACLE_PRECON_RAS) {
    oracle_precon_free(p->inner);
    free(p->ras_src); free(p->ras_dst); free(p->ras_re); free(p->ras_ze);
    free(p);
    return;
  }
  free(p->blo); free(p->bhi); free(p->dinv);
  free(p->F64); free(p->F32);
  free(p->lstart); free(p->lend); free(p->ustart); free(p->uend);
  if (p->lptr) {
    for (int b = 0; b < p->nblocks; ++b) {
      free(p->lptr[b]); free(p->lcol[b]); free(p->lval[b]);
      free(p->uptr[b]); free(p->ucol[b]); free(p->uval[b]);
    }
    free(p->lptr); free(p->lcol); free(p->lval);
    free(p->uptr); free(p->ucol); free(p->uval);
  }
  free(p);
}

/* ------------------------------------------------------------------------ */
/* KSP_pCG (src/pcg.f90:224-892).                                            */
/* ------------------------------------------------------------------------ */
static double g_times[2];
void oracle_last_times(double* t) { t[0] = g_times[0]; t[1] = g_times[1]; }

static int pcg_core(const oracle_precon* M, int n, const int* rowptr, const int* col, const double* val,
                    const double* b, double* x, int itrmax, double rtol, const oracle_red* red,
                    double* hist, int* niter, int* iret, double t0);

int oracle_pcg(int n, const int* rowptr, const int* col, const double* val,
               const double* b, double* x, int ipreflag, int ilu_method,
               int nblocks, int independ_nvec, int itrmax, double rtol,
               const oracle_red* red, double* hist, int* niter, int* iret) {
  int err = 0;
  double t0 = omp_get_wtime();
  oracle_precon* M = oracle_precon_create(ipreflag, ilu_method, nblocks, independ_nvec,
                                          n, rowptr, col, val, &err);
  if (!M) return err;
  int rc = pcg_core(M, n, rowptr, col, val, b, x, itrmax, rtol, red, hist, niter, iret, t0);
  oracle_precon_free(M);
  return rc;
}

int oracle_pcg_with(const oracle_precon* M, int n, const int* rowptr, const int* col, const double* val,
                    const double* b, double* x, int itrmax, double rtol, const oracle_red* red,
                    double* hist, int* niter, int* iret) {
  return pcg_core(M, n, rowptr, col, val, b, x, itrmax, rtol, red, hist, niter, iret, omp_get_wtime());
}

static int pcg_core(const oracle_precon* M, int n, const int* rowptr, const int* col, const double* val,
                    const double* b, double* x, int itrmax, double rtol, const oracle_red* red,
                    double* hist, int* niter, int* iret, double t0) {
  double t1 = omp_get_wtime();
  size_t bytes = sizeof(double) * (size_t)(n > 0 ? n : 1);
  double* r = (double*)calloc(1, bytes);
  double* p = (double*)calloc(1, bytes);
  double* q = (double*)calloc(1, bytes);
  double* s = (double*)calloc(1, bytes);
  /* r = b - A x  (calc_res_CRS, krylov_h.f90:343: r = Ax; r = 1*b + (-1)*r) */
  oracle_spmv(n, rowptr, col, val, x, r);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) r[i] = b[i] - r[i];
  double bnorm = sqrt(oracle_dot(red, n, b, b));
  double rnorm = sqrt(oracle_dot(red, n, r, r));
  double alpha = 0.0, beta = 0.0;
  oracle_precon_apply(M, r, s);
  double rho0 = oracle_dot(red, n, s, r);
  int itr;
  for (itr = 1; itr <= itrmax; ++itr) {
    hist[itr - 1] = rnorm / bnorm;                 /* pcg.f90:512-518 */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {                  /* CG_axpy2, pcg.f90:8 */
      x[i] = fma(alpha, p[i], x[i]);
      p[i] = fma(beta, p[i], s[i]);
    }
    oracle_spmv(n, rowptr, col, val, p, q);        /* q = A p */
    double gamma = oracle_dot(red, n, p, q);       /* gamma = <p,q> */
    alpha = rho0 / gamma;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) r[i] = fma(-alpha, q[i], r[i]); /* r -= alpha q */
    oracle_precon_apply(M, r, s);                  /* s = M^-1 r */
    double rho1 = oracle_dot(red, n, s, r);        /* CG_calc_dot2 */
    double rr = oracle_dot(red, n, r, r);
    rnorm = sqrt(rr);
    if (rnorm / bnorm < rtol) break;               /* pcg.f90:657 */
    if (rnorm / bnorm > 1.0e8) break;              /* pcg.f90:663 */
    beta = rho1 / rho0;
    rho0 = rho1;
  }
  for (int i = 0; i < n; ++i) x[i] = fma(alpha, p[i], x[i]); /* pcg.f90:733 */
  *niter = itr < itrmax ? itr : itrmax;             /* pcg.f90:850 */
  g_times[0] = t1 - t0;
  g_times[1] = omp_get_wtime() - t1;
  *iret = (itrmax <= itr) ? 1 : 0;                  /* pcg.f90:851-855 */
  free(r); free(p); free(q); free(s);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* KSP_pBiCG (src/pbicg.f90:280-698).                                        */
/* ------------------------------------------------------------------------ */
int oracle_bicgstab(int n, const int* rowptr, const int* col, const double* val,
                    const double* b, double* x, int ipreflag, int ilu_method,
                    int nblocks, int independ_nvec, int itrmax, double rtol,
                    const oracle_red* red, double* hist, int* niter, int* iret) {
  int err = 0;
  oracle_precon* M = oracle_precon_create(ipreflag, ilu_method, nblocks, independ_nvec,
                                          n, rowptr, col, val, &err);
  if (!M) return err;
  size_t bytes = sizeof(double) * (size_t)(n > 0 ? n : 1);
  double* r = (double*)calloc(1, bytes);
  double* r0 = (double*)calloc(1, bytes);
  double* p = (double*)calloc(1, bytes);
  double* q = (double*)calloc(1, bytes);
  double* pp = (double*)calloc(1, bytes);
  double* e = (double*)calloc(1, bytes);
  double* pe = (double*)calloc(1, bytes);
  double* v = (double*)calloc(1, bytes);
  if (itrmax <= 0) { *niter = 0; *iret = 1; goto done; }
  /* r = b - A x, bnorm, rnorm (pbicg.f90:430-445) */
  oracle_spmv(n, rowptr, col, val, x, r);
  for (int i = 0; i < n; ++i) r[i] = b[i] - r[i];
  {
    double bnorm = sqrt(oracle_dot(red, n, b, b));
    double rnorm = sqrt(oracle_dot(red, n, r, r));
    memcpy(r0, r, bytes);                                 /* xcpy r -> r0 */
    memcpy(p, r, bytes);                                  /* xcpy r -> p */
    double c1 = oracle_dot(red, n, r0, r0), c2 = 0.0, c3 = 0.0, alpha = 0.0, beta = 0.0;
    int itr;
    for (itr = 1; itr <= itrmax; ++itr) {
      hist[itr - 1] = rnorm / bnorm;                      /* pbicg.f90:470 */
      oracle_precon_apply(M, p, pp);                      /* pp = M^-1 p */
      oracle_spmv(n, rowptr, col, val, pp, q);            /* q = A pp */
      c2 = oracle_dot(red, n, r0, q);
      alpha = c1 / c2;
      for (int i = 0; i < n; ++i) e[i] = fma(-alpha, q[i], r[i]);
      oracle_precon_apply(M, e, pe);                      /* pe = M^-1 e */
      oracle_spmv(n, rowptr, col, val, pe, v);            /* v = A pe */
      double ev = oracle_dot(red, n, e, v), vv = oracle_dot(red, n, v, v);
      c3 = ev / vv;
      for (int i = 0; i < n; ++i) x[i] = fma(c3, pe[i], fma(alpha, pp[i], x[i]));
      for (int i = 0; i < n; ++i) r[i] = fma(-c3, v[i], e[i]);   /* BiCG_axpy3_dot */
      c1 = oracle_dot(red, n, r, r0);
      rnorm = sqrt(oracle_dot(red, n, r, r));
      if (rnorm / bnorm < rtol) break;                    /* pbicg.f90:598 */
      beta = c1 / (c3 * c2);
      double bc = -(beta * c3);
      for (int i = 0; i < n; ++i) p[i] = fma(beta, p[i], fma(bc, q[i], r[i]));
    }
    *niter = itr < itrmax ? itr : itrmax;
    *iret = (itrmax <= itr) ? 1 : 0;
  }
done:
  free(r); free(r0); free(p); free(q); free(pp); free(e); free(pe); free(v);
  oracle_precon_free(M);
  return 0;
}

/* Givens rotation of Hessenberg column i (1-based) and of e
 * (GivensRotations, src/pcagmres.f90:190-242); plain arithmetic. */
static void givens(int m, double* H, int i, double* c, double* s, double* e) {
  double* hc = H + (size_t)(i - 1) * (m + 1);
  for (int k = 0; k < i - 1; ++k) {
    double t1 = hc[k] * c[k] - hc[k + 1] * s[k];
    double t2 = hc[k] * s[k] + hc[k + 1] * c[k];
    hc[k] = t1;
    hc[k + 1] = t2;
  }
  double t;
  if (hc[i - 1] == 0.0) t = hc[i] / 1.0 - 300; /* the reference's `/1.0d0-300` */
  else t = hc[i] / hc[i - 1];
  double ci = 1.0 / sqrt(1.0 + t * t);
  double si = -ci * t;
  c[i - 1] = ci;
  s[i - 1] = si;
  double e1 = ci * e[i - 1] - si * e[i];
  double e2 = si * e[i - 1] + ci * e[i];
  e[i - 1] = e1;
  e[i] = e2;
  hc[i - 1] = ci * hc[i - 1] - si * hc[i];
  hc[i] = 0.0;
}

int oracle_gmres(int n, const int* rowptr, const int* col, const double* val,
                 const double* b, double* x, int ipreflag, int ilu_method,
                 int nblocks, int independ_nvec, int itrmax, double rtol, int m,
                 int gsflag, const oracle_red* red, double* hist, int* niter, int* iret) {
  if (m < 1) return -1;
  int err = 0;
  oracle_precon* M = oracle_precon_create(ipreflag, ilu_method, nblocks, independ_nvec,
                                          n, rowptr, col, val, &err);
  if (!M) return err;
  const size_t nn = (size_t)(n > 0 ? n : 1);
  double* r = (double*)calloc(nn, sizeof(double));
  double* V = (double*)calloc(nn * (size_t)(m + 1), sizeof(double));
  double* pv = (double*)calloc(nn, sizeof(double));
  double* w = (double*)calloc(nn, sizeof(double));
  double* x0 = (double*)calloc(nn, sizeof(double));
  double* v0 = (double*)calloc(nn, sizeof(double));
  double* H = (double*)calloc((size_t)(m + 1) * m, sizeof(double));
  double* c = (double*)calloc((size_t)m + 1, sizeof(double));
  double* s = (double*)calloc((size_t)m + 1, sizeof(double));
  double* e = (double*)calloc((size_t)m + 1, sizeof(double));
  double* y = (double*)calloc((size_t)m + 1, sizeof(double));
  double* h = (double*)calloc((size_t)m + 1, sizeof(double));
  const int ld = m + 1;
  int itr = 1, inner = 1, mm = 0;                  /* pgmres.f90:248 inner = 1 */
  double bnorm = sqrt(oracle_dot(red, n, b, b));
  double rnorm = 0.0;
  for (itr = 1; itr <= itrmax; itr += m) {         /* do itr = 1,itrmax,m */
    memset(H, 0, sizeof(double) * (size_t)(m + 1) * m);
    memset(c, 0, sizeof(double) * (size_t)(m + 1));
    memset(s, 0, sizeof(double) * (size_t)(m + 1));
    memset(e, 0, sizeof(double) * (size_t)(m + 1));
    oracle_spmv(n, rowptr, col, val, x, r);        /* r = b - A x */
    for (int i = 0; i < n; ++i) r[i] = b[i] - r[i];
    rnorm = sqrt(oracle_dot(red, n, r, r));
    double a = -1.0 / rnorm;                        /* v1 = (-1/||r||) r */
    for (int i = 0; i < n; ++i) V[i] = a * r[i];
    e[0] = -rnorm;
    if (rnorm / bnorm < rtol) break;                /* pgmres.f90:293 (abstol = 0) */
    if (itr == 1) hist[0] = rnorm / bnorm;
    int stop = 0;
    for (inner = 1; inner <= m; ++inner) {
      const double* vj = V + (size_t)(inner - 1) * nn;
      oracle_precon_apply(M, vj, pv);               /* pv = M^-1 v_j */
      oracle_spmv(n, rowptr, col, val, pv, w);      /* w = A pv */
      double* hc = H + (size_t)(inner - 1) * ld;
      if (gsflag == 1) {                            /* MGS, pgmres.f90:336-360 */
        for (int k = 1; k <= inner; ++k) {
          const double* vk = V + (size_t)(k - 1) * nn;
          double hk = oracle_dot(red, n, w, vk);
          hc[k - 1] = hk;
          for (int i = 0; i < n; ++i) w[i] = fma(-hk, vk[i], w[i]);
        }
      } else {                                      /* CGS, pcagmres.f90:5-51 */
        for (int k = 1; k <= inner; ++k) h[k - 1] = oracle_dot(red, n, w, V + (size_t)(k - 1) * nn);
        for (int i = 0; i < n; ++i) {
          double wi = w[i];
          for (int k = 1; k <= inner; ++k) wi = fma(-h[k - 1], V[(size_t)(k - 1) * nn + i], wi);
          w[i] = wi;
        }
        for (int k = 1; k <= inner; ++k) hc[k - 1] = h[k - 1];
      }
      double hn = sqrt(oracle_dot(red, n, w, w));   /* h(j+1,j) = ||w|| */
      hc[inner] = hn;
      double sc = 1.0 / hn;
      double* vn = V + (size_t)inner * nn;
      for (int i = 0; i < n; ++i) vn[i] = sc * w[i];
      givens(m, H, inner, c, s, e);
      rnorm = fabs(e[inner]);
      mm = inner;
      hist[itr + inner - 1] = rnorm / bnorm;       /* reshistory(itr+inner) */
      if (rnorm / bnorm < rtol) { stop = 1; break; }
      if (itr + inner == itrmax) { stop = 1; break; }
    }
    (void)stop;                                     /* a completed DO leaves inner = m+1 */
    if (m < mm) mm = m;
    /* update_xVec: y = H^-1 e (back substitution), v0 = V y, x += M^-1 v0 */
    for (int i = mm; i >= 1; --i) {
      double yi = e[i - 1];
      for (int j = mm; j >= i + 1; --j) yi = yi - H[(i - 1) + (size_t)(j - 1) * ld] * y[j - 1];
      double d = H[(i - 1) + (size_t)(i - 1) * ld];
      y[i - 1] = (d == 0.0) ? yi / 1.0e-300 : yi / d;
    }
    for (int i = 0; i < n; ++i) {                   /* sum_vec_FP64_CPU, pcagmres.f90:54-90 */
      double acc = 0.0;
      for (int k = 0; k < mm; ++k) acc = fma(y[k], V[(size_t)k * nn + i], acc);
      v0[i] = acc;
    }
    oracle_precon_apply(M, v0, x0);
    for (int i = 0; i < n; ++i) x[i] = x0[i] + x[i];
    if (rnorm / bnorm < rtol) break;                /* pgmres.f90:457-462 */
    if (itr + inner == itrmax) break;
  }
  {
    int it = itr + inner;                           /* pgmres.f90:500 */
    *niter = it < itrmax ? it : itrmax;
    *iret = (itrmax <= itr) ? 1 : 0;
  }
  free(r); free(V); free(pv); free(w); free(x0); free(v0);
  free(H); free(c); free(s); free(e); free(y); free(h);
  oracle_precon_free(M);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Mixed-precision iterative refinement (no reference counterpart): FP64     */
/* outer loop r = b - A x, x += d with d from an FP32 PCG on A d = r.        */
/* ------------------------------------------------------------------------ */
static void spmv32(int n, const int* rowptr, const int* col, const double* val, const float* x, float* y) {
#pragma omp parallel for schedule(static)
  for (int k = 0; k < n; ++k) {
    float t = 0.0f;
    for (int j = rowptr[k] - 1; j < rowptr[k + 1] - 1; ++j) t = fmaf((float)val[j], x[col[j] - 1], t);
    y[k] = t;
  }
}

/* <x,y> of float vectors, products and sums in double, tiled order */
static double dot32(const oracle_red* red, int n, const float* x, const float* y, double* wx, double* wy) {
  for (int i = 0; i < n; ++i) { wx[i] = (double)x[i]; wy[i] = (double)y[i]; }
  return oracle_dot(red, n, wx, wy);
}

/* FP32 PCG on A d = b32 from d = 0 (the GPU's cg_run<float>); returns
 * the iterations it counts the way oracle_pcg does. */
static int pcg32(int n, const int* rowptr, const int* col, const double* val, const oracle_precon* M,
                 const float* b, float* x, int itrmax, double rtol, const oracle_red* red, float* r, float* p,
                 float* q, float* s, double* wx, double* wy) {
  for (int i = 0; i < n; ++i) { x[i] = 0.0f; p[i] = 0.0f; }
  spmv32(n, rowptr, col, val, x, r);
  for (int i = 0; i < n; ++i) r[i] = b[i] - r[i];
  const double bnorm = sqrt(dot32(red, n, b, b, wx, wy));
  double rnorm = sqrt(dot32(red, n, r, r, wx, wy));
  (void)rnorm;
  double alpha = 0.0, beta = 0.0;
  oracle_precon_apply32(M, r, s);
  double rho0 = dot32(red, n, s, r, wx, wy);
  int itr;
  for (itr = 1; itr <= itrmax; ++itr) {
    for (int i = 0; i < n; ++i) {
      x[i] = fmaf((float)alpha, p[i], x[i]);
      p[i] = fmaf((float)beta, p[i], s[i]);
    }
    spmv32(n, rowptr, col, val, p, q);
    const double gamma = dot32(red, n, p, q, wx, wy);
    alpha = rho0 / gamma;
    for (int i = 0; i < n; ++i) r[i] = fmaf(-(float)alpha, q[i], r[i]);
    oracle_precon_apply32(M, r, s);
    const double rho1 = dot32(red, n, s, r, wx, wy);
    rnorm = sqrt(dot32(red, n, r, r, wx, wy));
    if (rnorm / bnorm < rtol) break;
    if (rnorm / bnorm > 1.0e8) break;
    beta = rho1 / rho0;
    rho0 = rho1;
  }
  for (int i = 0; i < n; ++i) x[i] = fmaf((float)alpha, p[i], x[i]);
  return itr < itrmax ? itr : itrmax;
}

int oracle_pcg_mixed(int n, const int* rowptr, const int* col, const double* val, const double* b, double* x,
                     int ipreflag, int ilu_method, int nblocks, int independ_nvec, int itrmax, double rtol,
                     int inner_itrmax, double inner_rtol, const oracle_red* red, double* hist, int* niter,
                     int* iret, int* inner_total) {
  int err = 0;
  oracle_precon* M = oracle_precon_create(ipreflag, ilu_method, nblocks, independ_nvec, n, rowptr, col, val, &err);
  if (!M) return err;
  int rc = oracle_pcg_mixed_with(M, n, rowptr, col, val, b, x, itrmax, rtol, inner_itrmax, inner_rtol, red, hist,
                                 niter, iret, inner_total);
  oracle_precon_free(M);
  return rc;
}

int oracle_pcg_mixed_with(const oracle_precon* M, int n, const int* rowptr, const int* col, const double* val,
                          const double* b, double* x, int itrmax, double rtol, int inner_itrmax, double inner_rtol,
                          const oracle_red* red, double* hist, int* niter, int* iret, int* inner_total) {
  const size_t nn = (size_t)(n > 0 ? n : 1);
  double* r = (double*)calloc(nn, sizeof(double));
  double* wx = (double*)calloc(nn, sizeof(double));
  double* wy = (double*)calloc(nn, sizeof(double));
  float* r32 = (float*)calloc(nn, sizeof(float));
  float* d32 = (float*)calloc(nn, sizeof(float));
  float* ri = (float*)calloc(nn, sizeof(float));
  float* pi = (float*)calloc(nn, sizeof(float));
  float* qi = (float*)calloc(nn, sizeof(float));
  float* si = (float*)calloc(nn, sizeof(float));
  int total = 0, it = 0;
  oracle_spmv(n, rowptr, col, val, x, r);
  for (int i = 0; i < n; ++i) r[i] = b[i] - r[i];
  const double bnorm = sqrt(oracle_dot(red, n, b, b));
  double rnorm = sqrt(oracle_dot(red, n, r, r));
  for (it = 1; it <= itrmax; ++it) {
    hist[it - 1] = rnorm / bnorm;
    for (int i = 0; i < n; ++i) r32[i] = (float)r[i];
    total += pcg32(n, rowptr, col, val, M, r32, d32, inner_itrmax, inner_rtol, red, ri, pi, qi, si, wx, wy);
    for (int i = 0; i < n; ++i) x[i] = x[i] + (double)d32[i];
    oracle_spmv(n, rowptr, col, val, x, r);
    for (int i = 0; i < n; ++i) r[i] = b[i] - r[i];
    rnorm = sqrt(oracle_dot(red, n, r, r));
    if (rnorm / bnorm < rtol) break;
  }
  *niter = it < itrmax ? it : itrmax;
  *iret = (itrmax <= it) ? 1 : 0;
  *inner_total = total;
  free(r); free(wx); free(wy); free(r32); free(d32); free(ri); free(pi); free(qi); free(si);
  return 0;
}

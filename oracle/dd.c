/*
 * dd.c -- CPU oracle for the double-double (dd) path.  TEST INFRASTRUCTURE
 * ONLY (see parcel_oracle.h).
 *
 * Reference: dd_add / dd_sub / dd_mul / dd_div / dd_sqrt (src/dd_func.f90:5-250,
 * inline forms src/ddadd.h, ddsub.h, ddmul.h, dddiv.h): Knuth TwoSum and a
 * Dekker split with 2^27+1, written without fused multiply-add; this file is
 * compiled with -ffp-contract=off so every operation rounds as written.
 * ddKSP_pCG (src/ddpcg.f90:247-563), ddCG_axpy2 / ddCG_calc_dot2
 * (src/ddpcg.f90:9-246), ddaxpy / ddcalc_dot (src/ddvec.f90:14-240),
 * ddMatVec_crs (src/ddmatrix.f90:1785-2358), preconditioners:
 *   point Jacobi     ddget_DInv_crs (ddmatrix.f90:290-330), ddpointJacobi (:3520-3565)
 *   block D-ILU      ddDILU_DInv_crs (:1451-1662), ddsolve_ILU_crs (:3737-4230)
 *   block ILU(0)     ddILU_LU_crs (:5686-5830), ddsolve_LU_crs (:6580-7132)
 * The factorisations are always computed in dd (ddkrylov.f90:365-395); with
 * PRECISION_PRECONDITION = 1 the apply uses only their hi parts, as the
 * reference's KSP_DOUBLE_PRECISION branches pass 0.0d0 for every lo part.
 *
 * Reduction order of the dd dot products: the tiled order of oracle_dot
 * with every `+` a dd_add(own, other) and every product a dd_mul(x_i, y_i).
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_internal.h"

typedef struct {
  double hi, lo;
} dd_t;

static inline dd_t mk(double h, double l) {
  dd_t r;
  r.hi = h;
  r.lo = l;
  return r;
}

static inline dd_t dd_add(dd_t a, dd_t b) {
  const double p1 = a.hi + b.hi;
  const double p2 = p1 - a.hi;
  double t1l = (a.hi - (p1 - p2)) + (b.hi - p2);
  const double t1h = p1;
  t1l = t1l + a.lo + b.lo;
  const double s1 = t1h + t1l;
  t1l = t1l - (s1 - t1h);
  return mk(s1, t1l);
}

static inline dd_t dd_sub(dd_t a, dd_t b) {
  const double p1 = a.hi - b.hi;
  const double p2 = p1 - a.hi;
  double t1l = (a.hi - (p1 - p2)) - (b.hi + p2);
  const double t1h = p1;
  t1l = t1l + a.lo - b.lo;
  const double s1 = t1h + t1l;
  t1l = t1l - (s1 - t1h);
  return mk(s1, t1l);
}

static inline dd_t dd_mul(dd_t a, dd_t b) {
  const double t1h = a.hi * b.hi;
  double u1 = 134217729.0 * a.hi;
  const double p1 = u1 - (u1 - a.hi);
  const double p2 = a.hi - p1;
  u1 = 134217729.0 * b.hi;
  const double r1 = u1 - (u1 - b.hi);
  const double r2 = b.hi - r1;
  double t1l = ((p1 * r1 - t1h) + p1 * r2 + p2 * r1) + p2 * r2;
  t1l = t1l + (a.hi * b.lo + a.lo * b.hi);
  const double s1 = t1h + t1l;
  t1l = t1l - (s1 - t1h);
  return mk(s1, t1l);
}

static inline dd_t dd_div(dd_t a, dd_t b) {
  const double q1 = a.hi / b.hi;
  double t1h = b.hi * q1;
  double u1 = 134217729.0 * b.hi;
  const double p1 = u1 - (u1 - b.hi);
  const double p2 = b.hi - p1;
  u1 = 134217729.0 * q1;
  double r1 = u1 - (u1 - q1);
  const double r2 = q1 - r1;
  double t1l = ((p1 * r1 - t1h) + p1 * r2 + p2 * r1) + p2 * r2;
  t1l = t1l + (b.lo * q1);
  double s1 = t1h + t1l;
  t1l = t1l - (s1 - t1h);
  t1h = s1;
  s1 = a.hi - t1h;
  r1 = s1 - a.hi;
  double s2 = (a.hi - (s1 - r1)) - (t1h + r1);
  s2 = s2 - t1l;
  s2 = s2 + a.lo;
  const double q2 = (s1 + s2) / b.hi;
  const double h = q1 + q2;
  return mk(h, q2 - (h - q1));
}

/* dd_sqrt (dd_func.f90:146-250).  A zero argument is not returned early in
 * the reference: the computation continues with 1/sqrt(0) and yields NaN. */
static inline dd_t dd_sqrt(dd_t a) {
  if (a.hi < 0.0 || (a.hi == 0.0 && a.lo < 0.0)) return mk(0.0, 0.0);
  const double x = 1.0 / sqrt(a.hi);
  const double ax = a.hi * x;
  const dd_t y = dd_mul(mk(x, 0.0), mk(0.5, 0.0));
  const dd_t z = dd_mul(mk(ax, 0.0), mk(ax, 0.0));
  const dd_t t2 = dd_sub(a, z);
  const dd_t t1 = dd_mul(t2, y);
  return dd_add(t1, mk(ax, 0.0));
}

void oracle_dd_op(int op, int n, const double* ah, const double* al, const double* bh, const double* bl,
                  double* ch, double* cl) {
  for (int i = 0; i < n; ++i) {
    const dd_t a = mk(ah[i], al[i]);
    const dd_t b = (op == 4) ? mk(0.0, 0.0) : mk(bh[i], bl[i]);
    dd_t c;
    switch (op) {
      case 0: c = dd_add(a, b); break;
      case 1: c = dd_sub(a, b); break;
      case 2: c = dd_mul(a, b); break;
      case 3: c = dd_div(a, b); break;
      default: c = dd_sqrt(a); break;
    }
    ch[i] = c.hi;
    cl[i] = c.lo;
  }
}

/* ---- reductions ------------------------------------------------------- */

static dd_t dd_wave(dd_t* v) {
  dd_t w[64];
  for (int o = 32; o >= 1; o >>= 1) {
    for (int l = 0; l < 64; ++l) w[l] = dd_add(v[l], v[l ^ o]);
    memcpy(v, w, sizeof(w));
  }
  return v[0];
}

static dd_t dd_block(dd_t* acc, int nt) {
  dd_t s = dd_wave(acc);
  for (int w = 1; w < nt / 64; ++w) s = dd_add(s, dd_wave(acc + 64 * w));
  return s;
}

static dd_t dd_reduce_partials(const dd_t* part, int nb) {
  dd_t acc[256];
  for (int t = 0; t < 256; ++t) {
    dd_t a = mk(0.0, 0.0);
    for (int j = t; j < nb; j += 256) a = dd_add(a, part[j]);
    acc[t] = a;
  }
  return dd_block(acc, 256);
}

static dd_t dd_dot(const oracle_red* red, int n, const double* xh, const double* xl, const double* yh,
                   const double* yl) {
  if (!red || red->mode == ORACLE_RED_SEQ) {
    dd_t s = mk(0.0, 0.0);
    for (int i = 0; i < n; ++i) s = dd_add(s, dd_mul(mk(xh[i], xl[i]), mk(yh[i], yl[i])));
    return s;
  }
  const int nt = red->nt;
  const int ntiles = (n + nt - 1) / nt;
  int nb = ntiles < red->nbmax ? ntiles : red->nbmax;
  if (nb < 1) nb = 1;
  dd_t* part = (dd_t*)malloc(sizeof(dd_t) * (size_t)nb);
#pragma omp parallel for schedule(static)
  for (int b = 0; b < nb; ++b) {
    dd_t acc[1024];
    for (int t = 0; t < nt; ++t) acc[t] = mk(0.0, 0.0);
    for (long k = b; k < ntiles; k += nb)
      for (int t = 0; t < nt; ++t) {
        const long i = k * nt + t;
        if (i < n) acc[t] = dd_add(acc[t], dd_mul(mk(xh[i], xl[i]), mk(yh[i], yl[i])));
      }
    part[b] = dd_block(acc, nt);
  }
  const dd_t s = dd_reduce_partials(part, nb);
  free(part);
  return s;
}

void oracle_dd_dot(const oracle_red* red, int n, const double* xh, const double* xl, const double* yh,
                   const double* yl, double* rh, double* rl) {
  const dd_t s = dd_dot(red, n, xh, xl, yh, yl);
  *rh = s.hi;
  *rl = s.lo;
}

/* ---- SpMV --------------------------------------------------------------- */

/* y = A x; row k: t = 0; t = dd_add(dd_mul(A_j, x_c), t) in stored order.
 * val_lo may be NULL (PRECISION_A = 1: A has no lo part). */
void oracle_dd_spmv(int n, const int* rowptr, const int* col, const double* vh, const double* vl,
                    const double* xh, const double* xl, double* yh, double* yl) {
#pragma omp parallel for schedule(static)
  for (int k = 0; k < n; ++k) {
    dd_t t = mk(0.0, 0.0);
    for (int j = rowptr[k] - 1; j < rowptr[k + 1] - 1; ++j) {
      const int c = col[j] - 1;
      t = dd_add(dd_mul(mk(vh[j], vl ? vl[j] : 0.0), mk(xh[c], xl[c])), t);
    }
    yh[k] = t.hi;
    yl[k] = t.lo;
  }
}

/* ---- preconditioners ------------------------------------------------------ */

struct oracle_ddprecon {
  oracle_precon* base;  /* structure: blocks, D-ILU ranges, ILU(0) patterns */
  int prec;             /* PRECISION_PRECONDITION: 1 double, 2 double-double */
  const double* vh;     /* matrix values (D-ILU applies with A's entries) */
  const double* vl;
  double *dh, *dl;      /* Dinv (point Jacobi, D-ILU) */
  double **lh, **ll, **uh, **ul; /* ILU(0) factor values per block */
};

/* ddDILU_DInv_crs for rows lo..hi (1-based) */
static void dd_dilu_block(struct oracle_ddprecon* P, int lo, int hi) {
  const oracle_precon* p = P->base;
  const int* rp = p->rowptr; const int* cl = p->col;
  const int nvec = p->independ_nvec;
  dd_t* D = (dd_t*)malloc(sizeof(dd_t) * (size_t)(hi - lo + 1));
  for (int k = lo; k <= hi; ++k) {
    const int ist = lo + ((k - lo) / nvec) * nvec;
    const int ien = ist + nvec - 1 < hi ? ist + nvec - 1 : hi;
    dd_t d = mk(0.0, 0.0);
    int jstart = -1;
    for (int j = rp[k - 1] - 1; j < rp[k] - 1; ++j) {
      const int c = cl[j];
      if (c < lo || hi < c) continue;
      if (jstart == -1) jstart = j;
      if (k <= c) {
        d = (k == c) ? mk(P->vh[j], P->vl ? P->vl[j] : 0.0) : mk(0.0, 0.0);
        for (int jr = jstart; jr <= j - 1; ++jr) {
          const int kr = cl[jr];
          if (ist <= kr && kr <= ien && kr != k) continue;
          for (int i = rp[kr - 1] - 1; i < rp[kr] - 1; ++i) {
            const int rc = cl[i];
            if (rc < lo) continue;
            if (hi < rc) break;
            const int upd = (p->ilu_method == 1) ? (k == rc) : (kr < rc);
            if (!upd) continue;
            dd_t c2 = dd_mul(mk(P->vh[jr], P->vl ? P->vl[jr] : 0.0), mk(P->vh[i], P->vl ? P->vl[i] : 0.0));
            const dd_t dk = D[kr - lo];
            c2 = (dk.hi == 0.0 && dk.lo == 0.0) ? dd_div(c2, mk(1.0e-100, 0.0)) : dd_div(c2, dk);
            d = dd_sub(d, c2);
            if (p->ilu_method == 1) break;
          }
        }
        break;
      }
    }
    D[k - lo] = d;
  }
  for (int k = lo; k <= hi; ++k) {
    const dd_t dk = D[k - lo];
    const dd_t v = (dk.hi == 0.0 && dk.lo == 0.0) ? dd_div(mk(1.0, 0.0), mk(1.0e-100, 0.0))
                                                   : dd_div(mk(1.0, 0.0), dk);
    P->dh[k - 1] = v.hi;
    P->dl[k - 1] = v.lo;
  }
  free(D);
}

/* ddILU_LU_crs for block b (rows lo..hi): the in-block pattern of the FP64
 * factorisation (same dropping rules), IKJ in dd, split into unit-lower L
 * (diagonal 1 stored last) and upper U (diagonal first). */
static int dd_ilu0_block(struct oracle_ddprecon* P, int b, int lo, int hi) {
  const oracle_precon* p = P->base;
  const int* rp = p->rowptr; const int* cl = p->col;
  const int nvec = p->independ_nvec;
  const int nb = hi - lo + 1;
  int* ptr = (int*)malloc(sizeof(int) * (size_t)(nb + 1));
  const int cap = rp[hi] - rp[lo - 1];
  int* c2 = (int*)malloc(sizeof(int) * (size_t)(cap > 0 ? cap : 1));
  dd_t* v2 = (dd_t*)malloc(sizeof(dd_t) * (size_t)(cap > 0 ? cap : 1));
  int nz = 0;
  for (int ii = lo; ii <= hi; ii += nvec) {
    const int iimin = ii + nvec - 1 < hi ? ii + nvec - 1 : hi;
    for (int i = ii; i <= iimin; ++i) {
      ptr[i - lo] = nz;
      for (int j = rp[i - 1] - 1; j < rp[i] - 1; ++j) {
        const int c = cl[j];
        if (lo <= c && c <= hi && (c < ii || iimin < c || c == i)) {
          c2[nz] = c - lo;
          v2[nz] = mk(P->vh[j], P->vl ? P->vl[j] : 0.0);
          nz++;
        }
      }
    }
  }
  ptr[nb] = nz;
  for (int i = 1; i < nb; ++i) {
    for (int kc = ptr[i]; kc < ptr[i + 1]; ++kc) {
      const int k = c2[kc];
      if (k > i - 1) break;
      int kd = -1;
      for (int q = ptr[k]; q < ptr[k + 1]; ++q)
        if (c2[q] == k) { kd = q; break; }
      if (kd < 0) { free(ptr); free(c2); free(v2); return -1; }
      v2[kc] = dd_div(v2[kc], v2[kd]);
      for (int jc = ptr[i]; jc < ptr[i + 1]; ++jc) {
        const int j = c2[jc];
        if (j < k + 1) continue;
        for (int q = ptr[k]; q < ptr[k + 1]; ++q)
          if (c2[q] == j) { v2[jc] = dd_sub(v2[jc], dd_mul(v2[kc], v2[q])); break; }
      }
    }
  }
  /* values in the split layout of the FP64 factorisation (same pattern) */
  const int* lp = p->lptr[b]; const int* up = p->uptr[b];
  P->lh[b] = (double*)malloc(sizeof(double) * (size_t)(lp[nb] > 0 ? lp[nb] : 1));
  P->ll[b] = (double*)malloc(sizeof(double) * (size_t)(lp[nb] > 0 ? lp[nb] : 1));
  P->uh[b] = (double*)malloc(sizeof(double) * (size_t)(up[nb] > 0 ? up[nb] : 1));
  P->ul[b] = (double*)malloc(sizeof(double) * (size_t)(up[nb] > 0 ? up[nb] : 1));
  int nl = 0, nu = 0;
  for (int i = 0; i < nb; ++i)
    for (int q = ptr[i]; q < ptr[i + 1]; ++q) {
      const int c = c2[q];
      if (c < i) { P->lh[b][nl] = v2[q].hi; P->ll[b][nl] = v2[q].lo; nl++; }
      if (i < c) { P->uh[b][nu] = v2[q].hi; P->ul[b][nu] = v2[q].lo; nu++; }
      if (c == i) {
        P->lh[b][nl] = 1.0; P->ll[b][nl] = 0.0; nl++;
        P->uh[b][nu] = v2[q].hi; P->ul[b][nu] = v2[q].lo; nu++;
      }
    }
  free(ptr); free(c2); free(v2);
  return 0;
}

oracle_ddprecon* oracle_ddprecon_create(int ipreflag, int ilu_method, int nblocks, int independ_nvec, int prec,
                                        int n, const int* rowptr, const int* col, const double* vh,
                                        const double* vl, int* err) {
  struct oracle_ddprecon* P = (struct oracle_ddprecon*)calloc(1, sizeof(*P));
  P->prec = prec;
  P->vh = vh;
  P->vl = vl;
  P->base = oracle_precon_create(ipreflag, ilu_method, nblocks, independ_nvec, n, rowptr, col, vh, err);
  if (!P->base) { free(P); return NULL; }
  const oracle_precon* p = P->base;
  if (ipreflag == 1) {
    /* ddget_DInv_crs: Dinv = 1/a_kk in dd, 1/1e-100 for a zero diagonal */
    P->dh = (double*)calloc((size_t)n, sizeof(double));
    P->dl = (double*)calloc((size_t)n, sizeof(double));
    for (int k = 0; k < n; ++k)
      for (int j = rowptr[k] - 1; j < rowptr[k + 1] - 1; ++j)
        if (col[j] - 1 == k) {
          const dd_t a = mk(vh[j], vl ? vl[j] : 0.0);
          const dd_t v = (a.hi == 0.0 && a.lo == 0.0) ? dd_div(mk(1.0, 0.0), mk(1.0e-100, 0.0))
                                                      : dd_div(mk(1.0, 0.0), a);
          P->dh[k] = v.hi;
          P->dl[k] = v.lo;
        }
  } else if (ipreflag == 2 && ilu_method != 0) {
    P->dh = (double*)calloc((size_t)n, sizeof(double));
    P->dl = (double*)calloc((size_t)n, sizeof(double));
#pragma omp parallel for schedule(dynamic)
    for (int b = 0; b < p->nblocks; ++b) dd_dilu_block(P, p->blo[b], p->bhi[b]);
  } else if (ipreflag == 2) {
    P->lh = (double**)calloc((size_t)p->nblocks, sizeof(double*));
    P->ll = (double**)calloc((size_t)p->nblocks, sizeof(double*));
    P->uh = (double**)calloc((size_t)p->nblocks, sizeof(double*));
    P->ul = (double**)calloc((size_t)p->nblocks, sizeof(double*));
    int bad = 0;
#pragma omp parallel for schedule(dynamic) reduction(| : bad)
    for (int b = 0; b < p->nblocks; ++b)
      if (dd_ilu0_block(P, b, p->blo[b], p->bhi[b]) != 0) bad |= 1;
    if (bad) { *err = -2; oracle_ddprecon_free(P); return NULL; }
  }
  return P;
}

void oracle_ddprecon_free(oracle_ddprecon* P) {
  if (!P) return;
  if (P->lh) {
    for (int b = 0; b < P->base->nblocks; ++b) { free(P->lh[b]); free(P->ll[b]); free(P->uh[b]); free(P->ul[b]); }
    free(P->lh); free(P->ll); free(P->uh); free(P->ul);
  }
  free(P->dh); free(P->dl);
  oracle_precon_free(P->base);
  free(P);
}

void oracle_ddprecon_apply(const oracle_ddprecon* P, const double* rh, const double* rl, double* sh, double* sl) {
  const oracle_precon* p = P->base;
  const int n = p->n;
  const int two = P->prec == 2;
  if (p->ipreflag == 0) {
    memcpy(sh, rh, sizeof(double) * (size_t)n);
    memcpy(sl, rl, sizeof(double) * (size_t)n);
  } else if (p->ipreflag == 1) {
    for (int i = 0; i < n; ++i) {
      const dd_t v = dd_mul(mk(P->dh[i], two ? P->dl[i] : 0.0), mk(rh[i], rl[i]));
      sh[i] = v.hi;
      sl[i] = v.lo;
    }
  } else if (p->ilu_method != 0) {
    const int* cl = p->col;
#pragma omp parallel for schedule(dynamic)
    for (int b = 0; b < p->nblocks; ++b) {
      const int lo = p->blo[b], hi = p->bhi[b];
      for (int k = lo; k <= hi; ++k) {
        dd_t zz = mk(rh[k - 1], rl[k - 1]);
        for (int j = p->lstart[k - 1]; j <= p->lend[k - 1]; ++j) {
          const int c = cl[j] - 1;
          zz = dd_sub(zz, dd_mul(mk(sh[c], sl[c]), mk(P->vh[j], two && P->vl ? P->vl[j] : 0.0)));
        }
        const dd_t v = dd_mul(zz, mk(P->dh[k - 1], two ? P->dl[k - 1] : 0.0));
        sh[k - 1] = v.hi;
        sl[k - 1] = v.lo;
      }
      for (int k = hi; k >= lo; --k) {
        dd_t zz = mk(0.0, 0.0);
        for (int j = p->ustart[k - 1]; j >= p->uend[k - 1]; --j) {
          const int c = cl[j] - 1;
          zz = dd_add(zz, dd_mul(mk(sh[c], sl[c]), mk(P->vh[j], two && P->vl ? P->vl[j] : 0.0)));
        }
        const dd_t c2 = dd_mul(mk(P->dh[k - 1], two ? P->dl[k - 1] : 0.0), zz);
        const dd_t v = dd_sub(mk(sh[k - 1], sl[k - 1]), c2);
        sh[k - 1] = v.hi;
        sl[k - 1] = v.lo;
      }
    }
  } else {
#pragma omp parallel for schedule(dynamic)
    for (int b = 0; b < p->nblocks; ++b) {
      const int lo = p->blo[b], hi = p->bhi[b], nb = hi - lo + 1;
      const int* lp = p->lptr[b]; const int* lc = p->lcol[b];
      const int* up = p->uptr[b]; const int* uc = p->ucol[b];
      const double *lh = P->lh[b], *ll = P->ll[b], *uh = P->uh[b], *ul = P->ul[b];
      double* ssh = sh + (lo - 1);
      double* ssl = sl + (lo - 1);
      for (int i = 0; i < nb; ++i) {
        dd_t z = mk(rh[lo - 1 + i], rl[lo - 1 + i]);
        for (int j = lp[i]; j < lp[i + 1] - 1; ++j)
          z = dd_sub(z, dd_mul(mk(ssh[lc[j]], ssl[lc[j]]), mk(lh[j], two ? ll[j] : 0.0)));
        const int d = lp[i + 1] - 1;
        const dd_t v = dd_div(z, mk(lh[d], two ? ll[d] : 0.0));
        ssh[i] = v.hi;
        ssl[i] = v.lo;
      }
      for (int i = nb - 1; i >= 0; --i) {
        dd_t z = mk(ssh[i], ssl[i]);
        for (int j = up[i] + 1; j < up[i + 1]; ++j)
          z = dd_sub(z, dd_mul(mk(ssh[uc[j]], ssl[uc[j]]), mk(uh[j], two ? ul[j] : 0.0)));
        const int d = up[i];
        const dd_t v = dd_div(z, mk(uh[d], two ? ul[d] : 0.0));
        ssh[i] = v.hi;
        ssl[i] = v.lo;
      }
    }
  }
}

/* ---- dd PCG (src/ddpcg.f90:247-563) --------------------------------------- */

int oracle_ddcg(int n, const int* rowptr, const int* col, const double* vh, const double* vl, const double* bh,
                const double* bl, double* xh, double* xl, int ipreflag, int ilu_method, int nblocks,
                int independ_nvec, int prec, int itrmax, double rtol, const oracle_red* red, double* hist,
                int* niter, int* iret) {
  int err = 0;
  oracle_ddprecon* M = oracle_ddprecon_create(ipreflag, ilu_method, nblocks, independ_nvec, prec, n, rowptr, col,
                                              vh, vl, &err);
  if (!M) return err;
  const size_t nn = (size_t)(n > 0 ? n : 1);
  double *rh = calloc(nn, 8), *rl = calloc(nn, 8), *ph = calloc(nn, 8), *pl = calloc(nn, 8);
  double *qh = calloc(nn, 8), *ql = calloc(nn, 8), *sh = calloc(nn, 8), *sl = calloc(nn, 8);
  /* r = b - A x: ddaxpy(1, -1, b, Ax) = dd_add(dd_mul(1, b), dd_mul(-1, Ax)) */
  oracle_dd_spmv(n, rowptr, col, vh, vl, xh, xl, rh, rl);
  for (int i = 0; i < n; ++i) {
    const dd_t c = dd_mul(mk(-1.0, 0.0), mk(rh[i], rl[i]));
    const dd_t v = dd_add(dd_mul(mk(1.0, 0.0), mk(bh[i], bl[i])), c);
    rh[i] = v.hi;
    rl[i] = v.lo;
  }
  const dd_t bnorm = dd_sqrt(dd_dot(red, n, bh, bl, bh, bl));
  dd_t rnorm = dd_sqrt(dd_dot(red, n, rh, rl, rh, rl));
  dd_t alpha = mk(0.0, 0.0), beta = mk(0.0, 0.0);
  oracle_ddprecon_apply(M, rh, rl, sh, sl);
  dd_t rho0 = dd_dot(red, n, sh, sl, rh, rl);
  int itr;
  for (itr = 1; itr <= itrmax; ++itr) {
    hist[itr - 1] = rnorm.hi / bnorm.hi;
    for (int i = 0; i < n; ++i) { /* ddCG_axpy2 */
      const dd_t p = mk(ph[i], pl[i]);
      const dd_t xv = dd_add(mk(xh[i], xl[i]), dd_mul(alpha, p));
      xh[i] = xv.hi;
      xl[i] = xv.lo;
      const dd_t pv = dd_add(mk(sh[i], sl[i]), dd_mul(beta, p));
      ph[i] = pv.hi;
      pl[i] = pv.lo;
    }
    oracle_dd_spmv(n, rowptr, col, vh, vl, ph, pl, qh, ql);
    const dd_t gamma = dd_dot(red, n, ph, pl, qh, ql);
    alpha = dd_div(rho0, gamma);
    const dd_t ma = dd_mul(mk(-1.0, 0.0), alpha);
    for (int i = 0; i < n; ++i) { /* ddaxpy(-alpha, 1, q, r) */
      const dd_t c = dd_mul(mk(1.0, 0.0), mk(rh[i], rl[i]));
      const dd_t v = dd_add(dd_mul(ma, mk(qh[i], ql[i])), c);
      rh[i] = v.hi;
      rl[i] = v.lo;
    }
    oracle_ddprecon_apply(M, rh, rl, sh, sl);
    const dd_t rho1 = dd_dot(red, n, rh, rl, sh, sl);     /* ddCG_calc_dot2: dd_mul(r, s) */
    rnorm = dd_sqrt(dd_dot(red, n, rh, rl, rh, rl));
    if (rnorm.hi / bnorm.hi < rtol) break;
    beta = dd_div(rho1, rho0);
    rho0 = rho1;
  }
  for (int i = 0; i < n; ++i) { /* ddaxpy(alpha, 1, p, x) */
    const dd_t c = dd_mul(mk(1.0, 0.0), mk(xh[i], xl[i]));
    const dd_t v = dd_add(dd_mul(alpha, mk(ph[i], pl[i])), c);
    xh[i] = v.hi;
    xl[i] = v.lo;
  }
  *niter = itr < itrmax ? itr : itrmax;
  *iret = (itrmax <= itr) ? 1 : 0;
  free(rh); free(rl); free(ph); free(pl); free(qh); free(ql); free(sh); free(sl);
  oracle_ddprecon_free(M);
  return 0;
}

/* as.c -- additive Schwarz preconditioner of the CRS path at one rank.
 * TEST INFRASTRUCTURE ONLY (see parcel_oracle.h).
 *
 * Restates get_DILU_DInv_AS_crs (src/matrix.f90:3809-4375), get_AS_region
 * and check_AS_region (:3146-3287), DILU_DInv_AS_crs (:3428-3480),
 * AdditiveSchwarz_crs (:4749-4858), AdditiveSchwarz_ILU_crs (:5907-6009)
 * and AS_Local (:5804-5904) for one rank (no neighbour ranks, so the rank
 * sleeves of set_AS_vec / get_AS_vec are empty):
 *   - the nblocks blocks of divide_region are the "threads"; block t owns
 *     rows [olo, ohi] and is extended by addL rows on both sides, clipped to
 *     [1, n] (get_AS_region);
 *   - if addL == 0, n < addL or some block has fewer than 2 addL rows
 *     (check_AS_region) the preconditioner is block Jacobi D-ILU instead;
 *   - region t is D-ILU factorised on its own rows and columns
 *     (DILU_DInv_crs with ILU_method 1 = ILUD, any other value the row-sum
 *     form) and solved with solve_ILU_crs;
 *   - iflagAS 1 BASIC / 2 RESTRICT: the region's input is r on all its
 *     rows; 3 INTERPOLATE / 4 NONE: r on the owned rows, 0 on the overlap;
 *   - 2 RESTRICT / 4 NONE: s = the region's result on its owned rows;
 *     1 BASIC / 3 INTERPOLATE: s = 0, then every region adds its result in
 *     three phases (lower overlap rows, owned rows, upper overlap rows). */
#include <stdlib.h>
#include <string.h>

#include "oracle_internal.h"

oracle_precon* oracle_precon_create_as(int ilu_method, int nblocks, int addL, int iflagAS, int independ_nvec, int n,
                                       const int* rowptr, const int* col, const double* val, int* err) {
  *err = 0;
  if (nblocks < 1) nblocks = 1;
  if (nblocks > n) nblocks = n;
  if (iflagAS < 1 || iflagAS > 4 || addL < 0) { *err = -1; return NULL; }
  const int dilu_method = ilu_method == 1 ? 1 : 2;
  int ok = addL > 0 && n >= addL;
  for (int t = 0; t < nblocks && ok; ++t) {
    int lo, hi;
    oracle_divide_region(t, nblocks, n, &lo, &hi);
    if ((hi - lo + 1) / 2 < addL) ok = 0;
  }
  if (!ok) /* block Jacobi with the D-ILU factor get_DILU_DInv_crs computes */
    return oracle_precon_create(2, dilu_method, nblocks, independ_nvec, n, rowptr, col, val, err);
  oracle_precon* p = (oracle_precon*)calloc(1, sizeof(oracle_precon));
  p->ipreflag = ORACLE_PRECON_AS;
  p->n = n;
  p->nblocks = nblocks;
  p->addL = addL;
  p->iflagAS = iflagAS;
  p->as_reg = (oracle_precon**)calloc((size_t)nblocks, sizeof(oracle_precon*));
  p->olo = (int*)malloc(sizeof(int) * (size_t)nblocks);
  p->ohi = (int*)malloc(sizeof(int) * (size_t)nblocks);
  p->as_in = (double**)calloc((size_t)nblocks, sizeof(double*));
  p->as_out = (double**)calloc((size_t)nblocks, sizeof(double*));
  int bad = 0;
  for (int t = 0; t < nblocks; ++t) {
    int lo, hi;
    oracle_divide_region(t, nblocks, n, &lo, &hi);
    p->olo[t] = lo;
    p->ohi[t] = hi;
    const int alo = lo - addL < 1 ? 1 : lo - addL;
    const int ahi = hi + addL > n ? n : hi + addL;
    const int len = ahi - alo + 1;
    oracle_precon* q = (oracle_precon*)calloc(1, sizeof(oracle_precon));
    q->ipreflag = 2;
    q->ilu_method = dilu_method;
    q->independ_nvec = independ_nvec < 1 ? 1 : independ_nvec;
    q->n = len;
    q->base = alo - 1;
    q->rowptr = rowptr; q->col = col; q->val = val;
    q->nblocks = 1;
    q->blo = (int*)malloc(sizeof(int));
    q->bhi = (int*)malloc(sizeof(int));
    q->blo[0] = alo;
    q->bhi[0] = ahi;
    q->dinv = (double*)calloc((size_t)len, sizeof(double));
    q->lstart = (int*)malloc(sizeof(int) * (size_t)len);
    q->lend = (int*)malloc(sizeof(int) * (size_t)len);
    q->ustart = (int*)malloc(sizeof(int) * (size_t)len);
    q->uend = (int*)malloc(sizeof(int) * (size_t)len);
    p->as_reg[t] = q;
    p->as_in[t] = (double*)calloc((size_t)len, sizeof(double));
    p->as_out[t] = (double*)calloc((size_t)len, sizeof(double));
  }
#pragma omp parallel for schedule(dynamic) reduction(| : bad)
  for (int t = 0; t < nblocks; ++t) {
    oracle_precon* q = p->as_reg[t];
    if (oracle_dilu_block(q, q->blo[0], q->bhi[0]) != 0) bad |= 1;
    else oracle_dilu_ranges(q, q->blo[0], q->bhi[0]);
  }
  if (bad) { *err = -2; oracle_as_free(p); return NULL; }
  return p;
}

void oracle_as_apply(const oracle_precon* p, const double* r, double* s) {
  const int nt = p->nblocks;
  const int own_only_in = p->iflagAS == 3 || p->iflagAS == 4;
  const int own_only_out = p->iflagAS == 2 || p->iflagAS == 4;
#pragma omp parallel for schedule(dynamic)
  for (int t = 0; t < nt; ++t) {
    const oracle_precon* q = p->as_reg[t];
    const int alo = q->blo[0], ahi = q->bhi[0];
    double* in = p->as_in[t];
    for (int k = alo; k <= ahi; ++k) {
      const int own = p->olo[t] <= k && k <= p->ohi[t];
      in[k - alo] = (own || !own_only_in) ? r[k - 1] : 0.0;
    }
    oracle_dilu_apply_block(q, alo, ahi, in, p->as_out[t]);
  }
  if (own_only_out) {
    for (int t = 0; t < nt; ++t) {
      const int alo = p->as_reg[t]->blo[0];
      for (int k = p->olo[t]; k <= p->ohi[t]; ++k) s[k - 1] = p->as_out[t][k - alo];
    }
    return;
  }
  for (int k = 0; k < p->n; ++k) s[k] = 0.0;
  for (int t = 0; t < nt; ++t) { /* lower overlap rows */
    const int alo = p->as_reg[t]->blo[0];
    for (int k = alo; k <= p->olo[t] - 1; ++k) s[k - 1] = s[k - 1] + p->as_out[t][k - alo];
  }
  for (int t = 0; t < nt; ++t) { /* owned rows */
    const int alo = p->as_reg[t]->blo[0];
    for (int k = p->olo[t]; k <= p->ohi[t]; ++k) s[k - 1] = s[k - 1] + p->as_out[t][k - alo];
  }
  for (int t = 0; t < nt; ++t) { /* upper overlap rows */
    const int alo = p->as_reg[t]->blo[0], ahi = p->as_reg[t]->bhi[0];
    for (int k = p->ohi[t] + 1; k <= ahi; ++k) s[k - 1] = s[k - 1] + p->as_out[t][k - alo];
  }
}

void oracle_as_free(oracle_precon* p) {
  if (!p) return;
  for (int t = 0; t < p->nblocks; ++t) {
    if (p->as_reg && p->as_reg[t]) {
      oracle_precon* q = p->as_reg[t];
      free(q->blo); free(q->bhi); free(q->dinv);
      free(q->lstart); free(q->lend); free(q->ustart); free(q->uend);
      free(q);
    }
    if (p->as_in) free(p->as_in[t]);
    if (p->as_out) free(p->as_out[t]);
  }
  free(p->as_reg); free(p->olo); free(p->ohi); free(p->as_in); free(p->as_out);
  free(p);
}

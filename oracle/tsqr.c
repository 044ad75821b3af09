/*
 * tsqr.c -- CPU oracle for the TSQR R factor (TSQR_Ronly_,
 * src/QRDecomposition.f90:887-1002; TSQR_local_Ronly_ and TSQR_GLOBAL,
 * src/TSQR.f90:114-498; House_local_Ronly_omp, src/TSQR.f90:836-927).
 * TEST INFRASTRUCTURE ONLY (see parcel_oracle.h).
 *
 * The HIP path's two levels, bit for bit:
 *  1. The n rows are cut into 64-row panels (the last zero-padded); panel p
 *     belongs to wavefront p mod W, W = min(#panels, 1024), taken in
 *     increasing p.  A wavefront keeps an nc x nc upper-triangular R (zero at
 *     the start) and folds in each panel P by Householder QR of [R; P]:
 *     for column i, with x = P(:,i) and alpha = R(i,i),
 *       s2 = sum_lanes x*x,  d_k = sum_lanes x*P(:,k)   (64-lane xor butterfly),
 *       nrm = sqrt(fma(alpha, alpha, s2)),  y1 = alpha > 0 ? alpha + nrm : alpha - nrm,
 *       den = fma(y1, y1, s2) (0 -> 1.0 - 300, the reference's guard),  t1 = 2/den,
 *       R(i,i) = fma(-(t1*y1), fma(y1, alpha, s2), alpha),
 *       for k > i: tk = t1*fma(y1, R(i,k), d_k),  P(:,k) = fma(-x, tk, P(:,k)),
 *                  R(i,k) = fma(-y1, tk, R(i,k)).
 *     This is House_local_Ronly_omp's reflector (v = [y1; x]) on the
 *     stacked matrix, whose rows below R(i,i) in column i are zero.
 *  2. The W factors stacked in wavefront order (W nc x nc) get
 *     House_local_Ronly_omp with sequential fma sums; TSQR_GLOBAL is the
 *     identity on one rank; rows with a negative diagonal are negated
 *     (QRDecomposition.f90:941-947).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "parcel_oracle.h"

#define IDX(i, j, ld) ((size_t)(i) + (size_t)(j) * (size_t)(ld))
#define TSQR_WAVES 1024

/* the 64-lane xor butterfly of wave_sum: every lane ends with lane 0's value */
static double wave64(const double* v) {
  double a[64], b[64];
  memcpy(a, v, sizeof(a));
  for (int o = 32; o >= 1; o >>= 1) {
    for (int l = 0; l < 64; ++l) b[l] = a[l] + a[l ^ o];
    memcpy(a, b, sizeof(a));
  }
  return a[0];
}

/* House_local_Ronly_omp (src/TSQR.f90:836-927) on the m x s matrix A (ld
 * lda, overwritten); R (s x s, ld s) gets its upper triangle. */
void oracle_house_r(int m, int s, double* A, int lda, double* R) {
  for (int i = 0; i < s; ++i) {
    double tmp2 = 0.0;
    for (int j = i + 1; j < m; ++j) tmp2 = fma(A[IDX(j, i, lda)], A[IDX(j, i, lda)], tmp2);
    const double aii = A[IDX(i, i, lda)];
    const double tmp3 = sqrt(fma(aii, aii, tmp2));
    const double y1 = aii > 0.0 ? aii + tmp3 : aii - tmp3;
    const double tmp = fma(y1, aii, tmp2);
    double den = fma(y1, y1, tmp2);
    if (den == 0.0) den = 1.0 - 300;
    const double t1 = 2.0 / den;
    A[IDX(i, i, lda)] = fma(-(t1 * y1), tmp, aii);
    for (int k = i + 1; k < s; ++k) {
      double d = y1 * A[IDX(i, k, lda)];
      for (int j = i + 1; j < m; ++j) d = fma(A[IDX(j, i, lda)], A[IDX(j, k, lda)], d);
      d = t1 * d;
      for (int j = i + 1; j < m; ++j) A[IDX(j, k, lda)] = fma(-A[IDX(j, i, lda)], d, A[IDX(j, k, lda)]);
      A[IDX(i, k, lda)] = fma(-y1, d, A[IDX(i, k, lda)]);
    }
  }
  for (int j = 0; j < s; ++j)
    for (int i = 0; i < s; ++i) R[IDX(i, j, s)] = (i <= j) ? A[IDX(i, j, lda)] : 0.0;
}

/* level 1 for one wavefront: fold panels w, w + W, ... into R (nc x nc) */
static void wave_tsqr(int n, int nc, const double* V, size_t ldv, int w, int W, double* R) {
  const int np = (n + 63) / 64;
  double P[64 * 16], prod[64];
  memset(R, 0, sizeof(double) * (size_t)nc * nc);
  for (int p = w; p < np; p += W) {
    for (int l = 0; l < 64; ++l) {
      const long row = (long)p * 64 + l;
      for (int c = 0; c < nc; ++c) P[l + 64 * c] = row < n ? V[(size_t)c * ldv + (size_t)row] : 0.0;
    }
    for (int i = 0; i < nc; ++i) {
      const double* x = P + 64 * i;
      for (int l = 0; l < 64; ++l) prod[l] = x[l] * x[l];
      const double s2 = wave64(prod);
      double d[16];
      for (int k = i + 1; k < nc; ++k) {
        for (int l = 0; l < 64; ++l) prod[l] = x[l] * P[l + 64 * k];
        d[k] = wave64(prod);
      }
      const double alpha = R[IDX(i, i, nc)];
      const double nrm = sqrt(fma(alpha, alpha, s2));
      const double y1 = alpha > 0.0 ? alpha + nrm : alpha - nrm;
      double den = fma(y1, y1, s2);
      if (den == 0.0) den = 1.0 - 300;
      const double t1 = 2.0 / den;
      R[IDX(i, i, nc)] = fma(-(t1 * y1), fma(y1, alpha, s2), alpha);
      for (int k = i + 1; k < nc; ++k) {
        const double rik = R[IDX(i, k, nc)];
        const double tk = t1 * fma(y1, rik, d[k]);
        for (int l = 0; l < 64; ++l) P[l + 64 * k] = fma(-x[l], tk, P[l + 64 * k]);
        R[IDX(i, k, nc)] = fma(-y1, tk, rik);
      }
    }
  }
}

int oracle_tsqr_waves(int n) {
  const int np = (n + 63) / 64;
  return np < TSQR_WAVES ? (np > 0 ? np : 1) : TSQR_WAVES;
}

/* R (nc x nc, ld nc, diagonal >= 0) of the n x nc matrix V (ld ldv) */
int oracle_tsqr_r(int n, int nc, const double* V, int ldv, double* R) {
  if (nc < 1 || nc > 16 || n < 1) return -1;
  const int W = oracle_tsqr_waves(n);
  const int m = W * nc;
  double* S = (double*)malloc(sizeof(double) * (size_t)m * nc);
#pragma omp parallel for schedule(static)
  for (int w = 0; w < W; ++w) {
    double Rw[256];
    wave_tsqr(n, nc, V, (size_t)ldv, w, W, Rw);
    for (int j = 0; j < nc; ++j)
      for (int i = 0; i < nc; ++i) S[IDX((size_t)w * nc + i, j, m)] = Rw[IDX(i, j, nc)];
  }
  oracle_house_r(m, nc, S, m, R);
  for (int j = 0; j < nc; ++j)
    if (R[IDX(j, j, nc)] < 0.0)
      for (int i = j; i < nc; ++i) R[IDX(j, i, nc)] = -R[IDX(j, i, nc)];
  free(S);
  return 0;
}

/*
 * cagmres.c -- CPU oracle for CA-GMRES(s,t) and the small dense algebra it
 * needs.  TEST INFRASTRUCTURE ONLY (see parcel_oracle.h).
 *
 * Reference: KSP_pCAGMRES (src/pcagmres.f90:1557-2517) with
 *   calc_NewtonBasis / calc_NewtonBasisVec   (src/pcagmres.f90:586-611,1480-1552)
 *   getHessenbergMat_Eigen (LAPACK DHSEQR)   (src/pcagmres.f90:613-739)
 *   modified_leja_ordering                   (src/LejaOrdering.f90:46-277)
 *   calc_HessenbergMatrix, k = 1             (src/pcagmres.f90:879-1051)
 *   GivensRotations                          (src/pcagmres.f90:190-242)
 *   update_xVec / update_xVec_cagmres        (src/pcagmres.f90:1136-1478)
 *   QRDecomposition_: MGS/CGS (GS_QR_), CholeskyQR, CholeskyQR2
 *                                            (src/QRDecomposition.f90:418-1132,1331-1551)
 *
 * Dense kernels follow the loop orders of the reference BLAS/LAPACK routines
 * the reference calls (DGEMM, DTRMM, DTRTI2, DPOTF2, DLAHQR, DLANV2), with
 * every a*b+c of an inner product or update evaluated as one fma.  LAPACK's
 * rounding is implementation dependent (MKL, OpenBLAS, netlib differ), so
 * these are pinned against numpy/scipy and closed forms, not bit for bit.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "parcel_oracle.h"

#define IDX(i, j, ld) ((size_t)(i) + (size_t)(j) * (size_t)(ld))

/* ---- dense kernels (column-major, 0-based) ----------------------------- */

/* C(m x n) = A(m x k) B(k x n)   (DGEMM 'N','N', alpha 1, beta 0) */
void oracle_gemm_nn(int m, int n, int k, const double* A, int lda, const double* B, int ldb, double* C,
                    int ldc) {
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i < m; ++i) C[IDX(i, j, ldc)] = 0.0;
    for (int l = 0; l < k; ++l) {
      const double t = B[IDX(l, j, ldb)];
      for (int i = 0; i < m; ++i) C[IDX(i, j, ldc)] = fma(t, A[IDX(i, l, lda)], C[IDX(i, j, ldc)]);
    }
  }
}

/* B := A B with A upper triangular, non-unit (DTRMM 'L','U','N','N', alpha 1) */
void oracle_trmm_lunn(int n, const double* A, int lda, double* B, int ldb) {
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < n; ++k) {
      double t = B[IDX(k, j, ldb)];
      if (t != 0.0) {
        for (int i = 0; i < k; ++i) B[IDX(i, j, ldb)] = fma(t, A[IDX(i, k, lda)], B[IDX(i, j, ldb)]);
        B[IDX(k, j, ldb)] = t * A[IDX(k, k, lda)];
      }
    }
}

/* In-place inverse of an upper triangular matrix (DTRTI2 'U','N'). */
void oracle_trinv_upper(int n, double* A, int lda) {
  for (int j = 0; j < n; ++j) {
    A[IDX(j, j, lda)] = 1.0 / A[IDX(j, j, lda)];
    const double ajj = -A[IDX(j, j, lda)];
    double* x = A + (size_t)j * lda;
    /* x(0:j) := T(0:j,0:j) x  (DTRMV upper, no transpose, non-unit) */
    for (int c = 0; c < j; ++c) {
      const double t = x[c];
      if (t != 0.0) {
        for (int i = 0; i < c; ++i) x[i] = fma(t, A[IDX(i, c, lda)], x[i]);
        x[c] = t * A[IDX(c, c, lda)];
      }
    }
    for (int i = 0; i < j; ++i) x[i] = ajj * x[i];
  }
}

/* Cholesky A = R^T R of the upper triangle (DPOTF2 'U').  Returns 0, or
 * j+1 if the j-th pivot is not positive (A(j,j) then holds it and later
 * columns are left untouched, as DPOTF2 leaves them). */
int oracle_chol_upper(int n, double* A, int lda) {
  for (int j = 0; j < n; ++j) {
    double d = 0.0;
    for (int k = 0; k < j; ++k) d = fma(A[IDX(k, j, lda)], A[IDX(k, j, lda)], d);
    double ajj = A[IDX(j, j, lda)] - d;
    if (ajj <= 0.0 || isnan(ajj)) {
      A[IDX(j, j, lda)] = ajj;
      return j + 1;
    }
    ajj = sqrt(ajj);
    A[IDX(j, j, lda)] = ajj;
    const double r = 1.0 / ajj;
    for (int c = j + 1; c < n; ++c) {
      double acc = 0.0;
      for (int k = 0; k < j; ++k) acc = fma(A[IDX(k, j, lda)], A[IDX(k, c, lda)], acc);
      A[IDX(j, c, lda)] = (A[IDX(j, c, lda)] - acc) * r;
    }
  }
  return 0;
}

/* Fortran SIGN(a, b): |a| with the sign bit of b */
static double sign_of(double a, double b) { return signbit(b) ? -fabs(a) : fabs(a); }

static double lapy2(double x, double y) {
  const double xa = fabs(x), ya = fabs(y);
  const double w = xa > ya ? xa : ya, z = xa < ya ? xa : ya;
  if (z == 0.0 || w > DBL_MAX) return w;
  const double q = z / w;
  return w * sqrt(1.0 + q * q);
}

/* Schur factorisation of a real 2x2 block in standard form (DLANV2):
 * returns the eigenvalue pair, complex pairs with the positive imaginary
 * part first. */
static void lanv2(double a, double b, double c, double d, double* rt1r, double* rt1i, double* rt2r,
                  double* rt2i) {
  const double eps = DBL_EPSILON;
  if (c == 0.0) {
  } else if (b == 0.0) {
    const double t = d;
    d = a;
    a = t;
    b = -c;
    c = 0.0;
  } else if ((a - d) == 0.0 && sign_of(1.0, b) != sign_of(1.0, c)) {
  } else {
    const double temp = a - d;
    double p = 0.5 * temp;
    const double bcmax = fabs(b) > fabs(c) ? fabs(b) : fabs(c);
    const double bcmis = (fabs(b) < fabs(c) ? fabs(b) : fabs(c)) * sign_of(1.0, b) * sign_of(1.0, c);
    const double scale = fabs(p) > bcmax ? fabs(p) : bcmax;
    double z = (p / scale) * p + (bcmax / scale) * bcmis;
    if (z >= 4.0 * eps) {
      /* real eigenvalues */
      z = p + sign_of(sqrt(scale) * sqrt(z), p);
      a = d + z;
      d = d - (bcmax / z) * bcmis;
      b = b - c;
      c = 0.0;
    } else {
      /* complex or real (almost) equal eigenvalues: make the diagonal equal */
      const double sigma = b + c;
      const double tau = lapy2(sigma, temp);
      double cs = sqrt(0.5 * (1.0 + fabs(sigma) / tau));
      double sn = -(p / (tau * cs)) * sign_of(1.0, sigma);
      const double aa = a * cs + b * sn, bb = -a * sn + b * cs;
      const double cc = c * cs + d * sn, dd = -c * sn + d * cs;
      a = aa * cs + cc * sn;
      b = bb * cs + dd * sn;
      c = -aa * sn + cc * cs;
      d = -bb * sn + dd * cs;
      const double t = 0.5 * (a + d);
      a = t;
      d = t;
      if (c != 0.0) {
        if (b != 0.0) {
          if (sign_of(1.0, b) == sign_of(1.0, c)) {
            /* real eigenvalues */
            const double sab = sqrt(fabs(b)), sac = sqrt(fabs(c));
            p = sign_of(sab * sac, c);
            a = t + p;
            d = t - p;
            b = b - c;
            c = 0.0;
          }
        } else {
          b = -c;
          c = 0.0;
        }
      }
    }
  }
  *rt1r = a;
  *rt2r = d;
  if (c == 0.0) {
    *rt1i = 0.0;
    *rt2i = 0.0;
  } else {
    *rt1i = sqrt(fabs(b)) * sqrt(fabs(c));
    *rt2i = -*rt1i;
  }
}

/* Householder reflector (DLARFG) for (alpha, x[0..nx)). */
static void larfg(int nx, double* alpha, double* x, double* tau) {
  double xnorm = 0.0;
  if (nx == 1) xnorm = fabs(x[0]);
  else if (nx == 2) xnorm = lapy2(x[0], x[1]);
  if (xnorm == 0.0) {
    *tau = 0.0;
    return;
  }
  const double beta = -sign_of(lapy2(*alpha, xnorm), *alpha);
  *tau = (beta - *alpha) / beta;
  const double sc = 1.0 / (*alpha - beta);
  for (int i = 0; i < nx; ++i) x[i] = sc * x[i];
  *alpha = beta;
}

/* Eigenvalues of the n x n upper Hessenberg H (overwritten) by the
 * double-shift QR algorithm (DLAHQR with WANTT = WANTZ = false, the path
 * DHSEQR takes for small matrices).  Returns 0, or the 1-based index of an
 * eigenvalue that failed to converge. */
int oracle_hqr_eigs(int n, double* H, int ldh, double* wr, double* wi) {
  if (n <= 0) return 0;
  if (n == 1) {
    wr[0] = H[0];
    wi[0] = 0.0;
    return 0;
  }
  for (int j = 0; j + 3 < n; ++j) {
    H[IDX(j + 2, j, ldh)] = 0.0;
    H[IDX(j + 3, j, ldh)] = 0.0;
  }
  if (n >= 3) H[IDX(n - 1, n - 3, ldh)] = 0.0;
  const double safmin = DBL_MIN;
  const double ulp = DBL_EPSILON;
  const double smlnum = safmin * ((double)n / ulp);
  const int itmax = 30 * (n > 10 ? n : 10);
  const double dat1 = 0.75, dat2 = -0.4375;
  const int kexsh = 10;
  int kdefl = 0;
  int i = n - 1;
  while (i >= 0) {
    int l = 0;
    int converged = 0;
    for (int its = 0; its <= itmax; ++its) {
      int k;
      for (k = i; k > l; --k) {
        if (fabs(H[IDX(k, k - 1, ldh)]) <= smlnum) break;
        double tst = fabs(H[IDX(k - 1, k - 1, ldh)]) + fabs(H[IDX(k, k, ldh)]);
        if (tst == 0.0) {
          if (k - 2 >= 0) tst += fabs(H[IDX(k - 1, k - 2, ldh)]);
          if (k + 1 <= n - 1) tst += fabs(H[IDX(k + 1, k, ldh)]);
        }
        if (fabs(H[IDX(k, k - 1, ldh)]) <= ulp * tst) {
          const double h1 = fabs(H[IDX(k, k - 1, ldh)]), h2 = fabs(H[IDX(k - 1, k, ldh)]);
          const double ab = h1 > h2 ? h1 : h2, ba = h1 < h2 ? h1 : h2;
          const double d1 = fabs(H[IDX(k, k, ldh)]);
          const double d2 = fabs(H[IDX(k - 1, k - 1, ldh)] - H[IDX(k, k, ldh)]);
          const double aa = d1 > d2 ? d1 : d2, bb = d1 < d2 ? d1 : d2;
          const double s = aa + ab;
          const double rhs = ulp * (bb * (aa / s));
          if (ba * (ab / s) <= (smlnum > rhs ? smlnum : rhs)) break;
        }
      }
      l = k;
      if (l > 0) H[IDX(l, l - 1, ldh)] = 0.0;
      if (l >= i - 1) {
        converged = 1;
        break;
      }
      ++kdefl;
      double h11, h12, h21, h22;
      if (kdefl % (2 * kexsh) == 0) {
        const double s = fabs(H[IDX(i, i - 1, ldh)]) + fabs(H[IDX(i - 1, i - 2, ldh)]);
        h11 = dat1 * s + H[IDX(i, i, ldh)];
        h12 = dat2 * s;
        h21 = s;
        h22 = h11;
      } else if (kdefl % kexsh == 0) {
        const double s = fabs(H[IDX(l + 1, l, ldh)]) + fabs(H[IDX(l + 2, l + 1, ldh)]);
        h11 = dat1 * s + H[IDX(l, l, ldh)];
        h12 = dat2 * s;
        h21 = s;
        h22 = h11;
      } else {
        h11 = H[IDX(i - 1, i - 1, ldh)];
        h21 = H[IDX(i, i - 1, ldh)];
        h12 = H[IDX(i - 1, i, ldh)];
        h22 = H[IDX(i, i, ldh)];
      }
      double rt1r, rt1i, rt2r, rt2i;
      const double s = fabs(h11) + fabs(h12) + fabs(h21) + fabs(h22);
      if (s == 0.0) {
        rt1r = rt1i = rt2r = rt2i = 0.0;
      } else {
        h11 /= s;
        h21 /= s;
        h12 /= s;
        h22 /= s;
        const double tr = (h11 + h22) / 2.0;
        const double det = (h11 - tr) * (h22 - tr) - h12 * h21;
        const double rtdisc = sqrt(fabs(det));
        if (det >= 0.0) {
          rt1r = tr * s;
          rt2r = rt1r;
          rt1i = rtdisc * s;
          rt2i = -rt1i;
        } else {
          rt1r = tr + rtdisc;
          rt2r = tr - rtdisc;
          if (fabs(rt1r - h22) <= fabs(rt2r - h22)) {
            rt1r *= s;
            rt2r = rt1r;
          } else {
            rt2r *= s;
            rt1r = rt2r;
          }
          rt1i = rt2i = 0.0;
        }
      }
      /* look for two consecutive small subdiagonal elements */
      int m;
      double v[3];
      for (m = i - 2; m >= l; --m) {
        double h21s = H[IDX(m + 1, m, ldh)];
        double sm = fabs(H[IDX(m, m, ldh)] - rt2r) + fabs(rt2i) + fabs(h21s);
        h21s = H[IDX(m + 1, m, ldh)] / sm;
        v[0] = h21s * H[IDX(m, m + 1, ldh)] + (H[IDX(m, m, ldh)] - rt1r) * ((H[IDX(m, m, ldh)] - rt2r) / sm) -
               rt1i * (rt2i / sm);
        v[1] = h21s * (H[IDX(m, m, ldh)] + H[IDX(m + 1, m + 1, ldh)] - rt1r - rt2r);
        v[2] = h21s * H[IDX(m + 2, m + 1, ldh)];
        sm = fabs(v[0]) + fabs(v[1]) + fabs(v[2]);
        v[0] /= sm;
        v[1] /= sm;
        v[2] /= sm;
        if (m == l) break;
        const double h00 = fabs(H[IDX(m, m - 1, ldh)]) * (fabs(v[1]) + fabs(v[2]));
        const double h01 = fabs(v[0]) * (fabs(H[IDX(m - 1, m - 1, ldh)]) + fabs(H[IDX(m, m, ldh)]) +
                                         fabs(H[IDX(m + 1, m + 1, ldh)]));
        if (h00 <= ulp * h01) break;
      }
      /* double-shift QR sweep over rows/columns l..i */
      for (int kk = m; kk <= i - 1; ++kk) {
        const int nr = (3 < i - kk + 1) ? 3 : i - kk + 1;
        if (kk > m)
          for (int q = 0; q < nr; ++q) v[q] = H[IDX(kk + q, kk - 1, ldh)];
        double t1;
        larfg(nr - 1, &v[0], &v[1], &t1);
        if (kk > m) {
          H[IDX(kk, kk - 1, ldh)] = v[0];
          H[IDX(kk + 1, kk - 1, ldh)] = 0.0;
          if (kk < i - 1) H[IDX(kk + 2, kk - 1, ldh)] = 0.0;
        } else if (m > l) {
          H[IDX(kk, kk - 1, ldh)] = H[IDX(kk, kk - 1, ldh)] * (1.0 - t1);
        }
        const double v2 = v[1], t2 = t1 * v2;
        if (nr == 3) {
          const double v3 = v[2], t3 = t1 * v3;
          for (int j = kk; j <= i; ++j) {
            const double sum = H[IDX(kk, j, ldh)] + v2 * H[IDX(kk + 1, j, ldh)] + v3 * H[IDX(kk + 2, j, ldh)];
            H[IDX(kk, j, ldh)] -= sum * t1;
            H[IDX(kk + 1, j, ldh)] -= sum * t2;
            H[IDX(kk + 2, j, ldh)] -= sum * t3;
          }
          const int jmax = (kk + 3 < i) ? kk + 3 : i;
          for (int j = l; j <= jmax; ++j) {
            const double sum = H[IDX(j, kk, ldh)] + v2 * H[IDX(j, kk + 1, ldh)] + v3 * H[IDX(j, kk + 2, ldh)];
            H[IDX(j, kk, ldh)] -= sum * t1;
            H[IDX(j, kk + 1, ldh)] -= sum * t2;
            H[IDX(j, kk + 2, ldh)] -= sum * t3;
          }
        } else if (nr == 2) {
          for (int j = kk; j <= i; ++j) {
            const double sum = H[IDX(kk, j, ldh)] + v2 * H[IDX(kk + 1, j, ldh)];
            H[IDX(kk, j, ldh)] -= sum * t1;
            H[IDX(kk + 1, j, ldh)] -= sum * t2;
          }
          for (int j = l; j <= i; ++j) {
            const double sum = H[IDX(j, kk, ldh)] + v2 * H[IDX(j, kk + 1, ldh)];
            H[IDX(j, kk, ldh)] -= sum * t1;
            H[IDX(j, kk + 1, ldh)] -= sum * t2;
          }
        }
      }
    }
    if (!converged) return i + 1;
    if (l == i) {
      wr[i] = H[IDX(i, i, ldh)];
      wi[i] = 0.0;
    } else if (l == i - 1) {
      lanv2(H[IDX(i - 1, i - 1, ldh)], H[IDX(i - 1, i, ldh)], H[IDX(i, i - 1, ldh)], H[IDX(i, i, ldh)], &wr[i - 1],
            &wi[i - 1], &wr[i], &wi[i]);
    }
    kdefl = 0;
    i = l - 1;
  }
  return 0;
}

/* modified_leja_ordering (src/LejaOrdering.f90:111-277), in place on
 * (re, im) of length n.  Restated with its exact control flow: stable
 * descending insertion sort by modulus, exponent mu = 1, capacity rescaling
 * by C/Cold each step, and the candidate product over the first L-1 chosen
 * points (mleja_prod loops j = 0..L-2). */
void oracle_leja(int n, double* re, double* im) {
  if (n <= 0) return;
  /* insertion_sort (:46-75) */
  for (int i = 1; i < n; ++i) {
    int j = i;
    while (j > 0) {
      if (sqrt(re[j - 1] * re[j - 1] + im[j - 1] * im[j - 1]) < sqrt(re[j] * re[j] + im[j] * im[j])) {
        double tr = re[j], ti = im[j];
        re[j] = re[j - 1];
        im[j] = im[j - 1];
        re[j - 1] = tr;
        im[j - 1] = ti;
        --j;
      } else {
        break;
      }
    }
  }
  double* ptmp = (double*)malloc(sizeof(double) * (size_t)n);
  double* thr = (double*)malloc(sizeof(double) * (size_t)(n + 1));
  double* thi = (double*)malloc(sizeof(double) * (size_t)(n + 1));
  int* outlist = (int*)malloc(sizeof(int) * (size_t)(n + 1));
  double* tmp = (double*)malloc(sizeof(double) * (size_t)n);
  double* pp = (double*)malloc(sizeof(double) * (size_t)n);
  for (int i = 0; i < n; ++i) ptmp[i] = 1.0;
  const double mu = 1.0;
  double C = 1.0;
  int L;
  for (int i = 0; i <= n; ++i) { thr[i] = 0.0; thi[i] = 0.0; }
  if (im[0] > 0.0) {
    thr[0] = re[0]; thi[0] = im[0];
    thr[1] = re[1]; thi[1] = im[1];
    outlist[0] = 0; outlist[1] = 1;
    L = 2;
  } else {
    thr[0] = re[0]; thi[0] = 0.0;
    outlist[0] = 0;
    L = 1;
  }
  while (L < n) {
    const double Cold = C;
    for (int j = 0; j <= L - 2; ++j) {
      const double tr = thr[L - 1] - thr[j], ti = thi[L - 1] - thi[j];
      ptmp[j + 1] = pow(sqrt(tr * tr + ti * ti), mu / (double)L);
    }
    C = 1.0;
    for (int j = 0; j < n; ++j) C = C * ptmp[j];
    for (int j = 0; j < n; ++j) {
      re[j] = re[j] / (C / Cold);
      im[j] = im[j] / (C / Cold);
      thr[j] = thr[j] / (C / Cold);
      thi[j] = thi[j] / (C / Cold);
    }
    for (int i = 0; i < n; ++i) tmp[i] = 0.0;
    for (int j = 0; j < n; ++j) {
      int contained = 0;
      for (int i = 0; i < L; ++i)
        if (outlist[i] == j) contained = 1;
      if (!contained) {
        /* mleja_prod (:77-99) */
        for (int q = 0; q <= L - 2; ++q) {
          const double tr = re[j] - thr[q], ti = im[j] - thi[q];
          pp[q] = pow(sqrt(tr * tr + ti * ti), mu);
        }
        double prod = 1.0;
        for (int q = 0; q <= L - 2; ++q) prod = prod * pp[q];
        tmp[j] = prod;
      }
    }
    double mx = fabs(tmp[0]);
    int k = 0;
    for (int j = 1; j < n; ++j)
      if (fabs(tmp[j]) > mx) {
        mx = fabs(tmp[j]);
        k = j;
      }
    if (fabs(im[k]) > 0.0) {
      if (im[k] < 0.0) {
        thr[L] = re[k - 1]; thi[L] = im[k - 1];
        thr[L + 1] = re[k]; thi[L + 1] = im[k];
        outlist[L] = k - 1; outlist[L + 1] = k;
      } else {
        thr[L] = re[k]; thi[L] = im[k];
        thr[L + 1] = re[k + 1]; thi[L + 1] = im[k + 1];
        outlist[L] = k; outlist[L + 1] = k + 1;
      }
      L += 2;
    } else {
      thr[L] = re[k]; thi[L] = 0.0;
      outlist[L] = k;
      L += 1;
    }
  }
  for (int i = 0; i < n; ++i) {
    re[i] = thr[i] * C;
    im[i] = thi[i] * C;
  }
  free(ptmp); free(thr); free(thi); free(outlist); free(tmp); free(pp);
}

/* Newton change-of-basis matrix update (calc_NewtonBasis, :586-611):
 * B is (s+1) x s, kept across restarts (entries it does not write keep
 * their previous values, as in the reference). */
static void newton_basis(int s, double* B, const double* er, const double* ei) {
  int i = 0;
  while (i < s) {
    if (i == s - 1) {
      B[IDX(i, i, s + 1)] = er[i];
    } else if (fabs(ei[i]) == 0.0) {
      B[IDX(i, i, s + 1)] = er[i];
    } else {
      B[IDX(i, i, s + 1)] = er[i];
      B[IDX(i + 1, i + 1, s + 1)] = er[i];
      B[IDX(i, i + 1, s + 1)] = -ei[i] * ei[i];
      ++i;
    }
    ++i;
  }
}

/* Givens rotation of column i (1-based) of the (m+1) x m Hessenberg H
 * (GivensRotations, :190-242). */
static void givens_m(int m, double* H, int i, double* c, double* s, double* e) {
  double* hc = H + (size_t)(i - 1) * (m + 1);
  for (int k = 0; k < i - 1; ++k) {
    double t1 = hc[k] * c[k] - hc[k + 1] * s[k];
    double t2 = hc[k] * s[k] + hc[k + 1] * c[k];
    hc[k] = t1;
    hc[k + 1] = t2;
  }
  double t;
  if (hc[i - 1] == 0.0) t = hc[i] / 1.0 - 300;
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

/* Upper-triangular Gram matrix G(k,j) = <V_k, V_j>, k <= j, of nc columns. */
static void gram_upper(const oracle_red* red, int n, int nc, const double* V, size_t ldv, double* G) {
  for (int j = 0; j < nc * nc; ++j) G[j] = 0.0;
  for (int j = 0; j < nc; ++j)
    for (int k = 0; k <= j; ++k) G[IDX(k, j, nc)] = oracle_dot(red, n, V + (size_t)k * ldv, V + (size_t)j * ldv);
}

/* CholeskyQR_Ronly_ (QRDecomposition.f90:717-852), GPU branch: R = chol(V^T V), lower part zeroed */
static void cholqr_r(const oracle_red* red, int n, int nc, const double* V, size_t ldv, double* R) {
  gram_upper(red, n, nc, V, ldv, R);
  (void)oracle_chol_upper(nc, R, nc);
  for (int j = 0; j < nc; ++j)
    for (int i = j + 1; i < nc; ++i) R[IDX(i, j, nc)] = 0.0;
}

/* Q = V R^{-1} (calcQ_VRinv_, QRDecomposition.f90:603-671; column i sums k = 1..i) */
static void q_vrinv(int n, int nc, const double* V, size_t ldv, const double* Rinv, double* Q, size_t ldq) {
  for (int i = 0; i < nc; ++i)
    for (int r = 0; r < n; ++r) {
      double acc = 0.0;
      for (int k = 0; k <= i; ++k) acc = fma(V[(size_t)k * ldv + r], Rinv[IDX(k, i, nc)], acc);
      Q[(size_t)i * ldq + r] = acc;
    }
}

/* QRDecomposition_ (QRDecomposition.f90:1331-1551) of nc columns of V into
 * R (nc x nc, ld nc) and, with qmat (or GS), Q = V R^-1.  nc == 1 always
 * uses MGS.  W: scratch of nc columns (CholeskyQR2's Q1). */
static void qr_cols(const oracle_red* red, int n, int nc, double* V, size_t ldv, double* Q, size_t ldq, double* R,
                    int qrflag, int qmat, double* W, double* R2, double* Rinv) {
  memset(R, 0, sizeof(double) * (size_t)nc * nc);
  if (nc == 1 || qrflag == 1 || qrflag == 2) {
    const int mgs = (nc == 1 || qrflag == 1);
    /* GS_QR_: MGS or CGS, R(k,i) and Q formed; V modified in place */
    for (int i = 0; i < nc; ++i) {
      double* vi = V + (size_t)i * ldv;
      if (mgs) {
        for (int k = 0; k < i; ++k) {
          const double* qk = Q + (size_t)k * ldq;
          const double rk = oracle_dot(red, n, vi, qk);
          R[IDX(k, i, nc)] = rk;
          for (int q = 0; q < n; ++q) vi[q] = fma(-rk, qk[q], vi[q]);
        }
      } else {
        for (int k = 0; k < i; ++k) R[IDX(k, i, nc)] = oracle_dot(red, n, Q + (size_t)k * ldq, vi);
        for (int q = 0; q < n; ++q) {
          double w = vi[q];
          for (int k = 0; k < i; ++k) w = fma(-R[IDX(k, i, nc)], Q[(size_t)k * ldq + q], w);
          vi[q] = w;
        }
      }
      const double rii = sqrt(oracle_dot(red, n, vi, vi));
      R[IDX(i, i, nc)] = rii;
      const double sc = 1.0 / rii;
      double* qi = Q + (size_t)i * ldq;
      for (int q = 0; q < n; ++q) qi[q] = sc * vi[q];
    }
    return;
  }
  if (qrflag == 3 && n > nc && n > 64) {           /* TSQR_Ronly_ (tsqr.c) */
    (void)oracle_tsqr_r(n, nc, V, (int)ldv, R);
  } else {
    cholqr_r(red, n, nc, V, ldv, R);
  }
  if (qrflag == 5 || (qrflag == 3 && !(n > nc && n > 64))) {   /* CholeskyQR2_Ronly_ */
    memcpy(Rinv, R, sizeof(double) * (size_t)nc * nc);
    oracle_trinv_upper(nc, Rinv, nc);
    q_vrinv(n, nc, V, ldv, Rinv, W, (size_t)n);   /* Q1 = V R^-1 */
    cholqr_r(red, n, nc, W, (size_t)n, R2);
    oracle_trmm_lunn(nc, R2, nc, R, nc);           /* R = R2 R */
  }
  if (qmat) {                                      /* calcQ_VRinv_ */
    memcpy(Rinv, R, sizeof(double) * (size_t)nc * nc);
    oracle_trinv_upper(nc, Rinv, nc);
    q_vrinv(n, nc, V, ldv, Rinv, Q, ldq);
  }
}

/* calc_HessenbergMatrix for block k >= 2 (pcagmres.f90:879-1051, with
 * calc_Rk_Bk_Rkinv :850 and calc_hke1eskRRinv :818).  Rp = R' (c1 x (s+1),
 * ld c1 = s(k-1)+1) from the block CGS, Rk the block's s x s R factor, H the
 * (st+1) x st Hessenberg; *hk_pre is read and updated.  The code's effect on
 * its local arrays: BlockR_ubar = R'(1:s(k-1), 1:s+1), BlockR = R'(1:s(k-1),
 * 1:s), Rtri0 = [R'(c1, 1:s+1); 0 Rk]. */
void oracle_hessenberg_block(int s, int t, int k, const double* B1, double* H, const double* Rk, const double* Rp,
                             double* hk_pre) {
  const int m = s * t, c0 = s * (k - 1), c1 = c0 + 1, s1 = s + 1, ldh = m + 1;
  double* Rt0 = (double*)calloc((size_t)s1 * s1, sizeof(double));
  double* Rinv = (double*)calloc((size_t)s * s, sizeof(double));
  double* Rub = (double*)calloc((size_t)c0 * s1, sizeof(double));
  double* BR = (double*)calloc((size_t)c0 * s, sizeof(double));
  double* BH = (double*)calloc((size_t)c0 * c0, sizeof(double));
  double* BHR = (double*)calloc((size_t)c0 * s, sizeof(double));
  double* RubB = (double*)calloc((size_t)c0 * s, sizeof(double));
  double* Hk = (double*)calloc((size_t)s * s, sizeof(double));
  double* RB = (double*)calloc((size_t)s * s, sizeof(double));
  double* B0 = (double*)calloc((size_t)s * s, sizeof(double));
  double* Rs = (double*)calloc((size_t)s * s, sizeof(double));
  double* E = (double*)calloc((size_t)s * s, sizeof(double));
  double* H0 = (double*)calloc((size_t)s * s, sizeof(double));
  for (int i = 0; i < s1; ++i) Rt0[IDX(0, i, s1)] = Rp[IDX(c0, i, c1)];
  for (int j = 0; j < s; ++j)
    for (int i = 0; i < s; ++i) Rt0[IDX(1 + i, j + 1, s1)] = Rk[IDX(i, j, s)];
  for (int j = 0; j < s1; ++j)
    for (int i = 0; i < c0; ++i) Rub[IDX(i, j, c0)] = Rp[IDX(i, j, c1)];
  for (int j = 0; j < s; ++j)
    for (int i = 0; i < c0; ++i) BR[IDX(i, j, c0)] = Rp[IDX(i, j, c1)];
  for (int j = 0; j < s; ++j)
    for (int i = 0; i < s; ++i) Rinv[IDX(i, j, s)] = Rt0[IDX(i, j, s1)];
  oracle_trinv_upper(s, Rinv, s);
  for (int j = 0; j < c0; ++j)
    for (int i = 0; i < c0; ++i) BH[IDX(i, j, c0)] = H[IDX(i, j, ldh)];
  oracle_gemm_nn(c0, s, c0, BH, c0, BR, c0, BHR, c0);
  oracle_gemm_nn(c0, s, s1, Rub, c0, B1, s1, RubB, c0);
  for (int i = 0; i < c0 * s; ++i) BHR[i] = -BHR[i] + RubB[i];
  oracle_gemm_nn(c0, s, s, BHR, c0, Rinv, s, RubB, c0);
  for (int i = 0; i < s; ++i)
    for (int j = 0; j < c0; ++j) H[IDX(j, c0 + i, ldh)] = RubB[IDX(j, i, c0)];
  const double bk = B1[IDX(s, s - 1, s1)];
  const double rho_tilde = Rt0[IDX(s - 1, s - 1, s1)], rho = Rt0[IDX(s, s, s1)];
  const double hk = rho / rho_tilde * bk;
  H[IDX(s * k, s * k - 1, ldh)] = hk;
  /* calc_Rk_Bk_Rkinv */
  for (int i = 0; i < s; ++i)
    for (int j = 0; j < s; ++j) {
      B0[IDX(j, i, s)] = B1[IDX(j, i, s1)];
      Rs[IDX(j, i, s)] = Rt0[IDX(j, i, s1)];
    }
  oracle_gemm_nn(s, s, s, Rs, s, B0, s, RB, s);
  oracle_gemm_nn(s, s, s, RB, s, Rinv, s, Hk, s);
  for (int i = 0; i < s; ++i) Hk[IDX(i, s - 1, s)] = fma(bk / rho_tilde, Rt0[IDX(i, s, s1)], Hk[IDX(i, s - 1, s)]);
  /* calc_hke1eskRRinv */
  for (int i = 0; i < s; ++i) E[IDX(0, i, s)] = BR[IDX(c0 - 1, i, c0)] * *hk_pre;
  oracle_gemm_nn(s, s, s, E, s, Rinv, s, H0, s);
  for (int i = 0; i < s * s; ++i) Hk[i] = Hk[i] - H0[i];
  for (int i = 0; i < s; ++i)
    for (int j = 0; j < s; ++j) H[IDX(c0 + j, c0 + i, ldh)] = Hk[IDX(j, i, s)];
  *hk_pre = hk;
  free(Rt0); free(Rinv); free(Rub); free(BR); free(BH); free(BHR); free(RubB); free(Hk); free(RB); free(B0);
  free(Rs); free(E); free(H0);
}

int oracle_cagmres(int n, const int* rowptr, const int* col, const double* val, const double* b, double* x,
                   int ipreflag, int ilu_method, int nblocks, int independ_nvec, int itrmax, double rtol, int s,
                   int t, int basis, int qrflag, const oracle_red* red, double* hist, int* niter, int* iret) {
  if (s < 1 || t < 1) return -4;
  if (qrflag < 1 || qrflag > 5) return -4;
  int err = 0;
  oracle_precon* M = oracle_precon_create(ipreflag, ilu_method, nblocks, independ_nvec, n, rowptr, col, val, &err);
  if (!M) return err;
  const int m = s * t, s1 = s + 1;
  const int c1max = s * (t - 1) + 1;
  const size_t nn = (size_t)(n > 0 ? n : 1);
  double* r = (double*)calloc(nn, sizeof(double));
  double* V = (double*)calloc(nn * (size_t)s1, sizeof(double));
  double* Q = (double*)calloc(nn * (size_t)(m + 1), sizeof(double));
  double* W = (double*)calloc(nn * (size_t)s1, sizeof(double));
  double* pv = (double*)calloc(nn, sizeof(double));
  double* v0 = (double*)calloc(nn, sizeof(double));
  double* x0 = (double*)calloc(nn, sizeof(double));
  double* B1 = (double*)calloc((size_t)s1 * s, sizeof(double));
  double* R = (double*)calloc((size_t)s1 * s1, sizeof(double));
  double* R2 = (double*)calloc((size_t)s1 * s1, sizeof(double));
  double* Rinv = (double*)calloc((size_t)s1 * s1, sizeof(double));
  double* Rp = (double*)calloc((size_t)c1max * s1, sizeof(double));
  double* RB = (double*)calloc((size_t)s1 * s, sizeof(double));
  double* Hes = (double*)calloc((size_t)(m + 1) * m, sizeof(double));
  double* Hg = (double*)calloc((size_t)(m + 1) * m, sizeof(double));
  double* Hs = (double*)calloc((size_t)m * m, sizeof(double));
  double* co = (double*)calloc((size_t)m + 1, sizeof(double));
  double* si = (double*)calloc((size_t)m + 1, sizeof(double));
  double* zeta = (double*)calloc((size_t)m + 1, sizeof(double));
  double* y = (double*)calloc((size_t)m + 1, sizeof(double));
  double* yt = (double*)calloc((size_t)m + 1, sizeof(double));
  double* er = (double*)calloc((size_t)m, sizeof(double));
  double* ei = (double*)calloc((size_t)m, sizeof(double));
  for (int i = 0; i < s; ++i) B1[IDX(i + 1, i, s1)] = 1.0; /* sub-diagonal 1 (monomial) */
  int iflag_exit = 0, iflag_basis = 0, itr, itr0 = 1, mm = 0, qmat = 0;
  const int newton = basis == 1 || basis == 2;
  double rnorm, hk_pre = 0.0;
  oracle_spmv(n, rowptr, col, val, x, r);
  for (int i = 0; i < n; ++i) r[i] = b[i] - r[i];
  const double bnorm = sqrt(oracle_dot(red, n, b, b));
  for (itr = 1; itr <= itrmax; itr += s * t) {
    itr0 = itr;
    memset(co, 0, sizeof(double) * (m + 1));
    memset(si, 0, sizeof(double) * (m + 1));
    memset(zeta, 0, sizeof(double) * (m + 1));
    memset(Hes, 0, sizeof(double) * (m + 1) * m);
    memset(Hg, 0, sizeof(double) * (m + 1) * m);
    memset(R, 0, sizeof(double) * s1 * s1);
    memset(Rp, 0, sizeof(double) * (size_t)c1max * s1);
    hk_pre = 0.0;
    oracle_spmv(n, rowptr, col, val, x, r);           /* r = b - A x */
    for (int i = 0; i < n; ++i) r[i] = b[i] - r[i];
    rnorm = sqrt(oracle_dot(red, n, r, r));
    if (rnorm / bnorm < rtol) break;                   /* :1977 (abstol = 0) */
    if (itr == 1) hist[itr - 1] = rnorm / bnorm;
    zeta[0] = -rnorm;
    const double a = -1.0 / rnorm;
    for (int i = 0; i < n; ++i) V[i] = a * r[i];       /* v1 = -r/||r|| */
    for (int k = 1; k <= t; ++k) {
      if (newton && iflag_basis) newton_basis(s, B1, er, ei);
      qmat = (k != t);
      for (int i = 1; i <= s; ++i) {
        const double* vi = V + (size_t)(i - 1) * nn;
        double* vn = V + (size_t)i * nn;
        oracle_precon_apply(M, vi, pv);
        oracle_spmv(n, rowptr, col, val, pv, vn);      /* V_{i+1} = A M^-1 V_i */
        if (iflag_basis && newton) {                   /* calc_NewtonBasisVec */
          const double alpha = -B1[IDX(i - 1, i - 1, s1)];
          if (i != 1) {
            /* basis 1: + B1(i-1,i) V_{i-2} as :1519-1530; basis 2: minus it,
               the recurrence the Hessenberg update (:879-1051) assumes */
            const double beta = basis == 2 ? -B1[IDX(i - 2, i - 1, s1)] : B1[IDX(i - 2, i - 1, s1)];
            const double* vp = V + (size_t)(i - 2) * nn;
            for (int q = 0; q < n; ++q) vn[q] = fma(beta, vp[q], fma(alpha, vi[q], vn[q]));
          } else {
            for (int q = 0; q < n; ++q) vn[q] = fma(alpha, vi[q], vn[q]);
          }
        }
      }
      if (k == 1) {
        qr_cols(red, n, s1, V, nn, Q, nn, R, qrflag, qmat, W, R2, Rinv);
        /* calc_HessenbergMatrix, k = 1: H = R B1 R(1:s,1:s)^-1 */
        for (int j = 0; j < s; ++j)
          for (int i = 0; i < s; ++i) Rinv[IDX(i, j, s)] = R[IDX(i, j, s1)];
        oracle_trinv_upper(s, Rinv, s);
        oracle_gemm_nn(s1, s, s1, R, s1, B1, s1, RB, s1);
        oracle_gemm_nn(s1, s, s, RB, s1, Rinv, s, Hes, m + 1);
        hk_pre = Hes[IDX(s, s - 1, m + 1)];
      } else {
        /* block CGS (:2049-2112): R' = Q(:,1:c1)^T V, V(:,2:s+1) -= Q(:,1:c1) R'(:,2:s+1) */
        const int c1 = s * (k - 1) + 1;
        for (int j = 0; j < s1; ++j)
          for (int i = 0; i < c1; ++i)
            Rp[IDX(i, j, c1)] = oracle_dot(red, n, Q + (size_t)i * nn, V + (size_t)j * nn);
        for (int j = 1; j < s1; ++j) {
          double* vj = V + (size_t)j * nn;
          for (int q = 0; q < n; ++q) {
            double w = vj[q];
            for (int i = 0; i < c1; ++i) w = fma(-Rp[IDX(i, j, c1)], Q[(size_t)i * nn + q], w);
            vj[q] = w;
          }
        }
        qr_cols(red, n, s, V + nn, nn, Q + (size_t)c1 * nn, nn, R, qrflag, qmat, W, R2, Rinv);
        oracle_hessenberg_block(s, t, k, B1, Hes, R, Rp, &hk_pre);
      }
      if (k != t) memcpy(V, Q + (size_t)(s * k) * nn, sizeof(double) * nn);
      for (int i = s * (k - 1); i < s * k; ++i)
        for (int j = 0; j <= s * k; ++j) Hg[IDX(j, i, m + 1)] = Hes[IDX(j, i, m + 1)];
      for (int i = s * (k - 1) + 1; i <= s * k; ++i) {
        mm = i;
        givens_m(m, Hg, i, co, si, zeta);
        rnorm = fabs(zeta[i]);
        itr0 = i + (itr - 1) + 1;
        hist[itr0 - 1] = rnorm / bnorm;
        if (rnorm / bnorm < rtol) { iflag_exit = 1; break; }
        if (itr0 == itrmax) { iflag_exit = 1; break; }
      }
      if (iflag_exit) break;
      if (newton) {
        iflag_basis = 1;
        const int sk = s * k;
        for (int j = 0; j < sk; ++j)
          for (int i = 0; i < sk; ++i) Hs[IDX(i, j, sk)] = Hes[IDX(i, j, m + 1)];
        memset(er, 0, sizeof(double) * m);
        memset(ei, 0, sizeof(double) * m);
        (void)oracle_hqr_eigs(sk, Hs, sk, er, ei);
        oracle_leja(sk, er, ei);
      }
    }
    /* solution update (:2287-2430) */
    const int gs = (qrflag == 1 || qrflag == 2);
    if (gs || qmat || (t > 1 && s == 1)) {
      /* update_xVec with the orthonormal Q (:1136-1236) */
      for (int i = mm; i >= 1; --i) {
        double yi = zeta[i - 1];
        for (int j = mm; j >= i + 1; --j) yi = yi - Hg[IDX(i - 1, j - 1, m + 1)] * y[j - 1];
        const double d = Hg[IDX(i - 1, i - 1, m + 1)];
        y[i - 1] = (d == 0.0) ? yi / 1.0e-300 : yi / d;
      }
      for (int q = 0; q < n; ++q) {
        double acc = 0.0;
        for (int k = 0; k < mm; ++k) acc = fma(y[k], Q[(size_t)k * nn + q], acc);
        v0[q] = acc;
      }
    } else {
      /* update_xVec_cagmres (:1338-1478): y = H^-1 zeta; the last Vstep
       * entries times R(1:sv-1,1:sv-1)^-1 (sv = s+1 for t = 1, else s);
       * Qy = sum of the first mm - Vstep Q columns and Vstep V columns */
      for (int i = mm; i >= 1; --i) {
        double yi = zeta[i - 1];
        for (int j = mm; j >= i + 1; --j) yi = yi - Hg[IDX(i - 1, j - 1, m + 1)] * y[j - 1];
        y[i - 1] = yi / Hg[IDX(i - 1, i - 1, m + 1)];
      }
      const int sv = (t == 1) ? s + 1 : s;
      const int vstep = (sv - 1) - (m - mm);
      for (int i = 0; i < sv - 1; ++i) yt[i] = (i < vstep) ? y[mm - vstep + i] : 0.0;
      memcpy(Rinv, R, sizeof(double) * sv * sv);
      oracle_trinv_upper(sv, Rinv, sv);
      oracle_gemm_nn(sv - 1, 1, sv - 1, Rinv, sv, yt, sv - 1, y + (mm - vstep), sv - 1);
      const double* Vv = (t == 1) ? V : V + nn;
      for (int q = 0; q < n; ++q) {
        double acc = 0.0;
        for (int k = 0; k < mm - vstep; ++k) acc = fma(y[k], Q[(size_t)k * nn + q], acc);
        for (int k = 0; k < vstep; ++k) acc = fma(y[mm - vstep + k], Vv[(size_t)k * nn + q], acc);
        v0[q] = acc;
      }
    }
    oracle_precon_apply(M, v0, x0);
    for (int q = 0; q < n; ++q) x[q] = x0[q] + x[q];
    if (iflag_exit) break;
  }
  *niter = itr0 < itrmax ? itr0 : itrmax;            /* :2466 */
  *iret = (itrmax <= itr) ? 1 : 0;
  free(r); free(V); free(Q); free(W); free(pv); free(v0); free(x0); free(B1); free(R); free(R2); free(Rinv);
  free(Rp); free(RB); free(Hes); free(Hg); free(Hs); free(co); free(si); free(zeta); free(y); free(yt); free(er);
  free(ei);
  oracle_precon_free(M);
  return 0;
}

// dense.hpp -- host-side small dense algebra of CA-GMRES (the part the
// reference does on the host with LAPACK/BLAS and its own Fortran):
// DGEMM/DTRMM/DTRTI2/DPOTF2 loop orders with every multiply-add of an inner
// product as one fma, Hessenberg eigenvalues by the double-shift QR of
// DLAHQR (with DLANV2 for 2x2 blocks, conjugate pairs positive imaginary part
// first), and modified_leja_ordering (src/LejaOrdering.f90:46-277) with its
// exact control flow.  Column-major, 0-based.  Compiled with
// -ffp-contract=off like the rest of the library.
#pragma once

#include <cfloat>
#include <cmath>
#include <vector>

namespace parcel {
namespace dense {

#define IDX(i, j, ld) ((size_t)(i) + (size_t)(j) * (size_t)(ld))

/* C(m x n) = A(m x k) B(k x n)   (DGEMM 'N','N', alpha 1, beta 0) */
inline void gemm_nn(int m, int n, int k, const double* A, int lda, const double* B, int ldb, double* C,
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
inline void trmm_lunn(int n, const double* A, int lda, double* B, int ldb) {
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
inline void trinv_upper(int n, double* A, int lda) {
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
inline int chol_upper(int n, double* A, int lda) {
  for (int j = 0; j < n; ++j) {
    double d = 0.0;
    for (int k = 0; k < j; ++k) d = fma(A[IDX(k, j, lda)], A[IDX(k, j, lda)], d);
    double ajj = A[IDX(j, j, lda)] - d;
    if (ajj <= 0.0 || std::isnan(ajj)) {
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
inline double sign_of(double a, double b) { return std::signbit(b) ? -fabs(a) : fabs(a); }

inline double lapy2(double x, double y) {
  const double xa = fabs(x), ya = fabs(y);
  const double w = xa > ya ? xa : ya, z = xa < ya ? xa : ya;
  if (z == 0.0 || w > DBL_MAX) return w;
  const double q = z / w;
  return w * sqrt(1.0 + q * q);
}

/* Schur factorisation of a real 2x2 block in standard form (DLANV2):
 * returns the eigenvalue pair, complex pairs with the positive imaginary
 * part first. */
inline void lanv2(double a, double b, double c, double d, double* rt1r, double* rt1i, double* rt2r,
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
inline void larfg(int nx, double* alpha, double* x, double* tau) {
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
inline int hqr_eigs(int n, double* H, int ldh, double* wr, double* wi) {
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
inline void leja_order(int n, double* re, double* im) {
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
  std::vector<double> ptmp(n), thr(n + 1), thi(n + 1), tmp(n), pp(n);
  std::vector<int> outlist(n + 1);
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
}

/* House_local_Ronly_omp (src/TSQR.f90:836-927) on the m x s matrix A (ld
 * lda, overwritten): reflector v = [y1; A(i+1:m, i)], y1 = a_ii + sign(a_ii)
 * ||A(i:m, i)||, t1 = 2 / v^T v (1.0 - 300 for a zero v^T v, the reference's
 * guard), sequential fma sums in row order; R (s x s, ld s) gets the upper
 * triangle (diagonal -sign(a_ii) ||.||). */
inline void house_r(int m, int s, double* A, int lda, double* R) {
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

/* calc_HessenbergMatrix for block k >= 2 (src/pcagmres.f90:879-1051, with
 * calc_Rk_Bk_Rkinv :850 and calc_hke1eskRRinv :818).  Rp = R' = Q(:,1:c1)^T V
 * (c1 x (s+1), ld c1 = s(k-1)+1) from the block CGS, Rk the block's s x s R
 * factor (ld s), H the (st+1) x st Hessenberg (ld st+1); *hk_pre is read and
 * updated.  What the reference's local arrays hold: BlockR_ubar =
 * R'(1:s(k-1), 1:s+1), BlockR = R'(1:s(k-1), 1:s), Rtri0 = [R'(c1, 1:s+1);
 * 0 Rk].  New columns of H:
 *   H(1:s(k-1), new) = (-H_prev BlockR + BlockR_ubar B1) Rtri0(1:s,1:s)^-1
 *   H(new, new)      = Rtri0 B1 Rinv + (bk/rho~) z e_s^T - hk_pre e1 BlockR(end,:) Rinv
 *   H(sk+1, sk)      = rho / rho~ * bk                                      */
inline void hessenberg_block(int s, int t, int k, const double* B1, double* H, const double* Rk, const double* Rp,
                             double* hk_pre) {
  const int m = s * t, c0 = s * (k - 1), c1 = c0 + 1, s1 = s + 1, ldh = m + 1;
  std::vector<double> Rt0((size_t)s1 * s1, 0.0), Rinv((size_t)s * s, 0.0), Rub((size_t)c0 * s1, 0.0),
      BR((size_t)c0 * s, 0.0), BH((size_t)c0 * c0, 0.0), BHR((size_t)c0 * s, 0.0), RubB((size_t)c0 * s, 0.0),
      Hk((size_t)s * s, 0.0), RB((size_t)s * s, 0.0), B0((size_t)s * s, 0.0), Rs((size_t)s * s, 0.0),
      E((size_t)s * s, 0.0), H0((size_t)s * s, 0.0);
  for (int i = 0; i < s1; ++i) Rt0[IDX(0, i, s1)] = Rp[IDX(c0, i, c1)];
  for (int j = 0; j < s; ++j)
    for (int i = 0; i < s; ++i) Rt0[IDX(1 + i, j + 1, s1)] = Rk[IDX(i, j, s)];
  for (int j = 0; j < s1; ++j)
    for (int i = 0; i < c0; ++i) Rub[IDX(i, j, c0)] = Rp[IDX(i, j, c1)];
  for (int j = 0; j < s; ++j)
    for (int i = 0; i < c0; ++i) BR[IDX(i, j, c0)] = Rp[IDX(i, j, c1)];
  for (int j = 0; j < s; ++j)
    for (int i = 0; i < s; ++i) Rinv[IDX(i, j, s)] = Rt0[IDX(i, j, s1)];
  trinv_upper(s, Rinv.data(), s);
  for (int j = 0; j < c0; ++j)
    for (int i = 0; i < c0; ++i) BH[IDX(i, j, c0)] = H[IDX(i, j, ldh)];
  gemm_nn(c0, s, c0, BH.data(), c0, BR.data(), c0, BHR.data(), c0);
  gemm_nn(c0, s, s1, Rub.data(), c0, B1, s1, RubB.data(), c0);
  for (int i = 0; i < c0 * s; ++i) BHR[i] = -BHR[i] + RubB[i];
  gemm_nn(c0, s, s, BHR.data(), c0, Rinv.data(), s, RubB.data(), c0);
  for (int i = 0; i < s; ++i)
    for (int j = 0; j < c0; ++j) H[IDX(j, c0 + i, ldh)] = RubB[IDX(j, i, c0)];
  const double bk = B1[IDX(s, s - 1, s1)];
  const double rho_tilde = Rt0[IDX(s - 1, s - 1, s1)], rho = Rt0[IDX(s, s, s1)];
  const double hk = rho / rho_tilde * bk;
  H[IDX(s * k, s * k - 1, ldh)] = hk;
  for (int i = 0; i < s; ++i)
    for (int j = 0; j < s; ++j) {
      B0[IDX(j, i, s)] = B1[IDX(j, i, s1)];
      Rs[IDX(j, i, s)] = Rt0[IDX(j, i, s1)];
    }
  gemm_nn(s, s, s, Rs.data(), s, B0.data(), s, RB.data(), s);
  gemm_nn(s, s, s, RB.data(), s, Rinv.data(), s, Hk.data(), s);
  for (int i = 0; i < s; ++i) Hk[IDX(i, s - 1, s)] = fma(bk / rho_tilde, Rt0[IDX(i, s, s1)], Hk[IDX(i, s - 1, s)]);
  for (int i = 0; i < s; ++i) E[IDX(0, i, s)] = BR[IDX(c0 - 1, i, c0)] * *hk_pre;
  gemm_nn(s, s, s, E.data(), s, Rinv.data(), s, H0.data(), s);
  for (int i = 0; i < s * s; ++i) Hk[i] = Hk[i] - H0[i];
  for (int i = 0; i < s; ++i)
    for (int j = 0; j < s; ++j) H[IDX(c0 + j, c0 + i, ldh)] = Hk[IDX(j, i, s)];
  *hk_pre = hk;
}

/* Pseudo-inverse of a symmetric n x n matrix (moore_penrose,
 * src/Moore_Penrose.f90:3-67: A+ = V diag(1/sigma_i, sigma_i > 0) U^T by
 * DGESVD).  For a symmetric A the singular triplets are (|lambda_i|, v_i,
 * sign(lambda_i) v_i), so A+ = sum over lambda_i != 0 of v_i v_i^T / lambda_i;
 * the eigenpairs come from cyclic Jacobi rotations (rows p < q in order; an
 * off-diagonal entry at or below 2^-53 sqrt(|a_pp a_qq|) is set to zero; stop
 * after a sweep without rotations, at most 64 sweeps).  Plain expressions
 * without fma.  A is read from the upper triangle; P may not alias A. */
inline void sym_pinv(int n, const double* A, int lda, double* P, int ldp) {
  std::vector<double> a((size_t)n * n), v((size_t)n * n, 0.0);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) a[IDX(i, j, n)] = (i <= j) ? A[IDX(i, j, lda)] : A[IDX(j, i, lda)];
  for (int i = 0; i < n; ++i) v[IDX(i, i, n)] = 1.0;
  for (int sweep = 0; sweep < 64; ++sweep) {
    bool rotated = false;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a[IDX(p, q, n)];
        if (apq == 0.0) continue;
        const double app = a[IDX(p, p, n)], aqq = a[IDX(q, q, n)];
        if (fabs(apq) <= 0x1p-53 * sqrt(fabs(app * aqq))) {
          a[IDX(p, q, n)] = 0.0;
          a[IDX(q, p, n)] = 0.0;
          continue;
        }
        rotated = true;
        const double theta = (aqq - app) / (2.0 * apq);
        double t;
        if (fabs(theta) > 1e150) {
          t = 0.5 / theta;
        } else {
          t = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
          if (theta < 0.0) t = -t;
        }
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int r = 0; r < n; ++r) {
          if (r == p || r == q) continue;
          const double arp = a[IDX(r, p, n)], arq = a[IDX(r, q, n)];
          const double np_ = c * arp - s * arq, nq_ = s * arp + c * arq;
          a[IDX(r, p, n)] = np_; a[IDX(p, r, n)] = np_;
          a[IDX(r, q, n)] = nq_; a[IDX(q, r, n)] = nq_;
        }
        a[IDX(p, p, n)] = app - t * apq;
        a[IDX(q, q, n)] = aqq + t * apq;
        a[IDX(p, q, n)] = 0.0;
        a[IDX(q, p, n)] = 0.0;
        for (int r = 0; r < n; ++r) {
          const double vrp = v[IDX(r, p, n)], vrq = v[IDX(r, q, n)];
          v[IDX(r, p, n)] = c * vrp - s * vrq;
          v[IDX(r, q, n)] = s * vrp + c * vrq;
        }
      }
    if (!rotated) break;
  }
  for (int c = 0; c < n; ++c)
    for (int r = 0; r < n; ++r) {
      double acc = 0.0;
      for (int i = 0; i < n; ++i) {
        const double lam = a[IDX(i, i, n)];
        if (lam != 0.0) acc = acc + v[IDX(r, i, n)] * (v[IDX(c, i, n)] / lam);
      }
      P[IDX(r, c, ldp)] = acc;
    }
}

#undef IDX

}  // namespace dense
}  // namespace parcel

/*
 * cbcg.c -- CPU oracle for Chebyshev-basis CG (KSP_pCBCG) and its eigenvalue
 * bounds.  TEST INFRASTRUCTURE ONLY (see parcel_oracle.h).
 *
 * Reference:
 *   KSP_pCBCG                         src/pcbcg.f90:1182-2025
 *   make_chebyshev_basis              src/pcbcg.f90:49-111 (CBCG_axpy2/3 :12-47)
 *   CBCG_gemm_dot_local               src/pcbcg.f90:204-295, CUDA kernel
 *                                     cu_pcbcg.cu:14-108 (x += (sum_l Q_l alpha_l))
 *   CBCG_ALLreduce_CPU (symmetrise)   src/pcbcg.f90:297-350
 *   moore_penrose                     src/Moore_Penrose.f90:3-67
 *   getEigenValue                     src/pcbcg.f90:179-196
 *   power_method_symMatrix            src/power_method.f90:188-437
 *     (oracle_power_emax, oracle_power_eigen)
 *
 * Arithmetic (the HIP path's, bit for bit):
 *   SS_0 = fma(eta, AS_0, -(zeta*r)),
 *   SS_j = fma(2eta, AS_j, fma(-2zeta, SS_{j-1}, -SS_{j-2}))  (SS_{-1} = r),
 *   r    = fma(-alpha_l, AQ_l, r) for l in order,
 *   x    = x + (sum_l fma(Q_l, alpha_l, .) from 0),
 *   S_l  = fma(-Q_i, beta(i,l), S_l) for i in order (AS_l likewise with AQ_i),
 *   every dot in the tiled order (oracle_dot), alpha = P Qtr and
 *   beta = P QtAS by oracle_gemm_nn, P the symmetric pseudo-inverse below.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "parcel_oracle.h"

#define IDX(i, j, ld) ((size_t)(i) + (size_t)(j) * (size_t)(ld))

/* A+ of a symmetric matrix (upper triangle read) from cyclic Jacobi
 * eigenpairs: the singular triplets of a symmetric A are (|l_i|, v_i,
 * sign(l_i) v_i), so DGESVD's sum over sigma_i > 0 of v_i u_i^T / sigma_i is
 * the sum over l_i != 0 of v_i v_i^T / l_i.  Rotation rule: rows p < q in
 * order; skip zeros; zero an entry at or below 2^-53 sqrt(|a_pp a_qq|);
 * t = sign(theta) / (|theta| + sqrt(theta^2 + 1)), theta = (a_qq - a_pp) /
 * (2 a_pq) (t = 1/(2 theta) for |theta| > 1e150); stop after a sweep without
 * rotations, at most 64 sweeps. */
void oracle_sym_pinv(int n, const double* A, int lda, double* P, int ldp) {
  double* a = (double*)malloc(sizeof(double) * (size_t)n * n);
  double* v = (double*)calloc((size_t)n * n, sizeof(double));
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) a[IDX(i, j, n)] = (i <= j) ? A[IDX(i, j, lda)] : A[IDX(j, i, lda)];
  for (int i = 0; i < n; ++i) v[IDX(i, i, n)] = 1.0;
  for (int sweep = 0; sweep < 64; ++sweep) {
    int rotated = 0;
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
        rotated = 1;
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
          a[IDX(r, p, n)] = np_;
          a[IDX(p, r, n)] = np_;
          a[IDX(r, q, n)] = nq_;
          a[IDX(q, r, n)] = nq_;
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
  free(a);
  free(v);
}

/* the loop of power_method_symMatrix (src/power_method.f90:290-334): y = x0
 * (all ones when ||x0|| = 0); itrmax times: x' = (1/sqrt(<y,y>)) y,
 * y = A M^-1 x', Emax = <x',y>.  Leaves the last y in y and its <y,y> in
 * *vnorm_out; returns Emax (0 when itrmax < 1). */
static double power_loop(const oracle_precon* M, int n, const int* rowptr, const int* col, const double* val,
                         const double* x0, int itrmax, const oracle_red* red, double* y, double* vnorm_out) {
  double* xp = (double*)calloc((size_t)n, sizeof(double));
  double* z = (double*)malloc(sizeof(double) * (size_t)n);
  memcpy(y, x0, sizeof(double) * (size_t)n);
  double vnorm = oracle_dot(red, n, y, y), emax = 0.0;
  if (vnorm == 0.0) {
    for (int i = 0; i < n; ++i) y[i] = 1.0;
    vnorm = oracle_dot(red, n, y, y);
  }
  for (int it = 1; it <= itrmax; ++it) {
    const double t1 = 1.0 / sqrt(vnorm);
    for (int i = 0; i < n; ++i) xp[i] = t1 * y[i];
    oracle_precon_apply(M, xp, z);
    oracle_spmv(n, rowptr, col, val, z, y);
    vnorm = oracle_dot(red, n, y, y);
    emax = oracle_dot(red, n, xp, y);
  }
  free(xp);
  free(z);
  *vnorm_out = vnorm;
  return emax;
}

/* power_method_symMatrix's Emax (src/power_method.f90:188-437), the
 * eigenvalue bound of KSP_pCBCG; Emin is 0. */
double oracle_power_emax(const oracle_precon* M, int n, const int* rowptr, const int* col, const double* val,
                         const double* x0, int itrmax, const oracle_red* red) {
  double* y = (double*)malloc(sizeof(double) * (size_t)n);
  double vnorm;
  const double emax = power_loop(M, n, rowptr, col, val, x0, itrmax, red, y, &vnorm);
  free(y);
  return emax;
}

/* power_method_symMatrix in full (src/power_method.f90:188-437): the loop,
 * then (:336-398) x = (1/sqrt(<y,y>)) y, y = A x (no preconditioner),
 * kz = Emax x, z = y - kz, errnorm = sqrt|<z,z>| / sqrt|<kz,kz>|; vec
 * receives x (EigenVector, :415-420). */
void oracle_power_eigen(const oracle_precon* M, int n, const int* rowptr, const int* col, const double* val,
                        const double* x0, int itrmax, const oracle_red* red, double* emax_out, double* errnorm_out,
                        double* vec) {
  double* y = (double*)malloc(sizeof(double) * (size_t)n);
  double* kz = (double*)malloc(sizeof(double) * (size_t)n);
  double vnorm;
  const double emax = power_loop(M, n, rowptr, col, val, x0, itrmax, red, y, &vnorm);
  const double t1 = 1.0 / sqrt(vnorm);
  for (int i = 0; i < n; ++i) vec[i] = t1 * y[i];
  oracle_spmv(n, rowptr, col, val, vec, y);
  for (int i = 0; i < n; ++i) kz[i] = emax * vec[i];
  const double kx = oracle_dot(red, n, kz, kz);
  for (int i = 0; i < n; ++i) y[i] = y[i] - kz[i];
  const double en = oracle_dot(red, n, y, y);
  *emax_out = emax;
  *errnorm_out = sqrt(fabs(en)) / sqrt(fabs(kx));
  free(y);
  free(kz);
}

/* S_0 = M^-1 r, AS_0 = A S_0, SS_j and S_{j+1} = M^-1 SS_j (see header) */
static void cheb_basis(const oracle_precon* M, int n, const int* rowptr, const int* col, const double* val,
                       int k, double eta, double zeta, const double* r, double* S, double* AS, double* SS) {
  const size_t N = (size_t)n;
  for (int j = 0; j < k; ++j) {
    const double* in = (j == 0) ? r : SS + (size_t)(j - 1) * N;
    oracle_precon_apply(M, in, S + (size_t)j * N);
    oracle_spmv(n, rowptr, col, val, S + (size_t)j * N, AS + (size_t)j * N);
    if (j + 1 >= k) continue;
    const double* as = AS + (size_t)j * N;
    double* ss = SS + (size_t)j * N;
    if (j == 0) {
      for (int i = 0; i < n; ++i) ss[i] = fma(eta, as[i], -(zeta * r[i]));
    } else {
      const double* s1 = SS + (size_t)(j - 1) * N;
      const double* s2 = (j == 1) ? r : SS + (size_t)(j - 2) * N;
      const double e2 = eta * 2.0, z2 = zeta * 2.0;
      for (int i = 0; i < n; ++i) ss[i] = fma(e2, as[i], fma(-z2, s1[i], -s2[i]));
    }
  }
}

/* Q^T AQ (upper: entry (i,l), i <= l, = <Q_i, AQ_l>), Q^T r, <r,r> */
static void cb_dots(int n, int k, const double* Q, const double* AQ, const double* r, const oracle_red* red,
                    double* QtAQ, double* Qtr, double* rr) {
  const size_t N = (size_t)n;
  for (int l = 0; l < k; ++l) {
    for (int i = 0; i <= l; ++i) {
      const double d = oracle_dot(red, n, Q + (size_t)i * N, AQ + (size_t)l * N);
      QtAQ[IDX(i, l, k)] = d;
      QtAQ[IDX(l, i, k)] = d;
    }
    Qtr[l] = oracle_dot(red, n, Q + (size_t)l * N, r);
  }
  *rr = oracle_dot(red, n, r, r);
}

int oracle_cbcg_with(const oracle_precon* M, int n, const int* rowptr, const int* col, const double* val,
                     const double* b, double* x, int itrmax, double rtol, int k, double Emax, double Emin,
                     const oracle_red* red, double* hist, int* niter, int* iret) {
  if (k < 1 || k > 8) return -1;
  /* guards of KSP_pCBCG (src/pcbcg.f90:1500-1512) */
  if (Emax == Emin) {
    if (Emax < 0.0) Emax = 0.0;
    else Emin = 0.0;
  }
  if (Emax == 0.0 && Emin == 0.0) Emax = 1.0;
  const double eta = 2.0 / (Emax - Emin), zeta = (Emax + Emin) / (Emax - Emin);
  const size_t N = (size_t)n;
  double* r = (double*)malloc(sizeof(double) * N);
  double* buf = (double*)malloc(sizeof(double) * N * (size_t)(5 * k));
  double *Q = buf, *S = buf + (size_t)k * N, *AQ = buf + 2 * (size_t)k * N, *AS = buf + 3 * (size_t)k * N;
  double* SS = buf + 4 * (size_t)k * N;
  double QtAQ[64], P[64], Qtr[8], QtAS[64], beta[64], alpha[8], rr;
  oracle_spmv(n, rowptr, col, val, x, r);
  for (int i = 0; i < n; ++i) r[i] = b[i] - r[i];
  const double bnorm = sqrt(oracle_dot(red, n, b, b));
  cheb_basis(M, n, rowptr, col, val, k, eta, zeta, r, Q, AQ, SS);
  cb_dots(n, k, Q, AQ, r, red, QtAQ, Qtr, &rr);
  const int nout = itrmax / k;
  int itr;
  for (itr = 1; itr <= nout; ++itr) {
    oracle_sym_pinv(k, QtAQ, k, P, k);
    oracle_gemm_nn(k, 1, k, P, k, Qtr, k, alpha, k);
    for (int i = 0; i < n; ++i) {
      double v = r[i];
      for (int l = 0; l < k; ++l) v = fma(-alpha[l], AQ[(size_t)l * N + i], v);
      r[i] = v;
    }
    const double rnorm = sqrt(rr);
    hist[itr - 1] = rnorm / bnorm;
    if (rnorm / bnorm < rtol) break;
    cheb_basis(M, n, rowptr, col, val, k, eta, zeta, r, S, AS, SS);
    for (int l = 0; l < k; ++l)
      for (int i = 0; i < k; ++i) QtAS[IDX(i, l, k)] = oracle_dot(red, n, Q + (size_t)i * N, AS + (size_t)l * N);
    oracle_gemm_nn(k, k, k, P, k, QtAS, k, beta, k);
    for (int j = 0; j < n; ++j) {
      double xs = 0.0;
      for (int l = 0; l < k; ++l) xs = fma(Q[(size_t)l * N + j], alpha[l], xs);
      x[j] = x[j] + xs;
      for (int l = 0; l < k; ++l) {
        double s = S[(size_t)l * N + j], a = AS[(size_t)l * N + j];
        for (int i = 0; i < k; ++i) {
          s = fma(-Q[(size_t)i * N + j], beta[IDX(i, l, k)], s);
          a = fma(-AQ[(size_t)i * N + j], beta[IDX(i, l, k)], a);
        }
        S[(size_t)l * N + j] = s;
        AS[(size_t)l * N + j] = a;
      }
    }
    double* t = Q; Q = S; S = t;
    t = AQ; AQ = AS; AS = t;
    cb_dots(n, k, Q, AQ, r, red, QtAQ, Qtr, &rr);
  }
  *niter = itrmax < itr ? itrmax : itr;     /* src/pcbcg.f90:1975-1980 */
  *iret = (itrmax <= itr) ? 1 : 0;
  for (int i = (itr <= nout ? itr : nout); i < *niter; ++i) hist[i] = 0.0;
  free(r);
  free(buf);
  return 0;
}

int oracle_cbcg(int n, const int* rowptr, const int* col, const double* val, const double* b, double* x,
                int ipreflag, int ilu_method, int nblocks, int independ_nvec, int itrmax, double rtol, int k,
                int eigflag, int pm_itrmax, const oracle_red* red, double* hist, int* niter, int* iret,
                double* emax_out, double* emin_out) {
  int err = 0;
  oracle_precon* M = oracle_precon_create(ipreflag, ilu_method, nblocks, independ_nvec, n, rowptr, col, val, &err);
  if (!M) return err ? err : -1;
  double Emax = 0.0, Emin = 0.0;
  /* flag 2 (CA-Arnoldi): the reference runs the power method there, its
   * KSP_pCAARNOLDI being an `#if 1` call of power_method_symMatrix
   * (src/pcagmres.f90:2647-2651) */
  /* set_KSP_power_method_itrmax: at least one iteration (src/type.f90:1331-1336) */
  if (eigflag == 1 || eigflag == 2)
    Emax = oracle_power_emax(M, n, rowptr, col, val, x, pm_itrmax < 1 ? 1 : pm_itrmax, red);
  if (emax_out) *emax_out = Emax;
  if (emin_out) *emin_out = Emin;
  int rc = oracle_cbcg_with(M, n, rowptr, col, val, b, x, itrmax, rtol, k, Emax, Emin, red, hist, niter, iret);
  oracle_precon_free(M);
  return rc;
}

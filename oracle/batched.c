/*
 * batched.c -- oracle of the batched solver's factor preconditioners.
 *
 * TEST INFRASTRUCTURE ONLY (see parcel_oracle.h).  No reference counterpart:
 * the reference solves one system per call and has no incomplete-Cholesky or
 * learned factor.  The PCG iteration around these preconditioners is
 * oracle_pcg's (src/pcg.f90:224-892) through oracle_pcg_with.
 *
 * Parity status:
 *   oracle_ic0_5pt       pinned: IC(0) of a tridiagonal (nx = n) equals the
 *                        exact Cholesky factor (numpy), and L L^T matches A
 *                        on the pattern for random five-point SPD matrices.
 *   oracle_lscan_apply   pinned: z equals (L L^T)^-1 r from scipy's sparse
 *                        triangular solves to rounding, on grids with nx not
 *                        a multiple of 64 and on nx = 64 E.
 *   the lane order       restates the GPU's scan exactly (bit-exact tests);
 *                        the mathematics above pins what it computes.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_internal.h"

static int lanes_e(int nx) {
  int e = 1;
  while (64 * e < nx) e <<= 1;
  return e;
}

/* IC(0) on the five-point pattern in natural order (rows sorted s, w, c, e, n):
 *   Ls_i = a(i,i-nx) / Ld_{i-nx},  Lw_i = a(i,i-1) / Ld_{i-1},
 *   Ld_i = sqrt(fma(-Lw, Lw, fma(-Ls, Ls, a_ii))).
 * L holds Ls, Lw, Ld (n each).  Returns 0, or -2 at a non-positive pivot. */
int oracle_ic0_5pt(int n, int nx, const int* rowptr, const double* val, double* L) {
  double* Ls = L;
  double* Lw = L + n;
  double* Ld = L + 2 * (long)n;
  for (int i = 0; i < n; ++i) {
    int e = rowptr[i] - 1;
    double ls = 0.0, lw = 0.0;
    if (i >= nx) { ls = val[e] / Ld[i - nx]; ++e; }
    if (i % nx > 0) { lw = val[e] / Ld[i - 1]; ++e; }
    const double t = fma(-lw, lw, fma(-ls, ls, val[e]));
    if (!(t > 0.0)) return -2;
    Ls[i] = ls;
    Lw[i] = lw;
    Ld[i] = sqrt(t);
  }
  return 0;
}

float oracle_bf16_to_f32(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

/* round to nearest even; NaN -> quiet NaN */
uint16_t oracle_f32_to_bf16(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

/* Apply arrays: ls = Ls, inv = 1/Ld, af = -Lw*inv, ab = -Lw(x+1)*inv (0 at the
 * end of a line), computed in the target precision from the factor values
 * converted to it.  src: 0 double factor, 1 bf16 factor. */
void oracle_derive(int n, int nx, int src, const void* L, double* F64, float* F32) {
  const double* Ld64 = (const double*)L;
  const uint16_t* Lb = (const uint16_t*)L;
  for (int i = 0; i < n; ++i) {
    const int x = i % nx;
    double ls, lw, ld, lwn;
    if (src == 0) {
      ls = Ld64[i]; lw = Ld64[n + i]; ld = Ld64[2 * (long)n + i];
      lwn = x + 1 < nx ? Ld64[n + i + 1] : 0.0;
    } else {
      ls = oracle_bf16_to_f32(Lb[i]); lw = oracle_bf16_to_f32(Lb[n + i]); ld = oracle_bf16_to_f32(Lb[2 * (long)n + i]);
      lwn = x + 1 < nx ? oracle_bf16_to_f32(Lb[n + i + 1]) : 0.0;
    }
    if (F64) {
      const double inv = 1.0 / ld;
      F64[i] = ls;
      F64[n + i] = inv;
      F64[2 * (long)n + i] = -lw * inv;
      F64[3 * (long)n + i] = -lwn * inv;
    }
    if (F32) {
      const float inv = 1.0f / (float)ld;
      F32[i] = (float)ls;
      F32[n + i] = inv;
      F32[2 * (long)n + i] = -(float)lw * inv;
      F32[3 * (long)n + i] = -(float)lwn * inv;
    }
  }
}

/* The GPU's wavefront scan, lane by lane.  Lane l owns cells l*E .. l*E+E-1
 * of a line; cells past nx are padding with a = c = 0.
 * Forward line: c = fma(-ls, zprev, r) * inv; lane map A, C composed in cell
 * order (C = fma(a, C, c); A = a*A); inclusive scan over lanes with offsets
 * 1, 2, .., 32 (C = fma(A, C[l-d], C); A = A * A[l-d], from the values before
 * the step); the lane's entry value is C[l-1] (0 for lane 0); cells re-run
 * v = fma(a, v, c).  Backward: mirrored, cells and lanes from the right. */
#define ORACLE_LSCAN(NAME, T, FMA)                                                           \
  static void NAME##_fwd(int nx, int ny, int E, const T* F, long n, const T* r, T* w) {      \
    const int W = 64 * E;                                                                    \
    T* zp = (T*)calloc((size_t)W, sizeof(T));                                               \
    T* a = (T*)calloc((size_t)W, sizeof(T));                                                \
    T* c = (T*)calloc((size_t)W, sizeof(T));                                                \
    T A[64], C[64], A2[64], C2[64];                                                          \
    const T *ls = F, *inv = F + n, *af = F + 2 * n;                                         \
    for (int y = 0; y < ny; ++y) {                                                           \
      const long row = (long)y * nx;                                                         \
      for (int x = 0; x < W; ++x) {                                                          \
        if (x < nx) {                                                                        \
          a[x] = af[row + x];                                                                \
          c[x] = FMA(-ls[row + x], zp[x], r[row + x]) * inv[row + x];                       \
        } else { a[x] = 0; c[x] = 0; }                                                       \
      }                                                                                      \
      for (int l = 0; l < 64; ++l) {                                                         \
        T Al = 1, Cl = 0;                                                                    \
        for (int k = 0; k < E; ++k) { const int x = l * E + k; Cl = FMA(a[x], Cl, c[x]); Al = a[x] * Al; } \
        A[l] = Al; C[l] = Cl;                                                                \
      }                                                                                      \
      for (int d = 1; d < 64; d <<= 1) {                                                     \
        for (int l = 0; l < 64; ++l) {                                                       \
          if (l >= d) { C2[l] = FMA(A[l], C[l - d], C[l]); A2[l] = A[l] * A[l - d]; }       \
          else { C2[l] = C[l]; A2[l] = A[l]; }                                               \
        }                                                                                    \
        memcpy(A, A2, sizeof(A)); memcpy(C, C2, sizeof(C));                                  \
      }                                                                                      \
      for (int l = 0; l < 64; ++l) {                                                         \
        T v = l > 0 ? C[l - 1] : (T)0;                                                       \
        for (int k = 0; k < E; ++k) {                                                        \
          const int x = l * E + k;                                                           \
          v = FMA(a[x], v, c[x]);                                                            \
          zp[x] = v;                                                                         \
          if (x < nx) w[row + x] = v;                                                        \
        }                                                                                    \
      }                                                                                      \
    }                                                                                        \
    free(zp); free(a); free(c);                                                              \
  }                                                                                          \
  static void NAME##_bwd(int nx, int ny, int E, const T* F, long n, const T* w, T* z) {      \
    const int W = 64 * E;                                                                    \
    T* yn = (T*)calloc((size_t)W, sizeof(T));                                               \
    T* lsn = (T*)calloc((size_t)W, sizeof(T));                                              \
    T* a = (T*)calloc((size_t)W, sizeof(T));                                                \
    T* c = (T*)calloc((size_t)W, sizeof(T));                                                \
    T A[64], C[64], A2[64], C2[64];                                                          \
    const T *ls = F, *inv = F + n, *ab = F + 3 * n;                                         \
    for (int y = ny - 1; y >= 0; --y) {                                                      \
      const long row = (long)y * nx;                                                         \
      for (int x = 0; x < W; ++x) {                                                          \
        if (x < nx) {                                                                        \
          a[x] = ab[row + x];                                                                \
          c[x] = FMA(-lsn[x], yn[x], w[row + x]) * inv[row + x];                            \
          lsn[x] = ls[row + x];                                                              \
        } else { a[x] = 0; c[x] = 0; lsn[x] = 0; }                                           \
      }                                                                                      \
      for (int l = 0; l < 64; ++l) {                                                         \
        T Al = 1, Cl = 0;                                                                    \
        for (int k = E - 1; k >= 0; --k) { const int x = l * E + k; Cl = FMA(a[x], Cl, c[x]); Al = a[x] * Al; } \
        A[l] = Al; C[l] = Cl;                                                                \
      }                                                                                      \
      for (int d = 1; d < 64; d <<= 1) {                                                     \
        for (int l = 0; l < 64; ++l) {                                                       \
          if (l + d < 64) { C2[l] = FMA(A[l], C[l + d], C[l]); A2[l] = A[l] * A[l + d]; }   \
          else { C2[l] = C[l]; A2[l] = A[l]; }                                               \
        }                                                                                    \
        memcpy(A, A2, sizeof(A)); memcpy(C, C2, sizeof(C));                                  \
      }                                                                                      \
      for (int l = 0; l < 64; ++l) {                                                         \
        T v = l < 63 ? C[l + 1] : (T)0;                                                      \
        for (int k = E - 1; k >= 0; --k) {                                                   \
          const int x = l * E + k;                                                           \
          v = FMA(a[x], v, c[x]);                                                            \
          yn[x] = v;                                                                         \
          if (x < nx) z[row + x] = v;                                                        \
        }                                                                                    \
      }                                                                                      \
    }                                                                                        \
    free(yn); free(lsn); free(a); free(c);                                                   \
  }

ORACLE_LSCAN(lscan64, double, fma)
ORACLE_LSCAN(lscan32, float, fmaf)

void oracle_lscan_apply(int nx, int ny, const double* F, const double* r, double* z) {
  const long n = (long)nx * ny;
  double* w = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  const int E = lanes_e(nx);
  lscan64_fwd(nx, ny, E, F, n, r, w);
  lscan64_bwd(nx, ny, E, F, n, w, z);
  free(w);
}

void oracle_lscan_apply32(int nx, int ny, const float* F, const float* r, float* z) {
  const long n = (long)nx * ny;
  float* w = (float*)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
  const int E = lanes_e(nx);
  lscan32_fwd(nx, ny, E, F, n, r, w);
  lscan32_bwd(nx, ny, E, F, n, w, z);
  free(w);
}

oracle_precon* oracle_precon_create_factor(int n, int nx, const double* F64, const float* F32) {
  oracle_precon* p = (oracle_precon*)calloc(1, sizeof(oracle_precon));
  p->ipreflag = ORACLE_PRECON_FACTOR;
  p->n = n;
  p->nx = nx;
  p->ny = n / nx;
  p->nblocks = 1;
  p->F64 = (double*)malloc(sizeof(double) * 4 * (size_t)n);
  p->F32 = (float*)malloc(sizeof(float) * 4 * (size_t)n);
  memcpy(p->F64, F64, sizeof(double) * 4 * (size_t)n);
  memcpy(p->F32, F32, sizeof(float) * 4 * (size_t)n);
  return p;
}

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

/* Block-Jacobi IC(0): the grid's ny lines cut into nstrip strips (strip k =
 * lines [ny k / nstrip, ny (k+1) / nstrip)), each strip's diagonal block
 * factorised on its own: the first line of a strip drops its couplings to
 * the strip below (Ls = 0), as the block-Jacobi set-up keeps in-block
 * entries only (src/matrix.f90:8249-8392).  For an SPD matrix this is the
 * block-Jacobi ILU(0) preconditioner in exact arithmetic (ILU(0) = L D L^T,
 * IC(0) = (L D^1/2)(L D^1/2)^T on the same pattern). */
int oracle_ic0_5pt_blocks(int n, int nx, int nstrip, const int* rowptr, const double* val, double* L) {
  double* Ls = L;
  double* Lw = L + n;
  double* Ld = L + 2 * (long)n;
  const int ny = n / nx;
  int strip = 0, first = 0;  /* current strip and its first line */
  for (int i = 0; i < n; ++i) {
    const int line = i / nx;
    while (strip + 1 < nstrip && line >= (int)((long)ny * (strip + 1) / nstrip)) {
      ++strip;
      first = (int)((long)ny * strip / nstrip);
    }
    int e = rowptr[i] - 1;
    double ls = 0.0, lw = 0.0;
    if (i >= nx) {
      if (line > first) ls = val[e] / Ld[i - nx];
      ++e;
    }
    if (i % nx > 0) { lw = val[e] / Ld[i - 1]; ++e; }
    const double t = fma(-lw, lw, fma(-ls, ls, val[e]));
    if (!(t > 0.0)) return -2;
    Ls[i] = ls;
    Lw[i] = lw;
    Ld[i] = sqrt(t);
  }
  return 0;
}

/* Block-Jacobi IC(0) of a general CRS matrix whose blocks are the strips of
 * whole grid lines of oracle_ic0_5pt_blocks: a row's couplings are found by
 * column (i - nx on the line below, in the same strip; i - 1 on the same
 * line), every other off-diagonal entry is outside the factor's pattern
 * (couplings between planes of a 3D grid whose strips are single planes).
 * The pivots follow oracle_ic0_5pt.  This is the factor of the FP32 inner
 * apply of the mixed solver for grid-line blocks (block ILU(0) of a
 * symmetric matrix, src/matrix.f90:8249-8392, in exact arithmetic). */
static double crs_entry(const int* rowptr, const int* col, const double* val, int i, int j) {
  for (int e = rowptr[i] - 1; e < rowptr[i + 1] - 1; ++e)
    if (col[e] - 1 == j) return val[e];
  return 0.0;
}

int oracle_ic0_grid_blocks(int n, int nx, int nstrip, const int* rowptr, const int* col, const double* val,
                           double* L) {
  double* Ls = L;
  double* Lw = L + n;
  double* Ld = L + 2 * (long)n;
  const int ny = n / nx;
  int strip = 0, first = 0;
  for (int i = 0; i < n; ++i) {
    const int line = i / nx;
    while (strip + 1 < nstrip && line >= (int)((long)ny * (strip + 1) / nstrip)) {
      ++strip;
      first = (int)((long)ny * strip / nstrip);
    }
    double ls = 0.0, lw = 0.0;
    if (line > first) ls = crs_entry(rowptr, col, val, i, i - nx) / Ld[i - nx];
    if (i % nx > 0) lw = crs_entry(rowptr, col, val, i, i - 1) / Ld[i - 1];
    const double t = fma(-lw, lw, fma(-ls, ls, crs_entry(rowptr, col, val, i, i)));
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

/* Wavefronts per line and cells per lane of the GPU's factor apply for lines
 * of nx cells: E1 = the smallest power of two with 64 E1 >= nx (the cells per
 * lane of a single wavefront); lines of E1 >= 8 are split over 4 wavefronts
 * (E = E1 / 4), E1 = 4 over 2 (E = 2), shorter lines take one wavefront. */
static void lscan_layout(int nx, int* E, int* NW) {
  const int e1 = lanes_e(nx);
  *NW = e1 >= 8 ? 4 : (e1 >= 4 ? 2 : 1);
  *E = e1 / *NW;
}

/* The GPU's wavefront scan, lane by lane.  A line is padded to WT = 64 E NW
 * cells in "slot order" u = 0 .. WT-1: forward x = u, backward
 * x = WT - 1 - u, so the right-to-left recurrence also runs in slot order.
 * Wavefront w owns slots [64 E w, 64 E (w+1)), its lane l the E slots
 * 64 E w + l E + k (k = 0 .. E-1).  Cells past nx are padding with
 * a = c = 0.
 * A cell's map: v = a v_prev + c, c = fma(-p, carry, q) * inv (forward:
 * p = ls, carry = w of the same cell on line y-1, q = r; backward: p = ls of
 * line y+1, carry = z of the same cell on line y+1, q = w).  In each
 * wavefront the lane's map (A, C) is composed in k order (C = fma(a, C, c);
 * A = a*A), then scanned in lane order: shifts 1, 2, 4, 8 inside rows of 16
 * lanes, then lane 16r-1 into rows 1 and 3, then lane 31 into rows 2 and 3;
 * a lane that receives nothing in a step keeps its map; the others combine
 * C = fma(A, Cp, C), A = A*Ap from the values before the step.  The value entering lane l: in wavefront 0, C[l-1] (0 for lane 0);
 * in wavefront w > 0, fma(A[l-1], vin, C[l-1]) with (A, C)[-1] = (1, 0) and
 * vin = the value of wavefront w-1's last slot on the same line.  The lane
 * then re-applies its maps v = fma(a, v, c) in k order. */
#define ORACLE_LSCAN(NAME, T, FMA)                                                              \
  static void NAME##_scan(T* A, T* C) {                                                         \
    T A2[64], C2[64];                                                                           \
    for (int step = 0; step < 6; ++step) {                                                      \
      for (int l = 0; l < 64; ++l) {                                                            \
        int src = -1;                                                                           \
        if (step < 4) {                                                                         \
          const int s = 1 << step;                                                              \
          if (l % 16 >= s) src = l - s;                                                         \
        } else if (step == 4) {                                                                 \
          if ((l / 16) % 2 == 1) src = 16 * (l / 16) - 1;                                       \
        } else {                                                                                \
          if (l / 16 >= 2) src = 31;                                                            \
        }                                                                                       \
        if (src >= 0) {                                                                         \
          C2[l] = FMA(A[l], C[src], C[l]);                                                      \
          A2[l] = A[l] * A[src];                                                                \
        } else {                                                                                \
          C2[l] = C[l];                                                                         \
          A2[l] = A[l];                                                                         \
        }                                                                                       \
      }                                                                                         \
      memcpy(A, A2, sizeof(A2));                                                                \
      memcpy(C, C2, sizeof(C2));                                                                \
    }                                                                                           \
  }                                                                                             \
  static void NAME##_sweep(int nx, int ny, int bwd, const T* a_, const T* ls, const T* q,       \
                           const T* inv, T* out) {                                              \
    int E, NW;                                                                                  \
    lscan_layout(nx, &E, &NW);                                                                  \
    const int W = 64 * E, WT = NW * W;                                                          \
    T* carry = (T*)calloc((size_t)WT, sizeof(T));                                               \
    T* pn = (T*)calloc((size_t)WT, sizeof(T));                                                  \
    T* a = (T*)calloc((size_t)WT, sizeof(T));                                                   \
    T* c = (T*)calloc((size_t)WT, sizeof(T));                                                   \
    T A[64], C[64];                                                                             \
    for (int t = 0; t < ny; ++t) {                                                              \
      const int y = bwd ? ny - 1 - t : t;                                                       \
      const long row = (long)y * nx;                                                            \
      for (int u = 0; u < WT; ++u) {                                                            \
        const int x = bwd ? WT - 1 - u : u;                                                     \
        if (x < nx) {                                                                           \
          a[u] = a_[row + x];                                                                   \
          const T p = bwd ? pn[u] : ls[row + x];                                                \
          c[u] = FMA(-p, carry[u], q[row + x]) * inv[row + x];                                  \
          if (bwd) pn[u] = ls[row + x];                                                         \
        } else {                                                                                \
          a[u] = 0; c[u] = 0;                                                                   \
          if (bwd) pn[u] = 0;                                                                   \
        }                                                                                       \
      }                                                                                         \
      T vin = 0;                                                                                \
      for (int w = 0; w < NW; ++w) {                                                            \
        const int u0 = w * W;                                                                   \
        for (int l = 0; l < 64; ++l) {                                                          \
          T Al = 1, Cl = 0;                                                                     \
          for (int k = 0; k < E; ++k) {                                                         \
            const int u = u0 + l * E + k;                                                       \
            Cl = FMA(a[u], Cl, c[u]);                                                           \
            Al = a[u] * Al;                                                                     \
          }                                                                                     \
          A[l] = Al; C[l] = Cl;                                                                 \
        }                                                                                       \
        NAME##_scan(A, C);                                                                      \
        T vlast = 0;                                                                            \
        for (int l = 0; l < 64; ++l) {                                                          \
          T v;                                                                                  \
          if (w == 0)                                                                           \
            v = l > 0 ? C[l - 1] : (T)0;                                                        \
          else                                                                                  \
            v = FMA(l > 0 ? A[l - 1] : (T)1, vin, l > 0 ? C[l - 1] : (T)0);                     \
          for (int k = 0; k < E; ++k) {                                                         \
            const int u = u0 + l * E + k;                                                       \
            const int x = bwd ? WT - 1 - u : u;                                                 \
            v = FMA(a[u], v, c[u]);                                                             \
            carry[u] = v;                                                                       \
            if (x < nx) out[row + x] = v;                                                       \
          }                                                                                     \
          vlast = v;                                                                            \
        }                                                                                       \
        vin = vlast;                                                                            \
      }                                                                                         \
    }                                                                                           \
    free(carry); free(pn); free(a); free(c);                                                    \
  }

ORACLE_LSCAN(lscan64, double, fma)
ORACLE_LSCAN(lscan32, float, fmaf)

/* Strips: lines [ny k / nstrip, ny (k+1) / nstrip) solved independently. */
void oracle_lscan_apply_strips(int nx, int ny, int nstrip, const double* F, const double* r, double* z) {
  const long n = (long)nx * ny;
  double* w = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
  for (int k = 0; k < nstrip; ++k) {
    const int y0 = (int)((long)ny * k / nstrip), y1 = (int)((long)ny * (k + 1) / nstrip);
    if (y1 <= y0) continue;
    const long o = (long)y0 * nx;
    lscan64_sweep(nx, y1 - y0, 0, F + 2 * n + o, F + o, r + o, F + n + o, w + o);
    lscan64_sweep(nx, y1 - y0, 1, F + 3 * n + o, F + o, w + o, F + n + o, z + o);
  }
  free(w);
}

void oracle_lscan_apply32_strips(int nx, int ny, int nstrip, const float* F, const float* r, float* z) {
  const long n = (long)nx * ny;
  float* w = (float*)calloc((size_t)(n > 0 ? n : 1), sizeof(float));
  for (int k = 0; k < nstrip; ++k) {
    const int y0 = (int)((long)ny * k / nstrip), y1 = (int)((long)ny * (k + 1) / nstrip);
    if (y1 <= y0) continue;
    const long o = (long)y0 * nx;
    lscan32_sweep(nx, y1 - y0, 0, F + 2 * n + o, F + o, r + o, F + n + o, w + o);
    lscan32_sweep(nx, y1 - y0, 1, F + 3 * n + o, F + o, w + o, F + n + o, z + o);
  }
  free(w);
}

void oracle_lscan_apply(int nx, int ny, const double* F, const double* r, double* z) {
  oracle_lscan_apply_strips(nx, ny, 1, F, r, z);
}

void oracle_lscan_apply32(int nx, int ny, const float* F, const float* r, float* z) {
  oracle_lscan_apply32_strips(nx, ny, 1, F, r, z);
}

/* FP32 apply for double vectors: r rounded to float, z widened back */
void oracle_lscan_apply32in64(int nx, int ny, int nstrip, const float* F, const double* r, double* z) {
  const long n = (long)nx * ny;
  float* r32 = (float*)calloc((size_t)(n > 0 ? n : 1), sizeof(float));
  float* z32 = (float*)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
  for (long i = 0; i < n; ++i) r32[i] = (float)r[i];
  oracle_lscan_apply32_strips(nx, ny, nstrip, F, r32, z32);
  for (long i = 0; i < n; ++i) z[i] = (double)z32[i];
  free(r32);
  free(z32);
}

oracle_precon* oracle_precon_create_factor(int n, int nx, const double* F64, const float* F32) {
  oracle_precon* p = (oracle_precon*)calloc(1, sizeof(oracle_precon));
  p->ipreflag = ORACLE_PRECON_FACTOR;
  p->n = n;
  p->nx = nx;
  p->ny = n / nx;
  p->nblocks = 1;
  p->nstrip = 1;
  p->F64 = (double*)malloc(sizeof(double) * 4 * (size_t)n);
  p->F32 = (float*)malloc(sizeof(float) * 4 * (size_t)n);
  memcpy(p->F64, F64, sizeof(double) * 4 * (size_t)n);
  memcpy(p->F32, F32, sizeof(float) * 4 * (size_t)n);
  return p;
}

void oracle_precon_factor_apply32(oracle_precon* p, int on) { p->apply32in64 = on; }
void oracle_precon_factor_strips(oracle_precon* p, int nstrip) { p->nstrip = nstrip < 1 ? 1 : nstrip; }

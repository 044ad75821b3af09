// cagmres.hip -- CA-GMRES(s,t), KSP_pCAGMRES (src/pcagmres.f90:1557-2517),
// with right preconditioning, monomial or Newton basis (calc_NewtonBasis /
// calc_NewtonBasisVec, :586-611,1480-1552) and the QR variants of
// QRDecomposition_ (src/QRDecomposition.f90:1331-1551): MGS / CGS (GS_QR_),
// TSQR (TSQR_Ronly_), CholeskyQR (CholeskyQR_Ronly_) and CholeskyQR2, Q = V R^-1 where a later
// block needs it (calcQ_VRinv_, :603-671).
//
// Device work per block of a restart cycle: s preconditioner applies and
// SpMVs that build the basis V (n x (s+1), the Newton recurrence fused into
// the SpMV's store); for blocks k >= 2 the block classical Gram-Schmidt
// against the s(k-1)+1 earlier Q columns (one pass computing every entry of
// Q^T V, one pass subtracting Q R'); one pass computing every entry of the
// Gram matrix V^T V at once (CholeskyQR; CholeskyQR2 adds one pass over
// V R^-1 formed on the fly, never stored); and after the cycle the solution
// update.  The small dense work -- Cholesky, the Hessenberg matrix
// (calc_HessenbergMatrix, :879-1051), Givens rotations, Ritz values and
// Leja ordering -- runs on the host as in the reference, after one
// device-to-host copy of the Gram matrix (and of R') per block.
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "dense.hpp"
#include "kernels.hpp"

namespace parcel {

namespace {

constexpr int kMaxCols = 16;  // s + 1 <= 16
constexpr int kMaxM = 128;    // s * t

// st[0] = bnorm, st[1] = rnorm, st[2] = -1/rnorm
__global__ void __launch_bounds__(kNT) k_ca_restart(const double* part_bb, const double* part_rr, int nb,
                                                    double* st, int first) {
  __shared__ double sh[kWaves];
  const double rr = reduce_part(part_rr, nb, sh);
  double bb = 0.0;
  if (first) bb = reduce_part(part_bb, nb, sh);
  if (threadIdx.x == 0) {
    if (first) st[0] = sqrt(bb);
    const double rn = sqrt(rr);
    st[1] = rn;
    st[2] = -1.0 / rn;
  }
}

__global__ void __launch_bounds__(kNT) k_ca_scale(int n, const double* __restrict__ r, double* __restrict__ v,
                                                  const double* st) {
  const double a = st[2];
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    v[i] = a * r[i];
}

// V_{i+1} = A pv, then the Newton recurrence of calc_NewtonBasisVec:
// mode 0: as is (monomial); 1: fma(alpha, V_i, .); 2: fma(beta, V_{i-1}, fma(alpha, V_i, .))
__global__ void __launch_bounds__(kNT) k_ca_basis(int n, SellArgs A, const double* __restrict__ pv,
                                                  double* __restrict__ vn, const double* __restrict__ vi,
                                                  const double* __restrict__ vp, double alpha, double beta,
                                                  int mode) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double t = sell_row(A, i, [&](int c) { return pv[c]; });
    if (mode >= 1) {
      t = fma(alpha, vi[i], t);
      if (mode == 2) t = fma(beta, vp[i], t);
    }
    vn[i] = t;
  }
}

// Partials of every upper-triangle Gram entry G(k,j) = <W_k, W_j>, k <= j,
// of the NC columns W = V (rinv == nullptr) or W = V R^-1 formed on the fly
// (column i of W sums fma(V_k, Rinv(k,i)) over k = 0..i, calcQ_VRinv_).
// Entry p = j(j+1)/2 + k goes to part[p * kNBMax + workgroup], reduced in
// the same tree as every other dot product.
template <int NC>
__global__ void __launch_bounds__(kNT) k_ca_gram(int n, const double* V, long ldv, const double* rinv,
                                                 double* part) {
  constexpr int NP = NC * (NC + 1) / 2;
  __shared__ double sh[kWaves * NP];
  __shared__ double sri[NC * NC];
  const bool use_r = rinv != nullptr;
  if (use_r)
    for (int q = threadIdx.x; q < NC * NC; q += blockDim.x) sri[q] = rinv[q];
  __syncthreads();
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  double acc[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) acc[p] = 0.0;
  for (long t = blockIdx.x; t < ntiles; t += nb) {
    const long i = t * kNT + threadIdx.x;
    if (i < n) {
      double w[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) w[c] = V[(long)c * ldv + i];
      if (use_r) {
        // column c of V R^-1: sum over k = 0..c, from the last column down so
        // that w[k] for k < c is still the V value when it is read
#pragma unroll
        for (int c = NC - 1; c >= 0; --c) {
          double a = 0.0;
#pragma unroll
          for (int k = 0; k <= c; ++k) a = fma(w[k], sri[k + c * NC], a);
          w[c] = a;
        }
      }
      int p = 0;
#pragma unroll
      for (int j = 0; j < NC; ++j)
#pragma unroll
        for (int k = 0; k <= j; ++k, ++p) acc[p] = fma(w[k], w[j], acc[p]);
    }
  }
  // wave butterflies, then the four waves left to right (block_sum's tree)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const double v = wave_sum(acc[p]);
    if (lane == 0) sh[wave * NP + p] = v;
  }
  __syncthreads();
  for (int p = threadIdx.x; p < NP; p += blockDim.x) {
    double v = sh[p];
    for (int w = 1; w < kWaves; ++w) v += sh[w * NP + p];
    part[(long)p * kNBMax + blockIdx.x] = v;
  }
}

// G(k,j) for the upper triangle; one workgroup per entry
__global__ void __launch_bounds__(kNT) k_ca_gram_reduce(const double* part, int nb, int nc, double* G) {
  __shared__ double sh[kWaves];
  const int p = blockIdx.x;
  const double v = reduce_part(part + (long)p * kNBMax, nb, sh);
  if (threadIdx.x == 0) {
    int j = 0;
    while ((j + 1) * (j + 2) / 2 <= p) ++j;
    const int k = p - j * (j + 1) / 2;
    G[k + j * nc] = v;
  }
}

// ---- GS_QR_ (MGS / CGS) with device-resident R -------------------------
__global__ void __launch_bounds__(kNT) k_ca_dot(int n, const double* __restrict__ x, const double* __restrict__ y,
                                                double* part) {
  __shared__ double sh[kWaves];
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  double acc = 0.0;
  for (long t = blockIdx.x; t < ntiles; t += nb) {
    const long i = t * kNT + threadIdx.x;
    if (i < n) acc = fma(x[i], y[i], acc);
  }
  const double v = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// MGS step: R(k,i) = reduce(part_in); v_i = fma(-R(k,i), q_k, v_i);
// partials of <v_i, q_next> (or <v_i, v_i> when q_next is null)
__global__ void __launch_bounds__(kNT) k_ca_mgs(int n, double* __restrict__ vi, const double* __restrict__ qk,
                                                const double* __restrict__ qnext, const double* part_in,
                                                double* part_out, int nb_prod, double* R, int ldr, int k, int i) {
  __shared__ double sh[kWaves];
  const double r = reduce_part(part_in, nb_prod, sh);
  if (blockIdx.x == 0 && threadIdx.x == 0) R[k + i * ldr] = r;
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  double acc = 0.0;
  for (long t = blockIdx.x; t < ntiles; t += nb) {
    const long j = t * kNT + threadIdx.x;
    if (j < n) {
      const double v = fma(-r, qk[j], vi[j]);
      vi[j] = v;
      acc = fma(v, qnext ? qnext[j] : v, acc);
    }
  }
  const double v = block_sum(acc, sh);
  if (threadIdx.x == 0) part_out[blockIdx.x] = v;
}

// CGS: partials of <q_k, v_i> for k = 0..i-1 (i <= 15), slot k
template <int CNT>
__global__ void __launch_bounds__(kNT) k_ca_cgs_dots(int n, const double* Q, long ldq, const double* __restrict__ vi,
                                                     double* part) {
  __shared__ double sh[kWaves];
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  double acc[CNT];
#pragma unroll
  for (int d = 0; d < CNT; ++d) acc[d] = 0.0;
  for (long t = blockIdx.x; t < ntiles; t += nb) {
    const long j = t * kNT + threadIdx.x;
    if (j < n) {
      const double v = vi[j];
#pragma unroll
      for (int d = 0; d < CNT; ++d) acc[d] = fma(Q[(long)d * ldq + j], v, acc[d]);
    }
  }
#pragma unroll
  for (int d = 0; d < CNT; ++d) {
    const double v = block_sum(acc[d], sh);
    if (threadIdx.x == 0) part[(long)d * kNBMax + blockIdx.x] = v;
  }
}

__global__ void __launch_bounds__(kNT) k_ca_cgs_reduce(const double* part, int nb, double* R, int ldr, int i) {
  __shared__ double sh[kWaves];
  const int k = blockIdx.x;
  const double v = reduce_part(part + (long)k * kNBMax, nb, sh);
  if (threadIdx.x == 0) R[k + i * ldr] = v;
}

// CGS update v_i = fma(-R(k,i), q_k, v_i), k = 0..i-1 in order; partials of <v_i,v_i>
__global__ void __launch_bounds__(kNT) k_ca_cgs_update(int n, const double* Q, long ldq, double* __restrict__ vi,
                                                       const double* R, int ldr, int i, double* part) {
  __shared__ double sh[kWaves];
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  double acc = 0.0;
  for (long t = blockIdx.x; t < ntiles; t += nb) {
    const long j = t * kNT + threadIdx.x;
    if (j < n) {
      double v = vi[j];
      for (int k = 0; k < i; ++k) v = fma(-R[k + i * ldr], Q[(long)k * ldq + j], v);
      vi[j] = v;
      acc = fma(v, v, acc);
    }
  }
  const double v = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// R(i,i) = sqrt(<v_i,v_i>); q_i = (1/R(i,i)) v_i
__global__ void __launch_bounds__(kNT) k_ca_qr_norm(int n, const double* __restrict__ vi, double* __restrict__ qi,
                                                    const double* part, int nb_prod, double* R, int ldr, int i) {
  __shared__ double sh[kWaves];
  const double rii = sqrt(reduce_part(part, nb_prod, sh));
  if (blockIdx.x == 0 && threadIdx.x == 0) R[i + i * ldr] = rii;
  const double sc = 1.0 / rii;
  for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (long)gridDim.x * blockDim.x)
    qi[j] = sc * vi[j];
}

// ---- block CGS between blocks (t > 1) ------------------------------------
// Partials of <Q_i, V_j> for Q columns i in [4*blockIdx.y, 4*blockIdx.y + 4)
// (below c1) and the NV columns of V; slot i + c1*j.
template <int NV>
__global__ void __launch_bounds__(kNT) k_ca_bdots(int n, const double* Q, long ldq, int c1, const double* V,
                                                  long ldv, double* part) {
  constexpr int NP = 4 * NV;
  __shared__ double sh[kWaves * NP];
  const int i0 = blockIdx.y * 4;
  const int ni = (c1 - i0) < 4 ? (c1 - i0) : 4;
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  double acc[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) acc[p] = 0.0;
  for (long t = blockIdx.x; t < ntiles; t += nb) {
    const long q = t * kNT + threadIdx.x;
    if (q < n) {
      double v[NV];
#pragma unroll
      for (int j = 0; j < NV; ++j) v[j] = V[(long)j * ldv + q];
#pragma unroll
      for (int ii = 0; ii < 4; ++ii) {
        if (ii < ni) {
          const double qv = Q[(long)(i0 + ii) * ldq + q];
#pragma unroll
          for (int j = 0; j < NV; ++j) acc[ii * NV + j] = fma(qv, v[j], acc[ii * NV + j]);
        }
      }
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const double v = wave_sum(acc[p]);
    if (lane == 0) sh[wave * NP + p] = v;
  }
  __syncthreads();
  for (int p = threadIdx.x; p < NP; p += blockDim.x) {
    const int ii = p / NV, j = p % NV;
    if (ii >= ni) continue;
    double v = sh[p];
    for (int w = 1; w < kWaves; ++w) v += sh[w * NP + p];
    part[(long)(i0 + ii + c1 * j) * kNBMax + blockIdx.x] = v;
  }
}

// out[slot] = reduction of slot's partials (one workgroup per slot)
__global__ void __launch_bounds__(kNT) k_ca_reduce_slots(const double* part, int nb, double* out) {
  __shared__ double sh[kWaves];
  const double v = reduce_part(part + (long)blockIdx.x * kNBMax, nb, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = v;
}

// V_j -= sum_i R'(i,j) Q_i for j = 1..nv (columns of V), i = 0..c1-1 in
// order, as fma(-R'(i,j), Q_i, .) (calc_gemm with the negated R',
// vec.f90:403-440); R' is c1 x (nv+1), ld c1, column 0 unused
__global__ void __launch_bounds__(kNT) k_ca_bupd(int n, const double* Q, long ldq, int c1, double* V, long ldv,
                                                 int nv, const double* Rp) {
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    double w[kMaxCols];
#pragma unroll
    for (int j = 0; j < kMaxCols; ++j)
      if (j < nv) w[j] = V[(long)(j + 1) * ldv + q];
    for (int i = 0; i < c1; ++i) {
      const double qv = Q[(long)i * ldq + q];
#pragma unroll
      for (int j = 0; j < kMaxCols; ++j)
        if (j < nv) w[j] = fma(-Rp[i + (long)c1 * (j + 1)], qv, w[j]);
    }
#pragma unroll
    for (int j = 0; j < kMaxCols; ++j)
      if (j < nv) V[(long)(j + 1) * ldv + q] = w[j];
  }
}

// Q = V R^-1 (calcQ_VRinv_, QRDecomposition.f90:603-671): column c sums
// fma(V_k, Rinv(k,c), .) over k = 0..c from 0; Rinv is cnt x cnt, ld cnt
__global__ void __launch_bounds__(kNT) k_ca_vrinv(int n, const double* V, long ldv, int cnt, const double* Rinv,
                                                  double* Q, long ldq) {
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
    double w[kMaxCols];
#pragma unroll
    for (int c = 0; c < kMaxCols; ++c)
      if (c < cnt) w[c] = V[(long)c * ldv + q];
#pragma unroll
    for (int c = 0; c < kMaxCols; ++c) {
      if (c < cnt) {
        double a = 0.0;
#pragma unroll
        for (int k = 0; k <= c; ++k) a = fma(w[k], Rinv[k + c * cnt], a);
        Q[(long)c * ldq + q] = a;
      }
    }
  }
}

// ---- TSQR, first level (TSQR_local_Ronly_, TSQR.f90:114-261) -------------
// 64-row panels; panel p belongs to wavefront p mod W.  A wavefront keeps an
// NC x NC upper-triangular R (lane l holds row l) and folds each panel P in
// by Householder QR of [R; P], House_local_Ronly_omp's reflector
// v = [y1; P(:,i)] on the stacked matrix: the 64-lane sums are xor
// butterflies, the R entries are broadcast from their row's lane.  The W
// factors go to Rout (NC x NC each, column-major) for the host's second level.
constexpr int kTsqrWaves = 1024;
inline int tsqr_waves(int n) {
  const int np = (n + 63) / 64;
  return np < kTsqrWaves ? (np > 0 ? np : 1) : kTsqrWaves;
}

template <int NC>
__global__ void __launch_bounds__(kNT) k_tsqr(int n, const double* V, long ldv, int W, double* Rout) {
  const int lane = threadIdx.x & 63;
  const int w = blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (w >= W) return;  // whole wavefronts only; no workgroup barrier below
  const int np = (n + 63) / 64;
  double Rrow[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) Rrow[c] = 0.0;
  for (int p = w; p < np; p += W) {
    const long row = (long)p * 64 + lane;
    double P[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) P[c] = row < n ? V[(long)c * ldv + row] : 0.0;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
      const double x = P[i];
      double d[NC];
#pragma unroll
      for (int k = i; k < NC; ++k) d[k] = x * P[k];
      // the xor butterfly (offsets 32..1) in registers
#define PARCEL_XSTEP(O) \
  for (int k = i; k < NC; ++k) d[k] += xor_lane<O>(d[k]);
      PARCEL_XSTEP(32) PARCEL_XSTEP(16) PARCEL_XSTEP(8) PARCEL_XSTEP(4) PARCEL_XSTEP(2) PARCEL_XSTEP(1)
#undef PARCEL_XSTEP
      const double alpha = __shfl(Rrow[i], i, 64);
      const double s2 = d[i];
      const double nrm = sqrt(fma(alpha, alpha, s2));
      const double y1 = alpha > 0.0 ? alpha + nrm : alpha - nrm;
      double den = fma(y1, y1, s2);
      if (den == 0.0) den = 1.0 - 300;
      const double t1 = 2.0 / den;
      const double na = fma(-(t1 * y1), fma(y1, alpha, s2), alpha);
#pragma unroll
      for (int k = i + 1; k < NC; ++k) {
        const double rik = __shfl(Rrow[k], i, 64);
        const double tk = t1 * fma(y1, rik, d[k]);
        P[k] = fma(-x, tk, P[k]);
        const double nr = fma(-y1, tk, rik);
        if (lane == i) Rrow[k] = nr;
      }
      if (lane == i) Rrow[i] = na;
    }
  }
  if (lane < NC) {
#pragma unroll
    for (int c = 0; c < NC; ++c) Rout[(long)w * NC * NC + lane + c * NC] = (c >= lane) ? Rrow[c] : 0.0;
  }
}

void tsqr_local(int nc, int n, const double* V, long ldv, int W, double* Rout, hipStream_t st) {
  const dim3 g((W + kWaves - 1) / kWaves), b(kNT);
  switch (nc) {
#define PARCEL_TS(C) \
  case C: hipLaunchKernelGGL(k_tsqr<C>, g, b, 0, st, n, V, ldv, W, Rout); break;
    PARCEL_TS(1) PARCEL_TS(2) PARCEL_TS(3) PARCEL_TS(4) PARCEL_TS(5) PARCEL_TS(6) PARCEL_TS(7) PARCEL_TS(8)
    PARCEL_TS(9) PARCEL_TS(10) PARCEL_TS(11) PARCEL_TS(12) PARCEL_TS(13) PARCEL_TS(14) PARCEL_TS(15)
    PARCEL_TS(16)
#undef PARCEL_TS
    default: throw hip_error("TSQR: column count must be in 1..16", PARCEL_HIP_ERR_ARG);
  }
}

// ---- solution update ---------------------------------------------------
// v0 = sum_{k < nq} fma(y_k, Q_k, .) then sum_{k < nv} fma(y_{nq+k}, W_k, .),
// from 0 (sum_vec_cagmres, cu_pcagmres.cu:9-60; nv = 0: sum_vec)
__global__ void __launch_bounds__(kNT) k_ca_sumvec(int n, const double* Q, long ldq, int nq, const double* W,
                                                   long ldw, int nv, const double* y, double* __restrict__ v0) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double a = 0.0;
    for (int k = 0; k < nq; ++k) a = fma(y[k], Q[(long)k * ldq + i], a);
    for (int k = 0; k < nv; ++k) a = fma(y[nq + k], W[(long)k * ldw + i], a);
    v0[i] = a;
  }
}

__global__ void __launch_bounds__(kNT) k_ca_xadd(int n, double* __restrict__ x, const double* __restrict__ x0) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    x[i] = x0[i] + x[i];
}

template <int NC>
void launch_gram(dim3 grid, hipStream_t st, int n, const double* V, long ldv, const double* rinv, double* part) {
  hipLaunchKernelGGL(k_ca_gram<NC>, grid, dim3(kNT), 0, st, n, V, ldv, rinv, part);
}

void gram(int nc, dim3 grid, hipStream_t st, int n, const double* V, long ldv, const double* rinv, double* part) {
  switch (nc) {
#define PARCEL_GRAM(C) \
  case C: launch_gram<C>(grid, st, n, V, ldv, rinv, part); break;
    PARCEL_GRAM(2) PARCEL_GRAM(3) PARCEL_GRAM(4) PARCEL_GRAM(5) PARCEL_GRAM(6) PARCEL_GRAM(7) PARCEL_GRAM(8)
    PARCEL_GRAM(9) PARCEL_GRAM(10) PARCEL_GRAM(11) PARCEL_GRAM(12) PARCEL_GRAM(13) PARCEL_GRAM(14)
    PARCEL_GRAM(15) PARCEL_GRAM(16)
#undef PARCEL_GRAM
    default: throw hip_error("CA-GMRES: s + 1 must be in 2..16", PARCEL_HIP_ERR_ARG);
  }
}

void cgs_dots(int cnt, dim3 grid, hipStream_t st, int n, const double* Q, long ldq, const double* vi,
              double* part) {
  switch (cnt) {
#define PARCEL_CGSD(C) \
  case C: hipLaunchKernelGGL(k_ca_cgs_dots<C>, grid, dim3(kNT), 0, st, n, Q, ldq, vi, part); break;
    PARCEL_CGSD(1) PARCEL_CGSD(2) PARCEL_CGSD(3) PARCEL_CGSD(4) PARCEL_CGSD(5) PARCEL_CGSD(6) PARCEL_CGSD(7)
    PARCEL_CGSD(8) PARCEL_CGSD(9) PARCEL_CGSD(10) PARCEL_CGSD(11) PARCEL_CGSD(12) PARCEL_CGSD(13)
    PARCEL_CGSD(14) PARCEL_CGSD(15)
#undef PARCEL_CGSD
    default: throw hip_error("CA-GMRES: CGS column count out of range", PARCEL_HIP_ERR_ARG);
  }
}

void bdots(int nv, dim3 grid, hipStream_t st, int n, const double* Q, long ldq, int c1, const double* V, long ldv,
           double* part) {
  switch (nv) {
#define PARCEL_BD(C) \
  case C: hipLaunchKernelGGL(k_ca_bdots<C>, grid, dim3(kNT), 0, st, n, Q, ldq, c1, V, ldv, part); break;
    PARCEL_BD(2) PARCEL_BD(3) PARCEL_BD(4) PARCEL_BD(5) PARCEL_BD(6) PARCEL_BD(7) PARCEL_BD(8) PARCEL_BD(9)
    PARCEL_BD(10) PARCEL_BD(11) PARCEL_BD(12) PARCEL_BD(13) PARCEL_BD(14) PARCEL_BD(15) PARCEL_BD(16)
#undef PARCEL_BD
    default: throw hip_error("CA-GMRES: s + 1 must be in 2..16", PARCEL_HIP_ERR_ARG);
  }
}

// Newton change-of-basis matrix (calc_NewtonBasis, pcagmres.f90:586-611);
// B is (s+1) x s and keeps across restarts the entries it does not rewrite.
void newton_basis(int s, std::vector<double>& B, const std::vector<double>& er, const std::vector<double>& ei) {
  const int ld = s + 1;
  int i = 0;
  while (i < s) {
    if (i == s - 1) {
      B[i + (size_t)i * ld] = er[i];
    } else if (std::fabs(ei[i]) == 0.0) {
      B[i + (size_t)i * ld] = er[i];
    } else {
      B[i + (size_t)i * ld] = er[i];
      B[(i + 1) + (size_t)(i + 1) * ld] = er[i];
      B[i + (size_t)(i + 1) * ld] = -ei[i] * ei[i];
      ++i;
    }
    ++i;
  }
}

// GivensRotations (pcagmres.f90:190-242) on column i (1-based) of the
// (m+1) x m matrix H
void givens(int m, double* H, int i, double* c, double* s, double* e) {
  double* hc = H + (size_t)(i - 1) * (m + 1);
  for (int k = 0; k < i - 1; ++k) {
    const double t1 = hc[k] * c[k] - hc[k + 1] * s[k];
    const double t2 = hc[k] * s[k] + hc[k + 1] * c[k];
    hc[k] = t1;
    hc[k + 1] = t2;
  }
  double t;
  if (hc[i - 1] == 0.0) t = hc[i] / 1.0 - 300;  // the reference's `/1.0d0-300`
  else t = hc[i] / hc[i - 1];
  const double ci = 1.0 / std::sqrt(1.0 + t * t);
  const double si = -ci * t;
  c[i - 1] = ci;
  s[i - 1] = si;
  const double e1 = ci * e[i - 1] - si * e[i];
  const double e2 = si * e[i - 1] + ci * e[i];
  e[i - 1] = e1;
  e[i] = e2;
  hc[i - 1] = ci * hc[i - 1] - si * hc[i];
  hc[i] = 0.0;
}

}  // namespace

// TSQR_Ronly_ (QRDecomposition.f90:887-1002): the first level on the device
// into dT (tsqr_waves(n) factors), House_local_Ronly_omp over the stacked
// factors on the host (hT: pinned, same size), TSQR_GLOBAL (the identity on
// one rank), then rows with a negative diagonal negated.  R: nc x nc, ld nc.
void tsqr_r(int n, int nc, const double* V, long ldv, double* dT, double* hT, double* R, hipStream_t st) {
  const int W = tsqr_waves(n);
  tsqr_local(nc, n, V, ldv, W, dT, st);
  PARCEL_HIP_CHECK(hipMemcpyAsync(hT, dT, sizeof(double) * (size_t)W * nc * nc, hipMemcpyDeviceToHost, st));
  PARCEL_HIP_CHECK(hipStreamSynchronize(st));
  const int mrows = W * nc;
  std::vector<double> S((size_t)mrows * nc);
  for (int w = 0; w < W; ++w)
    for (int j = 0; j < nc; ++j)
      for (int i = 0; i < nc; ++i) S[(size_t)w * nc + i + (size_t)j * mrows] = hT[(size_t)w * nc * nc + i + (size_t)j * nc];
  dense::house_r(mrows, nc, S.data(), mrows, R);
  for (int j = 0; j < nc; ++j)
    if (R[j + (size_t)j * nc] < 0.0)
      for (int i = j; i < nc; ++i) R[j + (size_t)i * nc] = -R[j + (size_t)i * nc];
}

int cagmres_solve(Plan& P, double* x, const double* b, int itrmax, double rtol, double* hist_host, int* niter,
                  int* iret) {
  const int n = P.A.n;
  const hipStream_t st = P.stream;
  const int s = P.prm.CAGMRES_sstep, t = P.prm.CAGMRES_tstep;
  const int basis = P.prm.CAGMRES_basis, qrflag = P.prm.CAGMRES_QRflag;
  const bool newton = basis == 1 || basis == 2;
  if (s < 1 || s + 1 > kMaxCols) throw hip_error("CAGMRES_sstep must be in 1..15", PARCEL_HIP_ERR_ARG);
  if (t < 1 || s * t > kMaxM) throw hip_error("CAGMRES_tstep must be >= 1 with s*t <= 128", PARCEL_HIP_ERR_ARG);
  if (qrflag < PARCEL_QR_MGS || qrflag > PARCEL_QR_CHOLESKYQR2)
    throw hip_error("CAGMRES_QRflag must be in 1..5", PARCEL_HIP_ERR_ARG);
  const bool gs = qrflag == PARCEL_QR_MGS || qrflag == PARCEL_QR_CGS;
  const int m = s * t, nc = s + 1;
  const int c1max = s * (t - 1) + 1;
  const bool needQ = gs || t > 1;
  plan_workspace(P, 1);
  const long ldv = n;
  if (P.basis.n < (size_t)nc * (size_t)n) P.basis.alloc((size_t)nc * (size_t)n);
  if (needQ && P.basis2.n < (size_t)(m + 1) * (size_t)n) P.basis2.alloc((size_t)(m + 1) * (size_t)n);
  ensure_vecs(P, 3);
  double* V = P.basis.p;
  double* Q = needQ ? P.basis2.p : nullptr;
  double* pv = P.vec[0].p;
  double* v0 = P.vec[1].p;
  double* x0 = P.vec[2].p;
  const int np = nc * (nc + 1) / 2;
  const size_t nslot = (size_t)std::max(std::max(np, 2), t > 1 ? c1max * nc : 0) * kNBMax;
  if (P.wpart.n < nslot) P.wpart.alloc(nslot);
  double* gp = P.wpart.p;
  // device scalars: st[0..3], G / R (nc x nc), Rinv (nc x nc), R' (c1max x nc), y (m + 1)
  const size_t nrp = t > 1 ? (size_t)c1max * nc : 0;
  const int tsw = tsqr_waves(n);
  const size_t ntsqr = qrflag == PARCEL_QR_TSQR ? (size_t)tsw * nc * nc : 0;
  const size_t dcount = 4 + 2 * (size_t)nc * nc + nrp + (size_t)m + 1 + ntsqr;
  if (P.solver_state.n < sizeof(double) * dcount) P.solver_state.alloc(sizeof(double) * dcount);
  double* dst = reinterpret_cast<double*>(P.solver_state.p);
  double* dG = dst + 4;
  double* dRinv = dG + (size_t)nc * nc;
  double* dRp = dRinv + (size_t)nc * nc;
  double* dY = dRp + nrp;
  double* dT = dY + m + 1;  // first-level TSQR factors
  double* hbuf = nullptr;  // pinned mirror of the device scalars
  PARCEL_HIP_CHECK(hipHostMalloc((void**)&hbuf, sizeof(double) * dcount, hipHostMallocDefault));
  double* hG = hbuf + 4;
  double* hRinv = hG + (size_t)nc * nc;
  double* hRp = hRinv + (size_t)nc * nc;
  double* hY = hRp + nrp;
  double* hT = hY + m + 1;
  const int nb = tiled_nb(n);
  const dim3 grid(nb), block(kNT);
  const SellArgs A = sell_args(P.A);
  auto t0 = std::chrono::steady_clock::now();
  hipEvent_t ev0, ev1;
  PARCEL_HIP_CHECK(hipEventCreate(&ev0));
  PARCEL_HIP_CHECK(hipEventCreate(&ev1));
  PARCEL_HIP_CHECK(hipEventRecord(ev0, st));
  int launches = 0;

  // host-side state (the reference's host arrays)
  std::vector<double> hist((size_t)std::max(itrmax, 0) + m + 1, 0.0);
  std::vector<double> B1((size_t)nc * s, 0.0), R((size_t)nc * nc), R2((size_t)nc * nc), Rinv((size_t)nc * nc);
  std::vector<double> RB((size_t)nc * s), Hes((size_t)(m + 1) * m), Hg((size_t)(m + 1) * m), Hs((size_t)m * m);
  std::vector<double> Rp(std::max(nrp, (size_t)1));
  std::vector<double> co(m + 1), si(m + 1), zeta(m + 1), y(m + 1), yt(m + 1), er(m, 0.0), ei(m, 0.0);
  for (int i = 0; i < s; ++i) B1[(i + 1) + (size_t)i * nc] = 1.0;  // monomial: sub-diagonal 1
  bool iflag_exit = false, iflag_basis = false, qmat = false;
  int itr = 1, itr0 = 1, mm = 0;
  double bnorm = 0.0, rnorm = 0.0, hk_pre = 0.0;
  auto fetch = [&](size_t count) {
    PARCEL_HIP_CHECK(hipMemcpyAsync(hbuf, dst, sizeof(double) * count, hipMemcpyDeviceToHost, st));
    PARCEL_HIP_CHECK(hipStreamSynchronize(st));
  };
  // Gram matrix of cnt columns (of V R^-1 with rinv_dev) -> Cholesky factor
  auto cholqr = [&](const double* Vb, int cnt, const double* rinv_dev, std::vector<double>& Rout) {
    const int npc = cnt * (cnt + 1) / 2;
    PARCEL_HIP_CHECK(hipMemsetAsync(dG, 0, sizeof(double) * (size_t)cnt * cnt, st));
    gram(cnt, grid, st, n, Vb, ldv, rinv_dev, gp);
    hipLaunchKernelGGL(k_ca_gram_reduce, dim3(npc), block, 0, st, (const double*)gp, nb, cnt, dG);
    launches += 2;
    fetch(4 + (size_t)cnt * cnt);
    for (int j = 0; j < cnt; ++j)
      for (int i = 0; i < cnt; ++i) Rout[i + (size_t)j * cnt] = (i <= j) ? hG[i + (size_t)j * cnt] : 0.0;
    (void)dense::chol_upper(cnt, Rout.data(), cnt);
    for (int j = 0; j < cnt; ++j)
      for (int i = j + 1; i < cnt; ++i) Rout[i + (size_t)j * cnt] = 0.0;
  };
  auto upload_rinv = [&](const std::vector<double>& Rsrc, int cnt) {
    std::memcpy(Rinv.data(), Rsrc.data(), sizeof(double) * (size_t)cnt * cnt);
    dense::trinv_upper(cnt, Rinv.data(), cnt);
    std::memcpy(hRinv, Rinv.data(), sizeof(double) * (size_t)cnt * cnt);
    PARCEL_HIP_CHECK(hipMemcpyAsync(dRinv, hRinv, sizeof(double) * (size_t)cnt * cnt, hipMemcpyHostToDevice, st));
  };
  // QRDecomposition_ (QRDecomposition.f90:1331-1551) of cnt columns Vb into
  // R (cnt x cnt) and, for GS or with qmat, Qb = the orthonormal factor
  auto qr_block = [&](double* Vb, int cnt, double* Qb, bool qm, std::vector<double>& Rout) {
    if (cnt == 1 || gs) {
      const bool mgs = (cnt == 1 || qrflag == PARCEL_QR_MGS);
      double* dR = dG;
      PARCEL_HIP_CHECK(hipMemsetAsync(dR, 0, sizeof(double) * (size_t)cnt * cnt, st));
      for (int i = 0; i < cnt; ++i) {
        double* vi = Vb + (long)i * ldv;
        double* sl = gp;
        int fin;  // slot holding the partials of <v_i,v_i>
        if (mgs) {
          if (i == 0) {
            hipLaunchKernelGGL(k_ca_dot, grid, block, 0, st, n, (const double*)vi, (const double*)vi, sl);
            fin = 0;
            ++launches;
          } else {
            hipLaunchKernelGGL(k_ca_dot, grid, block, 0, st, n, (const double*)vi, (const double*)Qb, sl);
            ++launches;
            for (int k = 0; k < i; ++k) {
              const double* qnext = (k + 1 < i) ? Qb + (long)(k + 1) * ldv : nullptr;
              hipLaunchKernelGGL(k_ca_mgs, grid, block, 0, st, n, vi, (const double*)(Qb + (long)k * ldv), qnext,
                                 (const double*)(gp + (size_t)(k & 1) * kNBMax),
                                 gp + (size_t)((k + 1) & 1) * kNBMax, nb, dR, cnt, k, i);
              ++launches;
            }
            fin = i & 1;
          }
        } else {
          if (i > 0) {
            cgs_dots(i, grid, st, n, Qb, ldv, vi, gp);
            hipLaunchKernelGGL(k_ca_cgs_reduce, dim3(i), block, 0, st, (const double*)gp, nb, dR, cnt, i);
            launches += 2;
          }
          // the reduce kernel has consumed slots 0..i-1 before this one writes slot 0
          hipLaunchKernelGGL(k_ca_cgs_update, grid, block, 0, st, n, (const double*)Qb, ldv, vi, (const double*)dR,
                             cnt, i, gp);
          ++launches;
          fin = 0;
        }
        const double* pfin = gp + (size_t)fin * kNBMax;
        hipLaunchKernelGGL(k_ca_qr_norm, grid, block, 0, st, n, (const double*)vi, Qb + (long)i * ldv, pfin, nb, dR,
                           cnt, i);
        ++launches;
      }
      fetch(4 + (size_t)cnt * cnt);
      for (size_t q = 0; q < (size_t)cnt * cnt; ++q) Rout[q] = hG[q];
      return;
    }
    if (qrflag == PARCEL_QR_TSQR && n > cnt && n > 64) {
      // TSQR_Ronly_ (QRDecomposition.f90:887-1002): first level on the
      // device, House_local_Ronly_omp over the stacked factors on the host,
      // TSQR_GLOBAL (the identity on one rank), then rows made non-negative
      tsqr_r(n, cnt, Vb, ldv, dT, hT, Rout.data(), st);
      ++launches;
      fetch(4);
    } else {
      cholqr(Vb, cnt, nullptr, Rout);
      if (qrflag == PARCEL_QR_CHOLESKYQR2 || qrflag == PARCEL_QR_TSQR) {
        // Q1 = V R^-1 (never stored), R2 = chol(Q1^T Q1), R = R2 R
        upload_rinv(Rout, cnt);
        cholqr(Vb, cnt, dRinv, R2);
        dense::trmm_lunn(cnt, R2.data(), cnt, Rout.data(), cnt);
      }
    }
    if (qm) {  // calcQ_VRinv_
      upload_rinv(Rout, cnt);
      hipLaunchKernelGGL(k_ca_vrinv, grid, block, 0, st, n, (const double*)Vb, ldv, cnt, (const double*)dRinv, Qb,
                         ldv);
      ++launches;
    }
  };

  for (itr = 1; itr <= itrmax; itr += s * t) {
    itr0 = itr;
    std::fill(co.begin(), co.end(), 0.0);
    std::fill(si.begin(), si.end(), 0.0);
    std::fill(zeta.begin(), zeta.end(), 0.0);
    std::fill(Hes.begin(), Hes.end(), 0.0);
    std::fill(Hg.begin(), Hg.end(), 0.0);
    std::fill(R.begin(), R.end(), 0.0);
    std::fill(Rp.begin(), Rp.end(), 0.0);
    hk_pre = 0.0;
    // r = b - A x, ||r|| (and ||b|| on the first cycle); v1 = -r/||r||
    launch_residual(P, x, b, P.r.p, P.part.p);
    hipLaunchKernelGGL(k_ca_restart, dim3(1), block, 0, st, (const double*)(P.part.p + PART_BB * kNBMax),
                       (const double*)(P.part.p + PART_RR0 * kNBMax), nb, dst, itr == 1 ? 1 : 0);
    hipLaunchKernelGGL(k_ca_scale, grid, block, 0, st, n, (const double*)P.r.p, V, (const double*)dst);
    launches += 3;
    bool converged = false;
    for (int k = 1; k <= t; ++k) {
      if (newton && iflag_basis) newton_basis(s, B1, er, ei);
      qmat = (k != t);
      for (int i = 1; i <= s; ++i) {
        const double* vi = V + (long)(i - 1) * ldv;
        precon_apply(P, vi, pv, nullptr);
        int mode = 0;
        double alpha = 0.0, beta = 0.0;
        if (iflag_basis && newton) {
          alpha = -B1[(i - 1) + (size_t)(i - 1) * nc];
          mode = 1;
          if (i != 1) {
            // basis 1 adds B1(i-1,i) V_{i-2} as calc_NewtonBasisVec does
            // (pcagmres.f90:1519-1530); basis 2 subtracts it, the recurrence
            // A V_{i-1} = V_i + B1(i,i) V_{i-1} + B1(i-1,i) V_{i-2} that the
            // Hessenberg update assumes
            beta = basis == 2 ? -B1[(i - 2) + (size_t)(i - 1) * nc] : B1[(i - 2) + (size_t)(i - 1) * nc];
            mode = 2;
          }
        }
        hipLaunchKernelGGL(k_ca_basis, grid, block, 0, st, n, A, (const double*)pv, V + (long)i * ldv, vi,
                           i >= 2 ? (const double*)(V + (long)(i - 2) * ldv) : nullptr, alpha, beta, mode);
        launches += 2;
      }
      if (k == 1) {
        qr_block(V, nc, Q, gs || qmat, R);
        // the restart test (pcagmres.f90:1977, abstol = 0) uses the norm
        // fetched with R; a converged restart leaves x untouched either way
        if (itr == 1) bnorm = hbuf[0];
        rnorm = hbuf[1];
        if (rnorm / bnorm < rtol) {
          converged = true;
          break;
        }
        if (itr == 1) hist[0] = rnorm / bnorm;
        zeta[0] = -rnorm;
        // calc_HessenbergMatrix, k = 1: H = R B1 R(1:s,1:s)^-1
        for (int j = 0; j < s; ++j)
          for (int i = 0; i < s; ++i) Rinv[i + (size_t)j * s] = R[i + (size_t)j * nc];
        dense::trinv_upper(s, Rinv.data(), s);
        dense::gemm_nn(nc, s, nc, R.data(), nc, B1.data(), nc, RB.data(), nc);
        dense::gemm_nn(nc, s, s, RB.data(), nc, Rinv.data(), s, Hes.data(), m + 1);
        hk_pre = Hes[s + (size_t)(s - 1) * (m + 1)];
      } else {
        // block CGS: R' = Q(:,1:c1)^T V, V(:,2:s+1) -= Q(:,1:c1) R'(:,2:s+1)
        const int c1 = s * (k - 1) + 1;
        bdots(nc, dim3(nb, (c1 + 3) / 4), st, n, Q, ldv, c1, V, ldv, gp);
        hipLaunchKernelGGL(k_ca_reduce_slots, dim3(c1 * nc), block, 0, st, (const double*)gp, nb, dRp);
        hipLaunchKernelGGL(k_ca_bupd, grid, block, 0, st, n, (const double*)Q, ldv, c1, V, ldv, s,
                           (const double*)dRp);
        launches += 3;
        PARCEL_HIP_CHECK(hipMemcpyAsync(hRp, dRp, sizeof(double) * (size_t)c1 * nc, hipMemcpyDeviceToHost, st));
        std::vector<double> Rk((size_t)s * s, 0.0);
        qr_block(V + ldv, s, Q + (long)c1 * ldv, gs || qmat, Rk);  // synchronises the stream
        for (size_t q = 0; q < (size_t)c1 * nc; ++q) Rp[q] = hRp[q];
        dense::hessenberg_block(s, t, k, B1.data(), Hes.data(), Rk.data(), Rp.data(), &hk_pre);
        std::memcpy(R.data(), Rk.data(), sizeof(double) * (size_t)s * s);
      }
      if (k != t)
        PARCEL_HIP_CHECK(hipMemcpyAsync(V, Q + (long)(s * k) * ldv, sizeof(double) * (size_t)n,
                                        hipMemcpyDeviceToDevice, st));
      for (int i = s * (k - 1); i < s * k; ++i)
        for (int j = 0; j <= s * k; ++j) Hg[j + (size_t)i * (m + 1)] = Hes[j + (size_t)i * (m + 1)];
      for (int i = s * (k - 1) + 1; i <= s * k; ++i) {
        mm = i;
        givens(m, Hg.data(), i, co.data(), si.data(), zeta.data());
        rnorm = std::fabs(zeta[i]);
        itr0 = i + (itr - 1) + 1;
        hist[itr0 - 1] = rnorm / bnorm;
        if (rnorm / bnorm < rtol || itr0 == itrmax) {
          iflag_exit = true;
          break;
        }
      }
      if (iflag_exit) break;
      if (newton) {
        iflag_basis = true;
        const int sk = s * k;
        for (int j = 0; j < sk; ++j)
          for (int i = 0; i < sk; ++i) Hs[i + (size_t)j * sk] = Hes[i + (size_t)j * (m + 1)];
        std::fill(er.begin(), er.end(), 0.0);
        std::fill(ei.begin(), ei.end(), 0.0);
        (void)dense::hqr_eigs(sk, Hs.data(), sk, er.data(), ei.data());
        dense::leja_order(sk, er.data(), ei.data());
      }
    }
    if (converged) break;
    // solution update (pcagmres.f90:2287-2430)
    int nq, nv;
    const double* W = nullptr;
    if (gs || qmat || (t > 1 && s == 1)) {
      // update_xVec (pcagmres.f90:1136-1236) with the orthonormal Q
      for (int i = mm; i >= 1; --i) {
        double yi = zeta[i - 1];
        for (int j = mm; j >= i + 1; --j) yi = yi - Hg[(i - 1) + (size_t)(j - 1) * (m + 1)] * y[j - 1];
        const double d = Hg[(i - 1) + (size_t)(i - 1) * (m + 1)];
        y[i - 1] = (d == 0.0) ? yi / 1.0e-300 : yi / d;
      }
      nq = mm;
      nv = 0;
    } else {
      // update_xVec_cagmres (pcagmres.f90:1338-1478): y = H^-1 zeta; its last
      // vstep entries times R(1:sv-1,1:sv-1)^-1 (sv = s+1 for t = 1, else s);
      // Qy = the first mm - vstep Q columns and vstep V columns
      for (int i = mm; i >= 1; --i) {
        double yi = zeta[i - 1];
        for (int j = mm; j >= i + 1; --j) yi = yi - Hg[(i - 1) + (size_t)(j - 1) * (m + 1)] * y[j - 1];
        y[i - 1] = yi / Hg[(i - 1) + (size_t)(i - 1) * (m + 1)];
      }
      const int sv = (t == 1) ? s + 1 : s;
      const int vstep = (sv - 1) - (m - mm);
      for (int i = 0; i < sv - 1; ++i) yt[i] = (i < vstep) ? y[mm - vstep + i] : 0.0;
      std::memcpy(Rinv.data(), R.data(), sizeof(double) * (size_t)sv * sv);
      dense::trinv_upper(sv, Rinv.data(), sv);
      dense::gemm_nn(sv - 1, 1, sv - 1, Rinv.data(), sv, yt.data(), sv - 1, y.data() + (mm - vstep), sv - 1);
      nq = mm - vstep;
      nv = vstep;
      W = (t == 1) ? V : V + ldv;
    }
    for (int k = 0; k < nq + nv; ++k) hY[k] = y[k];
    PARCEL_HIP_CHECK(hipMemcpyAsync(dY, hY, sizeof(double) * (size_t)std::max(nq + nv, 1), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_ca_sumvec, grid, block, 0, st, n, (const double*)Q, ldv, nq, W, ldv, nv,
                       (const double*)dY, v0);
    precon_apply(P, v0, x0, nullptr);
    hipLaunchKernelGGL(k_ca_xadd, grid, block, 0, st, n, x, (const double*)x0);
    launches += 3;
    // hY is rewritten by the next cycle only after the stream has drained
    PARCEL_HIP_CHECK(hipStreamSynchronize(st));
    PARCEL_HIP_CHECK(hipGetLastError());
    if (iflag_exit) break;
  }
  PARCEL_HIP_CHECK(hipEventRecord(ev1, st));
  PARCEL_HIP_CHECK(hipStreamSynchronize(st));
  (void)hipHostFree(hbuf);
  const int nit = itr0 < itrmax ? itr0 : itrmax;  // pcagmres.f90:2466
  const int ret = (itrmax <= itr) ? 1 : 0;
  if (nit > 0 && hist_host) std::memcpy(hist_host, hist.data(), sizeof(double) * (size_t)nit);
  float ms = 0.f;
  PARCEL_HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  P.stats.solve_device_ms = ms;
  P.stats.solve_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  P.stats.niter = nit;
  P.stats.launches = launches;
  *niter = nit;
  *iret = ret;
  return PARCEL_HIP_OK;
}

}  // namespace parcel

extern "C" int parcel_hip_tsqr_r(int n, int nc, const double* V, int ldv, double* R) {
  using namespace parcel;
  if (n < 1 || nc < 1 || nc > 16 || ldv < n || !V || !R) return api_fail(PARCEL_HIP_ERR_ARG, "tsqr_r: bad arguments");
  try {
    if (int rc = api_ensure_init()) return rc;
    hipStream_t st = nullptr;
    const bool dev = api_is_device_ptr(V);
    DevBuf<double> dv;
    const double* Vd = V;
    if (!dev) {
      dv.alloc((size_t)ldv * nc);
      PARCEL_HIP_CHECK(hipMemcpy(dv.p, V, sizeof(double) * (size_t)ldv * nc, hipMemcpyHostToDevice));
      Vd = dv.p;
    }
    const int W = tsqr_waves(n);
    DevBuf<double> dT((size_t)W * nc * nc);
    double* hT = nullptr;
    PARCEL_HIP_CHECK(hipHostMalloc((void**)&hT, sizeof(double) * (size_t)W * nc * nc, hipHostMallocDefault));
    std::vector<double> Rh((size_t)nc * nc);
    try {
      tsqr_r(n, nc, Vd, ldv, dT.p, hT, Rh.data(), st);
    } catch (...) {
      (void)hipHostFree(hT);
      throw;
    }
    (void)hipHostFree(hT);
    std::memcpy(R, Rh.data(), sizeof(double) * (size_t)nc * nc);
    return PARCEL_HIP_OK;
  } catch (const hip_error& e) {
    return api_fail(e.code < 0 ? e.code : PARCEL_HIP_ERR_HIP, e.what());
  } catch (const std::exception& e) {
    return api_fail(PARCEL_HIP_ERR_HIP, e.what());
  }
}

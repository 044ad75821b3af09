// cagmres.hip -- CA-GMRES(s,t), KSP_pCAGMRES (src/pcagmres.f90:1557-2517),
// for t = 1, with right preconditioning, monomial or Newton basis
// (calc_NewtonBasis / calc_NewtonBasisVec, :586-611,1480-1552) and the QR
// variants of QRDecomposition_ (src/QRDecomposition.f90:1331-1551):
// MGS / CGS (GS_QR_), CholeskyQR (CholeskyQR_Ronly_) and CholeskyQR2.
//
// Device work per restart cycle: the residual, s preconditioner applies and
// SpMVs that build the basis V (n x (s+1), the Newton recurrence fused into
// the SpMV's store), one pass computing every entry of the Gram matrix V^T V
// at once (CholeskyQR; CholeskyQR2 adds one pass over V R^-1 formed on the
// fly, never stored), and the solution update.  The (s+1) x (s+1) dense
// work -- Cholesky, Hessenberg matrix R B R^-1 (calc_HessenbergMatrix, k=1,
// :879-1051), Givens rotations, Ritz values and Leja ordering -- runs on the
// host as in the reference, after one device-to-host copy of the Gram matrix
// and the residual norm per cycle.
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "dense.hpp"
#include "kernels.hpp"

namespace parcel {

namespace {

constexpr int kMaxCols = 16;  // s + 1 <= 16

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

// ---- solution update ---------------------------------------------------
struct Coef {
  double c[kMaxCols];
};

// v0 = sum_{k < mm} fma(c_k, W_k, .) from 0 (sum_vec_cagmres / sum_vec_FP64)
__global__ void __launch_bounds__(kNT) k_ca_sumvec(int n, const double* W, long ldw, Coef y, int mm,
                                                   double* __restrict__ v0) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double a = 0.0;
    for (int k = 0; k < mm; ++k) a = fma(y.c[k], W[(long)k * ldw + i], a);
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

int cagmres_solve(Plan& P, double* x, const double* b, int itrmax, double rtol, double* hist_host, int* niter,
                  int* iret) {
  const int n = P.A.n;
  const hipStream_t st = P.stream;
  const int s = P.prm.CAGMRES_sstep, t = P.prm.CAGMRES_tstep;
  const int basis = P.prm.CAGMRES_basis, qrflag = P.prm.CAGMRES_QRflag;
  if (s < 1 || s + 1 > kMaxCols) throw hip_error("CAGMRES_sstep must be in 1..15", PARCEL_HIP_ERR_ARG);
  if (t != 1) throw hip_error("CAGMRES_tstep > 1 (block CGS between blocks) is not supported",
                              PARCEL_HIP_ERR_UNSUPPORTED);
  if (qrflag != PARCEL_QR_MGS && qrflag != PARCEL_QR_CGS && qrflag != PARCEL_QR_CHOLESKYQR &&
      qrflag != PARCEL_QR_CHOLESKYQR2)
    throw hip_error("CAGMRES_QRflag: TSQR is not supported", PARCEL_HIP_ERR_UNSUPPORTED);
  const bool gs = qrflag == PARCEL_QR_MGS || qrflag == PARCEL_QR_CGS;
  const int m = s * t, nc = s + 1;
  plan_workspace(P, 1);
  const long ldv = n;
  if (P.basis.n < (size_t)nc * (size_t)n) P.basis.alloc((size_t)nc * (size_t)n);
  if (gs && P.basis2.n < (size_t)nc * (size_t)n) P.basis2.alloc((size_t)nc * (size_t)n);
  ensure_vecs(P, 3);
  double* V = P.basis.p;
  double* Q = gs ? P.basis2.p : nullptr;
  double* pv = P.vec[0].p;
  double* v0 = P.vec[1].p;
  double* x0 = P.vec[2].p;
  const int np = nc * (nc + 1) / 2;
  const size_t nslot = (size_t)std::max(np, 2) * kNBMax;
  if (P.wpart.n < nslot) P.wpart.alloc(nslot);
  double* gp = P.wpart.p;
  // device scalars: st[0..2], then G / R (nc x nc), then Rinv (nc x nc)
  const size_t dbytes = sizeof(double) * (4 + 2 * (size_t)nc * nc);
  if (P.solver_state.n < dbytes) P.solver_state.alloc(dbytes);
  double* dst = reinterpret_cast<double*>(P.solver_state.p);
  double* dG = dst + 4;
  double* dRinv = dG + (size_t)nc * nc;
  double* hbuf = nullptr;  // pinned: st (4) + G (nc^2) + Rinv (nc^2)
  PARCEL_HIP_CHECK(hipHostMalloc((void**)&hbuf, dbytes, hipHostMallocDefault));
  double* hG = hbuf + 4;
  double* hRinv = hG + (size_t)nc * nc;
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
  std::vector<double> co(m + 1), si(m + 1), zeta(m + 1), y(m + 1), yt(m + 1), er(m, 0.0), ei(m, 0.0);
  for (int i = 0; i < s; ++i) B1[(i + 1) + (size_t)i * nc] = 1.0;  // monomial: sub-diagonal 1
  bool iflag_exit = false, iflag_basis = false;
  int itr = 1, itr0 = 1, mm = 0;
  double bnorm = 0.0, rnorm = 0.0;
  auto fetch = [&](size_t count) {
    PARCEL_HIP_CHECK(hipMemcpyAsync(hbuf, dst, sizeof(double) * count, hipMemcpyDeviceToHost, st));
    PARCEL_HIP_CHECK(hipStreamSynchronize(st));
  };
  auto cholqr = [&](const double* rinv_dev, std::vector<double>& Rout) {
    gram(nc, grid, st, n, V, ldv, rinv_dev, gp);
    hipLaunchKernelGGL(k_ca_gram_reduce, dim3(np), block, 0, st, (const double*)gp, nb, nc, dG);
    launches += 2;
    fetch(4 + (size_t)nc * nc);
    for (int j = 0; j < nc; ++j)
      for (int i = 0; i < nc; ++i) Rout[i + (size_t)j * nc] = (i <= j) ? hG[i + (size_t)j * nc] : 0.0;
    (void)dense::chol_upper(nc, Rout.data(), nc);
    for (int j = 0; j < nc; ++j)
      for (int i = j + 1; i < nc; ++i) Rout[i + (size_t)j * nc] = 0.0;
  };
  PARCEL_HIP_CHECK(hipMemsetAsync(dG, 0, sizeof(double) * (size_t)nc * nc, st));
  for (itr = 1; itr <= itrmax; itr += s * t) {
    itr0 = itr;
    std::fill(co.begin(), co.end(), 0.0);
    std::fill(si.begin(), si.end(), 0.0);
    std::fill(zeta.begin(), zeta.end(), 0.0);
    std::fill(Hes.begin(), Hes.end(), 0.0);
    std::fill(Hg.begin(), Hg.end(), 0.0);
    std::fill(R.begin(), R.end(), 0.0);
    // r = b - A x, ||r|| (and ||b|| on the first cycle); v1 = -r/||r||
    launch_residual(P, x, b, P.r.p, P.part.p);
    hipLaunchKernelGGL(k_ca_restart, dim3(1), block, 0, st, (const double*)(P.part.p + PART_BB * kNBMax),
                       (const double*)(P.part.p + PART_RR0 * kNBMax), nb, dst, itr == 1 ? 1 : 0);
    hipLaunchKernelGGL(k_ca_scale, grid, block, 0, st, n, (const double*)P.r.p, V, (const double*)dst);
    launches += 3;
    // basis (the restart test below uses the norm fetched with the Gram
    // matrix; a converged restart leaves x untouched either way)
    if (basis == 1 && iflag_basis) newton_basis(s, B1, er, ei);
    for (int i = 1; i <= s; ++i) {
      const double* vi = V + (long)(i - 1) * ldv;
      precon_apply(P, vi, pv, nullptr);
      int mode = 0;
      double alpha = 0.0, beta = 0.0;
      if (iflag_basis && basis == 1) {
        alpha = -B1[(i - 1) + (size_t)(i - 1) * nc];
        mode = 1;
        if (i != 1) {
          beta = B1[(i - 2) + (size_t)(i - 1) * nc];
          mode = 2;
        }
      }
      hipLaunchKernelGGL(k_ca_basis, grid, block, 0, st, n, A, (const double*)pv, V + (long)i * ldv, vi,
                         i >= 2 ? (const double*)(V + (long)(i - 2) * ldv) : nullptr, alpha, beta, mode);
      launches += 2;
    }
    // QR of V
    if (gs) {
      double* dR = dG;
      PARCEL_HIP_CHECK(hipMemsetAsync(dR, 0, sizeof(double) * (size_t)nc * nc, st));
      for (int i = 0; i < nc; ++i) {
        double* vi = V + (long)i * ldv;
        double* sl = gp;
        int fin;  // slot holding the partials of <v_i,v_i>
        if (qrflag == PARCEL_QR_MGS) {
          if (i == 0) {
            hipLaunchKernelGGL(k_ca_dot, grid, block, 0, st, n, (const double*)vi, (const double*)vi, sl);
            fin = 0;
            ++launches;
          } else {
            hipLaunchKernelGGL(k_ca_dot, grid, block, 0, st, n, (const double*)vi, (const double*)Q, sl);
            ++launches;
            for (int k = 0; k < i; ++k) {
              const double* qnext = (k + 1 < i) ? Q + (long)(k + 1) * ldv : nullptr;
              hipLaunchKernelGGL(k_ca_mgs, grid, block, 0, st, n, vi, (const double*)(Q + (long)k * ldv), qnext,
                                 (const double*)(gp + (size_t)(k & 1) * kNBMax),
                                 gp + (size_t)((k + 1) & 1) * kNBMax, nb, dR, nc, k, i);
              ++launches;
            }
            fin = i & 1;
          }
        } else {
          if (i > 0) {
            cgs_dots(i, grid, st, n, Q, ldv, vi, gp);
            hipLaunchKernelGGL(k_ca_cgs_reduce, dim3(i), block, 0, st, (const double*)gp, nb, dR, nc, i);
            launches += 2;
          }
          // the reduce kernel has consumed slots 0..i-1 before this one writes slot 0
          hipLaunchKernelGGL(k_ca_cgs_update, grid, block, 0, st, n, (const double*)Q, ldv, vi, (const double*)dR,
                             nc, i, gp);
          ++launches;
          fin = 0;
        }
        const double* pfin = gp + (size_t)fin * kNBMax;
        hipLaunchKernelGGL(k_ca_qr_norm, grid, block, 0, st, n, (const double*)vi, Q + (long)i * ldv, pfin, nb, dR,
                           nc, i);
        ++launches;
      }
      fetch(4 + (size_t)nc * nc);
      for (size_t q = 0; q < (size_t)nc * nc; ++q) R[q] = hG[q];
    } else {
      cholqr(nullptr, R);
      if (qrflag == PARCEL_QR_CHOLESKYQR2) {
        // Q1 = V R^-1 (never stored), R2 = chol(Q1^T Q1), R = R2 R
        std::memcpy(Rinv.data(), R.data(), sizeof(double) * (size_t)nc * nc);
        dense::trinv_upper(nc, Rinv.data(), nc);
        std::memcpy(hRinv, Rinv.data(), sizeof(double) * (size_t)nc * nc);
        PARCEL_HIP_CHECK(hipMemcpyAsync(dRinv, hRinv, sizeof(double) * (size_t)nc * nc, hipMemcpyHostToDevice, st));
        cholqr(dRinv, R2);
        dense::trmm_lunn(nc, R2.data(), nc, R.data(), nc);
      }
    }
    if (itr == 1) bnorm = hbuf[0];
    rnorm = hbuf[1];
    if (rnorm / bnorm < rtol) break;  // pcagmres.f90:1977 (abstol = 0)
    if (itr == 1) hist[0] = rnorm / bnorm;
    zeta[0] = -rnorm;
    // calc_HessenbergMatrix, k = 1: H = R B1 R(1:s,1:s)^-1
    for (int j = 0; j < s; ++j)
      for (int i = 0; i < s; ++i) Rinv[i + (size_t)j * s] = R[i + (size_t)j * nc];
    dense::trinv_upper(s, Rinv.data(), s);
    dense::gemm_nn(nc, s, nc, R.data(), nc, B1.data(), nc, RB.data(), nc);
    dense::gemm_nn(nc, s, s, RB.data(), nc, Rinv.data(), s, Hes.data(), m + 1);
    for (int i = 0; i < s; ++i)
      for (int j = 0; j <= s; ++j) Hg[j + (size_t)i * (m + 1)] = Hes[j + (size_t)i * (m + 1)];
    for (int i = 1; i <= s; ++i) {
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
    if (!iflag_exit && basis == 1) {
      iflag_basis = true;
      for (int j = 0; j < s; ++j)
        for (int i = 0; i < s; ++i) Hs[i + (size_t)j * s] = Hes[i + (size_t)j * (m + 1)];
      std::fill(er.begin(), er.end(), 0.0);
      std::fill(ei.begin(), ei.end(), 0.0);
      (void)dense::hqr_eigs(s, Hs.data(), s, er.data(), ei.data());
      dense::leja_order(s, er.data(), ei.data());
    }
    // solution update
    Coef cy{};
    const double* W;
    if (gs) {
      // update_xVec (pcagmres.f90:1136-1236) with the orthonormal Q
      for (int i = mm; i >= 1; --i) {
        double yi = zeta[i - 1];
        for (int j = mm; j >= i + 1; --j) yi = yi - Hg[(i - 1) + (size_t)(j - 1) * (m + 1)] * y[j - 1];
        const double d = Hg[(i - 1) + (size_t)(i - 1) * (m + 1)];
        y[i - 1] = (d == 0.0) ? yi / 1.0e-300 : yi / d;
      }
      for (int k = 0; k < mm; ++k) cy.c[k] = y[k];
      W = Q;
    } else {
      // update_xVec_cagmres (pcagmres.f90:1338-1478): y = H^-1 zeta, then R(1:s,1:s)^-1 y
      for (int i = mm; i >= 1; --i) {
        double yi = zeta[i - 1];
        for (int j = mm; j >= i + 1; --j) yi = yi - Hg[(i - 1) + (size_t)(j - 1) * (m + 1)] * y[j - 1];
        y[i - 1] = yi / Hg[(i - 1) + (size_t)(i - 1) * (m + 1)];
      }
      for (int i = 0; i < s; ++i) yt[i] = (i < mm) ? y[i] : 0.0;
      std::memcpy(Rinv.data(), R.data(), sizeof(double) * (size_t)nc * nc);
      dense::trinv_upper(nc, Rinv.data(), nc);
      dense::gemm_nn(s, 1, s, Rinv.data(), nc, yt.data(), s, y.data(), s);
      for (int k = 0; k < mm; ++k) cy.c[k] = y[k];
      W = V;
    }
    hipLaunchKernelGGL(k_ca_sumvec, grid, block, 0, st, n, W, ldv, cy, mm, v0);
    precon_apply(P, v0, x0, nullptr);
    hipLaunchKernelGGL(k_ca_xadd, grid, block, 0, st, n, x, (const double*)x0);
    launches += 3;
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

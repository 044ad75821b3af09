// cbcg.hip -- Chebyshev-basis CG, KSP_pCBCG (src/pcbcg.f90:1182-2025), with
// the eigenvalue bounds of getEigenValue (src/pcbcg.f90:179-196): the power
// method (power_method_symMatrix, src/power_method.f90:188-437) for
// CBCG_Eigenflag 1 and, as the reference compiles it, for 2 (CA-Arnoldi,
// whose body is the power method under an `#if 1`, src/pcagmres.f90:2647-2651).
//
// One outer step of k-step CBCG:
//   r   -= AQ alpha                                       (calc_gemv)
//   S    = M^-1 [T_0 r, T_1 r, ..., T_{k-1} r],  AS = A S  (make_chebyshev_basis)
//   QtAS = Q^T AS                                         (gemm_dot)
//   beta = (Q^T AQ)^+ QtAS
//   x += Q alpha;  Q = S - Q beta;  AQ = AS - AQ beta;  Q^T AQ, Q^T r, <r,r>
//                                                         (CBCG_gemm_dot_local)
//   alpha = (Q^T AQ)^+ Q^T r
// with T_j the Chebyshev polynomials of the first kind on [Emin, Emax].
//
// Device work per outer step: one residual update, k preconditioner applies
// and k SpMVs (the three-term Chebyshev recurrence fused into the SpMV's
// store), one pass for the k x k block Q^T AS, and one fused pass that
// updates x, Q, AQ and produces every entry of Q^T AQ (upper triangle), Q^T r
// and <r,r> from registers.  The k x k pseudo-inverse and the two small GEMMs
// run on the host as in the reference (two device-to-host copies per step).
// Dot products use the library's tiled order (common.hpp).
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "dense.hpp"
#include "kernels.hpp"

namespace parcel {

namespace {

constexpr int kMaxK = 8;

struct KVec {
  double c[kMaxK];
};
struct KMat {
  double c[kMaxK * kMaxK];
};

// y = A s with, by mode, the Chebyshev recurrence stored from the product:
//   0: AS only
//   1: SS = eta*AS - zeta*r           (CBCG_axpy2, src/pcbcg.f90:12-27)
//   2: SS = 2eta*AS - 2zeta*SS1 - SS2 (CBCG_axpy3, src/pcbcg.f90:30-47)
// evaluated as fma(eta, AS, -(zeta*r)) and fma(eta2, AS, fma(-zeta2, SS1, -SS2)).
__global__ void __launch_bounds__(kNT) k_cb_spmv(int n, SellArgs A, const double* __restrict__ s,
                                                 double* __restrict__ as, double* __restrict__ ss,
                                                 const double* __restrict__ ss1, const double* __restrict__ ss2,
                                                 double eta, double zeta, int mode) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double t = sell_row(A, i, [&](int c) { return s[c]; });
    as[i] = t;
    if (mode == 1) {
      ss[i] = fma(eta, t, -(zeta * ss1[i]));
    } else if (mode == 2) {
      ss[i] = fma(eta, t, fma(-zeta, ss1[i], -ss2[i]));
    }
  }
}

// r = fma(-alpha_l, AQ_l, r), l = 0..k-1 in order (gemv with -alpha,
// src/vec.f90:368-399)
__global__ void __launch_bounds__(kNT) k_cb_gemv(int n, int k, const double* AQ, long ld, KVec al,
                                                 double* __restrict__ r) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double v = r[i];
    for (int l = 0; l < k; ++l) v = fma(-al.c[l], AQ[(long)l * ld + i], v);
    r[i] = v;
  }
}

// partials of <Q_i, AS_l> for i, l < K; slot i + K*l
template <int K>
__global__ void __launch_bounds__(kNT) k_cb_qtas(int n, const double* Q, const double* AS, long ld, double* part) {
  __shared__ double sh[kWaves];
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  double acc[K * K];
#pragma unroll
  for (int d = 0; d < K * K; ++d) acc[d] = 0.0;
  for (long t = blockIdx.x; t < ntiles; t += nb) {
    const long j = t * kNT + threadIdx.x;
    if (j < n) {
      double q[K], a[K];
#pragma unroll
      for (int l = 0; l < K; ++l) {
        q[l] = Q[(long)l * ld + j];
        a[l] = AS[(long)l * ld + j];
      }
#pragma unroll
      for (int l = 0; l < K; ++l)
#pragma unroll
        for (int i = 0; i < K; ++i) acc[i + K * l] = fma(q[i], a[l], acc[i + K * l]);
    }
  }
#pragma unroll
  for (int d = 0; d < K * K; ++d) {
    const double v = block_sum(acc[d], sh);
    if (threadIdx.x == 0) part[(long)d * kNBMax + blockIdx.x] = v;
  }
}

// CBCG_gemm_dot_local (src/pcbcg.f90:555-777; the CUDA kernel
// cu_pcbcg.cu:14-108): with UPD
//   x += sum_l Q_l alpha_l  (the sum formed first, from 0)
//   S_l -= sum_i Q_i beta(i,l);  AS_l -= sum_i AQ_i beta(i,l)   (i in order)
// in place (S, AS become the new Q, AQ); then partials of
//   <S_i, AS_l> (i <= l, slot l(l+1)/2 + i), <S_l, r> (slot K(K+1)/2 + l),
//   <r, r> (slot K(K+1)/2 + K).
// Without UPD only the partials (the first Q^T AQ, Q^T r and <r,r>).
template <int K, bool UPD>
__global__ void __launch_bounds__(kNT) k_cb_update(int n, double* __restrict__ x, const double* Q, const double* AQ,
                                                   double* S, double* AS, long ld, const double* __restrict__ r,
                                                   KVec al, KMat be, double* part) {
  __shared__ double sh[kWaves];
  constexpr int NU = K * (K + 1) / 2;
  constexpr int ND = NU + K + 1;
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  double acc[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) acc[d] = 0.0;
  for (long t = blockIdx.x; t < ntiles; t += nb) {
    const long j = t * kNT + threadIdx.x;
    if (j < n) {
      double s[K], a[K];
#pragma unroll
      for (int l = 0; l < K; ++l) {
        s[l] = S[(long)l * ld + j];
        a[l] = AS[(long)l * ld + j];
      }
      if (UPD) {
        double q[K], aq[K];
#pragma unroll
        for (int l = 0; l < K; ++l) {
          q[l] = Q[(long)l * ld + j];
          aq[l] = AQ[(long)l * ld + j];
        }
        double xs = 0.0;
#pragma unroll
        for (int l = 0; l < K; ++l) xs = fma(q[l], al.c[l], xs);
        x[j] = x[j] + xs;
#pragma unroll
        for (int l = 0; l < K; ++l) {
#pragma unroll
          for (int i = 0; i < K; ++i) {
            const double b = be.c[i + K * l];
            s[l] = fma(-q[i], b, s[l]);
            a[l] = fma(-aq[i], b, a[l]);
          }
          S[(long)l * ld + j] = s[l];
          AS[(long)l * ld + j] = a[l];
        }
      }
      const double rj = r[j];
#pragma unroll
      for (int l = 0; l < K; ++l) {
#pragma unroll
        for (int i = 0; i <= l; ++i) acc[l * (l + 1) / 2 + i] = fma(s[i], a[l], acc[l * (l + 1) / 2 + i]);
        acc[NU + l] = fma(s[l], rj, acc[NU + l]);
      }
      acc[NU + K] = fma(rj, rj, acc[NU + K]);
    }
  }
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    const double v = block_sum(acc[d], sh);
    if (threadIdx.x == 0) part[(long)d * kNBMax + blockIdx.x] = v;
  }
}

// out[slot] = reduction of slot's nb partials (one workgroup per slot)
__global__ void __launch_bounds__(kNT) k_cb_reduce(const double* part, int nb, double* out) {
  __shared__ double sh[kWaves];
  const double v = reduce_part(part + (long)blockIdx.x * kNBMax, nb, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = v;
}

// ---- power method ------------------------------------------------------
__global__ void __launch_bounds__(kNT) k_pm_fill_dot(int n, double* __restrict__ y, const double* __restrict__ x,
                                                     int fill_one, double* part) {
  __shared__ double sh[kWaves];
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  double acc = 0.0;
  for (long t = blockIdx.x; t < ntiles; t += nb) {
    const long j = t * kNT + threadIdx.x;
    if (j < n) {
      const double v = fill_one ? 1.0 : x[j];
      y[j] = v;
      acc = fma(v, v, acc);
    }
  }
  const double v = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ void __launch_bounds__(kNT) k_pm_scale(int n, double a, const double* __restrict__ y,
                                                  double* __restrict__ xp) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    xp[i] = a * y[i];
}

// y = A z; partials of <y,y> (slot 0) and <xp,y> (slot 1)
__global__ void __launch_bounds__(kNT) k_pm_spmv(int n, SellArgs A, const double* __restrict__ z,
                                                 double* __restrict__ y, const double* __restrict__ xp, double* part) {
  __shared__ double sh[kWaves];
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  double a0 = 0.0, a1 = 0.0;
  for (long t = blockIdx.x; t < ntiles; t += nb) {
    const long j = t * kNT + threadIdx.x;
    if (j < n) {
      const double v = sell_row(A, j, [&](int c) { return z[c]; });
      y[j] = v;
      a0 = fma(v, v, a0);
      a1 = fma(xp[j], v, a1);
    }
  }
  const double s0 = block_sum(a0, sh);
  const double s1 = block_sum(a1, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s0;
    part[kNBMax + blockIdx.x] = s1;
  }
}

template <int K>
void launch_qtas(dim3 g, hipStream_t st, int n, const double* Q, const double* AS, long ld, double* part) {
  hipLaunchKernelGGL(k_cb_qtas<K>, g, dim3(kNT), 0, st, n, Q, AS, ld, part);
}
template <int K>
void launch_update(bool upd, dim3 g, hipStream_t st, int n, double* x, const double* Q, const double* AQ, double* S,
                   double* AS, long ld, const double* r, const KVec& al, const KMat& be, double* part) {
  if (upd)
    hipLaunchKernelGGL((k_cb_update<K, true>), g, dim3(kNT), 0, st, n, x, Q, AQ, S, AS, ld, r, al, be, part);
  else
    hipLaunchKernelGGL((k_cb_update<K, false>), g, dim3(kNT), 0, st, n, x, Q, AQ, S, AS, ld, r, al, be, part);
}

void qtas(int k, dim3 g, hipStream_t st, int n, const double* Q, const double* AS, long ld, double* part) {
  switch (k) {
    case 1: launch_qtas<1>(g, st, n, Q, AS, ld, part); break;
    case 2: launch_qtas<2>(g, st, n, Q, AS, ld, part); break;
    case 3: launch_qtas<3>(g, st, n, Q, AS, ld, part); break;
    case 4: launch_qtas<4>(g, st, n, Q, AS, ld, part); break;
    case 5: launch_qtas<5>(g, st, n, Q, AS, ld, part); break;
    case 6: launch_qtas<6>(g, st, n, Q, AS, ld, part); break;
    case 7: launch_qtas<7>(g, st, n, Q, AS, ld, part); break;
    default: launch_qtas<8>(g, st, n, Q, AS, ld, part); break;
  }
}
void update(int k, bool upd, dim3 g, hipStream_t st, int n, double* x, const double* Q, const double* AQ, double* S,
            double* AS, long ld, const double* r, const KVec& al, const KMat& be, double* part) {
  switch (k) {
    case 1: launch_update<1>(upd, g, st, n, x, Q, AQ, S, AS, ld, r, al, be, part); break;
    case 2: launch_update<2>(upd, g, st, n, x, Q, AQ, S, AS, ld, r, al, be, part); break;
    case 3: launch_update<3>(upd, g, st, n, x, Q, AQ, S, AS, ld, r, al, be, part); break;
    case 4: launch_update<4>(upd, g, st, n, x, Q, AQ, S, AS, ld, r, al, be, part); break;
    case 5: launch_update<5>(upd, g, st, n, x, Q, AQ, S, AS, ld, r, al, be, part); break;
    case 6: launch_update<6>(upd, g, st, n, x, Q, AQ, S, AS, ld, r, al, be, part); break;
    case 7: launch_update<7>(upd, g, st, n, x, Q, AQ, S, AS, ld, r, al, be, part); break;
    default: launch_update<8>(upd, g, st, n, x, Q, AQ, S, AS, ld, r, al, be, part); break;
  }
}

}  // namespace

// power_method_symMatrix (src/power_method.f90:188-437): from y = x (or all
// ones when x = 0), repeat x' = y/||y||, y = A M^-1 x', Emax = <x',y>.
// Returns Emax; vnorm_out (when given) receives the last <y,y>.
double power_method_emax(Plan& P, const double* x0, int itrmax, int* launches, double* vnorm_out) {
  const int n = P.A.n;
  const hipStream_t st = P.stream;
  ensure_vecs(P, 3);
  double* y = P.vec[0].p;
  double* xp = P.vec[1].p;
  double* z = P.vec[2].p;
  const int nb = tiled_nb(n);
  const dim3 grid(nb), block(kNT);
  const SellArgs A = sell_args(P.A);
  if (P.wpart.n < 2 * (size_t)kNBMax + 8) P.wpart.alloc(2 * (size_t)kNBMax + 8);
  double* part = P.wpart.p;
  double* out = P.wpart.p + 2 * kNBMax;
  double h[2];
  auto fetch = [&](int cnt) {
    hipLaunchKernelGGL(k_cb_reduce, dim3(cnt), block, 0, st, (const double*)part, nb, out);
    PARCEL_HIP_CHECK(hipMemcpyAsync(h, out, sizeof(double) * cnt, hipMemcpyDeviceToHost, st));
    PARCEL_HIP_CHECK(hipStreamSynchronize(st));
    ++*launches;
  };
  PARCEL_HIP_CHECK(hipMemsetAsync(xp, 0, sizeof(double) * (size_t)n, st));
  hipLaunchKernelGGL(k_pm_fill_dot, grid, block, 0, st, n, y, x0, 0, part);
  ++*launches;
  fetch(1);
  double vnorm = h[0], emax = 0.0;
  if (vnorm == 0.0) {
    hipLaunchKernelGGL(k_pm_fill_dot, grid, block, 0, st, n, y, x0, 1, part);
    ++*launches;
    fetch(1);
    vnorm = h[0];
  }
  for (int it = 1; it <= itrmax; ++it) {
    const double t1 = 1.0 / sqrt(vnorm);
    hipLaunchKernelGGL(k_pm_scale, grid, block, 0, st, n, t1, (const double*)y, xp);
    precon_apply(P, xp, z, nullptr);
    hipLaunchKernelGGL(k_pm_spmv, grid, block, 0, st, n, A, (const double*)z, y, (const double*)xp, part);
    *launches += 3;
    fetch(2);
    vnorm = h[0];
    emax = h[1];
  }
  PARCEL_HIP_CHECK(hipGetLastError());
  if (vnorm_out) *vnorm_out = vnorm;
  return emax;
}

namespace {
// the eigenpair check of power_method_symMatrix after its loop
// (src/power_method.f90:366-398): with x the normalised iterate and y = A x,
// kz = Emax x (axpy with beta 0) and z = y - kz (axpy 1, -1); partials of
// <kz,kz> (slot 0) and <z,z> (slot 1)
__global__ void __launch_bounds__(kNT) k_pm_err(int n, double emax, const double* __restrict__ x,
                                                const double* __restrict__ y, double* part) {
  __shared__ double sh[kWaves];
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  double a0 = 0.0, a1 = 0.0;
  for (long t = blockIdx.x; t < ntiles; t += nb) {
    const long j = t * kNT + threadIdx.x;
    if (j < n) {
      const double kz = emax * x[j];
      const double z = y[j] - kz;
      a0 = fma(kz, kz, a0);
      a1 = fma(z, z, a1);
    }
  }
  const double s0 = block_sum(a0, sh);
  const double s1 = block_sum(a1, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s0;
    part[kNBMax + blockIdx.x] = s1;
  }
}
}  // namespace

// power_method_symMatrix in full (src/power_method.f90:188-437), the body of
// parcel_dpower and, as the reference compiles it (an `#if 1` branch), of
// parcel_dcaarnoldi
// (KSP_pCAARNOLDI, src/pcagmres.f90:2647-2651): the loop of
// power_method_emax, then x = y / ||y||, y = A x, and the eigenpair error
// ||y - Emax x|| / ||Emax x||.  vec (device, n) receives x.
void power_method_eigen(Plan& P, const double* x0, int itrmax, double* emax, double* errnorm, double* vec) {
  int launches = 0;
  double vnorm = 0.0;
  const double e = power_method_emax(P, x0, itrmax, &launches, &vnorm);
  const int n = P.A.n;
  const hipStream_t st = P.stream;
  double* y = P.vec[0].p;
  const int nb = tiled_nb(n);
  const dim3 grid(nb), block(kNT);
  double* part = P.wpart.p;
  double* out = P.wpart.p + 2 * kNBMax;
  double h[2];
  auto fetch = [&](int cnt) {
    hipLaunchKernelGGL(k_cb_reduce, dim3(cnt), block, 0, st, (const double*)part, nb, out);
    PARCEL_HIP_CHECK(hipMemcpyAsync(h, out, sizeof(double) * cnt, hipMemcpyDeviceToHost, st));
    PARCEL_HIP_CHECK(hipStreamSynchronize(st));
  };
  // vnorm: the last <y,y> of the loop (of the start vector when itrmax < 1)
  const double t1 = 1.0 / sqrt(vnorm);
  hipLaunchKernelGGL(k_pm_scale, grid, block, 0, st, n, t1, (const double*)y, vec);
  hipLaunchKernelGGL(k_pm_spmv, grid, block, 0, st, n, sell_args(P.A), (const double*)vec, y, (const double*)vec,
                     part);
  hipLaunchKernelGGL(k_pm_err, grid, block, 0, st, n, e, (const double*)vec, (const double*)y, part);
  PARCEL_HIP_CHECK(hipGetLastError());
  fetch(2);
  *emax = e;
  *errnorm = sqrt(fabs(h[1])) / sqrt(fabs(h[0]));
}

int cbcg_solve(Plan& P, double* x, const double* b, int itrmax, double rtol, double* hist_host, int* niter,
               int* iret) {
  const int n = P.A.n;
  const hipStream_t st = P.stream;
  const int k = P.prm.CBCG_kstep;
  if (k < 1 || k > kMaxK) throw hip_error("CBCG_kstep must be in 1..8", PARCEL_HIP_ERR_ARG);
  auto t0 = std::chrono::steady_clock::now();
  hipEvent_t ev0, ev1;
  PARCEL_HIP_CHECK(hipEventCreate(&ev0));
  PARCEL_HIP_CHECK(hipEventCreate(&ev1));
  PARCEL_HIP_CHECK(hipEventRecord(ev0, st));
  int launches = 0;

  // eigenvalue bounds (getEigenValue, src/pcbcg.f90:179-196, and the guards
  // of KSP_pCBCG, :1500-1512)
  double Emax = 0.0, Emin = 0.0;
  // flag 2 (CA-Arnoldi) is the power method as the reference compiles it:
  // KSP_pCAARNOLDI's `#if 1` branch calls power_method_symMatrix
  // (src/pcagmres.f90:2647-2651); SET_KSP_PARAM keeps power_method_itrmax
  // at least 1 (src/krylov.f90:32, src/type.f90:1331-1336)
  if (P.prm.CBCG_Eigenflag == 1 || P.prm.CBCG_Eigenflag == 2)
    Emax = power_method_emax(P, x, std::max(1, P.prm.power_method_itrmax), &launches, nullptr);
  if (Emax == Emin) {
    if (Emax < 0.0) Emax = 0.0;
    else Emin = 0.0;
  }
  if (Emax == 0.0 && Emin == 0.0) Emax = 1.0;
  const double eta = 2.0 / (Emax - Emin), zeta = (Emax + Emin) / (Emax - Emin);
  const double eta2 = eta * 2.0, zeta2 = zeta * 2.0;

  plan_workspace(P, 1);
  const long ld = n;
  // Q, S, AQ, AS (k each) and SS (k - 1)
  const size_t nvec = (size_t)(5 * k);
  if (P.basis.n < nvec * (size_t)n) P.basis.alloc(nvec * (size_t)n);
  double* Qb = P.basis.p;
  double* Sb = Qb + (size_t)k * n;
  double* AQb = Sb + (size_t)k * n;
  double* ASb = AQb + (size_t)k * n;
  double* SS = ASb + (size_t)k * n;
  const int nslot = std::max(k * k, k * (k + 1) / 2 + k + 1);
  if (P.wpart.n < (size_t)nslot * kNBMax + 2 * (size_t)nslot)
    P.wpart.alloc((size_t)nslot * kNBMax + 2 * (size_t)nslot);
  double* part = P.wpart.p;
  double* dout = P.wpart.p + (size_t)nslot * kNBMax;
  double* hout = nullptr;
  PARCEL_HIP_CHECK(hipHostMalloc((void**)&hout, sizeof(double) * (size_t)nslot, hipHostMallocDefault));
  const int nb = tiled_nb(n);
  const dim3 grid(nb), block(kNT);
  const SellArgs A = sell_args(P.A);
  auto fetch = [&](int cnt) {
    hipLaunchKernelGGL(k_cb_reduce, dim3(cnt), block, 0, st, (const double*)part, nb, dout);
    PARCEL_HIP_CHECK(hipMemcpyAsync(hout, dout, sizeof(double) * cnt, hipMemcpyDeviceToHost, st));
    PARCEL_HIP_CHECK(hipStreamSynchronize(st));
    ++launches;
  };
  // make_chebyshev_basis (src/pcbcg.f90:49-111): S_0 = M^-1 r,
  // SS_0 = eta A S_0 - zeta r, SS_j = 2eta A S_j - 2zeta SS_{j-1} - SS_{j-2}
  // (SS_{-1} = r), S_{j+1} = M^-1 SS_j
  auto basis = [&](double* S, double* AS) {
    for (int j = 0; j < k; ++j) {
      const double* in = (j == 0) ? P.r.p : SS + (size_t)(j - 1) * n;
      precon_apply(P, in, S + (size_t)j * n, nullptr);
      int mode = 0;
      double e = 0.0, z = 0.0;
      const double *ss1 = nullptr, *ss2 = nullptr;
      if (j + 1 < k) {
        if (j == 0) {
          mode = 1; e = eta; z = zeta; ss1 = P.r.p;
        } else {
          mode = 2; e = eta2; z = zeta2;
          ss1 = SS + (size_t)(j - 1) * n;
          ss2 = (j == 1) ? P.r.p : SS + (size_t)(j - 2) * n;
        }
      }
      hipLaunchKernelGGL(k_cb_spmv, grid, block, 0, st, n, A, (const double*)(S + (size_t)j * n),
                         AS + (size_t)j * n, mode ? SS + (size_t)j * n : nullptr, ss1, ss2, e, z, mode);
      launches += 2;
    }
  };

  std::vector<double> hist((size_t)std::max(itrmax / k, 0) + 1, 0.0);
  std::vector<double> QtAQ((size_t)k * k), Pinv((size_t)k * k), Qtr(k), QtAS((size_t)k * k), beta((size_t)k * k),
      alpha(k);
  const int NU = k * (k + 1) / 2;
  KVec al{};
  KMat be{};

  // r = b - A x; ||b||
  launch_residual(P, x, b, P.r.p, part);
  ++launches;
  hipLaunchKernelGGL(k_cb_reduce, dim3(1), block, 0, st, (const double*)(part + PART_BB * kNBMax), nb, dout);
  PARCEL_HIP_CHECK(hipMemcpyAsync(hout, dout, sizeof(double), hipMemcpyDeviceToHost, st));
  PARCEL_HIP_CHECK(hipStreamSynchronize(st));
  ++launches;
  const double bnorm = sqrt(hout[0]);
  double* Q = Qb;
  double* S = Sb;
  double* AQ = AQb;
  double* AS = ASb;
  basis(Q, AQ);
  update(k, false, grid, st, n, x, nullptr, nullptr, Q, AQ, ld, P.r.p, al, be, part);
  ++launches;
  fetch(NU + k + 1);
  int itr;
  const int nout = itrmax / k;
  for (itr = 1; itr <= nout; ++itr) {
    for (int l = 0; l < k; ++l)
      for (int i = 0; i <= l; ++i) {
        QtAQ[i + (size_t)k * l] = hout[l * (l + 1) / 2 + i];
        QtAQ[l + (size_t)k * i] = hout[l * (l + 1) / 2 + i];
      }
    for (int l = 0; l < k; ++l) Qtr[l] = hout[NU + l];
    double rnorm = hout[NU + k];
    dense::sym_pinv(k, QtAQ.data(), k, Pinv.data(), k);
    dense::gemm_nn(k, 1, k, Pinv.data(), k, Qtr.data(), k, alpha.data(), k);
    for (int l = 0; l < k; ++l) al.c[l] = alpha[l];
    hipLaunchKernelGGL(k_cb_gemv, grid, block, 0, st, n, k, (const double*)AQ, ld, al, P.r.p);
    ++launches;
    rnorm = sqrt(rnorm);
    hist[itr - 1] = rnorm / bnorm;
    if (rnorm / bnorm < rtol) break;
    basis(S, AS);
    qtas(k, grid, st, n, Q, AS, ld, part);
    ++launches;
    fetch(k * k);
    for (int q = 0; q < k * k; ++q) QtAS[q] = hout[q];
    dense::gemm_nn(k, k, k, Pinv.data(), k, QtAS.data(), k, beta.data(), k);
    for (int q = 0; q < k * k; ++q) be.c[q] = beta[q];
    update(k, true, grid, st, n, x, Q, AQ, S, AS, ld, P.r.p, al, be, part);
    ++launches;
    std::swap(Q, S);
    std::swap(AQ, AS);
    fetch(NU + k + 1);
    PARCEL_HIP_CHECK(hipGetLastError());
  }
  PARCEL_HIP_CHECK(hipEventRecord(ev1, st));
  PARCEL_HIP_CHECK(hipStreamSynchronize(st));
  (void)hipHostFree(hout);
  const int nit = std::min(itrmax, itr);  // src/pcbcg.f90:1975-1980
  const int ret = (itrmax <= itr) ? 1 : 0;
  if (nit > 0 && hist_host) {
    for (int i = 0; i < nit; ++i) hist_host[i] = (i < (int)hist.size()) ? hist[i] : 0.0;
  }
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

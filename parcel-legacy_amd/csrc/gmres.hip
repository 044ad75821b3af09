// gmres.hip -- restarted GMRES(m), KSP_pGMRES (src/pgmres.f90:62-525), with
// right preconditioning, MGS or CGS orthogonalisation, Givens rotations
// (GivensRotations, src/pcagmres.f90:190-242) and the restart update
// update_xVec (src/pcagmres.f90:1136-1236).
//
// Vector work stays on the device; the Hessenberg matrix, the rotations and
// the back-substitution live in a device-resident GmState updated by single
// threads, so the host never reads a scalar inside a cycle.  The host
// enqueues one restart cycle at a time (all its kernels return at once when
// the cycle or the solve has stopped) and reads the state once per cycle.
//
// Operation order (the oracle restates the same):
//   restart : r = b - A x ; rnorm = ||r|| ; v1 = (-1/rnorm) r ; e1 = -rnorm
//   inner j : pv = M^-1 v_j ; w = A pv ;
//             MGS: for k = 1..j: h_kj = <w, v_k>; w = fma(-h_kj, v_k, w)
//             CGS: h_kj = <w, v_k> for all k; then w = fma(-h_kj, v_k, w), k = 1..j
//             h_{j+1,j} = sqrt(<w,w>) ; v_{j+1} = (1/h_{j+1,j}) w ; Givens
//   update  : y = H^-1 e (back substitution) ; v0 = sum_k fma(y_k, v_k) ;
//             x0 = M^-1 v0 ; x = x0 + x
// The small dense operations (Givens, back substitution) use plain
// multiply/add as written in the reference, with no contraction.
#include <chrono>
#include <cmath>

#include "kernels.hpp"

namespace parcel {

namespace {

constexpr int kMaxM = 256;  // largest restart length handled on the device

struct GmState {
  double bnorm;
  double rnorm;
  double alpha;     // -1/rnorm of the restart (scale of v1)
  int inner;        // the reference's `inner` after the last executed step
  int cycle_stop;   // the inner loop of the current cycle has exited
  int converged;    // rnorm/bnorm < rtol at the last check
  int done;         // the outer loop has exited
  int niter;
  int iret;
};

struct GmArrays {
  double* H;   // (m+1) x m, column-major, leading dimension m+1
  double* c;   // m+1
  double* s;   // m+1
  double* e;   // m+1
  double* y;   // m+1
  double* h;   // m: CGS coefficients of the current column
};

__device__ __forceinline__ bool gm_skip(const GmState* st) { return st->done || st->cycle_stop; }

// bnorm = sqrt(<b,b>): partials of the residual kernel, slot 0
__global__ void __launch_bounds__(kNT) k_gm_init(const double* part, int nb, GmState* st) {
  __shared__ double sh[kWaves];
  const double bb = reduce_part(part, nb, sh);
  if (threadIdx.x == 0) {
    st->bnorm = sqrt(bb);
    st->rnorm = 0.0;
    st->inner = 1;  // pgmres.f90:230 (inner = 1 before the first cycle)
    st->cycle_stop = 0;
    st->converged = 0;
    st->done = 0;
    st->niter = 0;
    st->iret = 0;
  }
}

// restart of cycle `itr`: rnorm, v1 scale, e, exit test (pgmres.f90:240-275)
__global__ void __launch_bounds__(kNT) k_gm_restart(const double* part, int nb, GmState* st, GmArrays G, int m,
                                                    int itr, int itrmax, double rtol, double* hist, int hist_cap) {
  if (st->done) return;
  __shared__ double sh[kWaves];
  const double rr = reduce_part(part, nb, sh);
  for (int i = threadIdx.x; i < (m + 1) * m; i += blockDim.x) G.H[i] = 0.0;
  for (int i = threadIdx.x; i <= m; i += blockDim.x) { G.c[i] = 0.0; G.s[i] = 0.0; G.e[i] = 0.0; }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double rnorm = sqrt(rr);
    st->rnorm = rnorm;
    st->alpha = -1.0 / rnorm;
    G.e[0] = -rnorm;
    st->cycle_stop = 0;
    if (rnorm / st->bnorm < rtol) {
      st->converged = 1;
      st->done = 1;
      st->cycle_stop = 1;
      const int it = itr + st->inner;
      st->niter = it < itrmax ? it : itrmax;
      st->iret = (itrmax <= itr) ? 1 : 0;
    } else if (itr == 1 && hist_cap > 0) {
      hist[0] = rnorm / st->bnorm;
    }
  }
}

// v1 = alpha r
__global__ void __launch_bounds__(kNT) k_gm_scale(int n, const double* __restrict__ r, double* __restrict__ v,
                                                  const GmState* st) {
  if (st->done) return;
  const double a = st->alpha;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    v[i] = a * r[i];
}

// w = A pv and the partials of <w, v1> (MGS) or nothing (CGS)
__global__ void __launch_bounds__(kNT) k_gm_spmv(int n, SellArgs A, const double* __restrict__ pv,
                                                 double* __restrict__ w, const double* __restrict__ v1, double* part,
                                                 const GmState* st) {
  if (gm_skip(st)) return;
  __shared__ double sh[kWaves];
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  double acc = 0.0;
  for (long k = blockIdx.x; k < ntiles; k += nb) {
    const long i = k * kNT + threadIdx.x;
    if (i < n) {
      const double t = sell_row(A, i, [&](int c) { return pv[c]; });
      w[i] = t;
      if (v1) acc = fma(t, v1[i], acc);
    }
  }
  if (v1) {
    const double v = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
  }
}

// MGS step k of column j: h_kj = reduce(partials of <w, v_k>); w -= h_kj v_k;
// partials of <w, v_{k+1}> (k < j) or <w, w> (k == j) into part_out.
__global__ void __launch_bounds__(kNT) k_gm_mgs(int n, double* __restrict__ w, const double* __restrict__ vk,
                                                const double* __restrict__ vnext, const double* part_in,
                                                double* part_out, int nb_prod, GmState* st, GmArrays G, int m,
                                                int k, int j) {
  if (gm_skip(st)) return;
  __shared__ double sh[kWaves];
  const double h = reduce_part(part_in, nb_prod, sh);
  if (blockIdx.x == 0 && threadIdx.x == 0) G.H[(k - 1) + (j - 1) * (m + 1)] = h;
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  double acc = 0.0;
  for (long t = blockIdx.x; t < ntiles; t += nb) {
    const long i = t * kNT + threadIdx.x;
    if (i < n) {
      const double wi = fma(-h, vk[i], w[i]);
      w[i] = wi;
      acc = fma(wi, vnext ? vnext[i] : wi, acc);
    }
  }
  const double v = block_sum(acc, sh);
  if (threadIdx.x == 0) part_out[blockIdx.x] = v;
}

// CGS: partials of <w, v_k> for k = k0 .. k0+cnt-1 (cnt <= 8), into slots 0..cnt-1
template <int CNT>
__global__ void __launch_bounds__(kNT) k_gm_cgs_dots(int n, const double* __restrict__ w, const double* V,
                                                     long ldv, int k0, double* part, const GmState* st) {
  if (gm_skip(st)) return;
  __shared__ double sh[kWaves];
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  double acc[CNT];
#pragma unroll
  for (int d = 0; d < CNT; ++d) acc[d] = 0.0;
  for (long t = blockIdx.x; t < ntiles; t += nb) {
    const long i = t * kNT + threadIdx.x;
    if (i < n) {
      const double wi = w[i];
#pragma unroll
      for (int d = 0; d < CNT; ++d) acc[d] = fma(wi, V[(long)(k0 - 1 + d) * ldv + i], acc[d]);
    }
  }
#pragma unroll
  for (int d = 0; d < CNT; ++d) {
    const double v = block_sum(acc[d], sh);
    if (threadIdx.x == 0) part[d * kNBMax + blockIdx.x] = v;
  }
}

// one workgroup per coefficient: h_k = reduce(partials of <w, v_k>)
__global__ void __launch_bounds__(kNT) k_gm_cgs_reduce(const double* part, int nb, int k0, GmState* st,
                                                       GmArrays G, int m, int j) {
  if (gm_skip(st)) return;
  __shared__ double sh[kWaves];
  const int d = blockIdx.x;
  const double h = reduce_part(part + d * kNBMax, nb, sh);
  if (threadIdx.x == 0) {
    G.h[k0 - 1 + d] = h;
    G.H[(k0 - 1 + d) + (j - 1) * (m + 1)] = h;
  }
}

// CGS update w = fma(-h_k, v_k, w) for k = 1..j in order, and partials of <w,w>
__global__ void __launch_bounds__(kNT) k_gm_cgs_update(int n, double* __restrict__ w, const double* V, long ldv,
                                                       int j, const double* __restrict__ hv, double* part,
                                                       const GmState* st) {
  if (gm_skip(st)) return;
  __shared__ double sh[kWaves];
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  double acc = 0.0;
  for (long t = blockIdx.x; t < ntiles; t += nb) {
    const long i = t * kNT + threadIdx.x;
    if (i < n) {
      double wi = w[i];
      for (int k = 0; k < j; ++k) wi = fma(-hv[k], V[(long)k * ldv + i], wi);
      w[i] = wi;
      acc = fma(wi, wi, acc);
    }
  }
  const double v = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// h_{j+1,j} = sqrt(<w,w>); v_{j+1} = (1/h) w; workgroup 0 applies the Givens
// rotations to column j and runs the convergence tests (pgmres.f90:385-420)
__global__ void __launch_bounds__(kNT) k_gm_norm(int n, const double* __restrict__ w, double* __restrict__ vnext,
                                                 const double* part, int nb_prod, GmState* st, GmArrays G, int m,
                                                 int j, int itr, int itrmax, double rtol, double* hist, int hist_cap) {
  if (gm_skip(st)) return;
  __shared__ double sh[kWaves];
  const double hn = sqrt(reduce_part(part, nb_prod, sh));
  const double sc = 1.0 / hn;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    vnext[i] = sc * w[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double* Hc = G.H + (long)(j - 1) * (m + 1);  // column j
    Hc[j] = hn;
    // GivensRotations (src/pcagmres.f90:190-242), column i = j
    for (int k = 0; k < j - 1; ++k) {
      const double t1 = Hc[k] * G.c[k] - Hc[k + 1] * G.s[k];
      const double t2 = Hc[k] * G.s[k] + Hc[k + 1] * G.c[k];
      Hc[k] = t1;
      Hc[k + 1] = t2;
    }
    double t;
    if (Hc[j - 1] == 0.0) t = Hc[j] / 1.0 - 300;  // the reference's expression `/1.0d0-300`
    else t = Hc[j] / Hc[j - 1];
    const double ci = 1.0 / sqrt(1.0 + t * t);
    const double si = -ci * t;
    G.c[j - 1] = ci;
    G.s[j - 1] = si;
    const double e1 = ci * G.e[j - 1] - si * G.e[j];
    const double e2 = si * G.e[j - 1] + ci * G.e[j];
    G.e[j - 1] = e1;
    G.e[j] = e2;
    Hc[j - 1] = ci * Hc[j - 1] - si * Hc[j];
    Hc[j] = 0.0;
    const double rnorm = fabs(G.e[j]);
    st->rnorm = rnorm;
    st->inner = j;
    if (itr + j - 1 < hist_cap) hist[itr + j - 1] = rnorm / st->bnorm;
    if (rnorm / st->bnorm < rtol) {
      st->converged = 1;
      st->cycle_stop = 1;
    } else if (itr + j == itrmax) {
      st->cycle_stop = 1;
    }
  }
}

// y = H^-1 e on the leading mm x mm block (update_xVec, pcagmres.f90:1180-1192)
__global__ void k_gm_backsub(GmState* st, GmArrays G, int m) {
  if (st->done) return;
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  // the inner loop ran to completion: the reference's inner is m+1, mm = m
  if (!st->cycle_stop) st->inner = m + 1;
  int mm = st->inner;
  if (m < mm) mm = m;
  for (int i = mm - 1; i >= 0; --i) {
    double yi = G.e[i];
    for (int jj = mm - 1; jj >= i + 1; --jj) yi = yi - G.H[i + jj * (m + 1)] * G.y[jj];
    const double d = G.H[i + i * (m + 1)];
    yi = (d == 0.0) ? yi / 1.0e-300 : yi / d;
    G.y[i] = yi;
  }
  G.y[m] = (double)mm;  // the column count handed to the combination kernel
}

// v0 = sum_{k=1..mm} y_k v_k  (sum_vec_FP64_CPU, pcagmres.f90:54-90)
__global__ void __launch_bounds__(kNT) k_gm_sumvec(int n, const double* V, long ldv, double* __restrict__ v0,
                                                   const GmArrays G, int m, const GmState* st) {
  if (st->done) return;
  const int mm = (int)G.y[m];
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double a = 0.0;
    for (int k = 0; k < mm; ++k) a = fma(G.y[k], V[(long)k * ldv + i], a);
    v0[i] = a;
  }
}

// x = x0 + x; end-of-cycle exit tests (pgmres.f90:455-475)
__global__ void __launch_bounds__(kNT) k_gm_xadd(int n, double* __restrict__ x, const double* __restrict__ x0,
                                                 const GmState* st) {
  if (st->done) return;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    x[i] = x0[i] + x[i];
}

__global__ void k_gm_endcycle(GmState* st, int itr, int itrmax) {
  if (st->done || threadIdx.x != 0 || blockIdx.x != 0) return;
  const int inner = st->inner;
  if (st->converged || itr + inner == itrmax) {
    st->done = 1;
    const int it = itr + inner;
    st->niter = it < itrmax ? it : itrmax;
    st->iret = (itrmax <= itr) ? 1 : 0;
  }
}

}  // namespace

int gmres_solve(Plan& P, double* x, const double* b, int itrmax, double rtol, double* hist_host, int* niter,
                int* iret) {
  const int n = P.A.n;
  const hipStream_t st = P.stream;
  const int m = P.prm.GMRES_m;
  const bool cgs = P.prm.GMRES_GSflag != PARCEL_QR_MGS;  // pgmres.f90:336
  if (m < 1 || m > kMaxM) throw hip_error("GMRES_m must be in 1..256", PARCEL_HIP_ERR_ARG);
  if (itrmax <= 0) {
    // the outer DO loop does not run: niter = min(itrmax, itr + inner) with itr = inner = 1
    *niter = itrmax < 2 ? itrmax : 2;
    *iret = 1;
    P.stats.niter = *niter;
    return PARCEL_HIP_OK;
  }
  plan_workspace(P, itrmax);
  // vectors: V (m+1 columns of n), pv, w, v0, x0
  const long ldv = n;
  if (P.basis.n < (size_t)(m + 1) * (size_t)n) P.basis.alloc((size_t)(m + 1) * (size_t)n);
  ensure_vecs(P, 4);
  double* V = P.basis.p;
  double* pv = P.vec[0].p;
  double* w = P.vec[1].p;
  double* v0 = P.vec[2].p;
  double* x0 = P.vec[3].p;
  // device scalars: state, then H, c, s, e, y, h
  const size_t arr = (size_t)(m + 1) * m + 5 * (size_t)(m + 1);
  const size_t soff = ((sizeof(GmState) + 15) / 16) * 16;
  const size_t bytes = soff + sizeof(double) * arr;
  if (P.solver_state.n < bytes) P.solver_state.alloc(bytes);
  GmState* S = reinterpret_cast<GmState*>(P.solver_state.p);
  double* base = reinterpret_cast<double*>(P.solver_state.p + soff);
  GmArrays G;
  G.H = base;
  G.c = G.H + (size_t)(m + 1) * m;
  G.s = G.c + (m + 1);
  G.e = G.s + (m + 1);
  G.y = G.e + (m + 1);
  G.h = G.y + (m + 1);
  // partial sums: MGS alternates slots 0 and 1; CGS uses 0..7 and 9
  if (P.wpart.n < (size_t)10 * kNBMax) P.wpart.alloc((size_t)10 * kNBMax);
  double* gp = P.wpart.p;
  const int nb = tiled_nb(n);
  const dim3 grid(nb), block(kNT);
  const SellArgs A = sell_args(P.A);
  auto t0 = std::chrono::steady_clock::now();
  hipEvent_t ev0, ev1;
  PARCEL_HIP_CHECK(hipEventCreate(&ev0));
  PARCEL_HIP_CHECK(hipEventCreate(&ev1));
  PARCEL_HIP_CHECK(hipEventRecord(ev0, st));
  // entries the reference never writes (reshistory(itr+m+1) when a completed
  // cycle ends the solve) read as zero
  PARCEL_HIP_CHECK(hipMemsetAsync(P.hist.p, 0, sizeof(double) * (size_t)P.hist_cap, st));
  int launches = 0;
  GmState* h = nullptr;
  PARCEL_HIP_CHECK(hipHostMalloc((void**)&h, sizeof(GmState), hipHostMallocDefault));
  const int* done = &S->done;
  int itr = 1;
  for (; itr <= itrmax; itr += m) {
    // restart
    launch_residual(P, x, b, P.r.p, P.part.p);  // partials of <b,b> and <r,r>
    if (itr == 1) {
      // bnorm = sqrt(<b,b>) (the reference's first <r,r> uses a zero r and is unused)
      hipLaunchKernelGGL(k_gm_init, dim3(1), block, 0, st, (const double*)(P.part.p + PART_BB * kNBMax), nb, S);
      ++launches;
    }
    hipLaunchKernelGGL(k_gm_restart, dim3(1), block, 0, st, (const double*)(P.part.p + PART_RR0 * kNBMax), nb, S,
                       G, m, itr, itrmax, rtol, P.hist.p, P.hist_cap);
    hipLaunchKernelGGL(k_gm_scale, grid, block, 0, st, n, (const double*)P.r.p, V, (const GmState*)S);
    launches += 3;
    for (int j = 1; j <= m; ++j) {
      double* vj = V + (long)(j - 1) * ldv;
      precon_apply(P, vj, pv, nullptr, &S->cycle_stop);
      if (!cgs) {
        hipLaunchKernelGGL(k_gm_spmv, grid, block, 0, st, n, A, (const double*)pv, w, (const double*)V,
                           gp, (const GmState*)S);
        for (int k = 1; k <= j; ++k) {
          const double* vk = V + (long)(k - 1) * ldv;
          const double* vnext = (k < j) ? V + (long)k * ldv : nullptr;
          double* pin = gp + ((k - 1) & 1) * kNBMax;
          double* pout = gp + (k & 1) * kNBMax;
          hipLaunchKernelGGL(k_gm_mgs, grid, block, 0, st, n, w, vk, vnext, (const double*)pin, pout, nb, S, G, m,
                             k, j);
        }
        hipLaunchKernelGGL(k_gm_norm, grid, block, 0, st, n, (const double*)w, V + (long)j * ldv,
                           (const double*)(gp + (j & 1) * kNBMax), nb, S, G, m, j, itr, itrmax, rtol, P.hist.p,
                           P.hist_cap);
        launches += j + 3;
      } else {
        hipLaunchKernelGGL(k_gm_spmv, grid, block, 0, st, n, A, (const double*)pv, w, (const double*)nullptr,
                           gp, (const GmState*)S);
        for (int k0 = 1; k0 <= j; k0 += 8) {
          const int cnt = std::min(8, j - k0 + 1);
#define PARCEL_CGS_DOTS(C)                                                                              \
  case C:                                                                                               \
    hipLaunchKernelGGL(k_gm_cgs_dots<C>, grid, block, 0, st, n, (const double*)w, (const double*)V, ldv, k0, \
                       gp, (const GmState*)S);                                                     \
    break;
          switch (cnt) {
            PARCEL_CGS_DOTS(1) PARCEL_CGS_DOTS(2) PARCEL_CGS_DOTS(3) PARCEL_CGS_DOTS(4)
            PARCEL_CGS_DOTS(5) PARCEL_CGS_DOTS(6) PARCEL_CGS_DOTS(7) PARCEL_CGS_DOTS(8)
          }
#undef PARCEL_CGS_DOTS
          hipLaunchKernelGGL(k_gm_cgs_reduce, dim3(cnt), block, 0, st, (const double*)gp, nb, k0, S, G, m, j);
          launches += 2;
        }
        hipLaunchKernelGGL(k_gm_cgs_update, grid, block, 0, st, n, w, (const double*)V, ldv, j,
                           (const double*)G.h, gp + 9 * kNBMax, (const GmState*)S);
        hipLaunchKernelGGL(k_gm_norm, grid, block, 0, st, n, (const double*)w, V + (long)j * ldv,
                           (const double*)(gp + 9 * kNBMax), nb, S, G, m, j, itr, itrmax, rtol, P.hist.p,
                           P.hist_cap);
        launches += 4;
      }
    }
    // update_xVec
    hipLaunchKernelGGL(k_gm_backsub, dim3(1), dim3(64), 0, st, S, G, m);
    hipLaunchKernelGGL(k_gm_sumvec, grid, block, 0, st, n, (const double*)V, ldv, v0, (const GmArrays)G, m,
                       (const GmState*)S);
    precon_apply(P, v0, x0, nullptr, done);
    hipLaunchKernelGGL(k_gm_xadd, grid, block, 0, st, n, x, (const double*)x0, (const GmState*)S);
    hipLaunchKernelGGL(k_gm_endcycle, dim3(1), dim3(64), 0, st, S, itr, itrmax);
    launches += 5;
    PARCEL_HIP_CHECK(hipGetLastError());
    PARCEL_HIP_CHECK(hipMemcpyAsync(h, S, sizeof(GmState), hipMemcpyDeviceToHost, st));
    PARCEL_HIP_CHECK(hipStreamSynchronize(st));
    if (h->done) break;
  }
  PARCEL_HIP_CHECK(hipEventRecord(ev1, st));
  PARCEL_HIP_CHECK(hipMemcpyAsync(h, S, sizeof(GmState), hipMemcpyDeviceToHost, st));
  PARCEL_HIP_CHECK(hipStreamSynchronize(st));
  int nit, ret;
  if (h->done) {
    nit = h->niter;
    ret = h->iret;
  } else {
    // the outer DO loop ran out (itr > itrmax): niter = min(itrmax, itr + inner)
    const int it = itr + h->inner;
    nit = it < itrmax ? it : itrmax;
    ret = (itrmax <= itr) ? 1 : 0;
  }
  (void)hipHostFree(h);
  if (nit > 0 && hist_host) {
    const int cnt = std::min(nit, P.hist_cap);
    PARCEL_HIP_CHECK(hipMemcpy(hist_host, P.hist.p, sizeof(double) * (size_t)cnt, hipMemcpyDeviceToHost));
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

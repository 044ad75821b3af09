// ddcg.hip -- double-double PCG, ddKSP_pCG (src/ddpcg.f90:247-563).
//
// Vectors are stored as separate hi / lo arrays (coalesced loads of each);
// the matrix is the SELL-64 FP64 matrix plus, when it carries lo parts
// (PRECISION_A = 2), a lo array in the same layout.  Every arithmetic step
// is the reference's dd operation in its operand order:
//   SpMV   row: t = dd_add(dd_mul(A_j, x_c), t)          (ddMatVec_crs)
//   axpy2  x = dd_add(x, dd_mul(alpha, p)); p = dd_add(s, dd_mul(beta, p))
//   r      r = dd_add(dd_mul(-alpha, q), dd_mul(1, r))     (ddaxpy)
//   dots   acc = dd_add(acc, dd_mul(x_i, y_i)), tiled tree with dd_add
// Scalars (bnorm, rnorm, rho, alpha, beta) stay in HBM as dd; the host polls
// a done flag once per chunk of iterations.
#include <chrono>

#include "dd.hpp"
#include "kernels.hpp"

namespace parcel {

namespace {

struct DDState {
  dd bnorm, rnorm, rho0, alpha, beta;
  int done, niter, iret, pad;
};

enum { DS_BB = 0, DS_RR = 1, DS_RHO = 2, DS_GAMMA = 3, DS_RHO1 = 4, DS_RR1 = 5, DS_N = 6 };

// the value of lane (lane ^ O), in registers
template <int O>
__device__ __forceinline__ dd xor_lane_dd(dd v) {
  return mkdd(xor_lane<O>(v.hi), xor_lane<O>(v.lo));
}

// the tiled tree of common.hpp with dd_add(own, other) at every step
__device__ __forceinline__ dd dd_block_sum(dd v, dd* sh) {
  v = dd_add(v, xor_lane_dd<32>(v));
  v = dd_add(v, xor_lane_dd<16>(v));
  v = dd_add(v, xor_lane_dd<8>(v));
  v = dd_add(v, xor_lane_dd<4>(v));
  v = dd_add(v, xor_lane_dd<2>(v));
  v = dd_add(v, xor_lane_dd<1>(v));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  dd s = sh[0];
#pragma unroll
  for (int k = 1; k < kWaves; ++k) s = dd_add(s, sh[k]);
  return s;
}

__device__ __forceinline__ dd dd_reduce_part(const dd* part, int nb, dd* sh) {
  dd a = mkdd(0.0, 0.0);
  for (int j = threadIdx.x; j < nb; j += kNT) a = dd_add(a, part[j]);
  return dd_block_sum(a, sh);
}

struct DDMat {
  SellArgs A;
  const double* vlo;  // lo parts in SELL layout, or null
};

__device__ __forceinline__ dd dd_row(const DDMat& M, long i, const double* xh, const double* xl) {
  const long s = i >> 6;
  const long base = M.A.sptr[s] + (i & 63);
  const int k = M.A.sk[s];
  dd t = mkdd(0.0, 0.0);
  for (int j = 0; j < k; ++j) {
    const long at = base + (long)j * 64;
    const int c = M.A.col[at];
    if (c >= 0) t = dd_add(dd_mul(mkdd(M.A.val[at], M.vlo ? M.vlo[at] : 0.0), mkdd(xh[c], xl[c])), t);
  }
  long e0, e1;
  sell_overflow(M.A, i, e0, e1);
  for (long e = e0; e < e1; ++e) {
    const int c = M.A.col[e];
    t = dd_add(dd_mul(mkdd(M.A.val[e], M.vlo ? M.vlo[e] : 0.0), mkdd(xh[c], xl[c])), t);
  }
  return t;
}

// r = dd_add(dd_mul(1, b), dd_mul(-1, A x)) (ddcalc_res_CRS) and partials of <b,b>, <r,r>
__global__ void __launch_bounds__(kNT) k_dd_residual(int n, DDMat M, const double* xh, const double* xl,
                                                     const double* bh, const double* bl, double* rh, double* rl,
                                                     dd* part) {
  __shared__ dd sh[kWaves];
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  dd a0 = mkdd(0.0, 0.0), a1 = mkdd(0.0, 0.0);
  for (long k = blockIdx.x; k < ntiles; k += nb) {
    const long i = k * kNT + threadIdx.x;
    if (i < n) {
      const dd ax = dd_row(M, i, xh, xl);
      const dd b = mkdd(bh[i], bl[i]);
      const dd r = dd_add(dd_mul(mkdd(1.0, 0.0), b), dd_mul(mkdd(-1.0, 0.0), ax));
      rh[i] = r.hi;
      rl[i] = r.lo;
      a0 = dd_add(a0, dd_mul(b, b));
      a1 = dd_add(a1, dd_mul(r, r));
    }
  }
  const dd s0 = dd_block_sum(a0, sh);
  const dd s1 = dd_block_sum(a1, sh);
  if (threadIdx.x == 0) {
    part[DS_BB * kNBMax + blockIdx.x] = s0;
    part[DS_RR * kNBMax + blockIdx.x] = s1;
  }
}

// partials of <x, y> and, when x2 is set, <x2, y2>
__global__ void __launch_bounds__(kNT) k_dd_dots(int n, const double* xh, const double* xl, const double* yh,
                                                 const double* yl, const double* x2h, const double* x2l,
                                                 const double* y2h, const double* y2l, dd* p1, dd* p2,
                                                 const DDState* st) {
  if (st && st->done) return;
  __shared__ dd sh[kWaves];
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  dd a0 = mkdd(0.0, 0.0), a1 = mkdd(0.0, 0.0);
  for (long k = blockIdx.x; k < ntiles; k += nb) {
    const long i = k * kNT + threadIdx.x;
    if (i < n) {
      a0 = dd_add(a0, dd_mul(mkdd(xh[i], xl[i]), mkdd(yh[i], yl[i])));
      if (x2h) a1 = dd_add(a1, dd_mul(mkdd(x2h[i], x2l[i]), mkdd(y2h[i], y2l[i])));
    }
  }
  const dd s0 = dd_block_sum(a0, sh);
  if (threadIdx.x == 0) p1[blockIdx.x] = s0;
  if (x2h) {
    const dd s1 = dd_block_sum(a1, sh);
    if (threadIdx.x == 0) p2[blockIdx.x] = s1;
  }
}

// bnorm, rnorm from the residual partials; rho0 = <s, r>; hist[0]
__global__ void __launch_bounds__(kNT) k_dd_init(const dd* part, int nb, DDState* st, double* hist) {
  __shared__ dd sh[kWaves];
  const dd bb = dd_reduce_part(part + DS_BB * kNBMax, nb, sh);
  const dd rr = dd_reduce_part(part + DS_RR * kNBMax, nb, sh);
  const dd rho = dd_reduce_part(part + DS_RHO * kNBMax, nb, sh);
  if (threadIdx.x == 0) {
    st->bnorm = dd_sqrt(bb);
    st->rnorm = dd_sqrt(rr);
    st->rho0 = rho;
    st->alpha = mkdd(0.0, 0.0);
    st->beta = mkdd(0.0, 0.0);
    st->done = 0;
    st->niter = 0;
    st->iret = 0;
    hist[0] = st->rnorm.hi / st->bnorm.hi;  // ddpcg.f90:391
  }
}

// ddCG_axpy2: x = dd_add(x, dd_mul(alpha, p)); p = dd_add(s, dd_mul(beta, p))
__global__ void __launch_bounds__(kNT) k_dd_axpy2(int n, double* xh, double* xl, double* ph, double* pl,
                                                  const double* sh_, const double* sl_, const DDState* st) {
  if (st->done) return;
  const dd alpha = st->alpha, beta = st->beta;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const dd p = mkdd(ph[i], pl[i]);
    const dd x = dd_add(mkdd(xh[i], xl[i]), dd_mul(alpha, p));
    xh[i] = x.hi;
    xl[i] = x.lo;
    const dd pn = dd_add(mkdd(sh_[i], sl_[i]), dd_mul(beta, p));
    ph[i] = pn.hi;
    pl[i] = pn.lo;
  }
}

// q = A p and partials of <p, q>
__global__ void __launch_bounds__(kNT) k_dd_spmv_dot(int n, DDMat M, const double* ph, const double* pl,
                                                     double* qh, double* ql, dd* part, const DDState* st) {
  if (st->done) return;
  __shared__ dd sh[kWaves];
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  dd acc = mkdd(0.0, 0.0);
  for (long k = blockIdx.x; k < ntiles; k += nb) {
    const long i = k * kNT + threadIdx.x;
    if (i < n) {
      const dd q = dd_row(M, i, ph, pl);
      qh[i] = q.hi;
      ql[i] = q.lo;
      acc = dd_add(acc, dd_mul(mkdd(ph[i], pl[i]), q));
    }
  }
  const dd s = dd_block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// alpha = rho0 / <p,q>; r = dd_add(dd_mul(-alpha, q), dd_mul(1, r))
__global__ void __launch_bounds__(kNT) k_dd_update_r(int n, double* rh, double* rl, const double* qh,
                                                     const double* ql, const dd* part, int nb_prod, DDState* st) {
  if (st->done) return;
  __shared__ dd sh[kWaves];
  const dd gamma = dd_reduce_part(part, nb_prod, sh);
  const dd alpha = dd_div(st->rho0, gamma);
  if (blockIdx.x == 0 && threadIdx.x == 0) st->alpha = alpha;
  const dd ma = dd_mul(mkdd(-1.0, 0.0), alpha);
  const dd one = mkdd(1.0, 0.0);
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const dd r = dd_add(dd_mul(ma, mkdd(qh[i], ql[i])), dd_mul(one, mkdd(rh[i], rl[i])));
    rh[i] = r.hi;
    rl[i] = r.lo;
  }
}

// rho1, ||r||; exit test on rnorm_hi / bnorm_hi; beta; next history entry
__global__ void __launch_bounds__(kNT) k_dd_check(const dd* p_rho, const dd* p_rr, int nb, DDState* st, int it,
                                                  int itrmax, double rtol, double* hist, int hist_cap) {
  if (st->done) return;
  __shared__ dd sh[kWaves];
  const dd rho1 = dd_reduce_part(p_rho, nb, sh);
  const dd rr = dd_reduce_part(p_rr, nb, sh);
  if (threadIdx.x != 0) return;
  const dd rnorm = dd_sqrt(rr);
  st->rnorm = rnorm;
  const double ratio = rnorm.hi / st->bnorm.hi;
  if (ratio < rtol) {  // ddpcg.f90:494
    st->done = 1;
    st->niter = it < itrmax ? it : itrmax;
    st->iret = (itrmax <= it) ? 1 : 0;
    return;
  }
  st->beta = dd_div(rho1, st->rho0);
  st->rho0 = rho1;
  if (it >= itrmax) {  // the DO loop ends with itr = itrmax + 1
    st->done = 1;
    st->niter = itrmax;
    st->iret = 1;
    return;
  }
  if (it < hist_cap) hist[it] = ratio;
}

// x = dd_add(dd_mul(alpha, p), dd_mul(1, x)) after the loop
__global__ void __launch_bounds__(kNT) k_dd_final(int n, double* xh, double* xl, const double* ph, const double* pl,
                                                  const DDState* st) {
  const dd alpha = st->alpha, one = mkdd(1.0, 0.0);
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const dd x = dd_add(dd_mul(alpha, mkdd(ph[i], pl[i])), dd_mul(one, mkdd(xh[i], xl[i])));
    xh[i] = x.hi;
    xl[i] = x.lo;
  }
}

}  // namespace

int ddcg_solve(Plan& P, double* xh, double* xl, const double* bh, const double* bl, int itrmax, double rtol,
               double* hist_host, int* niter, int* iret) {
  const int n = P.A.n;
  const hipStream_t st = P.stream;
  if (itrmax <= 0) {
    *niter = itrmax < 0 ? itrmax : 0;
    *iret = 1;
    P.stats.niter = *niter;
    return PARCEL_HIP_OK;
  }
  plan_workspace(P, itrmax);
  ensure_vecs(P, 8);
  double *rh = P.vec[0].p, *rl = P.vec[1].p, *ph = P.vec[2].p, *pl = P.vec[3].p;
  double *qh = P.vec[4].p, *ql = P.vec[5].p, *sh = P.vec[6].p, *sl = P.vec[7].p;
  const size_t pbytes = sizeof(dd) * (size_t)DS_N * kNBMax;
  if (P.wpart.n * sizeof(double) < pbytes) P.wpart.alloc(pbytes / sizeof(double));
  dd* part = reinterpret_cast<dd*>(P.wpart.p);
  if (P.solver_state.n < sizeof(DDState)) P.solver_state.alloc(sizeof(DDState));
  DDState* S = reinterpret_cast<DDState*>(P.solver_state.p);
  DDMat M{sell_args(P.A), P.A.val_lo.n ? (const double*)P.A.val_lo.p : nullptr};
  const int nb = tiled_nb(n);
  const dim3 grid(nb), block(kNT);
  auto t0 = std::chrono::steady_clock::now();
  hipEvent_t ev0, ev1;
  PARCEL_HIP_CHECK(hipEventCreate(&ev0));
  PARCEL_HIP_CHECK(hipEventCreate(&ev1));
  PARCEL_HIP_CHECK(hipEventRecord(ev0, st));
  int launches = 0;
  PARCEL_HIP_CHECK(hipMemsetAsync(ph, 0, sizeof(double) * (size_t)n, st));
  PARCEL_HIP_CHECK(hipMemsetAsync(pl, 0, sizeof(double) * (size_t)n, st));
  hipLaunchKernelGGL(k_dd_residual, grid, block, 0, st, n, M, (const double*)xh, (const double*)xl, bh, bl, rh, rl,
                     part);
  precon_apply_dd(P, rh, rl, sh, sl);
  hipLaunchKernelGGL(k_dd_dots, grid, block, 0, st, n, (const double*)sh, (const double*)sl, (const double*)rh,
                     (const double*)rl, (const double*)nullptr, (const double*)nullptr, (const double*)nullptr,
                     (const double*)nullptr, part + DS_RHO * kNBMax, (dd*)nullptr, (const DDState*)nullptr);
  hipLaunchKernelGGL(k_dd_init, dim3(1), block, 0, st, (const dd*)part, nb, S, P.hist.p);
  launches += 4;
  DDState* h = nullptr;
  PARCEL_HIP_CHECK(hipHostMalloc((void**)&h, sizeof(DDState), hipHostMallocDefault));
  const int* done = &S->done;
  int it = 1, chunk = 8;
  while (it <= itrmax) {
    const int last = std::min(itrmax, it + chunk - 1);
    for (; it <= last; ++it) {
      hipLaunchKernelGGL(k_dd_axpy2, grid, block, 0, st, n, xh, xl, ph, pl, (const double*)sh, (const double*)sl,
                         (const DDState*)S);
      hipLaunchKernelGGL(k_dd_spmv_dot, grid, block, 0, st, n, M, (const double*)ph, (const double*)pl, qh, ql,
                         part + DS_GAMMA * kNBMax, (const DDState*)S);
      hipLaunchKernelGGL(k_dd_update_r, grid, block, 0, st, n, rh, rl, (const double*)qh, (const double*)ql,
                         (const dd*)(part + DS_GAMMA * kNBMax), nb, S);
      precon_apply_dd(P, rh, rl, sh, sl, done);
      // ddCG_calc_dot2: rho1 = <r, s>, <r, r>
      hipLaunchKernelGGL(k_dd_dots, grid, block, 0, st, n, (const double*)rh, (const double*)rl, (const double*)sh,
                         (const double*)sl, (const double*)rh, (const double*)rl, (const double*)rh,
                         (const double*)rl, part + DS_RHO1 * kNBMax, part + DS_RR1 * kNBMax, (const DDState*)S);
      hipLaunchKernelGGL(k_dd_check, dim3(1), block, 0, st, (const dd*)(part + DS_RHO1 * kNBMax),
                         (const dd*)(part + DS_RR1 * kNBMax), nb, S, it, itrmax, rtol, P.hist.p, P.hist_cap);
      launches += 6;
    }
    PARCEL_HIP_CHECK(hipGetLastError());
    PARCEL_HIP_CHECK(hipMemcpyAsync(h, S, sizeof(DDState), hipMemcpyDeviceToHost, st));
    PARCEL_HIP_CHECK(hipStreamSynchronize(st));
    if (h->done) break;
    if (chunk < 64) chunk *= 2;
  }
  hipLaunchKernelGGL(k_dd_final, grid, block, 0, st, n, xh, xl, (const double*)ph, (const double*)pl,
                     (const DDState*)S);
  ++launches;
  PARCEL_HIP_CHECK(hipEventRecord(ev1, st));
  PARCEL_HIP_CHECK(hipMemcpyAsync(h, S, sizeof(DDState), hipMemcpyDeviceToHost, st));
  PARCEL_HIP_CHECK(hipStreamSynchronize(st));
  const int nit = h->niter, ret = h->iret;
  (void)hipHostFree(h);
  if (nit > 0 && hist_host)
    PARCEL_HIP_CHECK(hipMemcpy(hist_host, P.hist.p, sizeof(double) * (size_t)nit, hipMemcpyDeviceToHost));
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

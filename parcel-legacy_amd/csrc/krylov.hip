// krylov.hip -- CSR SpMV, reductions and the preconditioned CG solver.
//
// KSP_pCG (src/pcg.f90:224-892) keeps alpha, beta, rho and the norms in host
// variables and moves each one through a 1-element memcpy plus a device
// synchronisation (src/parcel_vec.f90:98-214), ~30 host round trips per
// iteration.  Here every scalar lives in a CGState in HBM: the last workgroup
// of each reducing kernel finishes the reduction and updates the scalars, the
// host only enqueues kernels and polls the `done` flag once per chunk of
// iterations.  Kernels launched after the loop has exited return at once.
//
// One CG iteration is two kernels with point Jacobi / no preconditioner:
//   K1 k_cg_spmv      x += alpha p_old; p_new = s + beta p_old; q = A p_new;
//                     gamma = <p_new, q>; alpha = rho0 / gamma
//   K2 k_cg_update_pj r -= alpha q; s = Dinv r; rho1 = <s,r>; rr = <r,r>;
//                     convergence test; beta = rho1 / rho0
// and four with block Jacobi (r update, forward sweep, backward sweep, dots).
// K1 reads p_new of neighbour rows as fma(beta, p_old[c], s[c]), which is the
// same rounding as the owner's p_new, so p never needs a separate pass.
#include <chrono>
#include <cmath>
#include <vector>

#include "internal.hpp"

namespace parcel {

namespace {

__device__ __forceinline__ void cg_check(CGState* st, double rho1, double rr, int itrmax, double rtol) {
  // pcg.f90:640-690: rnorm = sqrt(<r,r>); exit tests; beta = rho1/rho0
  const double rnorm = sqrt(rr);
  st->rnorm = rnorm;
  st->rho1 = rho1;
  st->rr = rr;
  const double ratio = rnorm / st->bnorm;
  const int itr = st->itr;
  if (ratio < rtol || ratio > 1.0e8 || itr >= itrmax) {
    st->done = 1;
    st->iret = (itr >= itrmax) ? 1 : 0;
  } else {
    st->beta = rho1 / st->rho0;
    st->rho0 = rho1;
  }
}

// r = b - A x ;  bnorm = sqrt(<b,b>) ; rnorm = sqrt(<r,r>)   (calc_res_CRS,
// src/krylov_h.f90:343 and pcg.f90:456-470)
__global__ void __launch_bounds__(kNT) k_residual(int n, const int* __restrict__ rowptr,
                                                  const int* __restrict__ col,
                                                  const double* __restrict__ val,
                                                  const double* __restrict__ x,
                                                  const double* __restrict__ b, double* __restrict__ r,
                                                  double* part, unsigned* counter, CGState* st) {
  __shared__ double sh[kWaves];
  __shared__ int flag;
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  double a0 = 0.0, a1 = 0.0;
  for (long k = blockIdx.x; k < ntiles; k += nb) {
    const long i = k * kNT + threadIdx.x;
    if (i < n) {
      double t = 0.0;
      for (int j = rowptr[i]; j < rowptr[i + 1]; ++j) t = fma(val[j], x[col[j]], t);
      const double bi = b[i];
      const double ri = bi - t;
      r[i] = ri;
      a0 = fma(bi, bi, a0);
      a1 = fma(ri, ri, a1);
    }
  }
  double s0 = block_sum(a0, sh);
  double s1 = block_sum(a1, sh);
  if (threadIdx.x == 0) { part[blockIdx.x] = s0; part[kNBMax + blockIdx.x] = s1; }
  if (arrive_last(counter, &flag)) {
    double bb = reduce_partials(part, nb, sh);
    double rr = reduce_partials(part + kNBMax, nb, sh);
    if (threadIdx.x == 0) {
      st->bnorm = sqrt(bb);
      st->rnorm = sqrt(rr);
    }
  }
}

// rho0 = <s,r>; alpha = beta = 0 (pcg.f90:478-490)
__global__ void __launch_bounds__(kNT) k_cg_init(int n, const double* __restrict__ s,
                                                 const double* __restrict__ r, double* part,
                                                 unsigned* counter, CGState* st) {
  __shared__ double sh[kWaves];
  __shared__ int flag;
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  double a0 = 0.0;
  for (long k = blockIdx.x; k < ntiles; k += nb) {
    const long i = k * kNT + threadIdx.x;
    if (i < n) a0 = fma(s[i], r[i], a0);
  }
  double s0 = block_sum(a0, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s0;
  if (arrive_last(counter, &flag)) {
    double rho0 = reduce_partials(part, nb, sh);
    if (threadIdx.x == 0) {
      st->rho0 = rho0;
      st->alpha = 0.0;
      st->beta = 0.0;
      st->gamma = 0.0;
      st->itr = 0;
      st->done = 0;
      st->iret = 0;
    }
  }
}

// K1: fused CG_axpy2 (pcg.f90:8) + SpMV (matrix.f90:9051) + gamma = <p,q>
__global__ void __launch_bounds__(kNT) k_cg_spmv(int n, const int* __restrict__ rowptr,
                                                 const int* __restrict__ col,
                                                 const double* __restrict__ val, double* __restrict__ x,
                                                 const double* __restrict__ p_old,
                                                 const double* __restrict__ s, double* __restrict__ p_new,
                                                 double* __restrict__ q, double* part, unsigned* counter,
                                                 CGState* st, double* hist, int hist_cap) {
  if (st->done) return;
  __shared__ double sh[kWaves];
  __shared__ int flag;
  const double alpha = st->alpha, beta = st->beta;
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  double acc = 0.0;
  for (long k = blockIdx.x; k < ntiles; k += nb) {
    const long i = k * kNT + threadIdx.x;
    if (i < n) {
      const double po = p_old[i];
      x[i] = fma(alpha, po, x[i]);
      const double pn = fma(beta, po, s[i]);
      p_new[i] = pn;
      double t = 0.0;
      const int j1 = rowptr[i + 1];
      for (int j = rowptr[i]; j < j1; ++j) {
        const int c = col[j];
        t = fma(val[j], fma(beta, p_old[c], s[c]), t);
      }
      q[i] = t;
      acc = fma(pn, t, acc);
    }
  }
  double s0 = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s0;
  if (arrive_last(counter, &flag)) {
    double gamma = reduce_partials(part, nb, sh);
    if (threadIdx.x == 0) {
      const int itr = st->itr + 1;
      if (itr - 1 < hist_cap) hist[itr - 1] = st->rnorm / st->bnorm;  // pcg.f90:512-518
      st->gamma = gamma;
      st->alpha = st->rho0 / gamma;
      st->itr = itr;
    }
  }
}

// K2 (point Jacobi / none): r -= alpha q; s = Dinv r; rho1, rr; exit test
__global__ void __launch_bounds__(kNT) k_cg_update_pj(int n, double* __restrict__ r,
                                                      const double* __restrict__ q,
                                                      const double* __restrict__ dinv,
                                                      double* __restrict__ s, double* part,
                                                      unsigned* counter, CGState* st, int itrmax,
                                                      double rtol) {
  if (st->done) return;
  __shared__ double sh[kWaves];
  __shared__ int flag;
  const double alpha = st->alpha;
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  double a0 = 0.0, a1 = 0.0;
  for (long k = blockIdx.x; k < ntiles; k += nb) {
    const long i = k * kNT + threadIdx.x;
    if (i < n) {
      const double ri = fma(-alpha, q[i], r[i]);
      r[i] = ri;
      const double si = dinv ? ri * dinv[i] : ri;
      s[i] = si;
      a0 = fma(si, ri, a0);
      a1 = fma(ri, ri, a1);
    }
  }
  double s0 = block_sum(a0, sh);
  double s1 = block_sum(a1, sh);
  if (threadIdx.x == 0) { part[blockIdx.x] = s0; part[kNBMax + blockIdx.x] = s1; }
  if (arrive_last(counter, &flag)) {
    double rho1 = reduce_partials(part, nb, sh);
    double rr = reduce_partials(part + kNBMax, nb, sh);
    if (threadIdx.x == 0) cg_check(st, rho1, rr, itrmax, rtol);
  }
}

// r -= alpha q (block-Jacobi path: the sweeps follow)
__global__ void __launch_bounds__(kNT) k_cg_update_r(int n, double* __restrict__ r,
                                                     const double* __restrict__ q, const CGState* st) {
  if (st->done) return;
  const double alpha = st->alpha;
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  for (long k = blockIdx.x; k < ntiles; k += nb) {
    const long i = k * kNT + threadIdx.x;
    if (i < n) r[i] = fma(-alpha, q[i], r[i]);
  }
}

// rho1 = <s,r>, rr = <r,r> (CG_calc_dot2, pcg.f90:30) + exit test
__global__ void __launch_bounds__(kNT) k_cg_dot2(int n, const double* __restrict__ s,
                                                 const double* __restrict__ r, double* part,
                                                 unsigned* counter, CGState* st, int itrmax, double rtol) {
  if (st->done) return;
  __shared__ double sh[kWaves];
  __shared__ int flag;
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  double a0 = 0.0, a1 = 0.0;
  for (long k = blockIdx.x; k < ntiles; k += nb) {
    const long i = k * kNT + threadIdx.x;
    if (i < n) {
      const double ri = r[i];
      a0 = fma(s[i], ri, a0);
      a1 = fma(ri, ri, a1);
    }
  }
  double s0 = block_sum(a0, sh);
  double s1 = block_sum(a1, sh);
  if (threadIdx.x == 0) { part[blockIdx.x] = s0; part[kNBMax + blockIdx.x] = s1; }
  if (arrive_last(counter, &flag)) {
    double rho1 = reduce_partials(part, nb, sh);
    double rr = reduce_partials(part + kNBMax, nb, sh);
    if (threadIdx.x == 0) cg_check(st, rho1, rr, itrmax, rtol);
  }
}

// final x += alpha p (pcg.f90:733) with p the buffer written last
__global__ void k_cg_final(int n, double* __restrict__ x, const double* __restrict__ p0,
                           const double* __restrict__ p1, const CGState* st) {
  const double* p = (st->itr & 1) ? p1 : p0;
  const double alpha = st->alpha;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    x[i] = fma(alpha, p[i], x[i]);
}

__global__ void __launch_bounds__(kNT) k_spmv(int n, const int* __restrict__ rowptr,
                                              const int* __restrict__ col, const double* __restrict__ val,
                                              const double* __restrict__ x, double* __restrict__ y) {
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  for (long k = blockIdx.x; k < ntiles; k += nb) {
    const long i = k * kNT + threadIdx.x;
    if (i < n) {
      double t = 0.0;
      for (int j = rowptr[i]; j < rowptr[i + 1]; ++j) t = fma(val[j], x[col[j]], t);
      y[i] = t;
    }
  }
}

__global__ void __launch_bounds__(kNT) k_dot(int n, const double* __restrict__ x,
                                             const double* __restrict__ y, double* part,
                                             unsigned* counter, double* out) {
  __shared__ double sh[kWaves];
  __shared__ int flag;
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  double a0 = 0.0;
  for (long k = blockIdx.x; k < ntiles; k += nb) {
    const long i = k * kNT + threadIdx.x;
    if (i < n) a0 = fma(x[i], y[i], a0);
  }
  double s0 = block_sum(a0, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s0;
  if (arrive_last(counter, &flag)) {
    double v = reduce_partials(part, nb, sh);
    if (threadIdx.x == 0) *out = v;
  }
}

}  // namespace

void plan_workspace(Plan& P, int itrmax) {
  const int n = P.A.n;
  if (P.r.n < (size_t)n) {
    P.r.alloc(n); P.q.alloc(n); P.s.alloc(n); P.p0.alloc(n); P.p1.alloc(n);
  }
  if (!P.cg.p) P.cg.alloc(1);
  if (!P.part.p) P.part.alloc(4 * kNBMax);
  if (!P.counter.p) {
    P.counter.alloc(8);
    PARCEL_HIP_CHECK(hipMemsetAsync(P.counter.p, 0, 8 * sizeof(unsigned), P.stream));
  }
  if (P.hist_cap < itrmax) {
    P.hist.alloc(itrmax > 0 ? itrmax : 1);
    P.hist_cap = itrmax;
  }
}

void spmv(Plan& P, const double* x, double* y) {
  const int n = P.A.n;
  hipLaunchKernelGGL(k_spmv, dim3(tiled_nb(n)), dim3(kNT), 0, P.stream, n, P.A.rowptr.p, P.A.col.p,
                     P.A.val.p, x, y);
  PARCEL_HIP_CHECK(hipGetLastError());
}

double dot_device(int n, const double* x, const double* y, hipStream_t st) {
  DevBuf<double> part(kNBMax + 1);
  DevBuf<unsigned> cnt(1);
  PARCEL_HIP_CHECK(hipMemsetAsync(cnt.p, 0, sizeof(unsigned), st));
  hipLaunchKernelGGL(k_dot, dim3(tiled_nb(n)), dim3(kNT), 0, st, n, x, y, part.p, cnt.p, part.p + kNBMax);
  PARCEL_HIP_CHECK(hipGetLastError());
  double h = 0.0;
  PARCEL_HIP_CHECK(hipMemcpyAsync(&h, part.p + kNBMax, sizeof(double), hipMemcpyDeviceToHost, st));
  PARCEL_HIP_CHECK(hipStreamSynchronize(st));
  return h;
}

namespace {

struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
  int it = 0;  // iteration the pair belongs to
};

struct Timer {
  bool on = false;
  std::vector<EventPair> spmv, prec, vec;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  hipEvent_t get() {
    if (used == pool.size()) {
      hipEvent_t e;
      PARCEL_HIP_CHECK(hipEventCreate(&e));
      pool.push_back(e);
    }
    return pool[used++];
  }
  ~Timer() {
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};

}  // namespace

int cg_solve(Plan& P, double* x, const double* b, int itrmax, double rtol, double* hist_host,
             int* niter, int* iret) {
  const int n = P.A.n;
  const hipStream_t st = P.stream;
  plan_workspace(P, itrmax);
  auto t0 = std::chrono::steady_clock::now();
  const int nb = tiled_nb(n);
  const dim3 grid(nb), block(kNT);
  CGState* S = P.cg.p;
  const int* done = &S->done;
  double* part = P.part.p;
  unsigned* cnt = P.counter.p;
  int launches = 0;
  Timer T;
  T.on = P.prm.profile == 1;
  hipEvent_t ev0, ev1;
  PARCEL_HIP_CHECK(hipEventCreate(&ev0));
  PARCEL_HIP_CHECK(hipEventCreate(&ev1));
  PARCEL_HIP_CHECK(hipEventRecord(ev0, st));

  PARCEL_HIP_CHECK(hipMemsetAsync(P.p0.p, 0, sizeof(double) * (size_t)n, st));
  hipLaunchKernelGGL(k_residual, grid, block, 0, st, n, P.A.rowptr.p, P.A.col.p, P.A.val.p, x, b,
                     P.r.p, part, cnt, S);
  precon_apply(P, P.r.p, P.s.p, nullptr);
  hipLaunchKernelGGL(k_cg_init, grid, block, 0, st, n, P.s.p, P.r.p, part, cnt, S);
  launches += 3;
  PARCEL_HIP_CHECK(hipGetLastError());

  CGState* h = nullptr;
  PARCEL_HIP_CHECK(hipHostMalloc((void**)&h, sizeof(CGState), hipHostMallocDefault));
  h->done = 0;
  h->itr = 0;
  int enq = 0;
  int chunk = 8;
  const bool bj = P.M.kind == PRECON_BJ;
  const double* dinv = P.M.kind == PRECON_PJ ? P.M.dinv.p : nullptr;
  while (enq < itrmax) {
    const int k = std::min(chunk, itrmax - enq);
    for (int j = 1; j <= k; ++j) {
      const int it = enq + j;
      double* pold = (it - 1) & 1 ? P.p1.p : P.p0.p;
      double* pnew = it & 1 ? P.p1.p : P.p0.p;
      hipEvent_t ea = nullptr, eb = nullptr, ec = nullptr, ed = nullptr;
      if (T.on) { ea = T.get(); PARCEL_HIP_CHECK(hipEventRecord(ea, st)); }
      hipLaunchKernelGGL(k_cg_spmv, grid, block, 0, st, n, P.A.rowptr.p, P.A.col.p, P.A.val.p, x, pold,
                         P.s.p, pnew, P.q.p, part, cnt, S, P.hist.p, P.hist_cap);
      if (T.on) { eb = T.get(); PARCEL_HIP_CHECK(hipEventRecord(eb, st)); T.spmv.push_back({ea, eb, it}); }
      if (!bj) {
        hipLaunchKernelGGL(k_cg_update_pj, grid, block, 0, st, n, P.r.p, P.q.p, dinv, P.s.p, part, cnt, S,
                           itrmax, rtol);
        if (T.on) { ec = T.get(); PARCEL_HIP_CHECK(hipEventRecord(ec, st)); T.vec.push_back({eb, ec, it}); }
        launches += 2;
      } else {
        hipLaunchKernelGGL(k_cg_update_r, grid, block, 0, st, n, P.r.p, P.q.p, S);
        if (T.on) { ec = T.get(); PARCEL_HIP_CHECK(hipEventRecord(ec, st)); T.vec.push_back({eb, ec, it}); }
        precon_apply(P, P.r.p, P.s.p, done);
        if (T.on) { ed = T.get(); PARCEL_HIP_CHECK(hipEventRecord(ed, st)); T.prec.push_back({ec, ed, it}); }
        hipLaunchKernelGGL(k_cg_dot2, grid, block, 0, st, n, P.s.p, P.r.p, part, cnt, S, itrmax, rtol);
        if (T.on) {
          hipEvent_t ee = T.get(); PARCEL_HIP_CHECK(hipEventRecord(ee, st));
          T.vec.push_back({ed, ee, it});
        }
        launches += 5;
      }
    }
    PARCEL_HIP_CHECK(hipGetLastError());
    enq += k;
    PARCEL_HIP_CHECK(hipMemcpyAsync(h, S, sizeof(CGState), hipMemcpyDeviceToHost, st));
    PARCEL_HIP_CHECK(hipStreamSynchronize(st));
    if (h->done) break;
    if (chunk < 64) chunk *= 2;
  }
  if (itrmax > 0) {
    hipLaunchKernelGGL(k_cg_final, grid, block, 0, st, n, x, P.p0.p, P.p1.p, S);
    launches++;
  }
  PARCEL_HIP_CHECK(hipEventRecord(ev1, st));
  PARCEL_HIP_CHECK(hipMemcpyAsync(h, S, sizeof(CGState), hipMemcpyDeviceToHost, st));
  PARCEL_HIP_CHECK(hipStreamSynchronize(st));
  int nit = itrmax > 0 ? h->itr : 0;
  int ret = itrmax > 0 ? h->iret : 1;
  if (itrmax > 0 && !h->done) { ret = 1; }
  if (nit > 0 && hist_host)
    PARCEL_HIP_CHECK(hipMemcpy(hist_host, P.hist.p, sizeof(double) * (size_t)nit, hipMemcpyDeviceToHost));
  float ms = 0.f;
  PARCEL_HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
  P.stats.solve_device_ms = ms;
  P.stats.solve_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  P.stats.niter = nit;
  P.stats.launches = launches;
  if (T.on) {
    auto sum = [&](std::vector<EventPair>& v, double& out, int& calls) {
      out = 0.0; calls = 0;
      for (size_t i = 0; i < v.size(); ++i) {
        if (v[i].it > nit) continue;
        float t = 0.f;
        PARCEL_HIP_CHECK(hipEventElapsedTime(&t, v[i].a, v[i].b));
        out += t; calls++;
      }
    };
    sum(T.spmv, P.stats.spmv_ms, P.stats.spmv_calls);
    sum(T.prec, P.stats.precon_ms, P.stats.precon_calls);
    sum(T.vec, P.stats.vec_ms, P.stats.vec_calls);
  }
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  (void)hipHostFree(h);
  *niter = nit;
  *iret = ret;
  return PARCEL_HIP_OK;
}

double time_kernel(Plan& P, int which, int reps) {
  const int n = P.A.n;
  const hipStream_t st = P.stream;
  plan_workspace(P, 1);
  const int nb = tiled_nb(n);
  CGState h{};
  h.alpha = 0.0; h.beta = 0.5; h.rho0 = 1.0; h.bnorm = 1.0; h.rnorm = 1.0;
  PARCEL_HIP_CHECK(hipMemcpyAsync(P.cg.p, &h, sizeof(CGState), hipMemcpyHostToDevice, st));
  PARCEL_HIP_CHECK(hipMemsetAsync(P.p0.p, 0, sizeof(double) * (size_t)n, st));
  PARCEL_HIP_CHECK(hipMemsetAsync(P.s.p, 0, sizeof(double) * (size_t)n, st));
  PARCEL_HIP_CHECK(hipMemsetAsync(P.r.p, 0, sizeof(double) * (size_t)n, st));
  if (P.x.n < (size_t)n) P.x.alloc(n);
  PARCEL_HIP_CHECK(hipMemsetAsync(P.x.p, 0, sizeof(double) * (size_t)n, st));
  hipEvent_t a, b;
  PARCEL_HIP_CHECK(hipEventCreate(&a));
  PARCEL_HIP_CHECK(hipEventCreate(&b));
  auto launch = [&](int i) {
    if (which == 1) {
      hipLaunchKernelGGL(k_cg_spmv, dim3(nb), dim3(kNT), 0, st, n, P.A.rowptr.p, P.A.col.p, P.A.val.p,
                         P.x.p, (i & 1) ? P.p1.p : P.p0.p, P.s.p, (i & 1) ? P.p0.p : P.p1.p, P.q.p,
                         P.part.p, P.counter.p, P.cg.p, P.hist.p, 0);
    } else {
      precon_apply(P, P.r.p, P.s.p, nullptr);
    }
  };
  launch(0);
  PARCEL_HIP_CHECK(hipStreamSynchronize(st));
  PARCEL_HIP_CHECK(hipEventRecord(a, st));
  for (int i = 0; i < reps; ++i) launch(i + 1);
  PARCEL_HIP_CHECK(hipEventRecord(b, st));
  PARCEL_HIP_CHECK(hipEventSynchronize(b));
  float ms = 0.f;
  PARCEL_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms / reps;
}

}  // namespace parcel

// bicg.hip -- preconditioned BiCGStab, KSP_pBiCG (src/pbicg.f90:280-698).
//
// Same device-resident scheme as the CG solver: vector kernels write fixed-
// order partial sums, one-workgroup scalar kernels reduce them and update
// the BiState in HBM, and every kernel returns at once when the loop has
// ended.  Per iteration (reference order, src/pbicg.f90:460-640):
//   pp = M^-1 p ; q = A pp ; c2 = <r0,q> ; alpha = c1/c2
//   e = r - alpha q ; pe = M^-1 e ; v = A pe ; c3 = <e,v>/<v,v>
//   x = x + alpha pp + c3 pe ; r = e - c3 v ; c1 = <r,r0> ; rnorm = ||r||
//   exit if rnorm/bnorm < rtol ; beta = c1/(c3 c2) ; p = r - beta c3 q + beta p
// The reference evaluates the 3-term updates with axpy_vec3
// (src/vec.f90:971); each is fused here as nested fma() in the order the
// oracle states.
#include <chrono>
#include <cmath>

#include "kernels.hpp"

namespace parcel {

namespace {

struct BiState {
  double bnorm, rnorm, c1, c2, c3, alpha, beta;
  int itr, done, niter, iret;
};

enum { BP_BB = 0, BP_RR0 = 1, BP_C2 = 2, BP_EV = 3, BP_VV = 4, BP_C1 = 5, BP_RR = 6 };

// y = A x (optionally with x = M^-1 input already applied) and the partials
// of <u, y> and (if two) <y, y>.
template <int ND>
__global__ void __launch_bounds__(kNT) k_spmv_dots(int n, SellArgs A, const double* __restrict__ x,
                                                   double* __restrict__ y, const double* __restrict__ u,
                                                   double* part, int slot0, int slot1, const int* skip) {
  if (*skip) return;
  __shared__ double sh[kWaves];
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  double acc[2] = {0.0, 0.0};
  for (long k = blockIdx.x; k < ntiles; k += nb) {
    const long i = k * kNT + threadIdx.x;
    if (i < n) {
      const double t = sell_row(A, i, [&](int c) { return x[c]; });
      y[i] = t;
      acc[0] = fma(u[i], t, acc[0]);
      if (ND > 1) acc[1] = fma(t, t, acc[1]);
    }
  }
  const int slots[2] = {slot0, slot1};
  store_partials<ND>(part, slots, acc, sh);
}

__global__ void __launch_bounds__(kNT) k_bi_init(const double* part, int nb, BiState* st) {
  __shared__ double sh[kWaves];
  const double bb = reduce_part(part + BP_BB * kNBMax, nb, sh);
  const double rr = reduce_part(part + BP_RR0 * kNBMax, nb, sh);
  if (threadIdx.x == 0) {
    st->bnorm = sqrt(bb);
    st->rnorm = sqrt(rr);
    st->c1 = rr;  // <r0,r0> with r0 = r: the same partials, the same order
    st->c2 = st->c3 = st->alpha = st->beta = 0.0;
    st->itr = 0;
    st->done = 0;
    st->niter = 0;
    st->iret = 0;
  }
}

// top of iteration: history, c2, alpha (pbicg.f90:470-500)
__global__ void __launch_bounds__(kNT) k_bi_alpha(const double* part, int nb, BiState* st, double* hist,
                                                  int hist_cap) {
  if (st->done) return;
  __shared__ double sh[kWaves];
  const double c2 = reduce_part(part + BP_C2 * kNBMax, nb, sh);
  if (threadIdx.x == 0) {
    const int itr = st->itr + 1;
    if (itr - 1 < hist_cap) hist[itr - 1] = st->rnorm / st->bnorm;
    st->itr = itr;
    st->c2 = c2;
    st->alpha = st->c1 / c2;
  }
}

// e = r - alpha q
__global__ void __launch_bounds__(kNT) k_bi_e(int n, const double* __restrict__ r, const double* __restrict__ q,
                                              double* __restrict__ e, const BiState* st) {
  if (st->done) return;
  const double alpha = st->alpha;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    e[i] = fma(-alpha, q[i], r[i]);
}

// c3 = <e,v> / <v,v>
__global__ void __launch_bounds__(kNT) k_bi_c3(const double* part, int nb, BiState* st) {
  if (st->done) return;
  __shared__ double sh[kWaves];
  const double ev = reduce_part(part + BP_EV * kNBMax, nb, sh);
  const double vv = reduce_part(part + BP_VV * kNBMax, nb, sh);
  if (threadIdx.x == 0) st->c3 = ev / vv;
}

// x += alpha pp + c3 pe ; r = e - c3 v ; partials <r,r0>, <r,r>  (BiCG_axpy3_dot)
__global__ void __launch_bounds__(kNT) k_bi_update(int n, double* __restrict__ x, const double* __restrict__ pp,
                                                   const double* __restrict__ pe, double* __restrict__ r,
                                                   const double* __restrict__ e, const double* __restrict__ v,
                                                   const double* __restrict__ r0, double* part, const BiState* st) {
  if (st->done) return;
  __shared__ double sh[kWaves];
  const double alpha = st->alpha, c3 = st->c3;
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  double acc[2] = {0.0, 0.0};
  for (long k = blockIdx.x; k < ntiles; k += nb) {
    const long i = k * kNT + threadIdx.x;
    if (i < n) {
      x[i] = fma(c3, pe[i], fma(alpha, pp[i], x[i]));
      const double ri = fma(-c3, v[i], e[i]);
      r[i] = ri;
      acc[0] = fma(ri, r0[i], acc[0]);
      acc[1] = fma(ri, ri, acc[1]);
    }
  }
  const int slots[2] = {BP_C1, BP_RR};
  store_partials<2>(part, slots, acc, sh);
}

// exit test and beta (pbicg.f90:590-625)
__global__ void __launch_bounds__(kNT) k_bi_beta(const double* part, int nb, BiState* st, int itrmax,
                                                 double rtol) {
  if (st->done) return;
  __shared__ double sh[kWaves];
  const double c1 = reduce_part(part + BP_C1 * kNBMax, nb, sh);
  const double rr = reduce_part(part + BP_RR * kNBMax, nb, sh);
  if (threadIdx.x == 0) {
    const double rnorm = sqrt(rr);
    st->rnorm = rnorm;
    st->c1 = c1;
    const int itr = st->itr;
    if (rnorm / st->bnorm < rtol || itr >= itrmax) {
      st->done = 1;
      st->niter = itr;                    // min(itrmax, itr)
      st->iret = (itrmax <= itr) ? 1 : 0;  // pbicg.f90:672-676
    } else {
      st->beta = c1 / (st->c3 * st->c2);
    }
  }
}

// p = r - beta c3 q + beta p
__global__ void __launch_bounds__(kNT) k_bi_p(int n, double* __restrict__ p, const double* __restrict__ r,
                                              const double* __restrict__ q, const BiState* st) {
  if (st->done) return;
  const double beta = st->beta;
  const double bc = -(beta * st->c3);
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    p[i] = fma(beta, p[i], fma(bc, q[i], r[i]));
}

}  // namespace

int bicg_solve(Plan& P, double* x, const double* b, int itrmax, double rtol, double* hist_host, int* niter,
               int* iret) {
  const int n = P.A.n;
  const hipStream_t st = P.stream;
  plan_workspace(P, itrmax);
  ensure_vecs(P, 8);
  if (P.solver_state.n < sizeof(BiState)) P.solver_state.alloc(sizeof(BiState));
  auto t0 = std::chrono::steady_clock::now();
  BiState* S = reinterpret_cast<BiState*>(P.solver_state.p);
  const int* done = &S->done;
  double* r = P.r.p;
  double* r0 = P.vec[0].p;
  double* p = P.vec[1].p;
  double* q = P.vec[2].p;
  double* pp = P.vec[3].p;
  double* e = P.vec[4].p;
  double* pe = P.vec[5].p;
  double* v = P.vec[6].p;
  double* part = P.part.p;
  const int nb = tiled_nb(n);
  const SellArgs A = sell_args(P.A);
  hipEvent_t ev0, ev1;
  PARCEL_HIP_CHECK(hipEventCreate(&ev0));
  PARCEL_HIP_CHECK(hipEventCreate(&ev1));
  PARCEL_HIP_CHECK(hipEventRecord(ev0, st));
  int launches = 0;
  if (itrmax <= 0) {
    *niter = 0;
    *iret = 1;
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    return PARCEL_HIP_OK;
  }
  // r = b - A x ; bnorm ; rnorm ; r0 = r ; p = r ; c1 = <r0,r0>
  launch_residual(P, x, b, r, part);
  PARCEL_HIP_CHECK(hipMemcpyAsync(r0, r, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, st));
  PARCEL_HIP_CHECK(hipMemcpyAsync(p, r, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_bi_init, dim3(1), dim3(kNT), 0, st, (const double*)part, nb, S);
  launches += 2;
  BiState* h = nullptr;
  PARCEL_HIP_CHECK(hipHostMalloc((void**)&h, sizeof(BiState), hipHostMallocDefault));
  int enq = 0, chunk = 8;
  const dim3 grid(nb), block(kNT);
  while (enq < itrmax) {
    const int k = std::min(chunk, itrmax - enq);
    for (int j = 0; j < k; ++j) {
      precon_apply(P, p, pp, nullptr, done);
      hipLaunchKernelGGL(k_spmv_dots<1>, grid, block, 0, st, n, A, (const double*)pp, q, (const double*)r0, part,
                         (int)BP_C2, 0, done);
      hipLaunchKernelGGL(k_bi_alpha, dim3(1), block, 0, st, (const double*)part, nb, S, P.hist.p, P.hist_cap);
      hipLaunchKernelGGL(k_bi_e, grid, block, 0, st, n, (const double*)r, (const double*)q, e, (const BiState*)S);
      precon_apply(P, e, pe, nullptr, done);
      hipLaunchKernelGGL(k_spmv_dots<2>, grid, block, 0, st, n, A, (const double*)pe, v, (const double*)e, part,
                         (int)BP_EV, (int)BP_VV, done);
      hipLaunchKernelGGL(k_bi_c3, dim3(1), block, 0, st, (const double*)part, nb, S);
      hipLaunchKernelGGL(k_bi_update, grid, block, 0, st, n, x, (const double*)pp, (const double*)pe, r,
                         (const double*)e, (const double*)v, (const double*)r0, part, (const BiState*)S);
      hipLaunchKernelGGL(k_bi_beta, dim3(1), block, 0, st, (const double*)part, nb, S, itrmax, rtol);
      hipLaunchKernelGGL(k_bi_p, grid, block, 0, st, n, p, (const double*)r, (const double*)q, (const BiState*)S);
      launches += 10;
    }
    PARCEL_HIP_CHECK(hipGetLastError());
    enq += k;
    PARCEL_HIP_CHECK(hipMemcpyAsync(h, S, sizeof(BiState), hipMemcpyDeviceToHost, st));
    PARCEL_HIP_CHECK(hipStreamSynchronize(st));
    if (h->done) break;
    if (chunk < 64) chunk *= 2;
  }
  PARCEL_HIP_CHECK(hipEventRecord(ev1, st));
  PARCEL_HIP_CHECK(hipMemcpyAsync(h, S, sizeof(BiState), hipMemcpyDeviceToHost, st));
  PARCEL_HIP_CHECK(hipStreamSynchronize(st));
  const int nit = h->niter, ret = h->iret;
  const bool ok = h->done != 0;
  (void)hipHostFree(h);
  if (!ok) throw hip_error("BiCGStab loop did not terminate", PARCEL_HIP_ERR_HIP);
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

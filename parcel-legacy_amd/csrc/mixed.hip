// mixed.hip -- mixed-precision iterative refinement (no counterpart in the
// reference; its FP64 PCG is KSP_pCG, src/pcg.f90:224-892).
//
// Outer loop in FP64: r = b - A x with the FP64 matrix, ||r|| / ||b|| test,
// x += d.  Each correction d solves A d = r with PCG in FP32 from d = 0 to a
// relative tolerance inner_rtol: SELL values, vectors and the block-Jacobi /
// point-Jacobi factors in float (the FP64 factors rounded to nearest), dot
// products and CG scalars in double.  The FP32 solve is the FP64 solver's
// kernels instantiated for float (cg_run<float>), so it keeps the same
// fused SpMV + update kernel and the fused block-Jacobi sweep, with half the
// bytes per matrix entry and per vector element.
#include <chrono>
#include <vector>

#include "kernels.hpp"

namespace parcel {

namespace {

__global__ void k_mx_to_f32(long n, const double* __restrict__ a, float* __restrict__ b) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

// x = x + (double)d
__global__ void k_mx_add(int n, double* __restrict__ x, const float* __restrict__ d) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    x[i] = x[i] + (double)d[i];
}

// st[0] = ||b||, st[1] = ||r|| from the residual kernel's partials
__global__ void __launch_bounds__(kNT) k_mx_norms(const double* part, int nb, double* st) {
  __shared__ double sh[kWaves];
  const double bb = reduce_part(part + PART_BB * kNBMax, nb, sh);
  const double rr = reduce_part(part + PART_RR0 * kNBMax, nb, sh);
  if (threadIdx.x == 0) {
    st[0] = sqrt(bb);
    st[1] = sqrt(rr);
  }
}

void convert(long n, const double* a, float* b, hipStream_t st) {
  const long blocks = std::min<long>((n + kNT - 1) / kNT, 8192);
  if (n > 0) hipLaunchKernelGGL(k_mx_to_f32, dim3((unsigned)blocks), dim3(kNT), 0, st, n, a, b);
}

void apply32(Plan& P, const float* r, float* s, const CGFuse32* f) { precon_apply32(P, r, s, f); }

}  // namespace

void mixed_setup(Plan& P) {
  DevSELL& A = P.A;
  if (A.val32.n != (size_t)A.entries) {
    A.val32.alloc((size_t)A.entries);
    convert(A.entries, A.val.p, A.val32.p, P.stream);
  }
  precon_setup32(P);
  const int n = A.n;
  while (P.vec32.size() < 7) P.vec32.emplace_back((size_t)n);
  PARCEL_HIP_CHECK(hipGetLastError());
  PARCEL_HIP_CHECK(hipStreamSynchronize(P.stream));
}

int mixed_solve(Plan& P, double* x, const double* b, int itrmax, double rtol, double* hist_host, int* niter,
                int* iret) {
  const int n = P.A.n;
  const hipStream_t st = P.stream;
  mixed_setup(P);
  plan_workspace(P, 1);
  const double inner_rtol = P.inner_rtol;
  const int inner_itrmax = P.inner_itrmax;
  const int nb = tiled_nb(n);
  float* r32 = P.vec32[0].p;
  float* d32 = P.vec32[1].p;
  CGVecs<float> V;
  V.A = sell_args32(P.A);
  for (int k = 0; k < 4; ++k) V.st7[k] = P.st7_a32[k].p;
  V.r = P.vec32[2].p;
  V.q = P.vec32[3].p;
  V.s = P.vec32[4].p;
  V.p0 = P.vec32[5].p;
  V.p1 = P.vec32[6].p;
  V.dinv = P.M.kind == PRECON_PJ ? P.M.dinv32.p : nullptr;
  V.hist = nullptr;
  V.hist_cap = 0;
  V.precon = apply32;
  V.precon_skip = [](Plan& Q, const float* r, float* s, const int* skip) { precon_apply32(Q, r, s, nullptr, skip); };
  if (P.solver_state.n < 2 * sizeof(double)) P.solver_state.alloc(2 * sizeof(double));
  double* dst = reinterpret_cast<double*>(P.solver_state.p);
  double hs[2];
  auto t0 = std::chrono::steady_clock::now();
  hipEvent_t ev0, ev1;
  PARCEL_HIP_CHECK(hipEventCreate(&ev0));
  PARCEL_HIP_CHECK(hipEventCreate(&ev1));
  PARCEL_HIP_CHECK(hipEventRecord(ev0, st));
  auto residual = [&]() {
    launch_residual(P, x, b, P.r.p, P.part.p);
    hipLaunchKernelGGL(k_mx_norms, dim3(1), dim3(kNT), 0, st, (const double*)P.part.p, nb, dst);
    PARCEL_HIP_CHECK(hipMemcpyAsync(hs, dst, sizeof(hs), hipMemcpyDeviceToHost, st));
    PARCEL_HIP_CHECK(hipStreamSynchronize(st));
  };
  std::vector<double> hist((size_t)std::max(itrmax, 1), 0.0);
  parcel_hip_stats acc{};
  int launches = 0, inner_total = 0, it = 1;
  residual();
  launches += 2;
  const double bnorm = hs[0];
  double rnorm = hs[1];
  for (it = 1; it <= itrmax; ++it) {
    hist[it - 1] = rnorm / bnorm;
    convert(n, P.r.p, r32, st);
    // every correction solves A d = r from d = 0
    PARCEL_HIP_CHECK(hipMemsetAsync(d32, 0, sizeof(float) * (size_t)n, st));
    int nit = 0, ret = 0;
    cg_run<float>(P, V, d32, r32, inner_itrmax, inner_rtol, nullptr, &nit, &ret);
    inner_total += nit;
    acc.launches += P.stats.launches;
    acc.spmv_ms += P.stats.spmv_ms;
    acc.spmv_calls += P.stats.spmv_calls;
    acc.precon_ms += P.stats.precon_ms;
    acc.precon_calls += P.stats.precon_calls;
    acc.vec_ms += P.stats.vec_ms;
    acc.vec_calls += P.stats.vec_calls;
    hipLaunchKernelGGL(k_mx_add, dim3(nb), dim3(kNT), 0, st, n, x, (const float*)d32);
    residual();
    launches += 4;
    rnorm = hs[1];
    if (rnorm / bnorm < rtol) break;
  }
  PARCEL_HIP_CHECK(hipEventRecord(ev1, st));
  PARCEL_HIP_CHECK(hipStreamSynchronize(st));
  const int nit = it < itrmax ? it : itrmax;
  const int ret = (itrmax <= it) ? 1 : 0;
  if (itrmax <= 0) {
    *niter = 0;
    *iret = 1;
  } else {
    *niter = nit;
    *iret = ret;
    if (hist_host) std::copy(hist.begin(), hist.begin() + nit, hist_host);
  }
  float ms = 0.f;
  PARCEL_HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  P.stats.solve_device_ms = ms;
  P.stats.solve_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  P.stats.niter = *niter;
  P.stats.inner_niter = inner_total;
  P.stats.final_rel_res = rnorm / bnorm;
  P.stats.launches = launches + acc.launches;
  P.stats.spmv_ms = acc.spmv_ms;
  P.stats.spmv_calls = acc.spmv_calls;
  P.stats.precon_ms = acc.precon_ms;
  P.stats.precon_calls = acc.precon_calls;
  P.stats.vec_ms = acc.vec_ms;
  P.stats.vec_calls = acc.vec_calls;
  return PARCEL_HIP_OK;
}

}  // namespace parcel

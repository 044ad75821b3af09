// krylov.hip -- SpMV (SELL-64), reductions and the preconditioned CG solver.
//
// KSP_pCG (src/pcg.f90:224-892) keeps alpha, beta, rho and the norms in host
// variables and moves each one through a 1-element memcpy plus a device
// synchronisation (src/parcel_vec.f90:98-214), ~30 host round trips per
// iteration.  Here the scalars never leave HBM while the solve runs:
//   * every reducing kernel writes one partial sum per workgroup (the fixed
//     tiled order of common.hpp);
//   * the NEXT kernel reduces those partials in its prologue, every
//     workgroup redundantly and identically (kernel boundaries make the
//     partials visible; no atomics, no fences);
//   * scalars that must survive a kernel go to CGState::slot[i & 1], written
//     by workgroup 0 of iteration i's kernels and read by iteration i+1.
// The host only enqueues kernels and polls CGState::done_at once per chunk.
//
// One iteration i (point Jacobi / none) is two kernels:
//   K1 k_cg_spmv(i): reduce (rho1, rr) of iteration i-1 -> exit test, beta;
//      x += alpha p_old; p_new = s + beta p_old; q = A p_new; gamma partials.
//      If the loop has ended it only applies the final x += alpha p.
//   K2 k_cg_update_pj(i): alpha = rho0 / gamma; r -= alpha q; s = Dinv r;
//      (rho1, rr) partials.
// With block Jacobi, K2 is the fused sweep kernel (precon.hip) followed by
// k_cg_dot2(i).  K1 reads p_new of neighbour rows as fma(beta, p_old[c], s[c]),
// the same rounding the owning row uses, so p never needs its own pass.
#include <type_traits>
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "cg_kernels.hpp"

namespace parcel {

// --- SELL-64 build (host) ----------------------------------------------------
// Width of a slice and its overflow rows: the slice is as wide as its
// (m+1)-th longest row and its m longest rows (m <= kOvfMax) keep the rest of
// their entries in the overflow area, m chosen to minimise the stored
// entries; overflow only where it saves at least 8 slice columns.
constexpr int kOvfMax = 16;
static int slice_split(const std::vector<int>& len, std::vector<int>& ovf_lanes) {
  std::vector<int> ord(len.size());
  for (size_t l = 0; l < len.size(); ++l) ord[l] = (int)l;
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return len[a] > len[b]; });
  const long full = 64L * len[ord[0]];
  long best = full, extra = 0;
  int bm = 0;
  for (int m = 1; m <= kOvfMax && m < (int)ord.size(); ++m) {
    extra += len[ord[m - 1]];
    const long w = len[ord[m]];
    const long cost = 64 * w + (extra - m * w) + 2L * m;
    if (cost < best) { best = cost; bm = m; }
  }
  ovf_lanes.clear();
  if (bm == 0 || full - best < 64L * 8) return len[ord[0]];
  for (int m = 0; m < bm; ++m) ovf_lanes.push_back(ord[m]);
  std::sort(ovf_lanes.begin(), ovf_lanes.end());
  return len[ord[bm]];
}

void sell_build(Plan& P) {
  const HostCSR& A = P.hA;
  const int n = A.n;
  DevSELL& S = P.A;
  S.n = n;
  S.nnz = A.rowptr[n];
  const long ns = ((long)n + 63) / 64;
  S.nslices = ns;
  std::vector<long> sptr(ns + 1, 0);
  std::vector<int> sk(ns, 0);
  std::vector<int> so(ns, -1);
  std::vector<unsigned long long> omask(ns, 0ull);
  std::vector<long> orow;  // overflow rows, by slice then lane
  std::vector<int> len(64), lanes;
  for (long s = 0; s < ns; ++s) {
    const long r0 = s * 64, r1 = std::min((long)n, r0 + 64);
    len.assign((size_t)(r1 - r0), 0);
    for (long r = r0; r < r1; ++r) len[(size_t)(r - r0)] = A.rowptr[r + 1] - A.rowptr[r];
    sk[s] = slice_split(len, lanes);
    if (!lanes.empty()) {
      so[s] = (int)orow.size();
      for (int l : lanes) {
        omask[s] |= 1ull << l;
        orow.push_back(r0 + l);
      }
    }
    sptr[s + 1] = sptr[s] + 64L * sk[s];
  }
  const long main = sptr[ns];
  std::vector<long> optr(orow.size() + 1, main);
  for (size_t k = 0; k < orow.size(); ++k) {
    const long r = orow[k];
    optr[k + 1] = optr[k] + (A.rowptr[r + 1] - A.rowptr[r]) - sk[r >> 6];
  }
  const long tot = optr.back();
  S.entries = tot;
  S.ovf_entries = tot - main;
  std::vector<int> col((size_t)tot);
  std::vector<double> val((size_t)tot);
  const bool lo = !A.val_lo.empty();
  std::vector<double> vlo(lo ? (size_t)tot : 0);
#pragma omp parallel for schedule(static)
  for (long s = 0; s < ns; ++s) {
    const long base = sptr[s];
    const int k = sk[s];
    for (int l = 0; l < 64; ++l) {
      const long r = s * 64 + l;
      const int len = r < n ? A.rowptr[r + 1] - A.rowptr[r] : 0;
      for (int j = 0; j < k; ++j) {
        const long at = base + (long)j * 64 + l;
        if (j < len) {
          col[at] = A.col[A.rowptr[r] + j];
          val[at] = A.val[A.rowptr[r] + j];
          if (lo) vlo[at] = A.val_lo[A.rowptr[r] + j];
        } else {
          col[at] = -1;
          val[at] = 0.0;
          if (lo) vlo[at] = 0.0;
        }
      }
    }
  }
  // an overflow row's entries after the first sk of its slice, in row order
  for (size_t q = 0; q < orow.size(); ++q) {
    const long r = orow[q];
    const int k = sk[r >> 6];
    for (long e = optr[q]; e < optr[q + 1]; ++e) {
      const long src = A.rowptr[r] + k + (e - optr[q]);
      col[(size_t)e] = A.col[src];
      val[(size_t)e] = A.val[src];
      if (lo) vlo[(size_t)e] = A.val_lo[src];
    }
  }
  S.sptr.upload(sptr, P.stream);
  S.sk.upload(sk, P.stream);
  S.col.upload(col, P.stream);
  S.val.upload(val, P.stream);
  if (lo) S.val_lo.upload(vlo, P.stream);
  if (!orow.empty()) {
    S.so.upload(so, P.stream);
    S.omask.upload(omask, P.stream);
    S.optr.upload(optr, P.stream);
  } else {
    S.so.release();
    S.omask.release();
    S.optr.release();
  }
  PARCEL_HIP_CHECK(hipStreamSynchronize(P.stream));
}


// Symmetric stencil form (St7) of a natural-order matrix whose every row i
// holds exactly its grid neighbours i-nxy, i-nx, i-1, i, i+1, i+nx, i+nxy
// (those inside the grid, columns ascending) with a(i,j) == a(j,i) bit for
// bit: nx is the smallest off-diagonal offset above 1, nxy the next one (n
// itself for a single plane).  Leaves P.st7 false for anything else.
void stencil_build(Plan& P) {
  const HostCSR& A = P.hA;
  const int n = A.n;
  P.st7 = false;
  if (n < 4 || !A.val_lo.empty()) return;
  long o1 = 0, o2 = 0;
  for (int i = 0; i < n; ++i)
    for (int e = A.rowptr[i]; e < A.rowptr[i + 1]; ++e) {
      const long off = (long)A.col[e] - i;
      if (off <= 1) continue;
      if (off == o1 || off == o2) continue;
      if (!o1) o1 = off;
      else if (!o2) o2 = off;
      else return;
    }
  if (!o1) return;
  if (o2 && o2 < o1) std::swap(o1, o2);
  const long nx = o1, nxy = o2 ? o2 : n;
  if (nxy % nx || n % nxy || nx < 2) return;
  const int ny = (int)(nxy / nx);
  std::vector<double> a[4];
  for (auto& v : a) v.assign((size_t)n, 0.0);
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (int i = 0; i < n; ++i) {
    if (bad) continue;
    const long l = i / nx;
    const int x = (int)(i - l * nx), y = (int)(l % ny);
    const long nb[7] = {i - nxy, i - nx, i - 1L, (long)i, i + 1L, i + nx, i + nxy};
    const bool has[7] = {i >= nxy, y > 0, x > 0, true, x < nx - 1, y < ny - 1, i + nxy < n};
    int e = A.rowptr[i];
    for (int k = 0; k < 7; ++k) {
      if (!has[k]) continue;
      if (e >= A.rowptr[i + 1] || A.col[e] != nb[k]) { bad = 1; break; }
      if (k == 3) a[0][i] = A.val[e];
      else if (k > 3) a[k - 3][i] = A.val[e];  // e, nn, t: the upper couplings
      ++e;
    }
    if (!bad && e != A.rowptr[i + 1]) bad = 1;
  }
  if (bad) return;
  // symmetry bit for bit: a(j, i) for each stored upper coupling a(i, j)
  auto lower = [&](long j, long i) -> double {
    for (int e = A.rowptr[j]; e < A.rowptr[j + 1]; ++e)
      if (A.col[e] == i) return A.val[e];
    return 0.0;
  };
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (int i = 0; i < n; ++i) {
    const long l = i / nx;
    const int x = (int)(i - l * nx), y = (int)(l % ny);
    if (x < nx - 1) {
      const double v = lower(i + 1, i);
      if (std::memcmp(&a[1][i], &v, 8) != 0) bad = 1;
    }
    if (y < ny - 1) {
      const double v = lower(i + nx, i);
      if (std::memcmp(&a[2][i], &v, 8) != 0) bad = 1;
    }
    if (i + nxy < n) {
      const double v = lower(i + nxy, i);
      if (std::memcmp(&a[3][i], &v, 8) != 0) bad = 1;
    }
  }
  if (bad) return;
  for (int k = 0; k < 4; ++k) P.st7_a[k].upload(a[k], P.stream);
  if (P.prm.solver == PARCEL_SOLVER_CG_MIXED) {
    for (int k = 0; k < 4; ++k) {
      std::vector<float> f((size_t)n);
      for (int i = 0; i < n; ++i) f[i] = (float)a[k][i];
      P.st7_a32[k].upload(f, P.stream);
    }
  }
  PARCEL_HIP_CHECK(hipStreamSynchronize(P.stream));
  P.st7_nx = (int)nx;
  P.st7_ny = ny;
  P.st7_nxy = nxy;
  P.st7 = true;
}

void ensure_vecs(Plan& P, int k) {
  const int n = P.A.n;
  while ((int)P.vec.size() < k) P.vec.emplace_back((size_t)n);
}

void launch_residual(Plan& P, const double* x, const double* b, double* r, double* part) {
  const int n = P.A.n;
  hipLaunchKernelGGL(k_residual<double>, dim3(tiled_nb(n)), dim3(kNT), 0, P.stream, n, sell_args(P.A), x, b, r, part, kOne);
}

void plan_workspace(Plan& P, int itrmax) {
  const int n = P.A.n;
  if (P.r.n < (size_t)n) {
    P.r.alloc(n); P.q.alloc(n); P.s.alloc(n); P.p0.alloc(n); P.p1.alloc(n);
  }
  if (!P.cg.p) P.cg.alloc(1);
  if (!P.part.p) P.part.alloc(PART_N * kNBMax);
  if (P.hist_cap < itrmax) {
    P.hist.alloc(itrmax > 0 ? itrmax : 1);
    P.hist_cap = itrmax;
  }
}

void spmv(Plan& P, const double* x, double* y) {
  const int n = P.A.n;
  hipLaunchKernelGGL(k_spmv<double>, dim3(tiled_nb(n)), dim3(kNT), 0, P.stream, n, sell_args(P.A), x, y, kOne);
  PARCEL_HIP_CHECK(hipGetLastError());
}

double dot_device(int n, const double* x, const double* y, hipStream_t st) {
  DevBuf<double> part(kNBMax + 1);
  const int nb = tiled_nb(n);
  hipLaunchKernelGGL(k_dot_part<double>, dim3(nb), dim3(kNT), 0, st, n, x, y, part.p, kOne);
  hipLaunchKernelGGL(k_reduce_one, dim3(1), dim3(kNT), 0, st, (const double*)part.p, nb, part.p + kNBMax);
  PARCEL_HIP_CHECK(hipGetLastError());
  double h = 0.0;
  PARCEL_HIP_CHECK(hipMemcpyAsync(&h, part.p + kNBMax, sizeof(double), hipMemcpyDeviceToHost, st));
  PARCEL_HIP_CHECK(hipStreamSynchronize(st));
  return h;
}

// --- allocation-free dot -----------------------------------------------------
// The dot's workspace (kNBMax workgroup partials and the reduced value) as a
// file-scope device array, so that parcel_hip_dot_noalloc allocates nothing.
// Every GPU holds its own copy for the life of the process.  In host code
// `scratch` names a host-side stand-in, never the device copy: the kernels
// and copies use the address hipGetSymbolAddress returns for the current
// device (noalloc_scratch).
constexpr int kScratchDoubles = kNBMax + 1;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kScratchDevices] = {};
}  // namespace

// Device address of `scratch` on `dev` (which must be the current device),
// resolved once per device and kept.
double* noalloc_scratch(int dev) {
  if (dev < 0 || dev >= kScratchDevices) throw hip_error("device ordinal out of range", PARCEL_HIP_ERR_ARG);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    double* ptr = nullptr;
    PARCEL_HIP_CHECK(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw hip_error("hipGetSymbolAddress returned no address for the dot scratch", PARCEL_HIP_ERR_HIP);
    g_scratch[dev] = ptr;
  }
  return g_scratch[dev];
}

// <x, y> of device vectors with the partials in `scratch` (the same kernels,
// launch shape and reduction tree as dot_device, so the same bits).  With a
// device `out` the value stays on the device and the call does not wait;
// otherwise it is copied to the host `out` after a stream synchronisation.
void dot_noalloc(int dev, int n, const double* x, const double* y, double* out, bool out_dev, hipStream_t st) {
  const int nb = tiled_nb(n);
  if (nb > kScratchDoubles - 1) throw hip_error("dot_noalloc: more partials than the scratch holds", PARCEL_HIP_ERR_ARG);
  double* part = noalloc_scratch(dev);
  hipLaunchKernelGGL(k_dot_part<double>, dim3(nb), dim3(kNT), 0, st, n, x, y, part, kOne);
  hipLaunchKernelGGL(k_reduce_one, dim3(1), dim3(kNT), 0, st, (const double*)part, nb,
                     out_dev ? out : part + kNBMax);
  PARCEL_HIP_CHECK(hipGetLastError());
  if (out_dev) return;
  PARCEL_HIP_CHECK(hipMemcpyAsync(out, part + kNBMax, sizeof(double), hipMemcpyDeviceToHost, st));
  PARCEL_HIP_CHECK(hipStreamSynchronize(st));
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
  hipEvent_t rec(hipStream_t st) {
    if (used == pool.size()) {
      hipEvent_t e;
      PARCEL_HIP_CHECK(hipEventCreate(&e));
      pool.push_back(e);
    }
    hipEvent_t e = pool[used++];
    PARCEL_HIP_CHECK(hipEventRecord(e, st));
    return e;
  }
  ~Timer() {
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};

}  // namespace

template <class T>
int cg_run(Plan& P, const CGVecs<T>& V, T* x, const T* b, int itrmax, double rtol, double* hist_host, int* niter,
           int* iret) {
  const int n = P.A.n;
  const hipStream_t st = P.stream;
  plan_workspace(P, itrmax);
  auto t0 = std::chrono::steady_clock::now();
  const int nb = tiled_nb(n);
  const dim3 grid(nb), block(kNT);
  CGState* S = P.cg.p;
  double* part = P.part.p;
  const SellArgsT<T> A = V.A;
  int launches = 0;
  Timer Tm;
  Tm.on = P.prm.profile == 1;
  hipEvent_t ev0, ev1;
  PARCEL_HIP_CHECK(hipEventCreate(&ev0));
  PARCEL_HIP_CHECK(hipEventCreate(&ev1));
  PARCEL_HIP_CHECK(hipEventRecord(ev0, st));

  if (itrmax <= 0) {
    // the loop body never runs (pcg.f90:488 with itrmax = 0): niter 0, iret 1
    PARCEL_HIP_CHECK(hipEventRecord(ev1, st));
    PARCEL_HIP_CHECK(hipStreamSynchronize(st));
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    *niter = 0;
    *iret = 1;
    P.stats.niter = 0;
    return PARCEL_HIP_OK;
  }

  PARCEL_HIP_CHECK(hipMemsetAsync(V.p0, 0, sizeof(T) * (size_t)n, st));
  hipLaunchKernelGGL(k_residual<T>, grid, block, 0, st, n, A, (const T*)x, b, V.r, part, kOne);
  V.precon(P, V.r, V.s, nullptr);
  hipLaunchKernelGGL(k_dot_part<T>, grid, block, 0, st, n, (const T*)V.s, (const T*)V.r, part + PART_RHO0 * kNBMax, kOne);
  hipLaunchKernelGGL(k_cg_init, dim3(1), block, 0, st, (const double*)part, nb, S);
  launches += 4;
  PARCEL_HIP_CHECK(hipGetLastError());

  P.hstate.ensure(1);
  CGState* h = P.hstate.p;
  h->done_at = 0;
  int enq = 0;  // K1 launches enqueued
  int chunk = 8;
  const bool bj = P.M.kind == PRECON_BJ;
  const bool ras = P.M.kind == PRECON_RAS;
  const bool lscan = std::is_same<T, float>::value && bj && P.M.lscan.on;
  const T* dinv = V.dinv;
  while (enq < itrmax + 1) {
    const int k = std::min(chunk, itrmax + 1 - enq);
    for (int j = 1; j <= k; ++j) {
      const int it = enq + j;
      T* pold = (it - 1) & 1 ? V.p1 : V.p0;
      T* pnew = it & 1 ? V.p1 : V.p0;
      hipEvent_t ea = Tm.on ? Tm.rec(st) : nullptr;
      bool dia = false;
      if constexpr (std::is_same<T, double>::value) {
        if (P.dia_nnd > 0) {
          const DiaArgsT<double> Ad{P.dia_val.p, P.dia_off.p, P.dia_nnd, (long)n};
          hipLaunchKernelGGL((k_cg_spmv<T, 1, DiaArgsT<double>>), grid, block, 0, st, n, it, itrmax, rtol, Ad, x,
                             (const T*)pold, (const T*)V.s, pnew, V.q, part, S, V.hist, V.hist_cap, kOne);
          dia = true;
        }
      }
      if (!dia && P.st7) {
        const St7<T> As{V.st7[0], V.st7[1], V.st7[2], V.st7[3], P.st7_nx, P.st7_ny, P.st7_nxy, (long)n,
                        FastDiv((unsigned)P.st7_nx), FastDiv((unsigned)P.st7_ny)};
        hipLaunchKernelGGL((k_cg_spmv<T, 1, St7<T>>), grid, block, 0, st, n, it, itrmax, rtol, As, x,
                           (const T*)pold, (const T*)V.s, pnew, V.q, part, S, V.hist, V.hist_cap, kOne);
      } else if (!dia) {
        hipLaunchKernelGGL(k_cg_spmv<T>, grid, block, 0, st, n, it, itrmax, rtol, A, x, (const T*)pold,
                           (const T*)V.s, pnew, V.q, part, S, V.hist, V.hist_cap, kOne);
      }
      launches++;
      hipEvent_t eb = Tm.on ? Tm.rec(st) : nullptr;
      if (Tm.on) Tm.spmv.push_back({ea, eb, it});
      if (it > itrmax) break;  // the last K1 only closes the loop
      if (lscan) {
        // r -= alpha q; s = M^-1 r by line scans of the block IC(0) factor; dots
        hipLaunchKernelGGL((k_cg_axpy_r<T, 1>), grid, block, 0, st, n, it, V.r, (const T*)V.q, part, S, nb, kOne);
        hipEvent_t ec0 = Tm.on ? Tm.rec(st) : nullptr;
        const LscanBJ& G = P.M.lscan;
        lscan_apply32(G.F.p, (const float*)V.r, G.w.p, (float*)V.s, G.nx, G.ny, G.nstrip, (long)n, S, it, G.sink.p,
                      st);
        hipEvent_t ec = Tm.on ? Tm.rec(st) : nullptr;
        if (Tm.on) Tm.prec.push_back({ec0, ec, it});
        hipLaunchKernelGGL(k_cg_dot2<T>, grid, block, 0, st, n, it, (const T*)V.s, (const T*)V.r, part,
                           (const CGState*)S, kOne);
        launches += 3;
        if (Tm.on) Tm.vec.push_back({ec, Tm.rec(st), it});
      } else if (ras) {
        // r -= alpha q; s = M^-1 r (gather, block solves, scatter); dots
        hipLaunchKernelGGL((k_cg_axpy_r<T, 1>), grid, block, 0, st, n, it, V.r, (const T*)V.q, part, S, nb, kOne);
        hipEvent_t ec0 = Tm.on ? Tm.rec(st) : nullptr;
        V.precon_skip(P, V.r, V.s, &S->done_at);
        hipEvent_t ec = Tm.on ? Tm.rec(st) : nullptr;
        if (Tm.on) Tm.prec.push_back({ec0, ec, it});
        hipLaunchKernelGGL(k_cg_dot2<T>, grid, block, 0, st, n, it, (const T*)V.s, (const T*)V.r, part,
                           (const CGState*)S, kOne);
        launches += 5;
        if (Tm.on) Tm.vec.push_back({ec, Tm.rec(st), it});
      } else if (!bj) {
        hipLaunchKernelGGL(k_cg_update_pj<T>, grid, block, 0, st, n, it, V.r, (const T*)V.q, dinv, V.s, part, S,
                           nb, kOne);
        launches++;
        if (Tm.on) Tm.vec.push_back({eb, Tm.rec(st), it});
      } else {
        CGFuseT<T> f;
        f.q = V.q;
        f.r = V.r;
        f.st = S;
        f.part_gamma = part + PART_GAMMA * kNBMax;
        f.nb_prod = nb;
        f.it = it;
        V.precon(P, V.r, V.s, &f);
        hipEvent_t ec = Tm.on ? Tm.rec(st) : nullptr;
        if (Tm.on) Tm.prec.push_back({eb, ec, it});
        hipLaunchKernelGGL(k_cg_dot2<T>, grid, block, 0, st, n, it, (const T*)V.s, (const T*)V.r, part,
                           (const CGState*)S, kOne);
        launches += 2;
        if (Tm.on) Tm.vec.push_back({ec, Tm.rec(st), it});
      }
    }
    PARCEL_HIP_CHECK(hipGetLastError());
    enq += k;
    PARCEL_HIP_CHECK(hipMemcpyAsync(h, S, sizeof(CGState), hipMemcpyDeviceToHost, st));
    PARCEL_HIP_CHECK(hipStreamSynchronize(st));
    if (h->done_at) break;
    if (chunk < 64) chunk *= 2;
  }
  PARCEL_HIP_CHECK(hipEventRecord(ev1, st));
  PARCEL_HIP_CHECK(hipMemcpyAsync(h, S, sizeof(CGState), hipMemcpyDeviceToHost, st));
  PARCEL_HIP_CHECK(hipStreamSynchronize(st));
  if (!h->done_at) throw hip_error("CG loop did not terminate", PARCEL_HIP_ERR_HIP);
  const int nit = h->niter;
  const int ret = h->iret;
  if (nit > 0 && hist_host)
    PARCEL_HIP_CHECK(hipMemcpy(hist_host, V.hist, sizeof(double) * (size_t)nit, hipMemcpyDeviceToHost));
  float ms = 0.f;
  PARCEL_HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
  P.stats.solve_device_ms = ms;
  P.stats.solve_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  P.stats.niter = nit;
  P.stats.launches = launches;
  if (Tm.on) {
    auto sum = [&](std::vector<EventPair>& v, double& out, int& calls) {
      out = 0.0; calls = 0;
      for (size_t i = 0; i < v.size(); ++i) {
        if (v[i].it > nit) continue;
        float t = 0.f;
        PARCEL_HIP_CHECK(hipEventElapsedTime(&t, v[i].a, v[i].b));
        out += t; calls++;
      }
    };
    sum(Tm.spmv, P.stats.spmv_ms, P.stats.spmv_calls);
    sum(Tm.prec, P.stats.precon_ms, P.stats.precon_calls);
    sum(Tm.vec, P.stats.vec_ms, P.stats.vec_calls);
  }
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
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
  h.bnorm = 1.0;
  h.slot[0].rho0 = 1.0; h.slot[1].rho0 = 1.0;
  h.slot[0].rnorm = 1.0; h.slot[1].rnorm = 1.0;
  PARCEL_HIP_CHECK(hipMemcpyAsync(P.cg.p, &h, sizeof(CGState), hipMemcpyHostToDevice, st));
  for (double* v : {P.p0.p, P.p1.p, P.s.p, P.r.p, P.q.p})
    PARCEL_HIP_CHECK(hipMemsetAsync(v, 0, sizeof(double) * (size_t)n, st));
  PARCEL_HIP_CHECK(hipMemsetAsync(P.part.p, 0, sizeof(double) * PART_N * kNBMax, st));
  if (P.x.n < (size_t)n) P.x.alloc(n);
  PARCEL_HIP_CHECK(hipMemsetAsync(P.x.p, 0, sizeof(double) * (size_t)n, st));
  hipEvent_t a, b;
  PARCEL_HIP_CHECK(hipEventCreate(&a));
  PARCEL_HIP_CHECK(hipEventCreate(&b));
  auto launch = [&](int i) {
    if (which == 1) {
      // iteration 1 semantics every time (no exit test), no history write
      hipLaunchKernelGGL(k_cg_spmv<double>, dim3(nb), dim3(kNT), 0, st, n, 1, 1 << 30, 0.0, sell_args(P.A), P.x.p,
                         (i & 1) ? P.p1.p : P.p0.p, P.s.p, (i & 1) ? P.p0.p : P.p1.p, P.q.p, P.part.p, P.cg.p,
                         P.hist.p, 0, kOne);
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

template int cg_run<float>(Plan&, const CGVecs<float>&, float*, const float*, int, double, double*, int*, int*);

int cg_solve(Plan& P, double* x, const double* b, int itrmax, double rtol, double* hist_host, int* niter,
             int* iret) {
  plan_workspace(P, itrmax);
  CGVecs<double> V;
  V.A = sell_args(P.A);
  for (int k = 0; k < 4; ++k) V.st7[k] = P.st7_a[k].p;
  V.r = P.r.p;
  V.q = P.q.p;
  V.s = P.s.p;
  V.p0 = P.p0.p;
  V.p1 = P.p1.p;
  V.dinv = P.M.kind == PRECON_PJ ? P.M.dinv.p : nullptr;
  V.hist = P.hist.p;
  V.hist_cap = P.hist_cap;
  V.precon = [](Plan& Q, const double* r, double* s, const CGFuse* f) { precon_apply(Q, r, s, f); };
  V.precon_skip = [](Plan& Q, const double* r, double* s, const int* skip) { precon_apply(Q, r, s, nullptr, skip); };
  return cg_run<double>(P, V, x, b, itrmax, rtol, hist_host, niter, iret);
}

}  // namespace parcel

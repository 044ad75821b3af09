// cg_kernels.hpp -- the CG kernels, shared by the single-system solver
// (krylov.hip) and the batched solver (batched.hip).
//
// Every kernel takes a Bat: the per-system strides of vectors (vs) and of
// matrix values (as).  A batched launch has gridDim.y systems; system
// y = blockIdx.y works on vectors at y*vs, matrix values at y*as, partial
// sums at y*PART_N*kNBMax, its own CGState and history.  A single-system
// launch passes {0, 0} and gridDim.y = 1, so every offset is zero.  The
// arithmetic of a system never depends on the others.
#pragma once

#include "kernels.hpp"

namespace parcel {

struct Bat {
  long vs;  // vector stride between systems (0: one system)
  long as;  // matrix-value stride between systems
};

namespace {

__device__ __forceinline__ long sys_id() { return (long)blockIdx.y; }

// Owner index of this workgroup in the tiled layout.  Workgroups are dealt to
// the 8 XCDs round-robin by linear id, so with nb a multiple of 8 the owners
// of XCD k are made the contiguous range [k nb/8, (k+1) nb/8): a stencil's
// neighbour tiles then sit in the same XCD's L2.  A bijection on [0, nb):
// owner o still holds tiles o, o+nb, ... and writes partial slot o, so every
// sum is unchanged.
__device__ __forceinline__ int tile_owner() {
  const int nb = gridDim.x, b = blockIdx.x;
  return (nb & 7) ? b : (b & 7) * (nb >> 3) + (b >> 3);
}

// N workgroup sums at once, each equal to block_sum's (butterfly per wave,
// then the wave totals left to right); sh must hold kWaves*N doubles.
template <int N>
__device__ __forceinline__ void block_sum_n(double (&v)[N], double* sh) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int d = 0; d < N; ++d) v[d] = wave_sum(v[d]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int d = 0; d < N; ++d) sh[w * N + d] = v[d];
  }
  __syncthreads();
#pragma unroll
  for (int d = 0; d < N; ++d) {
    double t = sh[d];
#pragma unroll
    for (int k = 1; k < kWaves; ++k) t = t + sh[k * N + d];
    v[d] = t;
  }
}

template <class T>
__device__ __forceinline__ SellArgsT<T> sys_matrix(SellArgsT<T> A, const Bat& B) {
  A.val += sys_id() * B.as;
  return A;
}

// Matrix access of the fused CG kernel: the SELL-64 matrix, or a five-point
// stencil in natural order with symmetric couplings stored once per row.
template <class T, class G>
__device__ __forceinline__ T mat_row(const SellArgsT<T>& A, long i, G g) {
  return sell_row(A, i, g);
}

// Five-point stencil: c = a(i,i), e = a(i,i+1) = a(i+1,i), nn = a(i,i+nx) =
// a(i+nx,i); one system's arrays are vs apart.  A row sums its entries in
// the CSR order s, w, c, e, n (sell_row's order for that pattern), so the
// product is bitwise the SELL product when the matrix is exactly symmetric.
// DIA: nnd diagonals of n values each, diagonal j of row i at val[j*n + i],
// coupling row i to column i + off[j] (MatVec_dia, src/matrix.f90:2309-2379:
// every diagonal in stored order; entries whose column falls outside
// [0, n) are skipped).  A zero entry adds fma(0, x, t) = t, so the row sum
// equals the CRS row's without the zeros (dia2crs drops them).
template <class T>
struct DiaArgsT {
  const T* val;
  const int* off;
  int nnd;
  long n;
};
template <class T, class G>
__device__ __forceinline__ T mat_row(const DiaArgsT<T>& A, long i, G g) {
  T t = 0;
  for (int j = 0; j < A.nnd; ++j) {
    const long c = i + A.off[j];
    if (c >= 0 && c < A.n) t = fma(A.val[(long)j * A.n + i], g(c), t);
  }
  return t;
}
template <class T>
__device__ __forceinline__ DiaArgsT<T> sys_matrix(DiaArgsT<T> A, const Bat& B) {
  A.val += sys_id() * B.as;
  return A;
}

template <class T>
struct St5 {
  const T* c;
  const T* e;
  const T* nn;
  int nx, ny;
  FastDiv dx;  // by nx (a system's rows < 2^31)
};

template <class T, class G>
__device__ __forceinline__ T mat_row(const St5<T>& A, long i, G g) {
  const unsigned ui = (unsigned)i, yy = A.dx.div(ui);
  const int x = (int)(ui - yy * (unsigned)A.nx), y = (int)yy;
  T t = 0;
  if (y > 0) t = fma(A.nn[i - A.nx], g(i - A.nx), t);
  if (x > 0) t = fma(A.e[i - 1], g(i - 1), t);
  t = fma(A.c[i], g(i), t);
  if (x < A.nx - 1) t = fma(A.e[i], g(i + 1), t);
  if (y < A.ny - 1) t = fma(A.nn[i], g(i + A.nx), t);
  return t;
}

template <class T>
__device__ __forceinline__ St5<T> sys_matrix(St5<T> A, const Bat& B) {
  const long o = sys_id() * B.vs;
  A.c += o;
  A.e += o;
  A.nn += o;
  return A;
}

// Seven-point (3D) or five-point (2D) stencil in natural order with
// symmetric couplings stored once per row, for one system: c = a(i,i),
// e = a(i,i+1) = a(i+1,i), nn = a(i,i+nx) = a(i+nx,i), t = a(i,i+nxy) =
// a(i+nxy,i) (nxy = n for a single plane; t is then never read).  A row
// sums its entries in CSR order b, s, w, c, e, n, t over the neighbours the
// grid has, which is sell_row's order when every such neighbour is stored:
// the product equals the SELL product bit for bit.
template <class T>
struct St7 {
  const T* c;
  const T* e;
  const T* nn;
  const T* t;
  int nx, ny;
  long nxy, n;
  FastDiv dx, dy;  // by nx and ny (rows < 2^31)
};

template <class T, class G>
__device__ __forceinline__ T mat_row(const St7<T>& A, long i, G g) {
  const unsigned ui = (unsigned)i, l = A.dx.div(ui), q = A.dy.div(l);
  const int x = (int)(ui - l * (unsigned)A.nx), y = (int)(l - q * (unsigned)A.ny);
  T t = 0;
  if (i >= A.nxy) t = fma(A.t[i - A.nxy], g(i - A.nxy), t);
  if (y > 0) t = fma(A.nn[i - A.nx], g(i - A.nx), t);
  if (x > 0) t = fma(A.e[i - 1], g(i - 1), t);
  t = fma(A.c[i], g(i), t);
  if (x < A.nx - 1) t = fma(A.e[i], g(i + 1), t);
  if (y < A.ny - 1) t = fma(A.nn[i], g(i + A.nx), t);
  if (i + A.nxy < A.n) t = fma(A.t[i], g(i + A.nxy), t);
  return t;
}

template <class MT>
struct IsSt7 : std::false_type {};
template <class T>
struct IsSt7<St7<T>> : std::true_type {};

// Row i (< A.n) of the stencil form times g = beta p_old + s, mat_row's
// sum, with every operand loaded before the first product: an absent
// neighbour reads row i instead (an address inside the arrays) and a select
// drops its term, so a row is one round of independent loads.  gc: g(i).
template <class T, class TZ>
__device__ __forceinline__ T st7_row_g(const St7<T>& A, long i, T beta, const T* __restrict__ po,
                                       const TZ* __restrict__ s, T gc) {
  const unsigned ui = (unsigned)i, l = A.dx.div(ui), q = A.dy.div(l);
  const int x = (int)(ui - l * (unsigned)A.nx), y = (int)(l - q * (unsigned)A.ny);
  const bool hb = i >= A.nxy, hs = y > 0, hw = x > 0, he = x < A.nx - 1, hn = y < A.ny - 1, ht = i + A.nxy < A.n;
  const long ib = hb ? i - A.nxy : i, is = hs ? i - A.nx : i, iw = hw ? i - 1 : i;
  const long ie = he ? i + 1 : i, in = hn ? i + A.nx : i, it = ht ? i + A.nxy : i;
  const T ab = A.t[ib], as = A.nn[is], aw = A.e[iw], ac = A.c[i], ae = A.e[i], an = A.nn[i], at = A.t[i];
  const T pb = po[ib], ps = po[is], pw = po[iw], pe = po[ie], pn = po[in], pt = po[it];
  const T zb = (T)s[ib], zs = (T)s[is], zw = (T)s[iw], ze = (T)s[ie], zn = (T)s[in], zt = (T)s[it];
  T t = 0;
  t = hb ? fma(ab, fma(beta, pb, zb), t) : t;
  t = hs ? fma(as, fma(beta, ps, zs), t) : t;
  t = hw ? fma(aw, fma(beta, pw, zw), t) : t;
  t = fma(ac, gc, t);
  t = he ? fma(ae, fma(beta, pe, ze), t) : t;
  t = hn ? fma(an, fma(beta, pn, zn), t) : t;
  t = ht ? fma(at, fma(beta, pt, zt), t) : t;
  return t;
}

template <class T>
__device__ __forceinline__ St7<T> sys_matrix(St7<T> A, const Bat& B) {
  (void)B;  // single-system plans only
  return A;
}

__device__ __forceinline__ double* sys_part(double* part) { return part + sys_id() * (PART_N * kNBMax); }

// r = b - A x and the partials of <b,b>, <r,r> (calc_res_CRS,
// src/krylov_h.f90:343; pcg.f90:456-470)
// (T: vector and matrix value type; sums of products accumulate in double)
template <class T>
__global__ void __launch_bounds__(kNT) k_residual(int n, SellArgsT<T> A, const T* __restrict__ x,
                                                  const T* __restrict__ b, T* __restrict__ r, double* part,
                                                  Bat B) {
  __shared__ double sh[kWaves];
  const long so = sys_id() * B.vs;
  A = sys_matrix(A, B);
  x += so; b += so; r += so;
  part = sys_part(part);
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  double a0 = 0.0, a1 = 0.0;
  for (long k = tile_owner(); k < ntiles; k += nb) {
    const long i = k * kNT + threadIdx.x;
    if (i < n) {
      const T t = sell_row(A, i, [&](int c) { return x[c]; });
      const T bi = b[i];
      const T ri = bi - t;
      r[i] = ri;
      a0 = fma((double)bi, (double)bi, a0);
      a1 = fma((double)ri, (double)ri, a1);
    }
  }
  double s0 = block_sum(a0, sh);
  double s1 = block_sum(a1, sh);
  if (threadIdx.x == 0) {
    part[PART_BB * kNBMax + tile_owner()] = s0;
    part[PART_RR0 * kNBMax + tile_owner()] = s1;
  }
}

// partials of <x,y> into part[slot] (slot offset already applied by the
// caller; systems are PART_N*kNBMax apart)
template <class T, class TX = T>
__global__ void __launch_bounds__(kNT) k_dot_part(int n, const TX* __restrict__ x, const T* __restrict__ y,
                                                  double* part, Bat B) {
  __shared__ double sh[kWaves];
  const long so = sys_id() * B.vs;
  x += so; y += so;
  part = sys_part(part);
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  double a0 = 0.0;
  for (long k = tile_owner(); k < ntiles; k += nb) {
    const long i = k * kNT + threadIdx.x;
    if (i < n) a0 = fma((double)x[i], (double)y[i], a0);
  }
  double s0 = block_sum(a0, sh);
  if (threadIdx.x == 0) part[tile_owner()] = s0;
}

__global__ void __launch_bounds__(kNT) k_reduce_one(const double* part, int nb, double* out) {
  __shared__ double sh[kWaves];
  double v = reduce_part(part, nb, sh);
  if (threadIdx.x == 0) *out = v;
}

// bnorm, rnorm, rho0 of the initial residual (pcg.f90:456-490); slot 0.
// One workgroup per system.
__global__ void __launch_bounds__(kNT) k_cg_init(const double* part, int nb, CGState* st) {
  __shared__ double sh[kWaves];
  part = sys_part(const_cast<double*>(part));
  st += sys_id();
  const double bb = reduce_part(part + PART_BB * kNBMax, nb, sh);
  const double rr = reduce_part(part + PART_RR0 * kNBMax, nb, sh);
  const double rho0 = reduce_part(part + PART_RHO0 * kNBMax, nb, sh);
  if (threadIdx.x == 0) {
    st->bnorm = sqrt(bb);
    st->done_at = 0;
    st->niter = 0;
    st->iret = 0;
    CGSlot& s = st->slot[0];
    s.rho0 = rho0;
    s.alpha = 0.0;
    s.beta = 0.0;
    s.rnorm = sqrt(rr);
    s.itr = 0;
  }
}

// K1 of iteration `it`.  LB logical owners per workgroup: the tiled layout
// has nb = LB * gridDim.x owners; workgroup w runs owners tile_owner(w)*LB
// + j, j < LB, round by round (tile o + r*nb of every owner o in round r),
// and writes each owner's partial, so every sum equals the LB = 1 launch's
// while the partial reduction at the start is paid once per workgroup and
// LB rows per thread are in flight.  TZ: the stored type of the
// preconditioned residual s (float when an FP32 apply produced it: widened
// at use, exactly).
// rounds per trip of the stencil-form loop
constexpr int kSt7U = 2;

template <class T, int LB = 1, class MT = SellArgsT<T>, class TZ = T>
__global__ void __launch_bounds__(kNT) k_cg_spmv(int n, int it, int itrmax, double rtol, MT A,
                                                 T* __restrict__ x, const T* __restrict__ p_old,
                                                 const TZ* __restrict__ s, T* __restrict__ p_new,
                                                 T* __restrict__ q, double* part, CGState* st,
                                                 double* hist, int hist_cap, Bat B) {
  __shared__ double sh[kWaves * LB];
  st += sys_id();
  {
    const int d = st->done_at;
    if (d != 0 && d < it) return;
  }
  const long so = sys_id() * B.vs;
  A = sys_matrix(A, B);
  x += so; p_old += so; s += so; p_new += so; q += so;
  part = sys_part(part);
  hist += sys_id() * hist_cap;
  const CGSlot& prev = st->slot[(it - 1) & 1];
  const double bnorm = st->bnorm;
  double alpha_prev, beta, rho0, rnorm;
  bool exit_loop = false;
  const int nb = gridDim.x * LB;
  const int o0 = tile_owner() * LB;
  if (it == 1) {
    alpha_prev = 0.0;
    beta = 0.0;
    rho0 = prev.rho0;
    rnorm = prev.rnorm;
  } else {
    // pcg.f90:640-700: rnorm = sqrt(<r,r>); exit tests; beta = rho1/rho0
    const double rho1 = reduce_part(part + PART_RHO1 * kNBMax, nb, sh);
    const double rr = reduce_part(part + PART_RR * kNBMax, nb, sh);
    rnorm = sqrt(rr);
    const double ratio = rnorm / bnorm;
    alpha_prev = prev.alpha;
    exit_loop = ratio < rtol || ratio > 1.0e8 || it - 1 >= itrmax;
    beta = rho1 / prev.rho0;
    rho0 = rho1;
  }
  const long ntiles = ((long)n + kNT - 1) / kNT;
  if (exit_loop) {
    // loop ended after iteration it-1: final x += alpha p (pcg.f90:733)
    for (long base = 0; base < ntiles; base += nb) {
#pragma unroll
      for (int j = 0; j < LB; ++j) {
        const long i = (base + o0 + j) * kNT + threadIdx.x;
        if (i < n) x[i] = fma((T)alpha_prev, p_old[i], x[i]);
      }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->done_at = it;
      st->niter = it - 1;                       // pcg.f90:850
      st->iret = (it - 1 >= itrmax) ? 1 : 0;    // pcg.f90:851-855
      st->slot[it & 1].rnorm = rnorm;
      st->slot[it & 1].itr = it - 1;
    }
    return;
  }
  double acc[LB];
#pragma unroll
  for (int j = 0; j < LB; ++j) acc[j] = 0.0;
  if constexpr (IsSt7<MT>::value && LB == 1) {
    // kSt7U rounds per trip, every load of the trip issued before its first
    // product; each row's values and the partial's order are the loop's below
    constexpr int U = kSt7U;
    const long stride = (long)nb * kNT;
    for (long base = 0; base < ntiles; base += (long)U * nb) {
      long i[U];
      T xo[U], po[U], pn[U], t[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        i[u] = (base + o0) * kNT + threadIdx.x + u * stride;
        const long ic = i[u] < n ? i[u] : (long)n - 1;
        po[u] = p_old[ic];
        xo[u] = x[ic];
        pn[u] = fma((T)beta, po[u], (T)s[ic]);
        t[u] = st7_row_g(A, ic, (T)beta, p_old, s, pn[u]);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (i[u] < n) {
          x[i[u]] = fma((T)alpha_prev, po[u], xo[u]);
          p_new[i[u]] = pn[u];
          q[i[u]] = t[u];
          acc[0] = fma((double)pn[u], (double)t[u], acc[0]);
        }
      }
    }
  } else {
    for (long base = 0; base < ntiles; base += nb) {
#pragma unroll
      for (int j = 0; j < LB; ++j) {
        const long i = (base + o0 + j) * kNT + threadIdx.x;
        if (i < n) {
          const T po = p_old[i];
          x[i] = fma((T)alpha_prev, po, x[i]);
          const T pn = fma((T)beta, po, (T)s[i]);
          p_new[i] = pn;
          const T t = mat_row(A, i, [&](long c) { return fma((T)beta, p_old[c], (T)s[c]); });
          q[i] = t;
          acc[j] = fma((double)pn, (double)t, acc[j]);
        }
      }
    }
  }
  block_sum_n<LB>(acc, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < LB; ++j) part[PART_GAMMA * kNBMax + o0 + j] = acc[j];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (it - 1 < hist_cap) hist[it - 1] = rnorm / bnorm;  // pcg.f90:512-518
    CGSlot& cur = st->slot[it & 1];
    cur.rho0 = rho0;
    cur.beta = beta;
    cur.rnorm = rnorm;
    cur.itr = it;
  }
}

// K2 of iteration `it` (point Jacobi or none): alpha; r -= alpha q;
// s = Dinv r; (rho1, rr) partials.  dinv == nullptr: s = r.
template <class T>
__global__ void __launch_bounds__(kNT) k_cg_update_pj(int n, int it, T* __restrict__ r, const T* __restrict__ q,
                                                      const T* __restrict__ dinv, T* __restrict__ s, double* part,
                                                      CGState* st, int nb_prod, Bat B) {
  __shared__ double sh[kWaves];
  st += sys_id();
  {
    const int d = st->done_at;
    if (d != 0 && d <= it) return;
  }
  const long so = sys_id() * B.vs;
  r += so; q += so; s += so;
  if (dinv) dinv += so;
  part = sys_part(part);
  CGSlot& cur = st->slot[it & 1];
  const double gamma = reduce_part(part + PART_GAMMA * kNBMax, nb_prod, sh);
  const double alpha = cur.rho0 / gamma;  // pcg.f90:590-600
  if (blockIdx.x == 0 && threadIdx.x == 0) cur.alpha = alpha;
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  double a0 = 0.0, a1 = 0.0;
  for (long k = tile_owner(); k < ntiles; k += nb) {
    const long i = k * kNT + threadIdx.x;
    if (i < n) {
      const T ri = fma(-(T)alpha, q[i], r[i]);
      r[i] = ri;
      const T si = dinv ? ri * dinv[i] : ri;
      s[i] = si;
      a0 = fma((double)si, (double)ri, a0);
      a1 = fma((double)ri, (double)ri, a1);
    }
  }
  double s0 = block_sum(a0, sh);
  double s1 = block_sum(a1, sh);
  if (threadIdx.x == 0) {
    part[PART_RHO1 * kNBMax + tile_owner()] = s0;
    part[PART_RR * kNBMax + tile_owner()] = s1;
  }
}

// rho1 = <s,r>, rr = <r,r> partials (CG_calc_dot2, pcg.f90:30); LB logical
// owners per workgroup as in k_cg_spmv
template <class T, int LB = 1, class TZ = T>
__global__ void __launch_bounds__(kNT) k_cg_dot2(int n, int it, const TZ* __restrict__ s, const T* __restrict__ r,
                                                 double* part, const CGState* st, Bat B) {
  __shared__ double sh[kWaves * 2 * LB];
  st += sys_id();
  {
    const int d = st->done_at;
    if (d != 0 && d <= it) return;
  }
  const long so = sys_id() * B.vs;
  s += so; r += so;
  part = sys_part(part);
  const int nb = gridDim.x * LB;
  const int o0 = tile_owner() * LB;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  double acc[2 * LB];
#pragma unroll
  for (int j = 0; j < 2 * LB; ++j) acc[j] = 0.0;
  for (long base = 0; base < ntiles; base += nb) {
#pragma unroll
    for (int j = 0; j < LB; ++j) {
      const long i = (base + o0 + j) * kNT + threadIdx.x;
      if (i < n) {
        const T ri = r[i];
        acc[j] = fma((double)s[i], (double)ri, acc[j]);
        acc[LB + j] = fma((double)ri, (double)ri, acc[LB + j]);
      }
    }
  }
  block_sum_n<2 * LB>(acc, sh);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < LB; ++j) {
      part[PART_RHO1 * kNBMax + o0 + j] = acc[j];
      part[PART_RR * kNBMax + o0 + j] = acc[LB + j];
    }
  }
}

// alpha = rho0 / <p,q> (reduced from the SpMV's partials); r -= alpha q
// (pcg.f90:590-620) for a preconditioner applied in a separate step; LB
// logical owners per workgroup as in k_cg_spmv
template <class T, int LB = 1>
__global__ void __launch_bounds__(kNT) k_cg_axpy_r(int n, int it, T* __restrict__ r, const T* __restrict__ q,
                                                   double* part, CGState* st, int nb_prod, Bat B) {
  __shared__ double sh[kWaves];
  st += sys_id();
  {
    const int d = st->done_at;
    if (d != 0 && d <= it) return;
  }
  const long so = sys_id() * B.vs;
  r += so; q += so;
  part = sys_part(part);
  CGSlot& cur = st->slot[it & 1];
  const double gamma = reduce_part(part + PART_GAMMA * kNBMax, nb_prod, sh);
  const double alpha = cur.rho0 / gamma;  // pcg.f90:590-600
  if (blockIdx.x == 0 && threadIdx.x == 0) cur.alpha = alpha;
  const int nb = gridDim.x * LB;
  const int o0 = tile_owner() * LB;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  for (long base = 0; base < ntiles; base += nb) {
#pragma unroll
    for (int j = 0; j < LB; ++j) {
      const long i = (base + o0 + j) * kNT + threadIdx.x;
      if (i < n) r[i] = fma(-(T)alpha, q[i], r[i]);
    }
  }
}

template <class T>
__global__ void __launch_bounds__(kNT) k_spmv(int n, SellArgsT<T> A, const T* __restrict__ x, T* __restrict__ y,
                                              Bat B) {
  const long so = sys_id() * B.vs;
  A = sys_matrix(A, B);
  x += so; y += so;
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  for (long k = tile_owner(); k < ntiles; k += nb) {
    const long i = k * kNT + threadIdx.x;
    if (i < n) y[i] = sell_row(A, i, [&](int c) { return x[c]; });
  }
}

constexpr Bat kOne{0, 0};

}  // namespace
}  // namespace parcel

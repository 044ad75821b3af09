// kernels.hpp -- device helpers shared by the solver translation units:
// SELL-64 row products and the fixed-order reduction of partial sums.
#pragma once

#include "internal.hpp"

namespace parcel {

// SELL-64 arrays as the kernels see them; T is the value type (double, or
// float for the FP32 copy used by the mixed-precision inner solve).
template <class T>
struct SellArgsT {
  const long* sptr;
  const int* sk;
  const int* col;
  const T* val;
  // long rows (DevSELL): null when the matrix has none
  const int* so = nullptr;
  const unsigned long long* omask = nullptr;
  const long* optr = nullptr;
};
using SellArgs = SellArgsT<double>;

inline SellArgs sell_args(const DevSELL& S) {
  return SellArgs{S.sptr.p, S.sk.p, S.col.p, S.val.p, S.so.p, S.omask.p, S.optr.p};
}
inline SellArgsT<float> sell_args32(const DevSELL& S) {
  return SellArgsT<float>{S.sptr.p, S.sk.p, S.col.p, S.val32.p, S.so.p, S.omask.p, S.optr.p};
}

// The overflow entries [e0, e1) of row i (e0 == e1: none): a long row's
// entries after the first sk[slice], stored after the slices.
template <class T>
__device__ __forceinline__ void sell_overflow(const SellArgsT<T>& A, long i, long& e0, long& e1) {
  e0 = e1 = 0;
  if (!A.so) return;
  const long s = i >> 6;
  const int ob = A.so[s];
  if (ob < 0) return;
  const unsigned long long m = A.omask[s];
  const int lane = (int)(i & 63);
  if (!((m >> lane) & 1ull)) return;
  const int r = ob + __popcll(m & ((1ull << lane) - 1ull));
  e0 = A.optr[r];
  e1 = A.optr[r + 1];
}

// Row i of the SELL matrix times the vector produced by `g(c)`, entries in
// stored order, accumulated in T.
template <class T, class G>
__device__ __forceinline__ T sell_row(const SellArgsT<T>& A, long i, G g) {
  const long s = i >> 6;
  const long base = A.sptr[s] + (i & 63);
  const int k = A.sk[s];
  T t = 0;
  for (int j = 0; j < k; ++j) {
    const int c = A.col[base + (long)j * 64];
    const T v = A.val[base + (long)j * 64];
    if (c >= 0) t = fma(v, g(c), t);
  }
  long e0, e1;
  sell_overflow(A, i, e0, e1);
  for (long e = e0; e < e1; ++e) t = fma(A.val[e], g(A.col[e]), t);
  return t;
}

// Reduce the nb partials of a producer kernel (reference order of
// common.hpp) in a workgroup of kNT threads.
__device__ __forceinline__ double reduce_part(const double* part, int nb, double* sh) {
  double a = 0.0;
  for (int j = threadIdx.x; j < nb; j += kNT) a = a + part[j];
  return block_sum(a, sh);
}


// Partials of up to three dot products over the tiled layout, written to
// part[slot0], part[slot1], part[slot2] (each kNBMax doubles apart).
template <int ND>
__device__ __forceinline__ void store_partials(double* part, const int* slots, const double* acc, double* sh) {
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    const double v = block_sum(acc[d], sh);
    if (threadIdx.x == 0) part[slots[d] * kNBMax + blockIdx.x] = v;
  }
}

}  // namespace parcel

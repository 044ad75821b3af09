// common.hpp -- shared device helpers for the PARCEL MI355X Krylov library.
//
// Reduction layout (every dot product in the library uses it, so results do
// not depend on launch-time device properties):
//   * a vector of length n is cut into tiles of kNT = 256 consecutive entries;
//   * tile k is owned by workgroup k mod nb, nb = min(ceil(n/kNT), kNBMax);
//   * thread t of a workgroup accumulates fma(x[i], y[i], acc) over
//     i = k*kNT + t for its tiles in increasing k, starting from 0;
//   * each 64-lane wavefront reduces with an xor butterfly (offsets 32..1),
//     then the wave totals of a workgroup are added left to right;
//   * the nb workgroup partials are reduced (by the kernel that consumes the
//     result, in each of its workgroups) with thread t adding partials
//     t, t+256, ... left to right, then the same butterfly and
//     left-to-right wave sum.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace parcel {

constexpr int kNT = 256;      // threads per workgroup for vector kernels
constexpr int kWaves = kNT / 64;
constexpr int kNBMax = 2048;  // max owner workgroups of the tiled layout

// Division of n < 2^31 by a runtime divisor d >= 1 without a division
// instruction: q = (umulhi(n, m) + n) >> s with s = ceil(log2 d) and
// m = floor(2^32 (2^s - d) / d) + 1 (Granlund and Montgomery's round-up
// method; the sum stays below 2^32 because umulhi(n, m) <= n < 2^31).
struct FastDiv {
  unsigned m = 1;
  int s = 0;
  unsigned d = 1;
  FastDiv() = default;
  explicit FastDiv(unsigned dv) : d(dv) {
    while ((1u << s) < dv) ++s;
    m = (unsigned)((((unsigned long long)1 << 32) * (((unsigned long long)1 << s) - dv)) / dv + 1);
  }
  __host__ __device__ __forceinline__ unsigned div(unsigned n) const {
#ifdef __HIP_DEVICE_COMPILE__
    return (__umulhi(n, m) + n) >> s;
#else
    return (unsigned)((((unsigned long long)n * m >> 32) + n) >> s);
#endif
  }
};

inline int tiled_nb(long n) {
  long t = (n + kNT - 1) / kNT;
  if (t < 1) t = 1;
  return (int)(t < kNBMax ? t : kNBMax);
}

struct hip_error : std::runtime_error {
  int code;
  hip_error(const std::string& m, int c) : std::runtime_error(m), code(c) {}
};

#define PARCEL_HIP_CHECK(expr)                                                  \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess)                                                       \
      throw ::parcel::hip_error(std::string(#expr) + ": " + hipGetErrorString(e_), \
                                (int)e_);                                       \
  } while (0)

// --- wavefront / workgroup reductions -------------------------------------

// The value of lane (lane ^ O) of a 64-lane wavefront, in registers: DPP
// quad permutes for O = 1, 2, row shifts by 4 (one up, one down, picked by
// lane bit 2) for O = 4, a row rotation by 8 for O = 8, and the gfx950 row /
// half swaps (v_permlane16_swap, v_permlane32_swap) for O = 16, 32.
template <int O>
__device__ __forceinline__ unsigned xor_lane(unsigned v) {
  const unsigned lane = threadIdx.x & 63;
  if constexpr (O == 1) return __builtin_amdgcn_update_dpp(0u, v, 0xB1, 0xf, 0xf, false);  // quad_perm 1,0,3,2
  if constexpr (O == 2) return __builtin_amdgcn_update_dpp(0u, v, 0x4E, 0xf, 0xf, false);  // quad_perm 2,3,0,1
  if constexpr (O == 4) {
    const unsigned up = __builtin_amdgcn_update_dpp(0u, v, 0x104, 0xf, 0xf, false);    // row_shl:4
    const unsigned down = __builtin_amdgcn_update_dpp(0u, v, 0x114, 0xf, 0xf, false);  // row_shr:4
    return (lane & 4) ? down : up;
  }
  if constexpr (O == 8) return __builtin_amdgcn_update_dpp(0u, v, 0x128, 0xf, 0xf, false);  // row_ror:8
  if constexpr (O == 16) {
    // rows (0, 1, 2, 3) -> first (0, 0, 2, 2), second (1, 1, 3, 3)
    const auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return ((lane >> 4) & 1) ? p[0] : p[1];
  }
  if constexpr (O == 32) {
    // halves (0, 1) -> first (0, 0), second (1, 1)
    const auto p = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return lane >= 32 ? p[0] : p[1];
  }
  return v;
}
template <int O>
__device__ __forceinline__ double xor_lane(double v) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = xor_lane<O>((unsigned)u), hi = xor_lane<O>((unsigned)(u >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Xor butterfly over the 64 lanes, offsets 32 .. 1: after every step lane l
// holds v[l] + v[l ^ o]; every lane ends with the total (the same bits).
__device__ __forceinline__ double wave_sum(double v) {
  v += xor_lane<32>(v);
  v += xor_lane<16>(v);
  v += xor_lane<8>(v);
  v += xor_lane<4>(v);
  v += xor_lane<2>(v);
  v += xor_lane<1>(v);
  return v;
}

// Workgroup of kNT threads; returns the total in every thread.
// `sh` must hold kWaves doubles.
__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double s = sh[0];
#pragma unroll
  for (int k = 1; k < kWaves; ++k) s = s + sh[k];
  return s;
}

}  // namespace parcel

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

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
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

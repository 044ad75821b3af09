// mpnn.hip -- the learned preconditioner: a message-passing network over the
// graph of a 2D five-point matrix predicts the lower factor L of M = L L^T on
// the matrix's own lower pattern.  No counterpart in the reference.
//
// Network (C = 32 channels, 6 layers; a_ij the Jacobi-scaled couplings
// A_ij / sqrt(A_ii A_jj) of the neighbours s, w, e, n of cell i):
//   h0_i = [a_s, a_w, a_e, a_n, 1 + a_s + a_w + a_e + a_n, 1, 0 ...]   (BF16)
//   layer l:  m_i = sum_j a_ij h_j            (FP32, order s w e n; BF16)
//             z_i = W_l [h_i; m_i] + b_l      (W_l 32x64 BF16, MFMA, FP32 acc)
//             h_i = BF16(h_i + relu(z_i))
//   head:     o_i = Wo h_i + bo               (5 outputs, FP32)
//   factor of the scaled matrix, clamp(t) = min(max(t, -10), 10):
//             Lh_ii = sqrt(exp(2 clamp(o0_i)) + 1e-6)
//             Lh_iw = a_iw exp(clamp(o1_i + o3_{i-1}))
//             Lh_is = a_is exp(clamp(o2_i + o4_{i-nx}))
//   L = diag(sqrt(A_ii)) Lh, stored as BF16.  Lh_ii^2 >= 1e-6 keeps M SPD
//   with a margin (the 1e-6 diag(A) shift placed inside the factor so that
//   applying M^-1 stays two triangular solves).
// Training loss per system (x solves A x = b; xh = sqrt(A_ii) x, bh = b /
// sqrt(A_ii)): ||Lh Lh^T xh - bh||^2 / ||bh||^2, minibatch mean; Adam.
//
// The dense part of every layer is a 32 x 64 by 64 x 32 product per tile of
// 32 cells on v_mfma_f32_32x32x16_bf16: operand A the weights (row = output
// channel), operand B the [h; m] columns of 32 cells, so the accumulator of
// lane l holds 16 output channels of cell l % 32.  Backward recomputes z,
// forms dz = dh * (z > 0) in BF16, and runs two more MFMA products per tile:
// W^T dz for d[h; m] and dz [h; m]^T (cells as the K dimension) for dW,
// transposing the tile through LDS.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "batch.hpp"

using namespace parcel;

namespace {

constexpr int kC = 32;
constexpr int kNL = 6;
constexpr int kNO = 8;
constexpr int kLayer = kC * 2 * kC + kC;
constexpr int kOffWo = kNL * kLayer;
constexpr int kOffBo = kOffWo + kNO * kC;
constexpr int kNParam = kOffBo + kNO;
constexpr float kClamp = 10.0f;
constexpr float kEps = 1e-6f;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// round to nearest even (NaN kept quiet)
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

__device__ __forceinline__ uint4 pack8(const float* v) {
  uint4 r;
  r.x = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
  r.y = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
  r.z = (uint32_t)f2bf(v[4]) | ((uint32_t)f2bf(v[5]) << 16);
  r.w = (uint32_t)f2bf(v[6]) | ((uint32_t)f2bf(v[7]) << 16);
  return r;
}

__device__ __forceinline__ void unpack8(uint4 v, float* o) {
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    o[2 * k] = __uint_as_float(w[k] << 16);
    o[2 * k + 1] = __uint_as_float(w[k] & 0xffff0000u);
  }
}

// Output channel held in accumulator register r of lane l (32x32 MFMA layout).
__device__ __forceinline__ int acc_row(int r, int half) { return 8 * (r >> 2) + 4 * half + (r & 3); }

struct Graph {
  const float4* ahat;  // S*n scaled couplings (s, w, e, n)
  int n, nx, S;
};

// Cell index of lane `lane` in tile t (tiles of 32 cells never straddle two
// systems; cells past n are padding: loads clamp to the last cell).
struct Cell {
  long g;       // global index (system*n + i), clamped
  int i;        // index within the system, clamped
  bool valid;
};

__device__ __forceinline__ Cell tile_cell(const Graph& G, long t, int lane) {
  const long tps = ((long)G.n + 31) / 32;
  const long s = t / tps;
  int i = (int)((t - s * tps) * 32 + (lane & 31));
  Cell c;
  c.valid = i < G.n;
  if (!c.valid) i = G.n - 1;
  c.i = i;
  c.g = s * G.n + i;
  return c;
}

__device__ __forceinline__ void neighbours(const Graph& G, const Cell& c, long nb[4]) {
  const int i = c.i, x = i % G.nx;
  const long base = c.g - i;
  nb[0] = base + (i >= G.nx ? i - G.nx : i);
  nb[1] = base + (x > 0 ? i - 1 : i);
  nb[2] = base + (x < G.nx - 1 ? i + 1 : i);
  nb[3] = base + (i + G.nx < G.n ? i + G.nx : i);
}

// B operand of the layer product for the lane's cell: K steps 0..3 hold
// channels [16t + 8 half, +8) of [h; m].
__device__ __forceinline__ void load_hm(const Graph& G, const uint16_t* __restrict__ H, const Cell& c, int half,
                                        bf16x8 hm[4]) {
  const int c0 = 8 * half, c1 = 16 + 8 * half;
  const uint4 h0 = *(const uint4*)(H + c.g * kC + c0);
  const uint4 h1 = *(const uint4*)(H + c.g * kC + c1);
  hm[0] = as_bf16x8(h0);
  hm[1] = as_bf16x8(h1);
  long nb[4];
  neighbours(G, c, nb);
  const float4 a = G.ahat[c.g];
  const float av[4] = {a.x, a.y, a.z, a.w};
  float m0[8], m1[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) { m0[k] = 0.f; m1[k] = 0.f; }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    float v0[8], v1[8];
    unpack8(*(const uint4*)(H + nb[j] * kC + c0), v0);
    unpack8(*(const uint4*)(H + nb[j] * kC + c1), v1);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      m0[k] = fmaf(av[j], v0[k], m0[k]);
      m1[k] = fmaf(av[j], v1[k], m1[k]);
    }
  }
  hm[2] = as_bf16x8(pack8(m0));
  hm[3] = as_bf16x8(pack8(m1));
}

// z = W [h; m] + b for the tile (accumulator layout).
__device__ __forceinline__ floatx16 layer_z(const bf16x8 w[4], const bf16x8 hm[4], const float* __restrict__ b,
                                            int half) {
  floatx16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = b[acc_row(r, half)];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w[t], hm[t], acc, 0, 0, 0);
  return acc;
}

__device__ __forceinline__ void load_w(const uint16_t* __restrict__ W, int lane, bf16x8 w[4]) {
  const int half = lane >> 5, row = lane & 31;
#pragma unroll
  for (int t = 0; t < 4; ++t) w[t] = as_bf16x8(*(const uint4*)(W + row * 64 + 16 * t + 8 * half));
}

// --- forward -------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_nn_prep(int S, int n, int nx, const double* __restrict__ val, long entries,
                                                 const int* __restrict__ dslot, const int4* __restrict__ nslot,
                                                 float4* __restrict__ ahat, float* __restrict__ sd) {
  const long tot = (long)S * n;
  for (long g = (long)blockIdx.x * blockDim.x + threadIdx.x; g < tot; g += (long)gridDim.x * blockDim.x) {
    const long s = g / n;
    const int i = (int)(g - s * n);
    const double* v = val + s * entries;
    const double aii = v[dslot[i]];
    const int4 sl = nslot[i];
    const int slots[4] = {sl.x, sl.y, sl.z, sl.w};
    const int nbr[4] = {i - nx, i - 1, i + 1, i + nx};
    float o[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      o[k] = 0.f;
      if (slots[k] >= 0) {
        const double ajj = v[dslot[nbr[k]]];
        o[k] = (float)(v[slots[k]] / sqrt(aii * ajj));
      }
    }
    ahat[g] = make_float4(o[0], o[1], o[2], o[3]);
    sd[g] = (float)sqrt(aii);
  }
}

__global__ void __launch_bounds__(256) k_nn_input(long tot, const float4* __restrict__ ahat, uint16_t* __restrict__ H) {
  for (long g = (long)blockIdx.x * blockDim.x + threadIdx.x; g < tot; g += (long)gridDim.x * blockDim.x) {
    const float4 a = ahat[g];
    float f[8] = {a.x, a.y, a.z, a.w, (((1.0f + a.x) + a.y) + a.z) + a.w, 1.0f, 0.f, 0.f};
    uint4* out = (uint4*)(H + g * kC);
    out[0] = pack8(f);
    const uint4 z = make_uint4(0, 0, 0, 0);
    out[1] = z; out[2] = z; out[3] = z;
  }
}

// One layer forward; 4 waves per workgroup, one 32-cell tile per wave.
__global__ void __launch_bounds__(256) k_nn_layer(Graph G, const uint16_t* __restrict__ Hin,
                                                  uint16_t* __restrict__ Hout, const uint16_t* __restrict__ W,
                                                  const float* __restrict__ b, long ntiles) {
  const int lane = threadIdx.x & 63, half = lane >> 5;
  const long t = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= ntiles) return;
  bf16x8 w[4];
  load_w(W, lane, w);
  const Cell c = tile_cell(G, t, lane);
  bf16x8 hm[4];
  load_hm(G, Hin, c, half, hm);
  const floatx16 z = layer_z(w, hm, b, half);
  if (!c.valid) return;
#pragma unroll
  for (int gq = 0; gq < 4; ++gq) {
    const int ch = 8 * gq + 4 * half;
    const uint2 hv = *(const uint2*)(Hin + c.g * kC + ch);
    const float h0 = __uint_as_float(hv.x << 16), h1 = __uint_as_float(hv.x & 0xffff0000u);
    const float h2 = __uint_as_float(hv.y << 16), h3 = __uint_as_float(hv.y & 0xffff0000u);
    const float o0 = h0 + fmaxf(z[4 * gq + 0], 0.f), o1 = h1 + fmaxf(z[4 * gq + 1], 0.f);
    const float o2 = h2 + fmaxf(z[4 * gq + 2], 0.f), o3 = h3 + fmaxf(z[4 * gq + 3], 0.f);
    uint2 ov;
    ov.x = (uint32_t)f2bf(o0) | ((uint32_t)f2bf(o1) << 16);
    ov.y = (uint32_t)f2bf(o2) | ((uint32_t)f2bf(o3) << 16);
    *(uint2*)(Hout + c.g * kC + ch) = ov;
  }
}

// head: o = Wo h + bo (FP32, channels in order), 8 floats per cell
__global__ void __launch_bounds__(256) k_nn_head(long tot, const uint16_t* __restrict__ H, const float* __restrict__ P,
                                                 float* __restrict__ O) {
  __shared__ float wo[kNO * kC + kNO];
  for (int k = threadIdx.x; k < kNO * kC + kNO; k += blockDim.x) wo[k] = P[kOffWo + k];
  __syncthreads();
  for (long g = (long)blockIdx.x * blockDim.x + threadIdx.x; g < tot; g += (long)gridDim.x * blockDim.x) {
    float h[kC];
#pragma unroll
    for (int q = 0; q < 4; ++q) unpack8(*(const uint4*)(H + g * kC + 8 * q), h + 8 * q);
    float o[kNO];
#pragma unroll
    for (int k = 0; k < kNO; ++k) {
      float acc = wo[kNO * kC + k];
#pragma unroll
      for (int ch = 0; ch < kC; ++ch) acc = fmaf(wo[k * kC + ch], h[ch], acc);
      o[k] = acc;
    }
    float4* out = (float4*)(O + g * kNO);
    out[0] = make_float4(o[0], o[1], o[2], o[3]);
    out[1] = make_float4(o[4], o[5], o[6], o[7]);
  }
}

__device__ __forceinline__ float clampf(float t, float c) { return fminf(fmaxf(t, -c), c); }

// Factor from the head outputs.  Lh (3*tot, FP32) and/or Lb (3*tot, BF16 of
// sd * Lh) may be null.
__global__ void __launch_bounds__(256) k_nn_factor(int S, int n, int nx, const float* __restrict__ O,
                                                   const float4* __restrict__ ahat, const float* __restrict__ sd,
                                                   float* __restrict__ Lh, uint16_t* __restrict__ Lb) {
  const long tot = (long)S * n;
  for (long g = (long)blockIdx.x * blockDim.x + threadIdx.x; g < tot; g += (long)gridDim.x * blockDim.x) {
    const long s = g / n;
    const int i = (int)(g - s * n), x = i % nx;
    const float* o = O + g * kNO;
    const float4 a = ahat[g];
    const float e0 = expf(clampf(2.0f * o[0], 2.0f * kClamp));
    const float ld = sqrtf(e0 + kEps);
    const float tw = o[1] + (x > 0 ? O[(g - 1) * kNO + 3] : 0.0f);
    const float ts = o[2] + (i >= nx ? O[(g - nx) * kNO + 4] : 0.0f);
    const float lw = a.y * expf(clampf(tw, kClamp));
    const float ls = a.x * expf(clampf(ts, kClamp));
    if (Lh) { Lh[g] = ls; Lh[tot + g] = lw; Lh[2 * tot + g] = ld; }
    if (Lb) {
      const float d = sd[g];
      Lb[g] = f2bf(d * ls);
      Lb[tot + g] = f2bf(d * lw);
      Lb[2 * tot + g] = f2bf(d * ld);
    }
  }
}

// --- loss ------------------------------------------------------------------------

// out = Lh^T v: Ld_j v_j + Lw_{j+1} v_{j+1} + Ls_{j+nx} v_{j+nx}; v = scale*src
__global__ void __launch_bounds__(256) k_nn_lt(int S, int n, int nx, const float* __restrict__ Lh,
                                               const float* __restrict__ v, const double* __restrict__ scale,
                                               float* __restrict__ out) {
  const long tot = (long)S * n;
  for (long g = (long)blockIdx.x * blockDim.x + threadIdx.x; g < tot; g += (long)gridDim.x * blockDim.x) {
    const long s = g / n;
    const int j = (int)(g - s * n);
    const float sc = scale ? (float)scale[s] : 1.0f;
    float acc = Lh[2 * tot + g] * (sc * v[g]);
    if (j % nx < nx - 1 && j + 1 < n) acc = fmaf(Lh[tot + g + 1], sc * v[g + 1], acc);
    if (j + nx < n) acc = fmaf(Lh[g + nx], sc * v[g + nx], acc);
    out[g] = acc;
  }
}

// r = Lh u - bh; per-system partial sums of r^2 (double, one per workgroup)
__global__ void __launch_bounds__(256) k_nn_resid(int S, int n, int nx, const float* __restrict__ Lh,
                                                  const float* __restrict__ u, const float* __restrict__ bh,
                                                  float* __restrict__ r, double* __restrict__ part, int nbs) {
  __shared__ double sh[4];
  const long tot = (long)S * n;
  const int s = blockIdx.y;
  double acc = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)nbs * blockDim.x) {
    const long g = (long)s * n + i;
    float v = Lh[2 * tot + g] * u[g];
    if (i % nx > 0) v = fmaf(Lh[tot + g], u[g - 1], v);
    if (i >= nx) v = fmaf(Lh[g], u[g - nx], v);
    const float ri = v - bh[g];
    r[g] = ri;
    acc += (double)ri * (double)ri;
  }
  const double tsum = block_sum(acc, sh);
  if (threadIdx.x == 0) part[(long)s * nbs + blockIdx.x] = tsum;
}

// Gradients of the head outputs (dO, 8 per cell) from g = gs*r (gs per
// system), u = Lh^T xh, w = Lh^T g.
__global__ void __launch_bounds__(256) k_nn_factor_bwd(int S, int n, int nx, const float* __restrict__ O,
                                                       const float* __restrict__ Lh, const float* __restrict__ r,
                                                       const double* __restrict__ gscale, const float* __restrict__ u,
                                                       const float* __restrict__ w, const float* __restrict__ xh,
                                                       float* __restrict__ dO) {
  const long tot = (long)S * n;
  for (long g = (long)blockIdx.x * blockDim.x + threadIdx.x; g < tot; g += (long)gridDim.x * blockDim.x) {
    const long s = g / n;
    const int i = (int)(g - s * n), x = i % nx;
    const float gsc = (float)gscale[s];
    const float* o = O + g * kNO;
    // own entries: G_ii, G_iw, G_is
    const float gi = gsc * r[g], xi = xh[g];
    const float Gd = fmaf(gi, u[g], xi * w[g]);
    const float t0 = 2.0f * o[0];
    const float e0 = expf(clampf(t0, 2.0f * kClamp));
    float d0 = fabsf(t0) < 2.0f * kClamp ? Gd * e0 / Lh[2 * tot + g] : 0.0f;
    float d1 = 0.f, d2 = 0.f, d3 = 0.f, d4 = 0.f;
    if (x > 0) {
      const float Gw = fmaf(gi, u[g - 1], xi * w[g - 1]);
      const float tw = o[1] + O[(g - 1) * kNO + 3];
      d1 = fabsf(tw) < kClamp ? Gw * Lh[tot + g] : 0.0f;
    }
    if (i >= nx) {
      const float Gs = fmaf(gi, u[g - nx], xi * w[g - nx]);
      const float ts = o[2] + O[(g - nx) * kNO + 4];
      d2 = fabsf(ts) < kClamp ? Gs * Lh[g] : 0.0f;
    }
    // entries of the neighbours that read this cell's o3 / o4
    if (x + 1 < nx) {
      const long e = g + 1;
      const float ge = gsc * r[e];
      const float Gw = fmaf(ge, u[g], xh[e] * w[g]);
      const float tw = O[e * kNO + 1] + o[3];
      d3 = fabsf(tw) < kClamp ? Gw * Lh[tot + e] : 0.0f;
    }
    if (i + nx < n) {
      const long e = g + nx;
      const float ge = gsc * r[e];
      const float Gs = fmaf(ge, u[g], xh[e] * w[g]);
      const float ts = O[e * kNO + 2] + o[4];
      d4 = fabsf(ts) < kClamp ? Gs * Lh[e] : 0.0f;
    }
    float4* out = (float4*)(dO + g * kNO);
    out[0] = make_float4(d0, d1, d2, d3);
    out[1] = make_float4(d4, 0.f, 0.f, 0.f);
  }
}

// --- backward ------------------------------------------------------------------

// dh = Wo^T dO per cell; per-workgroup partials of dWo (8x32) and dbo (8)
__global__ void __launch_bounds__(256) k_nn_head_bwd(long tot, const uint16_t* __restrict__ H,
                                                     const float* __restrict__ dO, const float* __restrict__ P,
                                                     float* __restrict__ dH, float* __restrict__ part) {
  __shared__ float wo[kNO * kC];
  __shared__ float sdo[256 * 5];
  __shared__ float sh[256 * (kC + 1)];
  for (int k = threadIdx.x; k < kNO * kC; k += blockDim.x) wo[k] = P[kOffWo + k];
  float pacc = 0.f;  // thread t < 165: pair (k = t / 32, c = t % 32), t >= 160: bias k - 160
  for (long base = (long)blockIdx.x * 256; base < tot; base += (long)gridDim.x * 256) {
    const long g = base + threadIdx.x;
    __syncthreads();
    float d[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
    float h[kC];
    if (g < tot) {
#pragma unroll
      for (int k = 0; k < 5; ++k) d[k] = dO[g * kNO + k];
#pragma unroll
      for (int q = 0; q < 4; ++q) unpack8(*(const uint4*)(H + g * kC + 8 * q), h + 8 * q);
      float out[kC];
#pragma unroll
      for (int ch = 0; ch < kC; ++ch) {
        float a = 0.f;
#pragma unroll
        for (int k = 0; k < 5; ++k) a = fmaf(wo[k * kC + ch], d[k], a);
        out[ch] = a;
      }
#pragma unroll
      for (int q = 0; q < 8; ++q)
        *(float4*)(dH + g * kC + 4 * q) = make_float4(out[4 * q], out[4 * q + 1], out[4 * q + 2], out[4 * q + 3]);
    } else {
#pragma unroll
      for (int ch = 0; ch < kC; ++ch) h[ch] = 0.f;
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) sdo[threadIdx.x * 5 + k] = d[k];
#pragma unroll
    for (int ch = 0; ch < kC; ++ch) sh[threadIdx.x * (kC + 1) + ch] = h[ch];
    __syncthreads();
    const int t = threadIdx.x;
    if (t < 5 * kC) {
      const int k = t / kC, ch = t % kC;
      for (int j = 0; j < 256; ++j) pacc = fmaf(sdo[j * 5 + k], sh[j * (kC + 1) + ch], pacc);
    } else if (t < 5 * kC + 5) {
      const int k = t - 5 * kC;
      for (int j = 0; j < 256; ++j) pacc += sdo[j * 5 + k];
    }
  }
  if (threadIdx.x < 5 * kC + 5) part[(long)blockIdx.x * (5 * kC + 5) + threadIdx.x] = pacc;
}

// One layer backward.  In: h^{l} (Hin), dh^{l+1} (dH).  Out: dHp = dH + W_h^T dz,
// dM = W_m^T dz (aggregated into dh^{l} by k_nn_agg_bwd), per-wave partials of
// dW (32x64) and db (32).  Every wave of a workgroup runs the same number of
// iterations (tiles past the end are masked) so the LDS barriers line up.
__global__ void __launch_bounds__(256) k_nn_layer_bwd(Graph G, const uint16_t* __restrict__ Hin,
                                                      const float* __restrict__ dH, const uint16_t* __restrict__ W,
                                                      const uint16_t* __restrict__ WT, const float* __restrict__ b,
                                                      float* __restrict__ dHp, float* __restrict__ dM,
                                                      float* __restrict__ dWpart, float* __restrict__ dbpart,
                                                      long ntiles) {
  __shared__ uint16_t lds_dzT[4][32 * 32];   // [cell][channel]
  __shared__ uint16_t lds_dzN[4][32 * 32];   // [channel][cell]
  __shared__ uint16_t lds_hmN[4][64 * 32];   // [channel][cell]
  const int lane = threadIdx.x & 63, half = lane >> 5, wv = threadIdx.x >> 6, col = lane & 31;
  uint16_t* dzT = lds_dzT[wv];
  uint16_t* dzN = lds_dzN[wv];
  uint16_t* hmN = lds_hmN[wv];
  bf16x8 w[4], wt[2][2];
  load_w(W, lane, w);
#pragma unroll
  for (int tl = 0; tl < 2; ++tl)
#pragma unroll
    for (int t = 0; t < 2; ++t)
      wt[tl][t] = as_bf16x8(*(const uint4*)(WT + (tl * 32 + col) * 32 + 16 * t + 8 * half));
  floatx16 dw0, dw1;
  float dbacc[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) { dw0[r] = 0.f; dw1[r] = 0.f; dbacc[r] = 0.f; }
  const long nwg_tiles = (long)gridDim.x * 4;
  for (long t0 = (long)blockIdx.x * 4; t0 < ntiles; t0 += nwg_tiles) {
    const long t = t0 + wv;
    const bool active = t < ntiles;
    bf16x8 hm[4];
    Cell c = tile_cell(G, active ? t : ntiles - 1, lane);
    const bool use = active && c.valid;
    load_hm(G, Hin, c, half, hm);
    const floatx16 z = layer_z(w, hm, b, half);
    float dhv[16];
#pragma unroll
    for (int gq = 0; gq < 4; ++gq) {
      const float4 v = *(const float4*)(dH + c.g * kC + 8 * gq + 4 * half);
      dhv[4 * gq] = v.x; dhv[4 * gq + 1] = v.y; dhv[4 * gq + 2] = v.z; dhv[4 * gq + 3] = v.w;
    }
    uint16_t dzb[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const float dz = (use && z[r] > 0.f) ? dhv[r] : 0.f;
      dbacc[r] += dz;
      dzb[r] = f2bf(dz);
    }
    // stage dz (both layouts) and [h; m] (channel-major) in LDS
#pragma unroll
    for (int gq = 0; gq < 4; ++gq) {
      const int ch = 8 * gq + 4 * half;
      uint2 v;
      v.x = (uint32_t)dzb[4 * gq] | ((uint32_t)dzb[4 * gq + 1] << 16);
      v.y = (uint32_t)dzb[4 * gq + 2] | ((uint32_t)dzb[4 * gq + 3] << 16);
      *(uint2*)(dzT + col * 32 + ch) = v;
#pragma unroll
      for (int e = 0; e < 4; ++e) dzN[(ch + e) * 32 + col] = dzb[4 * gq + e];
    }
#pragma unroll
    for (int tk = 0; tk < 4; ++tk) {
      const uint4 v = __builtin_bit_cast(uint4, hm[tk]);
      const uint32_t wd[4] = {v.x, v.y, v.z, v.w};
      const int ch = 16 * tk + 8 * half;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        // padding cells contribute nothing to dW: their dz is zero already
        hmN[(ch + 2 * e) * 32 + col] = (uint16_t)(wd[e] & 0xffffu);
        hmN[(ch + 2 * e + 1) * 32 + col] = (uint16_t)(wd[e] >> 16);
      }
    }
    __syncthreads();
    // d[h; m] = W^T dz: two 32x32 tiles (input channels 0-31 and 32-63), K = 32
    floatx16 a0, a1;
#pragma unroll
    for (int r = 0; r < 16; ++r) { a0[r] = 0.f; a1[r] = 0.f; }
#pragma unroll
    for (int tk = 0; tk < 2; ++tk) {
      const bf16x8 bz = as_bf16x8(*(const uint4*)(dzT + col * 32 + 16 * tk + 8 * half));
      a0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wt[0][tk], bz, a0, 0, 0, 0);
      a1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wt[1][tk], bz, a1, 0, 0, 0);
    }
    // dW += dz [h; m]^T with the tile's cells as K
#pragma unroll
    for (int tk = 0; tk < 2; ++tk) {
      const bf16x8 az = as_bf16x8(*(const uint4*)(dzN + col * 32 + 16 * tk + 8 * half));
      const bf16x8 bh0 = as_bf16x8(*(const uint4*)(hmN + col * 32 + 16 * tk + 8 * half));
      const bf16x8 bh1 = as_bf16x8(*(const uint4*)(hmN + (32 + col) * 32 + 16 * tk + 8 * half));
      dw0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(az, bh0, dw0, 0, 0, 0);
      dw1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(az, bh1, dw1, 0, 0, 0);
    }
    if (use) {
#pragma unroll
      for (int gq = 0; gq < 4; ++gq) {
        const int ch = 8 * gq + 4 * half;
        *(float4*)(dHp + c.g * kC + ch) =
            make_float4(dhv[4 * gq] + a0[4 * gq], dhv[4 * gq + 1] + a0[4 * gq + 1], dhv[4 * gq + 2] + a0[4 * gq + 2],
                        dhv[4 * gq + 3] + a0[4 * gq + 3]);
        *(float4*)(dM + c.g * kC + ch) = make_float4(a1[4 * gq], a1[4 * gq + 1], a1[4 * gq + 2], a1[4 * gq + 3]);
      }
    }
    __syncthreads();
  }
  // per-wave partials: dW[out][in] (out = accumulator row, in = column)
  const long gw = (long)blockIdx.x * 4 + wv;
  float* dw = dWpart + gw * (kC * 2 * kC);
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int out = acc_row(r, half);
    dw[out * 64 + col] = dw0[r];
    dw[out * 64 + 32 + col] = dw1[r];
  }
  // db: sum over the 32 cells of each half-wave
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    float v = dbacc[r];
#pragma unroll
    for (int o = 1; o < 32; o <<= 1) v += __shfl_xor(v, o, 64);
    dbacc[r] = v;
  }
  if (col == 0) {
#pragma unroll
    for (int r = 0; r < 16; ++r) dbpart[gw * kC + acc_row(r, half)] = dbacc[r];
  }
}

// dh^{l}_i = dHp_i + sum_j a_ij dM_j  (order s, w, e, n)
__global__ void __launch_bounds__(256) k_nn_agg_bwd(Graph G, const float* __restrict__ dHp,
                                                    const float* __restrict__ dM, float* __restrict__ dH) {
  const long tot = (long)G.S * G.n * (kC / 4);
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < tot; q += (long)gridDim.x * blockDim.x) {
    const long g = q / (kC / 4);
    const int ch = (int)(q - g * (kC / 4)) * 4;
    const long s = g / G.n;
    Cell c;
    c.g = g;
    c.i = (int)(g - s * G.n);
    c.valid = true;
    long nb[4];
    neighbours(G, c, nb);
    const float4 a = G.ahat[g];
    const float av[4] = {a.x, a.y, a.z, a.w};
    float4 acc = *(const float4*)(dHp + g * kC + ch);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float4 m = *(const float4*)(dM + nb[j] * kC + ch);
      acc.x = fmaf(av[j], m.x, acc.x);
      acc.y = fmaf(av[j], m.y, acc.y);
      acc.z = fmaf(av[j], m.z, acc.z);
      acc.w = fmaf(av[j], m.w, acc.w);
    }
    *(float4*)(dH + g * kC + ch) = acc;
  }
}

// out[k] = sum_p part[p*len + k] (fixed order, double accumulation)
__global__ void k_nn_reduce(const float* __restrict__ part, long nparts, int len, float* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= len) return;
  double acc = 0.0;
  for (long p = 0; p < nparts; ++p) acc += (double)part[p * len + k];
  out[k] = (float)acc;
}

// head partial layout (5*32 pairs then 5 biases) -> grad
__global__ void k_nn_reduce_head(const float* __restrict__ part, long nparts, float* __restrict__ grad) {
  const int t = threadIdx.x;
  if (t >= 5 * kC + 5) return;
  double acc = 0.0;
  for (long p = 0; p < nparts; ++p) acc += (double)part[p * (5 * kC + 5) + t];
  if (t < 5 * kC) grad[kOffWo + t] = (float)acc;  // rows 0..4 of Wo (row-major [k][c])
  else grad[kOffBo + (t - 5 * kC)] = (float)acc;
}

// BF16 copies of the layer weights, row-major W[l][out][in] and W^T[l][in][out]
__global__ void k_nn_pack(const float* __restrict__ P, uint16_t* __restrict__ W, uint16_t* __restrict__ WT) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= kNL * kC * 2 * kC) return;
  const int l = k / (kC * 2 * kC), rem = k % (kC * 2 * kC), out = rem / (2 * kC), in = rem % (2 * kC);
  const uint16_t v = f2bf(P[l * kLayer + out * 2 * kC + in]);
  W[k] = v;
  WT[l * kC * 2 * kC + in * kC + out] = v;
}

__global__ void k_nn_adam(int np, float* __restrict__ P, const float* __restrict__ g, float* __restrict__ m,
                          float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= np) return;
  const float gk = g[k];
  const float mk = b1 * m[k] + (1.0f - b1) * gk;
  const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
  m[k] = mk;
  v[k] = vk;
  P[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

unsigned grid_for(long work, int per_block = 256) {
  return (unsigned)std::max<long>(1, std::min<long>((work + per_block - 1) / per_block, 8192));
}

}  // namespace

namespace parcel {

// Network state on one device: FP32 master parameters and their BF16 copies.
struct Net {
  int device = 0;
  hipStream_t stream = nullptr;
  DevBuf<float> P;            // kNParam
  DevBuf<uint16_t> W, WT;     // kNL * 32 * 64
  void pack() {
    hipLaunchKernelGGL(k_nn_pack, dim3((kNL * kC * 2 * kC + 255) / 256), dim3(256), 0, stream, (const float*)P.p,
                       W.p, WT.p);
  }
};

// Activations and scratch of one forward/backward over S systems.
struct Work {
  int S = 0, n = 0, nx = 0;
  bool train = false;
  DevBuf<float4> ahat;
  DevBuf<float> sd;
  std::vector<DevBuf<uint16_t>> H;  // 7 (training) or 2 (inference) activations
  DevBuf<float> O;
  void ensure(int S_, int n_, int nx_, bool train_) {
    const long tot = (long)S_ * n_;
    if (S_ == S && n_ == n && nx_ == nx && train_ == train) return;
    S = S_; n = n_; nx = nx_; train = train_;
    ahat.alloc(tot);
    sd.alloc(tot);
    H.clear();
    const int nh = train ? kNL + 1 : 2;
    for (int k = 0; k < nh; ++k) H.emplace_back((size_t)tot * kC);
    O.alloc((size_t)tot * kNO);
  }
};

// Forward over the S systems held in w (ahat, sd filled).  Returns the index
// of the last activation buffer.
int net_forward(Net& N, Work& w) {
  const long tot = (long)w.S * w.n;
  const Graph G{w.ahat.p, w.n, w.nx, w.S};
  const long ntiles = (long)w.S * (((long)w.n + 31) / 32);
  hipLaunchKernelGGL(k_nn_input, dim3(grid_for(tot)), dim3(256), 0, N.stream, tot, (const float4*)w.ahat.p, w.H[0].p);
  int cur = 0;
  for (int l = 0; l < kNL; ++l) {
    const int in = w.train ? l : cur, out = w.train ? l + 1 : 1 - cur;
    hipLaunchKernelGGL(k_nn_layer, dim3((unsigned)((ntiles + 3) / 4)), dim3(256), 0, N.stream, G,
                       (const uint16_t*)w.H[in].p, w.H[out].p, (const uint16_t*)(N.W.p + (long)l * kC * 2 * kC),
                       (const float*)(N.P.p + (long)l * kLayer + kC * 2 * kC), ntiles);
    cur = out;
  }
  hipLaunchKernelGGL(k_nn_head, dim3(grid_for(tot)), dim3(256), 0, N.stream, tot, (const uint16_t*)w.H[cur].p,
                     (const float*)N.P.p, w.O.p);
  PARCEL_HIP_CHECK(hipGetLastError());
  return cur;
}

}  // namespace parcel

struct parcel_hip_mpnn {
  Net N;
  Work w;
};

// --- trainer ------------------------------------------------------------------------

namespace parcel {

struct Trainer {
  Net N;
  int n = 0, nx = 0, cap = 0, count = 0, mb = 0;
  DevBuf<float4> pool_ahat;
  DevBuf<float> pool_sd, pool_xh, pool_bh;
  std::vector<double> bnorm2;  // ||bh||^2 per pooled system
  Work w;
  DevBuf<float> xh, bh, Lh, u, r, wv, dO, dH, dHp, dM, grad, m, v, part_w, part_b, part_h;
  DevBuf<double> rpart, gsc;
  long adam_t = 0;
  int nbs = 0;
};

}  // namespace parcel

struct parcel_hip_mpnn_trainer {
  Trainer T;
};

namespace {

#define PARCEL_API_BEGIN try {
#define PARCEL_API_END                                                   \
  }                                                                      \
  catch (const hip_error& e) {                                           \
    return api_fail(e.code < 0 ? e.code : PARCEL_HIP_ERR_HIP, e.what()); \
  }                                                                      \
  catch (const std::exception& e) {                                      \
    return api_fail(PARCEL_HIP_ERR_HIP, e.what());                       \
  }

void upload_params(Net& N, const float* params) {
  N.P.alloc(kNParam);
  N.W.alloc((size_t)kNL * kC * 2 * kC);
  N.WT.alloc((size_t)kNL * kC * 2 * kC);
  PARCEL_HIP_CHECK(hipMemcpyAsync(N.P.p, params, sizeof(float) * kNParam,
                                  api_is_device_ptr(params) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                  N.stream));
  N.pack();
  PARCEL_HIP_CHECK(hipGetLastError());
}

// ahat, sd of the batch's systems [s0, s0+S) into w
void prep_from_batch(Net& N, Work& w, const Batch& B, int s0, int S) {
  const long tot = (long)S * B.n;
  hipLaunchKernelGGL(k_nn_prep, dim3(grid_for(tot)), dim3(256), 0, N.stream, S, B.n, B.nx,
                     (const double*)(B.A.val.p + (long)s0 * B.A.entries), B.A.entries, (const int*)B.diag_slot.p,
                     (const int4*)B.nbr_slot.p, w.ahat.p, w.sd.p);
  PARCEL_HIP_CHECK(hipGetLastError());
}

int check_batch(const Batch& B) {
  if (!B.nx) return api_fail(PARCEL_HIP_ERR_UNSUPPORTED, "mpnn: the batch is not a 2D five-point grid");
  return PARCEL_HIP_OK;
}

// xh = sd * x, bh = b / sd (FP32) for systems of a batch into the pool
__global__ void k_nn_scale_xb(long tot, const float* __restrict__ sd, const double* __restrict__ x,
                              const double* __restrict__ b, float* __restrict__ xh, float* __restrict__ bh) {
  for (long g = (long)blockIdx.x * blockDim.x + threadIdx.x; g < tot; g += (long)gridDim.x * blockDim.x) {
    const double d = (double)sd[g];
    xh[g] = (float)(d * x[g]);
    bh[g] = (float)(b[g] / d);
  }
}

// minibatch copy from the pool
void gather(Trainer& T, const int* idx, int S) {
  const long n = T.n;
  for (int k = 0; k < S; ++k) {
    const long src = (long)idx[k] * n, dst = (long)k * n;
    PARCEL_HIP_CHECK(hipMemcpyAsync(T.w.ahat.p + dst, T.pool_ahat.p + src, sizeof(float4) * n,
                                    hipMemcpyDeviceToDevice, T.N.stream));
    PARCEL_HIP_CHECK(hipMemcpyAsync(T.w.sd.p + dst, T.pool_sd.p + src, sizeof(float) * n, hipMemcpyDeviceToDevice,
                                    T.N.stream));
    PARCEL_HIP_CHECK(hipMemcpyAsync(T.xh.p + dst, T.pool_xh.p + src, sizeof(float) * n, hipMemcpyDeviceToDevice,
                                    T.N.stream));
    PARCEL_HIP_CHECK(hipMemcpyAsync(T.bh.p + dst, T.pool_bh.p + src, sizeof(float) * n, hipMemcpyDeviceToDevice,
                                    T.N.stream));
  }
}

// Loss (minibatch mean) and gradient into T.grad for the minibatch idx[0..S).
double loss_grad(Trainer& T, const int* idx, int S) {
  Net& N = T.N;
  const hipStream_t st = N.stream;
  const int n = T.n, nx = T.nx;
  const long tot = (long)S * n;
  T.w.ensure(S, n, nx, true);
  if (T.xh.n < (size_t)tot) {
    for (DevBuf<float>* b : {&T.xh, &T.bh, &T.u, &T.r, &T.wv}) b->alloc(tot);
    T.Lh.alloc(3 * tot);
    T.dO.alloc(tot * kNO);
    for (DevBuf<float>* b : {&T.dH, &T.dHp, &T.dM}) b->alloc(tot * kC);
  }
  gather(T, idx, S);
  const int last = net_forward(N, T.w);
  (void)last;
  hipLaunchKernelGGL(k_nn_factor, dim3(grid_for(tot)), dim3(256), 0, st, S, n, nx, (const float*)T.w.O.p,
                     (const float4*)T.w.ahat.p, (const float*)T.w.sd.p, T.Lh.p, (uint16_t*)nullptr);
  // u = Lh^T xh; r = Lh u - bh; loss partials
  hipLaunchKernelGGL(k_nn_lt, dim3(grid_for(tot)), dim3(256), 0, st, S, n, nx, (const float*)T.Lh.p,
                     (const float*)T.xh.p, (const double*)nullptr, T.u.p);
  T.nbs = (int)std::min<long>(((long)n + 255) / 256, 256);
  if (T.rpart.n < (size_t)S * T.nbs) T.rpart.alloc((size_t)S * T.nbs);
  hipLaunchKernelGGL(k_nn_resid, dim3(T.nbs, S), dim3(256), 0, st, S, n, nx, (const float*)T.Lh.p,
                     (const float*)T.u.p, (const float*)T.bh.p, T.r.p, T.rpart.p, T.nbs);
  std::vector<double> hp((size_t)S * T.nbs);
  PARCEL_HIP_CHECK(hipMemcpyAsync(hp.data(), T.rpart.p, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, st));
  PARCEL_HIP_CHECK(hipStreamSynchronize(st));
  double loss = 0.0;
  std::vector<double> gs(S);
  for (int k = 0; k < S; ++k) {
    double rr = 0.0;
    for (int j = 0; j < T.nbs; ++j) rr += hp[(size_t)k * T.nbs + j];
    const double bb = T.bnorm2[idx[k]];
    loss += rr / bb / S;
    gs[k] = 2.0 / (bb * S);
  }
  if (T.gsc.n < (size_t)S) T.gsc.alloc(S);
  PARCEL_HIP_CHECK(hipMemcpyAsync(T.gsc.p, gs.data(), sizeof(double) * S, hipMemcpyHostToDevice, st));
  // w = Lh^T g, g = gs * r
  hipLaunchKernelGGL(k_nn_lt, dim3(grid_for(tot)), dim3(256), 0, st, S, n, nx, (const float*)T.Lh.p,
                     (const float*)T.r.p, (const double*)T.gsc.p, T.wv.p);
  hipLaunchKernelGGL(k_nn_factor_bwd, dim3(grid_for(tot)), dim3(256), 0, st, S, n, nx, (const float*)T.w.O.p,
                     (const float*)T.Lh.p, (const float*)T.r.p, (const double*)T.gsc.p, (const float*)T.u.p,
                     (const float*)T.wv.p, (const float*)T.xh.p, T.dO.p);
  // head backward
  const unsigned hb = grid_for(tot);
  if (T.part_h.n < (size_t)hb * (5 * kC + 5)) T.part_h.alloc((size_t)hb * (5 * kC + 5));
  if (!T.grad.p) T.grad.alloc(kNParam);
  PARCEL_HIP_CHECK(hipMemsetAsync(T.grad.p, 0, sizeof(float) * kNParam, st));
  hipLaunchKernelGGL(k_nn_head_bwd, dim3(hb), dim3(256), 0, st, tot, (const uint16_t*)T.w.H[kNL].p,
                     (const float*)T.dO.p, (const float*)N.P.p, T.dH.p, T.part_h.p);
  hipLaunchKernelGGL(k_nn_reduce_head, dim3(1), dim3(256), 0, st, (const float*)T.part_h.p, (long)hb, T.grad.p);
  // layers, last to first
  const Graph G{T.w.ahat.p, n, nx, S};
  const long ntiles = (long)S * (((long)n + 31) / 32);
  const unsigned lb = (unsigned)std::max<long>(1, std::min<long>((ntiles + 3) / 4, 1024));
  const long nw = (long)lb * 4;
  if (T.part_w.n < (size_t)nw * kC * 2 * kC) T.part_w.alloc((size_t)nw * kC * 2 * kC);
  if (T.part_b.n < (size_t)nw * kC) T.part_b.alloc((size_t)nw * kC);
  for (int l = kNL - 1; l >= 0; --l) {
    hipLaunchKernelGGL(k_nn_layer_bwd, dim3(lb), dim3(256), 0, st, G, (const uint16_t*)T.w.H[l].p,
                       (const float*)T.dH.p, (const uint16_t*)(N.W.p + (long)l * kC * 2 * kC),
                       (const uint16_t*)(N.WT.p + (long)l * kC * 2 * kC),
                       (const float*)(N.P.p + (long)l * kLayer + kC * 2 * kC), T.dHp.p, T.dM.p, T.part_w.p,
                       T.part_b.p, ntiles);
    hipLaunchKernelGGL(k_nn_reduce, dim3((kC * 2 * kC + 255) / 256), dim3(256), 0, st, (const float*)T.part_w.p, nw,
                       kC * 2 * kC, T.grad.p + (long)l * kLayer);
    hipLaunchKernelGGL(k_nn_reduce, dim3(1), dim3(256), 0, st, (const float*)T.part_b.p, nw, kC,
                       T.grad.p + (long)l * kLayer + kC * 2 * kC);
    if (l > 0)
      hipLaunchKernelGGL(k_nn_agg_bwd, dim3(grid_for(tot * (kC / 4))), dim3(256), 0, st, G, (const float*)T.dHp.p,
                         (const float*)T.dM.p, T.dH.p);
  }
  PARCEL_HIP_CHECK(hipGetLastError());
  return loss;
}

}  // namespace

extern "C" {

int parcel_hip_mpnn_nparams(void) { return kNParam; }

int parcel_hip_mpnn_create(parcel_hip_mpnn** out, const float* params, int nparams) {
  PARCEL_API_BEGIN
  if (!out || !params) return api_fail(PARCEL_HIP_ERR_ARG, "mpnn_create: null argument");
  if (nparams != kNParam) return api_fail(PARCEL_HIP_ERR_ARG, "mpnn_create: wrong parameter count");
  *out = nullptr;
  int rc = api_ensure_init();
  if (rc != PARCEL_HIP_OK) return rc;
  auto* h = new parcel_hip_mpnn;
  h->N.device = api_device();
  try {
    PARCEL_HIP_CHECK(hipSetDevice(h->N.device));
    PARCEL_HIP_CHECK(hipStreamCreateWithFlags(&h->N.stream, hipStreamNonBlocking));
    upload_params(h->N, params);
    PARCEL_HIP_CHECK(hipStreamSynchronize(h->N.stream));
  } catch (...) {
    if (h->N.stream) (void)hipStreamDestroy(h->N.stream);
    delete h;
    throw;
  }
  *out = h;
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

int parcel_hip_mpnn_factor(parcel_hip_mpnn* h, parcel_hip_batch* hb) {
  PARCEL_API_BEGIN
  if (!h || !hb) return api_fail(PARCEL_HIP_ERR_ARG, "mpnn_factor: null argument");
  Batch& B = hb->B;
  if (B.precon != PARCEL_BATCH_FACTOR)
    return api_fail(PARCEL_HIP_ERR_ARG, "mpnn_factor: the batch was not created with PARCEL_BATCH_FACTOR");
  int rc = check_batch(B);
  if (rc != PARCEL_HIP_OK) return rc;
  PARCEL_HIP_CHECK(hipSetDevice(h->N.device));
  Net& N = h->N;
  // order after any work already queued on the batch's stream
  PARCEL_HIP_CHECK(hipStreamSynchronize(B.stream));
  h->w.ensure(B.nsys, B.n, B.nx, false);
  prep_from_batch(N, h->w, B, 0, B.nsys);
  net_forward(N, h->w);
  const long nv = B.nv();
  B.Lbf.alloc((size_t)(3 * nv));
  hipLaunchKernelGGL(k_nn_factor, dim3(grid_for(nv)), dim3(256), 0, N.stream, B.nsys, B.n, B.nx,
                     (const float*)h->w.O.p, (const float4*)h->w.ahat.p, (const float*)h->w.sd.p, (float*)nullptr,
                     B.Lbf.p);
  PARCEL_HIP_CHECK(hipGetLastError());
  PARCEL_HIP_CHECK(hipStreamSynchronize(N.stream));
  B.factor_src = 2;
  derive_factor(B);
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

// Head outputs (S*n*8 floats) and the BF16 factor of every system of the
// batch, for tests; Lbits: 3*nsys*n (Ls, Lw, Ld), o: nsys*n*8, h: the last
// activation (nsys*n*32 BF16 bits); each may be NULL.
int parcel_hip_mpnn_inspect(parcel_hip_mpnn* h, parcel_hip_batch* hb, float* o, unsigned short* hlast,
                            float* ahat, float* sd) {
  PARCEL_API_BEGIN
  if (!h || !hb) return api_fail(PARCEL_HIP_ERR_ARG, "mpnn_inspect: null argument");
  Batch& B = hb->B;
  int rc = check_batch(B);
  if (rc != PARCEL_HIP_OK) return rc;
  PARCEL_HIP_CHECK(hipSetDevice(h->N.device));
  h->w.ensure(B.nsys, B.n, B.nx, false);
  prep_from_batch(h->N, h->w, B, 0, B.nsys);
  const int last = net_forward(h->N, h->w);
  const long nv = B.nv();
  const hipStream_t st = h->N.stream;
  if (o) PARCEL_HIP_CHECK(hipMemcpyAsync(o, h->w.O.p, sizeof(float) * nv * kNO, hipMemcpyDeviceToHost, st));
  if (hlast)
    PARCEL_HIP_CHECK(hipMemcpyAsync(hlast, h->w.H[last].p, sizeof(uint16_t) * nv * kC, hipMemcpyDeviceToHost, st));
  if (ahat) PARCEL_HIP_CHECK(hipMemcpyAsync(ahat, h->w.ahat.p, sizeof(float4) * nv, hipMemcpyDeviceToHost, st));
  if (sd) PARCEL_HIP_CHECK(hipMemcpyAsync(sd, h->w.sd.p, sizeof(float) * nv, hipMemcpyDeviceToHost, st));
  PARCEL_HIP_CHECK(hipStreamSynchronize(st));
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

int parcel_hip_mpnn_destroy(parcel_hip_mpnn* h) {
  if (!h) return PARCEL_HIP_OK;
  (void)hipSetDevice(h->N.device);
  (void)hipStreamSynchronize(h->N.stream);
  hipStream_t s = h->N.stream;
  delete h;
  if (s) (void)hipStreamDestroy(s);
  return PARCEL_HIP_OK;
}

int parcel_hip_mpnn_trainer_create(parcel_hip_mpnn_trainer** out, int n, int nx, int capacity,
                                   const float* params) {
  PARCEL_API_BEGIN
  if (!out || !params || n < 1 || nx < 1 || n % nx != 0 || capacity < 1)
    return api_fail(PARCEL_HIP_ERR_ARG, "mpnn_trainer_create: bad argument");
  *out = nullptr;
  int rc = api_ensure_init();
  if (rc != PARCEL_HIP_OK) return rc;
  auto* h = new parcel_hip_mpnn_trainer;
  Trainer& T = h->T;
  T.N.device = api_device();
  try {
    PARCEL_HIP_CHECK(hipSetDevice(T.N.device));
    PARCEL_HIP_CHECK(hipStreamCreateWithFlags(&T.N.stream, hipStreamNonBlocking));
    upload_params(T.N, params);
    T.n = n; T.nx = nx; T.cap = capacity;
    const long tot = (long)capacity * n;
    T.pool_ahat.alloc(tot);
    T.pool_sd.alloc(tot);
    T.pool_xh.alloc(tot);
    T.pool_bh.alloc(tot);
    T.m.alloc(kNParam);
    T.v.alloc(kNParam);
    PARCEL_HIP_CHECK(hipMemsetAsync(T.m.p, 0, sizeof(float) * kNParam, T.N.stream));
    PARCEL_HIP_CHECK(hipMemsetAsync(T.v.p, 0, sizeof(float) * kNParam, T.N.stream));
    PARCEL_HIP_CHECK(hipStreamSynchronize(T.N.stream));
  } catch (...) {
    if (T.N.stream) (void)hipStreamDestroy(T.N.stream);
    delete h;
    throw;
  }
  *out = h;
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

// Add the systems of a batch with their solutions x and right-hand sides b
// (nsys*n each, host or device) to the training pool.
int parcel_hip_mpnn_trainer_add(parcel_hip_mpnn_trainer* h, parcel_hip_batch* hb, const double* x, const double* b) {
  PARCEL_API_BEGIN
  if (!h || !hb || !x || !b) return api_fail(PARCEL_HIP_ERR_ARG, "mpnn_trainer_add: null argument");
  Trainer& T = h->T;
  Batch& B = hb->B;
  int rc = check_batch(B);
  if (rc != PARCEL_HIP_OK) return rc;
  if (B.n != T.n || B.nx != T.nx) return api_fail(PARCEL_HIP_ERR_ARG, "mpnn_trainer_add: grid differs");
  if (T.count + B.nsys > T.cap) return api_fail(PARCEL_HIP_ERR_ARG, "mpnn_trainer_add: pool full");
  PARCEL_HIP_CHECK(hipSetDevice(T.N.device));
  const hipStream_t st = T.N.stream;
  PARCEL_HIP_CHECK(hipStreamSynchronize(B.stream));
  const long nv = B.nv(), off = (long)T.count * T.n;
  Work tmp;
  tmp.ensure(B.nsys, B.n, B.nx, false);
  prep_from_batch(T.N, tmp, B, 0, B.nsys);
  DevBuf<double> dx, db;
  const double* px = x;
  const double* pb = b;
  if (!api_is_device_ptr(x)) { dx.alloc(nv); dx.upload(x, nv, st); px = dx.p; }
  if (!api_is_device_ptr(b)) { db.alloc(nv); db.upload(b, nv, st); pb = db.p; }
  PARCEL_HIP_CHECK(hipMemcpyAsync(T.pool_ahat.p + off, tmp.ahat.p, sizeof(float4) * nv, hipMemcpyDeviceToDevice, st));
  PARCEL_HIP_CHECK(hipMemcpyAsync(T.pool_sd.p + off, tmp.sd.p, sizeof(float) * nv, hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_nn_scale_xb, dim3(grid_for(nv)), dim3(256), 0, st, nv, (const float*)tmp.sd.p, px, pb,
                     T.pool_xh.p + off, T.pool_bh.p + off);
  PARCEL_HIP_CHECK(hipGetLastError());
  std::vector<float> hbh(nv);
  PARCEL_HIP_CHECK(hipMemcpyAsync(hbh.data(), T.pool_bh.p + off, sizeof(float) * nv, hipMemcpyDeviceToHost, st));
  PARCEL_HIP_CHECK(hipStreamSynchronize(st));
  for (int s = 0; s < B.nsys; ++s) {
    double bb = 0.0;
    for (long i = 0; i < T.n; ++i) bb += (double)hbh[(long)s * T.n + i] * (double)hbh[(long)s * T.n + i];
    T.bnorm2.push_back(bb);
  }
  T.count += B.nsys;
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

// Loss and gradient (host, kNParam floats) of the minibatch idx[0..S).
int parcel_hip_mpnn_trainer_grad(parcel_hip_mpnn_trainer* h, const int* idx, int S, float* grad, double* loss) {
  PARCEL_API_BEGIN
  if (!h || !idx || S < 1) return api_fail(PARCEL_HIP_ERR_ARG, "mpnn_trainer_grad: bad argument");
  Trainer& T = h->T;
  for (int k = 0; k < S; ++k)
    if (idx[k] < 0 || idx[k] >= T.count) return api_fail(PARCEL_HIP_ERR_ARG, "mpnn_trainer_grad: index out of pool");
  PARCEL_HIP_CHECK(hipSetDevice(T.N.device));
  const double l = loss_grad(T, idx, S);
  if (grad)
    PARCEL_HIP_CHECK(hipMemcpyAsync(grad, T.grad.p, sizeof(float) * kNParam, hipMemcpyDeviceToHost, T.N.stream));
  PARCEL_HIP_CHECK(hipStreamSynchronize(T.N.stream));
  if (loss) *loss = l;
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

// One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the minibatch idx.
int parcel_hip_mpnn_trainer_step(parcel_hip_mpnn_trainer* h, const int* idx, int S, float lr, double* loss) {
  PARCEL_API_BEGIN
  if (!h || !idx || S < 1 || !(lr > 0.f)) return api_fail(PARCEL_HIP_ERR_ARG, "mpnn_trainer_step: bad argument");
  Trainer& T = h->T;
  for (int k = 0; k < S; ++k)
    if (idx[k] < 0 || idx[k] >= T.count) return api_fail(PARCEL_HIP_ERR_ARG, "mpnn_trainer_step: index out of pool");
  PARCEL_HIP_CHECK(hipSetDevice(T.N.device));
  const double l = loss_grad(T, idx, S);
  T.adam_t++;
  const float c1 = 1.0f - powf(0.9f, (float)T.adam_t), c2 = 1.0f - powf(0.999f, (float)T.adam_t);
  hipLaunchKernelGGL(k_nn_adam, dim3((kNParam + 255) / 256), dim3(256), 0, T.N.stream, kNParam, T.N.P.p,
                     (const float*)T.grad.p, T.m.p, T.v.p, lr, 0.9f, 0.999f, 1e-8f, c1, c2);
  T.N.pack();
  PARCEL_HIP_CHECK(hipGetLastError());
  PARCEL_HIP_CHECK(hipStreamSynchronize(T.N.stream));
  if (loss) *loss = l;
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

int parcel_hip_mpnn_trainer_params(parcel_hip_mpnn_trainer* h, float* params) {
  PARCEL_API_BEGIN
  if (!h || !params) return api_fail(PARCEL_HIP_ERR_ARG, "mpnn_trainer_params: null argument");
  PARCEL_HIP_CHECK(hipSetDevice(h->T.N.device));
  PARCEL_HIP_CHECK(hipMemcpyAsync(params, h->T.N.P.p, sizeof(float) * kNParam, hipMemcpyDeviceToHost, h->T.N.stream));
  PARCEL_HIP_CHECK(hipStreamSynchronize(h->T.N.stream));
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

int parcel_hip_mpnn_trainer_destroy(parcel_hip_mpnn_trainer* h) {
  if (!h) return PARCEL_HIP_OK;
  (void)hipSetDevice(h->T.N.device);
  (void)hipStreamSynchronize(h->T.N.stream);
  hipStream_t s = h->T.N.stream;
  delete h;
  if (s) (void)hipStreamDestroy(s);
  return PARCEL_HIP_OK;
}

}  // extern "C"

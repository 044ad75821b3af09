// batched.hip -- many systems with one sparsity pattern, solved together.
//
// The reference solves one system per call (KSP_pCG, src/pcg.f90:224-892).
// A batch keeps nsys systems resident at once: the pattern in SELL-64 form
// once, the values of system s at s*entries, every vector of system s at
// s*n.  Each CG kernel of cg_kernels.hpp runs with gridDim.y = nsys, so one
// launch advances every system by one step with exactly the single-system
// arithmetic; a system that has converged returns at the top of every kernel.
//
// Factor preconditioners (IC(0) and a supplied lower factor L, M = L L^T)
// need a 2D five-point grid in natural order: row i = x + nx*y couples to
// i-1 and i-nx below the diagonal.  The forward solve L w = r on grid line y
//     w_x = a_x w_{x-1} + c_x,  a_x = -Lw_x / Ld_x,
//     c_x = (r_x - Ls_x w_x^{(y-1)}) / Ld_x
// is a first-order linear recurrence along the line once line y-1 is known,
// i.e. a scan of affine maps.  A workgroup of up to 4 wavefronts per strip of
// lines walks the lines; each wavefront owns a piece of every line, its lane
// l E consecutive cells: the lane composes its E maps, an inclusive 64-lane
// scan (DPP row shifts and broadcasts) gives each lane the value entering
// its cells from the piece's start, the piece's entering value comes from
// the wavefront before (an LDS ring, the wavefronts pipelined a few lines
// apart), and the lane re-applies its maps.  The previous line's values a
// lane needs are its own (registers), and later lines' loads are issued
// ahead.  The backward solve L^T z = w runs the lines top-down with the
// mirrored layout.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "batch.hpp"
#include "cg_kernels.hpp"

using namespace parcel;

namespace {

// --- bf16 -------------------------------------------------------------------
__host__ __device__ inline float bf16_to_float(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

template <class S, class T>
__device__ __forceinline__ T cvt(S v) {
  return (T)v;
}
template <>
__device__ __forceinline__ float cvt<uint16_t, float>(uint16_t v) {
  return bf16_to_float(v);
}
template <>
__device__ __forceinline__ double cvt<uint16_t, double>(uint16_t v) {
  return (double)bf16_to_float(v);
}

// --- setup kernels ------------------------------------------------------------

// SELL values of every system from CSR values: slot e takes CSR entry idx[e]
// (-1: padding, 0).
__global__ void k_b_gather(long entries, long nnz, const int* __restrict__ idx, const double* __restrict__ csr,
                           double* __restrict__ sell, float* __restrict__ sell32) {
  const long s = blockIdx.y;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < entries; e += (long)gridDim.x * blockDim.x) {
    const int j = idx[e];
    const double v = j >= 0 ? csr[s * nnz + j] : 0.0;
    sell[s * entries + e] = v;
    sell32[s * entries + e] = (float)v;
  }
}

// Five-point stencil arrays c = a(i,i), e = a(i,i+1), n = a(i,i+nx) (0 where
// the entry is absent) of every system, and bad = 1 when a system is not
// exactly symmetric (a(i,i-1) != a(i-1,i) or a(i,i-nx) != a(i-nx,i), bitwise).
__global__ void k_b_st5(int n, int nx, int ny, long entries, long nv, const int* __restrict__ dslot,
                        const int4* __restrict__ nslot, const double* __restrict__ sell, double* __restrict__ out,
                        int* bad) {
  const long s = blockIdx.y;
  const double* v = sell + s * entries;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int x = (int)(i % nx), y = (int)(i / nx);
    const int4 sl = nslot[i];
    out[s * n + i] = v[dslot[i]];
    out[nv + s * n + i] = x < nx - 1 ? v[sl.z] : 0.0;
    out[2 * nv + s * n + i] = y < ny - 1 ? v[sl.w] : 0.0;
    bool ok = true;
    if (x > 0) ok = ok && __double_as_longlong(v[sl.y]) == __double_as_longlong(v[nslot[i - 1].z]);
    if (y > 0) ok = ok && __double_as_longlong(v[sl.x]) == __double_as_longlong(v[nslot[i - nx].w]);
    if (!ok) *bad = 1;
  }
}

// Point Jacobi: dinv = 1 / a_ii (the single-system setup), FP32 copy rounded.
__global__ void k_b_dinv(int n, long entries, const int* __restrict__ dslot, const double* __restrict__ sell,
                         double* __restrict__ dinv, float* __restrict__ dinv32) {
  const long s = blockIdx.y;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const double d = 1.0 / sell[s * entries + dslot[i]];
    dinv[s * n + i] = d;
    dinv32[s * n + i] = (float)d;
  }
}

// Apply arrays of the factor solves, in T from factor values of type S:
//   ls = Ls, inv = 1/Ld, af = -Lw * inv, ab = -Lw(x+1) * inv (0 on the last
// cell of a line).  L holds Ls, Lw, Ld (each nv apart); F gets ls, inv, af, ab.
template <class S, class T>
__global__ void k_b_derive(int n, int nx, long nv, const S* __restrict__ L, T* __restrict__ F) {
  const long s = blockIdx.y;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long g = s * n + i;
    const T ld = cvt<S, T>(L[2 * nv + g]);
    const T inv = (T)1 / ld;
    const T lw = cvt<S, T>(L[nv + g]);
    const int x = (int)(i % nx);
    const T lwn = x + 1 < nx ? cvt<S, T>(L[nv + g + 1]) : (T)0;
    F[g] = cvt<S, T>(L[g]);
    F[nv + g] = inv;
    F[2 * nv + g] = -lw * inv;
    F[3 * nv + g] = -lwn * inv;
  }
}

// --- factor solves: z = (L L^T)^-1 r, one workgroup per strip of grid lines -----

__device__ __forceinline__ bool sys_done(const CGState* st, int sys, int it) {
  if (!st) return false;
  const int d = st[sys].done_at;
  return d != 0 && d <= it;
}

// Registers of one grid line in flight.  The loads of line t + ahead are
// issued as soon as line t is done with its registers, so each line's load
// latency overlaps the recurrences of the lines before it.  q stays in its
// load type until use: a conversion at load time would wait for the load.
// A line takes E (3 sizeof(T) + sizeof(TQ)) / 4 VGPRs per lane; up to about
// 160 of them stay in flight.
template <class T, class TQ, int E>
constexpr int lines_ahead() {
  constexpr int regs = E * (3 * (int)sizeof(T) + (int)sizeof(TQ)) / 4;
  return regs >= 96 ? 1 : (regs >= 48 ? 2 : (regs >= 24 ? 4 : 8));
}
template <class T, class TQ, int E>
struct LineRegs {
  T a[E], p[E], i[E];
  TQ q[E];
};

// base + i elements with a 32-bit byte offset (a strip's arrays are far
// below 4 GiB): the access takes the uniform base in scalar registers and a
// 32-bit lane offset
template <class T>
__device__ __forceinline__ T* at(T* base, unsigned i) {
  using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
  return (T*)((B*)base + i * (unsigned)sizeof(T));
}

// The lane's E cells of one line: elements i0 .. i0+E-1 of the strip's
// arrays (i0 = line start + x0); MIRROR stores them in reverse (slot k holds
// cell x0+E-1-k), which the backward sweep uses.  Padding cells (x >= nx)
// load the line's last cell and take zeros (no branch).
template <class T, class TQ, int E, bool FULL, bool MIRROR>
__device__ __forceinline__ void load_line(LineRegs<T, TQ, E>& L, const T* __restrict__ a, const T* __restrict__ p,
                                          const TQ* __restrict__ q, const T* __restrict__ inv, unsigned i0, int x0,
                                          int nx) {
#pragma unroll
  for (int j = 0; j < E; ++j) {
    const int k = MIRROR ? E - 1 - j : j;
    if (FULL) {
      L.a[k] = at(a, i0)[j]; L.p[k] = at(p, i0)[j]; L.q[k] = at(q, i0)[j]; L.i[k] = at(inv, i0)[j];
    } else {
      const bool in = x0 + j < nx;
      const unsigned c = in ? i0 + j : i0 - x0 + nx - 1;
      const T av = a[c], pv = p[c], iv = inv[c];
      const TQ qv = q[c];
      L.a[k] = in ? av : (T)0; L.p[k] = in ? pv : (T)0; L.q[k] = in ? qv : (TQ)0; L.i[k] = in ? iv : (T)0;
    }
  }
}

// DPP lane moves (GFX9 controls): row_shr:n = 0x110+n, row_bcast:15 = 0x142,
// row_bcast:31 = 0x143, wave_shr:1 = 0x138.  Lanes without a source, or in
// rows outside RM, receive `old`.
template <int CTRL, int RM>
__device__ __forceinline__ float dpp(float old, float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), CTRL, RM, 0xf, false));
}
template <int CTRL, int RM>
__device__ __forceinline__ double dpp(double old, double v) {
  const long long o = __double_as_longlong(old), x = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp((int)o, (int)x, CTRL, RM, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(x >> 32), CTRL, RM, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Whether lane `lane` receives a map in the scan step with DPP control CTRL
// (row_shr:n inside rows of 16; row_bcast:15 into rows 1 and 3;
// row_bcast:31 into rows 2 and 3).
template <int CTRL>
__device__ __forceinline__ bool scan_has_source(int lane) {
  if (CTRL == 0x142) return (lane >> 4) & 1;
  if (CTRL == 0x143) return lane >= 32;
  return (lane & 15) >= (CTRL - 0x110);
}

// One scan step: the lane's map (A, C) after the map (Ap, Cp) moved in by
// CTRL; a lane that receives nothing keeps its map.
template <int CTRL, int RM, class T>
__device__ __forceinline__ void scan_step(T& A, T& C) {
  const T Ap = dpp<CTRL, RM>((T)1, A), Cp = dpp<CTRL, RM>((T)0, C);
  const T Cn = fma(A, Cp, C);
  A = A * Ap;
  C = scan_has_source<CTRL>(threadIdx.x & 63) ? Cn : C;
}

// Inclusive scan of affine maps over the 64 lanes in lane order: shifts
// 1, 2, 4, 8 inside each row of 16, then row 0's last map into rows 1 and 3
// (row_bcast:15) and lane 31's into rows 2 and 3 (row_bcast:31).
template <class T>
__device__ __forceinline__ void scan_inclusive(T& A, T& C) {
  scan_step<0x111, 0xf>(A, C);
  scan_step<0x112, 0xf>(A, C);
  scan_step<0x114, 0xf>(A, C);
  scan_step<0x118, 0xf>(A, C);
  scan_step<0x142, 0xa>(A, C);
  scan_step<0x143, 0xc>(A, C);
}

// FP32: each step is one FMA and one multiply that take the moved-in operand
// straight from DPP (C += C(src) A, A *= A(src)); a lane without a source
// (bound_ctrl off, or its row masked) is not written and keeps its map, as
// above, with half the dependent instructions.  A DPP read needs two wait
// states after a VALU write of its register: the order fma, mul, s_nop 0
// gives each of C and A exactly two between its write and its next read.
template <>
__device__ __forceinline__ void scan_inclusive<float>(float& A, float& C) {
#define PARCEL_SCAN_STEP(CTL)                                           \
  "v_fmac_f32_dpp %0, %0, %1 " CTL " bank_mask:0xf\n\t"                \
  "v_mul_f32_dpp %1, %1, %1 " CTL " bank_mask:0xf\n\t"                 \
  "s_nop 0\n\t"
  asm volatile(
      "s_nop 1\n\t"
      PARCEL_SCAN_STEP("row_shr:1 row_mask:0xf")
      PARCEL_SCAN_STEP("row_shr:2 row_mask:0xf")
      PARCEL_SCAN_STEP("row_shr:4 row_mask:0xf")
      PARCEL_SCAN_STEP("row_shr:8 row_mask:0xf")
      PARCEL_SCAN_STEP("row_bcast:15 row_mask:0xa")
      PARCEL_SCAN_STEP("row_bcast:31 row_mask:0xc")
      "s_nop 0"
      : "+v"(C), "+v"(A));
#undef PARCEL_SCAN_STEP
}

// Returns the value entering the lane's cells: the scanned maps of the lanes
// before it applied to vin when `chained` (lane 0: vin itself), else to 0.
template <class T>
__device__ __forceinline__ T scan_maps(T A, T C, bool chained, T vin) {
  scan_inclusive(A, C);
  const T Ce = dpp<0x138, 0xf>((T)0, C);
  const T Cv = fma(dpp<0x138, 0xf>((T)1, A), vin, Ce);
  return chained ? Cv : Ce;  // a select: no branch inside the line
}

// One line of a sweep: v = a v_prev + c along the lane's cells, with
// c = fma(-p, carry, q) * inv, carry = the value of the same cell on the
// previous line, p = ls of that previous line for the backward sweep.
// Returns the value of the lane's last cell.  A line that is not `valid`
// (a pipeline filler) leaves carry and pn as they are (selects, no branch);
// the caller points its stores at the sink, where padding cells store too.
template <class T, class TQ, class TO, int E, bool FULL, bool BWD>
__device__ __forceinline__ T sweep_line(const LineRegs<T, TQ, E>& L, T (&carry)[E], T (&pn)[E], int x0, int nx,
                                        TO* obase, unsigned orow, TO* sink, bool chained, T vin, bool valid) {
  T c[E];
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int x = BWD ? x0 + E - 1 - k : x0 + k;
    if (FULL || x < nx) {
      // forward: p is this line's ls; backward: the previous (upper) line's
      const T pk = BWD ? pn[k] : L.p[k];
      c[k] = fma(-pk, carry[k], (T)L.q[k]) * L.i[k];
      if (BWD) pn[k] = valid ? L.p[k] : pn[k];
    } else {
      c[k] = 0;
      if (BWD) pn[k] = 0;
    }
  }
  T A = 1, C = 0;
#pragma unroll
  for (int k = 0; k < E; ++k) { C = fma(L.a[k], C, c[k]); A = L.a[k] * A; }
  T v = scan_maps(A, C, chained, vin);
#pragma unroll
  for (int k = 0; k < E; ++k) {
    const int x = BWD ? x0 + E - 1 - k : x0 + k;
    v = fma(L.a[k], v, c[k]);
    carry[k] = valid ? v : carry[k];
    // the cell is element orow + x of obase; a padding cell
    // (x < 64 E NW <= 1024) stores to the sink
    if (FULL)
      at(obase, orow + (unsigned)x0)[x - x0] = (TO)v;
    else
      *(x < nx ? obase + orow + x : sink + x) = (TO)v;
  }
  return v;
}

// Lines per pipeline step of a multi-wavefront sweep (a multiple of every
// lines_ahead value).
constexpr int kLineBlk = 8;

// One sweep over the lines of a strip by the NW wavefronts of the workgroup
// (forward: y = 0 .. ny-1, cells in increasing x; backward: y = ny-1 .. 0,
// cells mirrored so the recurrence still runs in lane order).  A line is cut
// into NW pieces of 64 E cells in recurrence order; wavefront wv owns piece
// wv and starts each line from the last value of piece wv-1 on the same
// line.  The wavefronts run a pipeline of kLineBlk-line steps: at step s
// wavefront wv takes lines (s - wv) kLineBlk + j, j < kLineBlk, and hands
// every lane's last values to wavefront wv+1 through the LDS ring (two steps
// deep, slot (s kLineBlk + j) mod 2 kLineBlk), a barrier between steps.
// Every step runs its kLineBlk lines without a branch, so the compiler's
// wait before a line's first use counts only that line's loads: lines
// outside [0, ny) (a wavefront's first wv and last NW-1-wv steps, the end of
// the last step) are fillers that load the nearest line, store to `sink` (a
// line of scratch no one reads) and leave the recurrence state alone.  The
// loads of line t + kAhead are issued as soon as line t is done with its
// registers, with a scheduling barrier between lines.
template <class T, class TQ, class TO, int E, int NW, bool FULL, bool BWD>
__device__ __forceinline__ void sweep(int nx, int ny, const T* __restrict__ a, const T* __restrict__ ls,
                                      const TQ* __restrict__ q, const T* __restrict__ inv, TO* __restrict__ out,
                                      TO* __restrict__ sink, T (*ring)[2 * kLineBlk][64]) {
  constexpr int kAhead = lines_ahead<T, TQ, E>();
  static_assert(kLineBlk % kAhead == 0, "pipeline steps hold whole prefetch groups");
  constexpr int W = 64 * E;  // cells per wavefront piece
  // the wavefront index is uniform in the wavefront: readfirstlane lets the
  // compiler keep it (and what derives from it) in scalar registers
  const int wv = NW > 1 ? __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) : 0;
  const int lane = threadIdx.x & 63;
  const int x0 = BWD ? NW * W - wv * W - (lane + 1) * E : wv * W + lane * E;
  const bool chained = wv > 0;
  // first element of line t (clamped to the strip), 32-bit
  auto row_of = [&](int t) -> unsigned {
    t = min(max(t, 0), ny - 1);
    return (unsigned)((BWD ? ny - 1 - t : t) * nx);
  };
  T carry[E], pn[E];
#pragma unroll
  for (int k = 0; k < E; ++k) { carry[k] = 0; pn[k] = 0; }
  LineRegs<T, TQ, E> buf[kAhead];
  const int tw = -wv * kLineBlk;  // this wavefront's first line (its fillers first)
  // issued in line order (the loop's first wait then counts the later
  // lines' loads, as on every later trip)
#pragma unroll
  for (int d = 0; d < kAhead; ++d) {
    load_line<T, TQ, E, FULL, BWD>(buf[d], a, ls, q, inv, row_of(tw + d) + (unsigned)x0, x0, nx);
    __builtin_amdgcn_sched_barrier(0);
  }
  const int nsteps = (ny + kLineBlk - 1) / kLineBlk + NW - 1;
  for (int step = 0; step < nsteps; ++step) {
    const int tb = (step - wv) * kLineBlk;
#pragma unroll
    for (int j = 0; j < kLineBlk; ++j) {
      const int t = tb + j;
      const bool valid = t >= 0 && t < ny;
      const int slot = (step * kLineBlk + j) % (2 * kLineBlk);
      // the piece before this one, written by wavefront wv-1 a step earlier
      // (wavefront 0 reads a slot too and ignores it)
      const T vin = NW > 1 ? ring[chained ? wv - 1 : 0][(slot + kLineBlk) % (2 * kLineBlk)][63] : (T)0;
      // a filler's stores go to the sink's first line: the store base is a
      // select of two global pointers and the row a select of two scalars
      TO* obase = valid ? out : sink;
      const unsigned orow = valid ? row_of(t) : 0u;
      const T vl = sweep_line<T, TQ, TO, E, FULL, BWD>(buf[j % kAhead], carry, pn, x0, nx, obase, orow, sink,
                                                       chained, vin, valid);
      if (NW > 1) ring[wv][slot][lane] = vl;
      load_line<T, TQ, E, FULL, BWD>(buf[j % kAhead], a, ls, q, inv, row_of(t + kAhead) + (unsigned)x0, x0, nx);
      // keep the next line's work (its q conversions above all) from being
      // hoisted into this one, where it would wait on loads just issued
      __builtin_amdgcn_sched_barrier(0);
    }
    if (NW > 1) __syncthreads();
  }
}

// z = (L L^T)^-1 r for every system: forward sweep into w, then the
// backward sweep into z, one workgroup of NW wavefronts per strip of grid
// lines.  With nstrip strips per system, strip k holds lines
// [ny k / nstrip, ny (k+1) / nstrip) and drops the factor's couplings to the
// neighbouring strips (block-diagonal M = diag(L_k L_k^T)); nstrip = 1 is
// the exact M = L L^T.  T: arithmetic and factor type; TV: the type of r
// and z (double with T = float: the FP32 apply of an FP64 PCG); TZ: the
// stored type of z (float there too when the CG keeps z in FP32: its values
// are FP32 results, widened exactly where they are read).
template <class T, class TV, int E, int NW, bool FULL, class TZ = TV>
__global__ void __launch_bounds__(64 * NW) k_lscan(int nx, int ny, int nstrip, long n, long nv,
                                                   const T* __restrict__ F, const TV* __restrict__ r,
                                                   T* __restrict__ w, TZ* __restrict__ z, const CGState* st, int it,
                                                   double* __restrict__ sink) {
  __shared__ T ring[NW][2 * kLineBlk][64];
  const int sys = blockIdx.x / nstrip, strip = blockIdx.x - sys * nstrip;
  if (sys_done(st, sys, it)) return;
  const int y0 = (int)((long)ny * strip / nstrip), y1 = (int)((long)ny * (strip + 1) / nstrip);
  if (y1 <= y0) return;
  const long so = (long)sys * n + (long)y0 * nx;
  const T* ls = F + so;
  const T* inv = F + nv + so;
  const T* af = F + 2 * nv + so;
  const T* ab = F + 3 * nv + so;
  sweep<T, TV, T, E, NW, FULL, false>(nx, y1 - y0, af, ls, r + so, inv, w + so, (T*)sink, ring);
  // the backward sweep reads w cells written by other lanes
  __syncthreads();
  sweep<T, T, TZ, E, NW, FULL, true>(nx, y1 - y0, ab, ls, w + so, inv, z + so, (TZ*)sink, ring);
}

// s = Dinv r (dinv null: s = r), every system
template <class T>
__global__ void __launch_bounds__(kNT) k_b_scale(int n, const T* __restrict__ r, const T* __restrict__ dinv,
                                                 T* __restrict__ s, Bat B) {
  const long so = sys_id() * B.vs;
  r += so; s += so;
  if (dinv) dinv += so;
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  for (long k = tile_owner(); k < ntiles; k += nb) {
    const long i = k * kNT + threadIdx.x;
    if (i < n) s[i] = dinv ? r[i] * dinv[i] : r[i];
  }
}

// K2 of a factor-preconditioned iteration: alpha = rho0 / gamma; r -= alpha q
// --- mixed-precision outer loop ---------------------------------------------------

// Before FP64 correction k (k = 1, 2, ...): norms of the current residual
// (partials of k_residual), exit test, history.  The order follows
// PARCEL_SOLVER_CG_MIXED: at least one correction; after correction k-1 the
// solve ends when ||r|| / ||b|| < rtol, or when k-1 = itrmax.
__global__ void __launch_bounds__(kNT) k_bo_check(int k, int itrmax, double rtol, double* part, int nb,
                                                  OuterState* os, double* hist, int hist_cap) {
  __shared__ double sh[kWaves];
  os += sys_id();
  if (os->done) return;
  part = sys_part(part);
  hist += sys_id() * hist_cap;
  const double rr = reduce_part(part + PART_RR0 * kNBMax, nb, sh);
  if (k == 1) {
    const double bb = reduce_part(part + PART_BB * kNBMax, nb, sh);
    if (threadIdx.x == 0) {
      os->bnorm = sqrt(bb);
      os->rnorm = sqrt(rr);
      if (hist_cap > 0) hist[0] = os->rnorm / os->bnorm;
    }
    return;
  }
  if (threadIdx.x != 0) return;
  const double rnorm = sqrt(rr);
  os->rnorm = rnorm;
  const double ratio = rnorm / os->bnorm;
  if (ratio < rtol) {
    os->done = 1;
    os->niter = k - 1 < itrmax ? k - 1 : itrmax;
    os->iret = itrmax <= k - 1 ? 1 : 0;
  } else if (k - 1 >= itrmax) {
    os->done = 1;
    os->niter = itrmax;
    os->iret = 1;
  } else if (k - 1 < hist_cap) {
    hist[k - 1] = ratio;
  }
}

// r32 = (float) r, d32 = 0 for systems still refining
__global__ void __launch_bounds__(kNT) k_bo_tof32(int n, const double* __restrict__ r, float* __restrict__ r32,
                                                  float* __restrict__ d32, const OuterState* os, Bat B) {
  if (os[sys_id()].done) return;
  const long so = sys_id() * B.vs;
  r += so; r32 += so; d32 += so;
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  for (long k = tile_owner(); k < ntiles; k += nb) {
    const long i = k * kNT + threadIdx.x;
    if (i < n) { r32[i] = (float)r[i]; d32[i] = 0.0f; }
  }
}

// inner solves of finished systems are marked ended before they start
__global__ void k_bo_mask(const OuterState* os, CGState* st, int nsys) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nsys && os[s].done) { st[s].done_at = -1; st[s].niter = 0; }
}

// x += (double) d for systems still refining
__global__ void __launch_bounds__(kNT) k_bo_xadd(int n, double* __restrict__ x, const float* __restrict__ d,
                                                 const OuterState* os, Bat B) {
  if (os[sys_id()].done) return;
  const long so = sys_id() * B.vs;
  x += so; d += so;
  const int nb = gridDim.x;
  const long ntiles = ((long)n + kNT - 1) / kNT;
  for (long k = tile_owner(); k < ntiles; k += nb) {
    const long i = k * kNT + threadIdx.x;
    if (i < n) x[i] = x[i] + (double)d[i];
  }
}

// --- host helpers ------------------------------------------------------------------

// nsys systems of ny lines of nx cells, nstrip strips each; F: the apply
// arrays (nv = nsys n apart); one workgroup per strip
template <class T, class TV, class TZ = TV>
void lscan_launch(int nx, int ny, int nstrip, int nsys, long n, long nv, const T* F, const TV* r, T* w, TZ* z,
                  const CGState* st, int it, double* sink, hipStream_t stream) {
  const dim3 g(nsys * nstrip);
  // line layout (the oracle's lscan_layout): E1 = cells per lane of one
  // wavefront, a power of two with 64 E1 >= nx; E1 >= 8 splits the line
  // over 4 wavefronts, E1 = 4 over 2
#define PARCEL_LSCAN(E, NW)                                                                                     \
  if (nx == 64 * (E) * (NW) && ((n * sizeof(T)) % 16 == 0))                                                    \
    hipLaunchKernelGGL((k_lscan<T, TV, E, NW, true, TZ>), g, dim3(64 * (NW)), 0, stream, nx, ny, nstrip, n, nv,  \
                       F, r, w, z, st, it, sink);                                                              \
  else                                                                                                          \
    hipLaunchKernelGGL((k_lscan<T, TV, E, NW, false, TZ>), g, dim3(64 * (NW)), 0, stream, nx, ny, nstrip, n, nv, \
                       F, r, w, z, st, it, sink);
  if (nx <= 64) { PARCEL_LSCAN(1, 1) }
  else if (nx <= 128) { PARCEL_LSCAN(2, 1) }
  else if (nx <= 256) { PARCEL_LSCAN(2, 2) }
  else if (nx <= 512) { PARCEL_LSCAN(2, 4) }
  else { PARCEL_LSCAN(4, 4) }
#undef PARCEL_LSCAN
  PARCEL_HIP_CHECK(hipGetLastError());
}

template <class T, class TV, class TZ = TV>
void launch_lscan(const Batch& B, const T* F, const TV* r, T* w, TZ* z, const CGState* st, int it) {
  lscan_launch<T, TV, TZ>(B.nx, B.ny, B.nstrip, B.nsys, (long)B.n, B.nv(), F, r, w, z, st, it, B.sink.p, B.stream);
}

struct EvPair {
  hipEvent_t a = nullptr, b = nullptr;
};

struct Prof {
  bool on = false;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<EvPair> apply, spmv;
  hipEvent_t rec(hipStream_t st) {
    if (!on) return nullptr;
    if (used == pool.size()) {
      hipEvent_t e;
      PARCEL_HIP_CHECK(hipEventCreate(&e));
      pool.push_back(e);
    }
    hipEvent_t e = pool[used++];
    PARCEL_HIP_CHECK(hipEventRecord(e, st));
    return e;
  }
  void sum(std::vector<EvPair>& v, double& ms, int& calls) {
    ms = 0.0; calls = 0;
    for (auto& p : v) {
      float t = 0.f;
      PARCEL_HIP_CHECK(hipEventElapsedTime(&t, p.a, p.b));
      ms += t; calls++;
    }
  }
  ~Prof() {
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};

constexpr int kLB = 4;  // logical tile owners per workgroup in the batched iteration kernels

template <class T>
struct BVecs {
  SellArgsT<T> A;
  T *r, *q, *s, *p0, *p1, *w;
  const T* dinv;  // point Jacobi (PARCEL_BATCH_JACOBI) or null
  const T* F;     // factor apply arrays (IC0 / FACTOR) or null
  const float* Fmix = nullptr;  // FP32 factor apply arrays used on T vectors (FP32 apply in FP64 PCG)
  float* wmix = nullptr;        // its FP32 intermediate vector
  float* zmix = nullptr;        // its output z = M^-1 r kept in FP32 (the CG widens it where it reads it)
  const T* st5 = nullptr;       // five-point stencil arrays (c, e, n; nv apart) or null
  double* hist;
  int hist_cap;
};

// One batched PCG: every system from its x to ||r||/||b|| < rtol.  mask (may
// be null): systems whose OuterState is done take no step.  Returns the
// final CGState of every system in out.  z: the preconditioned residual,
// V.s or (FP32 apply) V.zmix.
template <class T, class TZ>
void batch_cg_loop(Batch& B, const BVecs<T>& V, TZ* z, T* x, const T* b, int itrmax, double rtol,
                   const OuterState* mask, std::vector<CGState>& out, Prof& pf, int& launches) {
  const int n = B.n, nsys = B.nsys;
  const hipStream_t st = B.stream;
  const int nb = tiled_nb(n);
  const dim3 grid(nb, nsys), block(kNT);
  // the per-iteration kernels run kLB owners per workgroup (same sums)
  const bool lb = nb % (8 * kLB) == 0;
  const dim3 gridl(lb ? nb / kLB : nb, nsys);
  const Bat bt{(long)n, B.A.entries};
  CGState* S = B.st.p;
  double* part = B.part.p;
  out.assign(nsys, CGState{});
  if (itrmax <= 0) {
    for (auto& c : out) { c.niter = 0; c.iret = 1; c.done_at = 1; }
    return;
  }
  PARCEL_HIP_CHECK(hipMemsetAsync(V.p0, 0, sizeof(T) * (size_t)B.nv(), st));
  hipLaunchKernelGGL(k_residual<T>, grid, block, 0, st, n, V.A, (const T*)x, b, V.r, part, bt);
  const bool fac = V.F || V.Fmix;
  // z = M^-1 r (TZ = float: the FP32 apply's output kept in FP32)
  auto apply = [&](const CGState* sst, int it) {
    if constexpr (std::is_same<TZ, T>::value) {
      if (V.F)
        launch_lscan<T, T>(B, V.F, (const T*)V.r, V.w, z, sst, it);
      else
        launch_lscan<float, T>(B, V.Fmix, (const T*)V.r, V.wmix, z, sst, it);
    } else {
      launch_lscan<float, T, TZ>(B, V.Fmix, (const T*)V.r, V.wmix, z, sst, it);
    }
  };
  if (fac) {
    apply(nullptr, 0);
  } else {
    if constexpr (std::is_same<TZ, T>::value)
      hipLaunchKernelGGL(k_b_scale<T>, grid, block, 0, st, n, (const T*)V.r, V.dinv, z, bt);
  }
  hipLaunchKernelGGL((k_dot_part<T, TZ>), grid, block, 0, st, n, (const TZ*)z, (const T*)V.r,
                     part + PART_RHO0 * kNBMax, bt);
  hipLaunchKernelGGL(k_cg_init, dim3(1, nsys), block, 0, st, (const double*)part, nb, S);
  launches += 4;
  if (mask) {
    hipLaunchKernelGGL(k_bo_mask, dim3((nsys + 63) / 64), dim3(64), 0, st, mask, S, nsys);
    launches++;
  }
  PARCEL_HIP_CHECK(hipGetLastError());
  B.hstate.ensure((size_t)nsys);
  CGState* h = B.hstate.p;
  int enq = 0, chunk = 8;
  bool all_done = false;
  while (enq < itrmax + 1) {
    const int k = std::min(chunk, itrmax + 1 - enq);
    for (int j = 1; j <= k; ++j) {
      const int it = enq + j;
      T* pold = (it - 1) & 1 ? V.p1 : V.p0;
      T* pnew = it & 1 ? V.p1 : V.p0;
      hipEvent_t ea = pf.rec(st);
      if (V.st5) {
        const St5<T> A5{V.st5, V.st5 + B.nv(), V.st5 + 2 * B.nv(), B.nx, B.ny, FastDiv((unsigned)B.nx)};
        if (lb)
          hipLaunchKernelGGL((k_cg_spmv<T, kLB, St5<T>, TZ>), gridl, block, 0, st, n, it, itrmax, rtol, A5, x,
                             (const T*)pold, (const TZ*)z, pnew, V.q, part, S, V.hist, V.hist_cap, bt);
        else
          hipLaunchKernelGGL((k_cg_spmv<T, 1, St5<T>, TZ>), grid, block, 0, st, n, it, itrmax, rtol, A5, x,
                             (const T*)pold, (const TZ*)z, pnew, V.q, part, S, V.hist, V.hist_cap, bt);
      } else if (lb)
        hipLaunchKernelGGL((k_cg_spmv<T, kLB, SellArgsT<T>, TZ>), gridl, block, 0, st, n, it, itrmax, rtol, V.A, x,
                           (const T*)pold, (const TZ*)z, pnew, V.q, part, S, V.hist, V.hist_cap, bt);
      else
        hipLaunchKernelGGL((k_cg_spmv<T, 1, SellArgsT<T>, TZ>), grid, block, 0, st, n, it, itrmax, rtol, V.A, x,
                           (const T*)pold, (const TZ*)z, pnew, V.q, part, S, V.hist, V.hist_cap, bt);
      launches++;
      if (pf.on) pf.spmv.push_back({ea, pf.rec(st)});
      if (it > itrmax) break;
      if (!fac) {
        if constexpr (std::is_same<TZ, T>::value)
          hipLaunchKernelGGL(k_cg_update_pj<T>, grid, block, 0, st, n, it, V.r, (const T*)V.q, V.dinv, z, part,
                             S, nb, bt);
        launches++;
      } else {
        if (lb)
          hipLaunchKernelGGL((k_cg_axpy_r<T, kLB>), gridl, block, 0, st, n, it, V.r, (const T*)V.q, part, S, nb, bt);
        else
          hipLaunchKernelGGL((k_cg_axpy_r<T, 1>), grid, block, 0, st, n, it, V.r, (const T*)V.q, part, S, nb, bt);
        hipEvent_t eb = pf.rec(st);
        apply(S, it);
        if (pf.on) pf.apply.push_back({eb, pf.rec(st)});
        if (lb)
          hipLaunchKernelGGL((k_cg_dot2<T, kLB, TZ>), gridl, block, 0, st, n, it, (const TZ*)z, (const T*)V.r,
                             part, (const CGState*)S, bt);
        else
          hipLaunchKernelGGL((k_cg_dot2<T, 1, TZ>), grid, block, 0, st, n, it, (const TZ*)z, (const T*)V.r, part,
                             (const CGState*)S, bt);
        launches += 3;
      }
    }
    PARCEL_HIP_CHECK(hipGetLastError());
    enq += k;
    PARCEL_HIP_CHECK(hipMemcpyAsync(h, S, sizeof(CGState) * nsys, hipMemcpyDeviceToHost, st));
    PARCEL_HIP_CHECK(hipStreamSynchronize(st));
    all_done = true;
    for (int s = 0; s < nsys; ++s) all_done = all_done && h[s].done_at != 0;
    if (all_done) break;
    if (chunk < 64) chunk *= 2;
  }
  if (!all_done) {
    PARCEL_HIP_CHECK(hipMemcpyAsync(h, S, sizeof(CGState) * nsys, hipMemcpyDeviceToHost, st));
    PARCEL_HIP_CHECK(hipStreamSynchronize(st));
    for (int s = 0; s < nsys; ++s)
      if (!h[s].done_at) throw hip_error("batched CG loop did not terminate", PARCEL_HIP_ERR_HIP);
  }
  for (int s = 0; s < nsys; ++s) out[s] = h[s];
}

template <class T>
void batch_cg_run(Batch& B, const BVecs<T>& V, T* x, const T* b, int itrmax, double rtol, const OuterState* mask,
                  std::vector<CGState>& out, Prof& pf, int& launches) {
  if constexpr (std::is_same<T, double>::value) {
    if (V.Fmix && V.zmix) {
      batch_cg_loop<T, float>(B, V, V.zmix, x, b, itrmax, rtol, mask, out, pf, launches);
      return;
    }
  }
  batch_cg_loop<T, T>(B, V, V.s, x, b, itrmax, rtol, mask, out, pf, launches);
}

void ensure_vectors(Batch& B, bool f32) {
  const size_t nv = (size_t)B.nv();
  if (B.r.n < nv) {
    for (DevBuf<double>* v : {&B.r, &B.q, &B.s, &B.p0, &B.p1, &B.w}) v->alloc(nv);
  }
  if (f32 && B.r32.n < nv) {
    for (DevBuf<float>* v : {&B.b32, &B.d32, &B.r32, &B.q32, &B.s32, &B.p032, &B.p132, &B.w32}) v->alloc(nv);
  }
  if (!B.part.p) B.part.alloc((size_t)B.nsys * PART_N * kNBMax);
  if (!B.st.p) B.st.alloc(B.nsys);
  if (!B.ost.p) B.ost.alloc(B.nsys);
}

// Five-point natural-order pattern check; sets nx, ny (0 when it is not one).
void detect_grid(Batch& B) {
  B.nx = B.ny = 0;
  const int n = B.n;
  if (n < 1) return;
  const int len0 = B.rowptr[1] - B.rowptr[0];
  int nx;
  if (n == 1) nx = 1;
  else if (len0 == 3) nx = B.col[B.rowptr[0] + 2];
  else if (len0 == 2) nx = n <= 1024 ? n : 1;
  else return;
  if (nx < 1 || n % nx != 0) return;
  const int ny = n / nx;
  for (int i = 0; i < n; ++i) {
    const int x = i % nx, y = i / nx;
    int expect[5], m = 0;
    if (y > 0) expect[m++] = i - nx;
    if (x > 0) expect[m++] = i - 1;
    expect[m++] = i;
    if (x < nx - 1) expect[m++] = i + 1;
    if (y < ny - 1) expect[m++] = i + nx;
    if (B.rowptr[i + 1] - B.rowptr[i] != m) return;
    for (int j = 0; j < m; ++j)
      if (B.col[B.rowptr[i] + j] != expect[j]) return;
  }
  if (nx > 1024) return;
  B.nx = nx;
  B.ny = ny;
}

// IC(0) on the five-point pattern, per system (natural order):
//   Ls_i = a(i,i-nx) / Ld_{i-nx},  Lw_i = a(i,i-1) / Ld_{i-1},
//   Ld_i = sqrt(fma(-Lw, Lw, fma(-Ls, Ls, a_ii)))
// (no other fill: rows i and i-1, i and i-nx share no earlier column).
// With nblk > 1 every strip of grid lines [ny k / nblk, ny (k+1) / nblk) is
// factorised on its own (block-Jacobi IC(0)): the first line of a strip
// drops its couplings to the strip below, as block-Jacobi set-ups keep
// in-block entries only.
int ic0_host(const Batch& B, const std::vector<double>& vals, std::vector<double>& L, int nblk = 1) {
  const int n = B.n, nx = B.nx, nsys = B.nsys, ny = B.ny;
  const long nv = (long)nsys * n;
  L.assign((size_t)(3 * nv), 0.0);
  // first[line]: the first line of the line's strip
  std::vector<int> first(ny > 0 ? ny : 1, 0);
  for (int k = 0; k < nblk; ++k)
    for (int l = (int)((long)ny * k / nblk); l < (int)((long)ny * (k + 1) / nblk); ++l)
      first[l] = (int)((long)ny * k / nblk);
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
  for (int s = 0; s < nsys; ++s) {
    const double* v = vals.data() + (long)s * B.nnz;
    double* Ls = L.data() + (long)s * n;
    double* Lw = Ls + nv;
    double* Ld = Ls + 2 * nv;
    for (int i = 0; i < n; ++i) {
      const int e0 = B.rowptr[i];
      int e = e0;
      double ls = 0.0, lw = 0.0;
      if (i >= nx) {
        if (i / nx > first[i / nx]) ls = v[e] / Ld[i - nx];
        ++e;
      }
      if (i % nx > 0) { lw = v[e] / Ld[i - 1]; ++e; }
      const double t = std::fma(-lw, lw, std::fma(-ls, ls, v[e]));
      if (!(t > 0.0)) { bad = 1; break; }
      Ls[i] = ls;
      Lw[i] = lw;
      Ld[i] = std::sqrt(t);
    }
  }
  return bad ? PARCEL_HIP_ERR_ZERO_DIAG : PARCEL_HIP_OK;
}

}  // namespace

void parcel::lscan_apply32(const float* F, const float* r, float* w, float* z, int nx, int ny, int nstrip, long n,
                           const CGState* st, int it, double* sink, hipStream_t stream) {
  lscan_launch<float, float, float>(nx, ny, nstrip, 1, n, n, F, r, w, z, st, it, sink, stream);
}

void parcel::derive_factor(Batch& B) {
  const long nv = B.nv();
  B.F64.alloc((size_t)(4 * nv));
  B.F32.alloc((size_t)(4 * nv));
  const dim3 g((unsigned)std::min<long>((B.n + 255) / 256, 4096), B.nsys), b(256);
  if (B.factor_src == 1) {
    hipLaunchKernelGGL((k_b_derive<double, double>), g, b, 0, B.stream, B.n, B.nx, nv, (const double*)B.L64.p,
                       B.F64.p);
    hipLaunchKernelGGL((k_b_derive<double, float>), g, b, 0, B.stream, B.n, B.nx, nv, (const double*)B.L64.p,
                       B.F32.p);
  } else {
    hipLaunchKernelGGL((k_b_derive<uint16_t, double>), g, b, 0, B.stream, B.n, B.nx, nv, (const uint16_t*)B.Lbf.p,
                       B.F64.p);
    hipLaunchKernelGGL((k_b_derive<uint16_t, float>), g, b, 0, B.stream, B.n, B.nx, nv, (const uint16_t*)B.Lbf.p,
                       B.F32.p);
  }
  PARCEL_HIP_CHECK(hipGetLastError());
  PARCEL_HIP_CHECK(hipStreamSynchronize(B.stream));
}

namespace {

template <class T>
void fetch_host(const T* src, size_t count, std::vector<T>& dst) {
  dst.resize(count);
  if (!count) return;
  if (api_is_device_ptr(src))
    PARCEL_HIP_CHECK(hipMemcpy(dst.data(), src, count * sizeof(T), hipMemcpyDeviceToHost));
  else
    std::memcpy(dst.data(), src, count * sizeof(T));
}

int batch_create(Batch& B, int nsys, int n, int nnz, const int* rowptr, const int* col, const double* vals,
                 int precon) {
  auto t0 = std::chrono::steady_clock::now();
  if (nsys < 1 || n < 1 || nnz < 1 || !rowptr || !col || !vals)
    return api_fail(PARCEL_HIP_ERR_ARG, "batch_create: empty batch or null array");
  if (precon < PARCEL_BATCH_NONE || precon > PARCEL_BATCH_FACTOR)
    return api_fail(PARCEL_HIP_ERR_ARG, "batch_create: unknown preconditioner");
  B.nsys = nsys; B.n = n; B.nnz = nnz; B.precon = precon;
  // a learned (FACTOR) batch falls back to FP64 PCG + block-Jacobi ILU(0)
  if (precon == PARCEL_BATCH_FACTOR) B.fb_itrmax = 20000;
  B.rowptr.resize(n + 1);
  B.col.resize(nnz);
  if (rowptr[0] != 1 || rowptr[n] != nnz + 1)
    return api_fail(PARCEL_HIP_ERR_ARG, "batch_create: rowptr must be 1-based with rowptr[n] = nnz + 1");
  for (int i = 0; i <= n; ++i) B.rowptr[i] = rowptr[i] - 1;
  for (int i = 0; i < n; ++i) {
    if (B.rowptr[i + 1] < B.rowptr[i]) return api_fail(PARCEL_HIP_ERR_ARG, "batch_create: rowptr decreases");
    bool diag = false;
    for (int e = B.rowptr[i]; e < B.rowptr[i + 1]; ++e) {
      const int c = col[e] - 1;
      if (c < 0 || c >= n) return api_fail(PARCEL_HIP_ERR_ARG, "batch_create: column out of range");
      B.col[e] = c;
      diag = diag || c == i;
    }
    if (!diag && precon != PARCEL_BATCH_NONE)
      return api_fail(PARCEL_HIP_ERR_ZERO_DIAG, "batch_create: row without a diagonal entry");
  }
  detect_grid(B);
  if ((precon == PARCEL_BATCH_IC0 || precon == PARCEL_BATCH_FACTOR) && !B.nx)
    return api_fail(PARCEL_HIP_ERR_UNSUPPORTED,
                    "batch_create: IC0/FACTOR need a 2D five-point pattern in natural order (lines <= 1024)");
  // the factor apply addresses a system's cells with 32-bit byte offsets
  if ((precon == PARCEL_BATCH_IC0 || precon == PARCEL_BATCH_FACTOR) && (long)n >= (1L << 29))
    return api_fail(PARCEL_HIP_ERR_UNSUPPORTED, "batch_create: IC0/FACTOR systems need n < 2^29");
  B.stats.nx = B.nx;
  B.stats.ny = B.ny;
  // SELL-64 layout of the pattern and the slot -> CSR entry map
  DevSELL& S = B.A;
  S.n = n;
  S.nnz = nnz;
  const long ns = ((long)n + 63) / 64;
  S.nslices = ns;
  std::vector<long> sptr(ns + 1, 0);
  std::vector<int> sk(ns, 0);
  for (long s = 0; s < ns; ++s) {
    int k = 0;
    for (long r = s * 64; r < std::min((long)n, s * 64 + 64); ++r) k = std::max(k, B.rowptr[r + 1] - B.rowptr[r]);
    sk[s] = k;
    sptr[s + 1] = sptr[s] + 64L * k;
  }
  const long tot = sptr[ns];
  S.entries = tot;
  std::vector<int> scol((size_t)tot), idx((size_t)tot);
  std::vector<int> dslot(n, 0);
  std::vector<int4> nslot(B.nx ? n : 0, int4{-1, -1, -1, -1});
  for (long s = 0; s < ns; ++s)
    for (int l = 0; l < 64; ++l) {
      const long r = s * 64 + l;
      const int len = r < n ? B.rowptr[r + 1] - B.rowptr[r] : 0;
      for (int j = 0; j < sk[s]; ++j) {
        const long at = sptr[s] + (long)j * 64 + l;
        if (j < len) {
          const int e = B.rowptr[r] + j;
          scol[at] = B.col[e];
          idx[at] = e;
          if (B.col[e] == r) dslot[r] = (int)at;
          if (B.nx) {
            const int d = B.col[e] - (int)r;
            if (d == -B.nx) nslot[r].x = (int)at;
            else if (d == -1) nslot[r].y = (int)at;
            else if (d == 1) nslot[r].z = (int)at;
            else if (d == B.nx) nslot[r].w = (int)at;
          }
        } else {
          scol[at] = -1;
          idx[at] = -1;
        }
      }
    }
  S.sptr.upload(sptr, B.stream);
  B.hsptr = sptr;
  S.sk.upload(sk, B.stream);
  S.col.upload(scol, B.stream);
  B.diag_slot.upload(dslot, B.stream);
  if (B.nx) {
    B.nbr_slot.upload(nslot, B.stream);
    B.sink.alloc(1024);  // one line (nx <= 1024) of the factor apply's filler and padding stores
  }
  DevBuf<int> didx;
  didx.upload(idx, B.stream);
  // values: CSR (host or device) -> SELL order on the device
  const long nvals = (long)nsys * nnz;
  DevBuf<double> csr;
  const double* csr_dev = vals;
  if (!api_is_device_ptr(vals)) {
    csr.alloc((size_t)nvals);
    csr.upload(vals, (size_t)nvals, B.stream);
    csr_dev = csr.p;
  }
  S.val.alloc((size_t)(tot * nsys));
  S.val32.alloc((size_t)(tot * nsys));
  {
    const dim3 g((unsigned)std::min<long>((tot + 255) / 256, 4096), nsys), b(256);
    hipLaunchKernelGGL(k_b_gather, g, b, 0, B.stream, tot, (long)nnz, (const int*)didx.p, csr_dev, S.val.p,
                       S.val32.p);
  }
  if (B.nx) {
    B.st5.alloc((size_t)(3 * B.nv()));
    DevBuf<int> bad(1);
    PARCEL_HIP_CHECK(hipMemsetAsync(bad.p, 0, sizeof(int), B.stream));
    const dim3 g((unsigned)std::min<long>((n + 255) / 256, 4096), nsys), b(256);
    hipLaunchKernelGGL(k_b_st5, g, b, 0, B.stream, n, B.nx, B.ny, tot, B.nv(), (const int*)B.diag_slot.p,
                       (const int4*)B.nbr_slot.p, (const double*)S.val.p, B.st5.p, bad.p);
    int hb = 1;
    PARCEL_HIP_CHECK(hipMemcpyAsync(&hb, bad.p, sizeof(int), hipMemcpyDeviceToHost, B.stream));
    PARCEL_HIP_CHECK(hipStreamSynchronize(B.stream));
    B.st5_ok = hb == 0;
    if (!B.st5_ok) B.st5.release();
    B.stats.stencil5 = B.st5_ok ? 1 : 0;
  }
  if (precon == PARCEL_BATCH_JACOBI) {
    B.dinv.alloc((size_t)B.nv());
    B.dinv32.alloc((size_t)B.nv());
    const dim3 g((unsigned)std::min<long>((n + 255) / 256, 4096), nsys), b(256);
    hipLaunchKernelGGL(k_b_dinv, g, b, 0, B.stream, n, tot, (const int*)B.diag_slot.p, (const double*)S.val.p,
                       B.dinv.p, B.dinv32.p);
  }
  PARCEL_HIP_CHECK(hipGetLastError());
  PARCEL_HIP_CHECK(hipStreamSynchronize(B.stream));
  if (precon == PARCEL_BATCH_IC0) {
    std::vector<double> hv, L;
    fetch_host(vals, (size_t)nvals, hv);
    const int rc = ic0_host(B, hv, L);
    if (rc != PARCEL_HIP_OK) return api_fail(rc, "batch_create: IC(0) breakdown (non-positive pivot)");
    B.L64.upload(L, B.stream);
    B.factor_src = 1;
    derive_factor(B);
  }
  B.stats.setup_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return PARCEL_HIP_OK;
}

template <class T>
BVecs<T> vecs_of(Batch& B);
template <>
BVecs<double> vecs_of<double>(Batch& B) {
  BVecs<double> V;
  V.A = SellArgsT<double>{B.A.sptr.p, B.A.sk.p, B.A.col.p, B.A.val.p};
  V.r = B.r.p; V.q = B.q.p; V.s = B.s.p; V.p0 = B.p0.p; V.p1 = B.p1.p; V.w = B.w.p;
  V.dinv = B.precon == PARCEL_BATCH_JACOBI ? B.dinv.p : nullptr;
  V.F = (B.precon == PARCEL_BATCH_IC0 || B.precon == PARCEL_BATCH_FACTOR) ? B.F64.p : nullptr;
  V.st5 = B.st5_ok ? B.st5.p : nullptr;
  V.hist = nullptr; V.hist_cap = 0;
  return V;
}
template <>
BVecs<float> vecs_of<float>(Batch& B) {
  BVecs<float> V;
  V.A = SellArgsT<float>{B.A.sptr.p, B.A.sk.p, B.A.col.p, B.A.val32.p};
  V.r = B.r32.p; V.q = B.q32.p; V.s = B.s32.p; V.p0 = B.p032.p; V.p1 = B.p132.p; V.w = B.w32.p;
  V.dinv = B.precon == PARCEL_BATCH_JACOBI ? B.dinv32.p : nullptr;
  V.F = (B.precon == PARCEL_BATCH_IC0 || B.precon == PARCEL_BATCH_FACTOR) ? B.F32.p : nullptr;
  V.hist = nullptr; V.hist_cap = 0;
  return V;
}

void ensure_hist(Batch& B, int cap) {
  if (B.hist_cap < cap || !B.hist.p) {
    B.hist.alloc((size_t)B.nsys * (cap > 0 ? cap : 1));
    B.hist_cap = cap;
  }
}

// CSR values of systems [s0, s0 + cnt) read back from their SELL slots.
void host_values(Batch& B, int s0, int cnt, std::vector<double>& hv) {
  const long ent = B.A.entries;
  std::vector<double> sell((size_t)ent * cnt);
  PARCEL_HIP_CHECK(hipMemcpyAsync(sell.data(), B.A.val.p + (long)s0 * ent, sizeof(double) * ent * cnt,
                                  hipMemcpyDeviceToHost, B.stream));
  PARCEL_HIP_CHECK(hipStreamSynchronize(B.stream));
  hv.resize((size_t)B.nnz * cnt);
  const int n = B.n;
#pragma omp parallel for schedule(static)
  for (int s = 0; s < cnt; ++s)
    for (int i = 0; i < n; ++i)
      for (int e = B.rowptr[i], j = 0; e < B.rowptr[i + 1]; ++e, ++j)
        hv[(size_t)s * B.nnz + e] = sell[(size_t)s * ent + B.hsptr[i / 64] + (long)j * 64 + i % 64];
}

// The learned path's fallback for system s: its CSR values are read back
// from the SELL slots, a single-system plan with the fallback
// preconditioner is set up, and parcel_dcg's FP64 PCG solves it from x = 0
// (src/pcg.f90:224-892 with x0 = 0), writing x in place on the device.
void batch_fallback(Batch& B, int s, double* dx, const double* db, double rtol, int& nit, int& ret,
                    std::vector<double>& hist) {
  const int n = B.n, nnz = B.nnz;
  std::vector<double> val;
  host_values(B, s, 1, val);
  PARCEL_HIP_CHECK(hipMemsetAsync(dx + (long)s * n, 0, sizeof(double) * n, B.stream));
  PARCEL_HIP_CHECK(hipStreamSynchronize(B.stream));
  std::vector<int> rp1(n + 1), col1(nnz);
  for (int i = 0; i < n; ++i) {
    rp1[i] = B.rowptr[i] + 1;
    for (int e = B.rowptr[i]; e < B.rowptr[i + 1]; ++e) col1[e] = B.col[e] + 1;
  }
  rp1[n] = nnz + 1;
  parcel_hip_param prm;
  parcel_hip_default_param(&prm);
  prm.solver = PARCEL_SOLVER_CG;
  prm.ipreflag = B.fb_ipreflag;
  prm.ILU_method = B.fb_ilu;
  prm.precon_thblock = B.fb_blocks;
  parcel_hip_plan* P = nullptr;
  int rc = parcel_hip_plan_create(&P, n, n, nnz, 1, val.data(), rp1.data(), col1.data(), &prm);
  if (rc != PARCEL_HIP_OK) throw hip_error(std::string("fallback setup: ") + parcel_hip_last_error(), rc);
  int it = B.fb_itrmax;
  hist.assign((size_t)it, 0.0);
  rc = parcel_hip_plan_solve(P, dx + (long)s * n, db + (long)s * n, &it, rtol, hist.data(), &ret);
  const std::string msg = rc == PARCEL_HIP_OK ? "" : parcel_hip_last_error();
  parcel_hip_plan_destroy(P);
  PARCEL_HIP_CHECK(hipSetDevice(B.device));
  if (rc != PARCEL_HIP_OK) throw hip_error("fallback solve: " + msg, rc);
  nit = it;
}

int batch_solve(Batch& B, double* x, const double* b, int itrmax, double rtol, int precision, double* hist_host,
                int* niter, int* iret) {
  if (!x || !b) return api_fail(PARCEL_HIP_ERR_ARG, "batch_solve: null vector");
  if (precision != PARCEL_BATCH_FP64 && precision != PARCEL_BATCH_REFINE32 && precision != PARCEL_BATCH_APPLY32)
    return api_fail(PARCEL_HIP_ERR_ARG, "batch_solve: precision must be 64, 32 or 48");
  if (precision == PARCEL_BATCH_APPLY32 && B.precon != PARCEL_BATCH_IC0 && B.precon != PARCEL_BATCH_FACTOR)
    return api_fail(PARCEL_HIP_ERR_ARG, "batch_solve: precision 48 needs a factor preconditioner");
  if (B.precon == PARCEL_BATCH_FACTOR && !B.factor_src)
    return api_fail(PARCEL_HIP_ERR_ARG, "batch_solve: FACTOR preconditioner without a factor (batch_set_factor)");
  auto t0 = std::chrono::steady_clock::now();
  const hipStream_t st = B.stream;
  const size_t nv = (size_t)B.nv();
  ensure_vectors(B, precision != PARCEL_BATCH_FP64);
  const int cap = itrmax > 0 ? itrmax : 1;
  ensure_hist(B, cap);
  PARCEL_HIP_CHECK(hipMemsetAsync(B.hist.p, 0, sizeof(double) * (size_t)B.nsys * B.hist_cap, st));
  // inputs resident on the device
  const bool xdev = api_is_device_ptr(x), bdev = api_is_device_ptr(b);
  if (!B.xb.p || B.xb.n < nv) { B.xb.alloc(nv); B.bb.alloc(nv); }
  double* dx = xdev ? x : B.xb.p;
  const double* db = bdev ? b : B.bb.p;
  if (!xdev) B.xb.upload(x, nv, st);
  if (!bdev) B.bb.upload(b, nv, st);
  Prof pf;
  pf.on = B.profile;
  hipEvent_t ev0, ev1;
  PARCEL_HIP_CHECK(hipEventCreate(&ev0));
  PARCEL_HIP_CHECK(hipEventCreate(&ev1));
  PARCEL_HIP_CHECK(hipEventRecord(ev0, st));
  int launches = 0;
  std::vector<int> nit(B.nsys, 0), ret(B.nsys, 0);
  B.stats.outer_steps = 0;
  B.stats.inner_total = 0;
  B.stats.inner_steps = 0;
  if (precision == PARCEL_BATCH_FP64 || precision == PARCEL_BATCH_APPLY32) {
    BVecs<double> V = vecs_of<double>(B);
    if (precision == PARCEL_BATCH_APPLY32) {
      V.F = nullptr;
      V.Fmix = B.F32.p;
      V.wmix = B.w32.p;
      V.zmix = B.s32.p;  // z stays FP32 (this path has no FP32 CG of its own)
    }
    V.hist = B.hist.p;
    V.hist_cap = B.hist_cap;
    std::vector<CGState> out;
    batch_cg_run<double>(B, V, dx, db, itrmax, rtol, nullptr, out, pf, launches);
    for (int s = 0; s < B.nsys; ++s) { nit[s] = out[s].niter; ret[s] = out[s].iret; }
  } else if (itrmax <= 0) {
    for (int s = 0; s < B.nsys; ++s) { nit[s] = 0; ret[s] = 1; }
  } else {
    const int n = B.n, nsys = B.nsys;
    const int nb = tiled_nb(n);
    const dim3 grid(nb, nsys), block(kNT);
    const Bat bt{(long)n, B.A.entries};
    const SellArgsT<double> A64{B.A.sptr.p, B.A.sk.p, B.A.col.p, B.A.val.p};
    BVecs<float> V = vecs_of<float>(B);
    PARCEL_HIP_CHECK(hipMemsetAsync(B.ost.p, 0, sizeof(OuterState) * nsys, st));
    hipLaunchKernelGGL(k_residual<double>, grid, block, 0, st, n, A64, (const double*)dx, db, B.r.p, B.part.p, bt);
    std::vector<OuterState> hos(nsys);
    std::vector<CGState> out;
    for (int k = 1; k <= itrmax + 1; ++k) {
      hipLaunchKernelGGL(k_bo_check, dim3(1, nsys), block, 0, st, k, itrmax, rtol, B.part.p, nb, B.ost.p, B.hist.p,
                         B.hist_cap);
      launches += 2;
      PARCEL_HIP_CHECK(hipMemcpyAsync(hos.data(), B.ost.p, sizeof(OuterState) * nsys, hipMemcpyDeviceToHost, st));
      PARCEL_HIP_CHECK(hipStreamSynchronize(st));
      bool all = true;
      for (auto& o : hos) all = all && o.done;
      if (all) break;
      hipLaunchKernelGGL(k_bo_tof32, grid, block, 0, st, n, (const double*)B.r.p, B.b32.p,
                         B.d32.p, (const OuterState*)B.ost.p, bt);
      batch_cg_run<float>(B, V, B.d32.p, B.b32.p, B.inner_itrmax, B.inner_rtol, B.ost.p, out, pf, launches);
      int most = 0;
      for (int s = 0; s < nsys; ++s) {
        B.stats.inner_total += out[s].niter;
        most = std::max(most, out[s].niter);
      }
      B.stats.inner_steps += most;
      B.stats.outer_steps++;
      hipLaunchKernelGGL(k_bo_xadd, grid, block, 0, st, n, dx, (const float*)B.d32.p, (const OuterState*)B.ost.p, bt);
      hipLaunchKernelGGL(k_residual<double>, grid, block, 0, st, n, A64, (const double*)dx, db, B.r.p, B.part.p, bt);
      launches += 3;
    }
    PARCEL_HIP_CHECK(hipGetLastError());
    for (int s = 0; s < nsys; ++s) { nit[s] = hos[s].niter; ret[s] = hos[s].iret; }
  }
  // fallback: systems the batched solve did not converge
  B.fell_back.assign(B.nsys, 0);
  std::vector<std::vector<double>> fb_hist(B.nsys);
  int nfb = 0;
  if (B.fb_itrmax > 0)
    for (int s = 0; s < B.nsys; ++s)
      if (ret[s] != 0) {
        batch_fallback(B, s, dx, db, rtol, nit[s], ret[s], fb_hist[s]);
        B.fell_back[s] = 1;
        ++nfb;
      }
  B.stats.fallback_systems = nfb;
  PARCEL_HIP_CHECK(hipEventRecord(ev1, st));
  if (!xdev) PARCEL_HIP_CHECK(hipMemcpyAsync(x, dx, sizeof(double) * nv, hipMemcpyDeviceToHost, st));
  PARCEL_HIP_CHECK(hipStreamSynchronize(st));
  float ms = 0.f;
  PARCEL_HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  int mx = 0;
  for (int s = 0; s < B.nsys; ++s) {
    if (niter) niter[s] = nit[s];
    if (iret) iret[s] = ret[s];
    mx = std::max(mx, nit[s]);
  }
  if (hist_host && itrmax > 0) {
    std::vector<double> h((size_t)B.nsys * B.hist_cap);
    PARCEL_HIP_CHECK(hipMemcpy(h.data(), B.hist.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    for (int s = 0; s < B.nsys; ++s)
      std::memcpy(hist_host + (long)s * itrmax, h.data() + (long)s * B.hist_cap, sizeof(double) * itrmax);
    // a fallen-back system's row: the fallback solve's history (its first
    // itrmax entries, zero-padded)
    for (int s = 0; s < B.nsys; ++s)
      if (B.fell_back[s]) {
        double* row = hist_host + (long)s * itrmax;
        std::memset(row, 0, sizeof(double) * itrmax);
        const size_t k = std::min((size_t)itrmax, (size_t)std::max(nit[s], 0));
        std::memcpy(row, fb_hist[s].data(), sizeof(double) * k);
      }
  }
  B.stats.solve_device_ms = ms;
  B.stats.solve_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  B.stats.max_niter = mx;
  B.stats.launches = launches;
  if (pf.on) {
    pf.sum(pf.apply, B.stats.apply_ms, B.stats.apply_calls);
    pf.sum(pf.spmv, B.stats.spmv_ms, B.stats.spmv_calls);
  } else {
    B.stats.apply_ms = 0.0; B.stats.apply_calls = 0;
    B.stats.spmv_ms = 0.0; B.stats.spmv_calls = 0;
  }
  return PARCEL_HIP_OK;
}

#define PARCEL_API_BEGIN try {
#define PARCEL_API_END                                                    \
  }                                                                       \
  catch (const hip_error& e) {                                            \
    return api_fail(e.code < 0 ? e.code : PARCEL_HIP_ERR_HIP, e.what());  \
  }                                                                       \
  catch (const std::exception& e) {                                       \
    return api_fail(PARCEL_HIP_ERR_HIP, e.what());                        \
  }

}  // namespace

extern "C" {

int parcel_hip_batch_create(parcel_hip_batch** out, int nsys, int n, int nnz, const int* rowptr, const int* col,
                            const double* vals, int precon) {
  PARCEL_API_BEGIN
  if (!out) return api_fail(PARCEL_HIP_ERR_ARG, "batch_create: null handle");
  *out = nullptr;
  int rc = api_ensure_init();
  if (rc != PARCEL_HIP_OK) return rc;
  auto* h = new parcel_hip_batch;
  h->B.device = api_device();
  PARCEL_HIP_CHECK(hipSetDevice(h->B.device));
  try {
    PARCEL_HIP_CHECK(hipStreamCreateWithFlags(&h->B.stream, hipStreamNonBlocking));
    rc = batch_create(h->B, nsys, n, nnz, rowptr, col, vals, precon);
  } catch (...) {
    if (h->B.stream) (void)hipStreamDestroy(h->B.stream);
    delete h;
    throw;
  }
  if (rc != PARCEL_HIP_OK) {
    (void)hipStreamDestroy(h->B.stream);
    delete h;
    return rc;
  }
  *out = h;
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

int parcel_hip_batch_set_factor(parcel_hip_batch* h, const unsigned short* Ls, const unsigned short* Lw,
                                const unsigned short* Ld) {
  PARCEL_API_BEGIN
  if (!h || !Ls || !Lw || !Ld) return api_fail(PARCEL_HIP_ERR_ARG, "batch_set_factor: null argument");
  Batch& B = h->B;
  if (B.precon != PARCEL_BATCH_FACTOR)
    return api_fail(PARCEL_HIP_ERR_ARG, "batch_set_factor: the batch was not created with PARCEL_BATCH_FACTOR");
  PARCEL_HIP_CHECK(hipSetDevice(B.device));
  const long nv = B.nv();
  B.Lbf.alloc((size_t)(3 * nv));
  const unsigned short* src[3] = {Ls, Lw, Ld};
  for (int k = 0; k < 3; ++k)
    PARCEL_HIP_CHECK(hipMemcpyAsync(B.Lbf.p + k * nv, src[k], sizeof(uint16_t) * nv,
                                    api_is_device_ptr(src[k]) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                    B.stream));
  B.factor_src = 2;
  derive_factor(B);
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

int parcel_hip_batch_solve(parcel_hip_batch* h, double* x, const double* b, int itrmax, double rtol, int precision,
                           double* hist, int* niter, int* iret) {
  PARCEL_API_BEGIN
  if (!h) return api_fail(PARCEL_HIP_ERR_ARG, "batch_solve: null handle");
  PARCEL_HIP_CHECK(hipSetDevice(h->B.device));
  return batch_solve(h->B, x, b, itrmax, rtol, precision, hist, niter, iret);
  PARCEL_API_END
}

int parcel_hip_batch_set_inner(parcel_hip_batch* h, double inner_rtol, int inner_itrmax) {
  if (!h || !(inner_rtol > 0.0) || inner_itrmax < 1) return api_fail(PARCEL_HIP_ERR_ARG, "batch_set_inner: bad argument");
  h->B.inner_rtol = inner_rtol;
  h->B.inner_itrmax = inner_itrmax;
  return PARCEL_HIP_OK;
}

int parcel_hip_batch_set_strips(parcel_hip_batch* h, int nstrip) {
  if (!h || nstrip < 1) return api_fail(PARCEL_HIP_ERR_ARG, "batch_set_strips: bad argument");
  if (h->B.ny && nstrip > h->B.ny) return api_fail(PARCEL_HIP_ERR_ARG, "batch_set_strips: more strips than lines");
  h->B.nstrip = nstrip;
  return PARCEL_HIP_OK;
}

int parcel_hip_batch_set_profile(parcel_hip_batch* h, int on) {
  if (!h) return api_fail(PARCEL_HIP_ERR_ARG, "batch_set_profile: null handle");
  h->B.profile = on != 0;
  return PARCEL_HIP_OK;
}

int parcel_hip_batch_set_ic0_blocks(parcel_hip_batch* h, int nblocks) {
  PARCEL_API_BEGIN
  if (!h || nblocks < 1) return api_fail(PARCEL_HIP_ERR_ARG, "batch_set_ic0_blocks: bad argument");
  Batch& B = h->B;
  if (B.precon != PARCEL_BATCH_IC0)
    return api_fail(PARCEL_HIP_ERR_ARG, "batch_set_ic0_blocks: the batch was not created with PARCEL_BATCH_IC0");
  if (nblocks > B.ny) return api_fail(PARCEL_HIP_ERR_ARG, "batch_set_ic0_blocks: more blocks than grid lines");
  PARCEL_HIP_CHECK(hipSetDevice(B.device));
  std::vector<double> hv, L;
  host_values(B, 0, B.nsys, hv);
  const int rc = ic0_host(B, hv, L, nblocks);
  if (rc != PARCEL_HIP_OK) return api_fail(rc, "batch_set_ic0_blocks: IC(0) breakdown (non-positive pivot)");
  B.L64.upload(L, B.stream);
  B.factor_src = 1;
  derive_factor(B);
  B.nstrip = nblocks;
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

int parcel_hip_batch_set_fallback(parcel_hip_batch* h, int ipreflag, int ILU_method, int precon_thblock,
                                  int itrmax) {
  if (!h) return api_fail(PARCEL_HIP_ERR_ARG, "batch_set_fallback: null handle");
  if (ipreflag < 0 || ipreflag > 2 || ILU_method < 0 || ILU_method > 2)
    return api_fail(PARCEL_HIP_ERR_ARG, "batch_set_fallback: ipreflag 0..2 and ILU_method 0..2");
  h->B.fb_ipreflag = ipreflag;
  h->B.fb_ilu = ILU_method;
  h->B.fb_blocks = precon_thblock;
  h->B.fb_itrmax = itrmax > 0 ? itrmax : 0;
  return PARCEL_HIP_OK;
}

int parcel_hip_batch_get_fallback(const parcel_hip_batch* h, int* fell_back) {
  if (!h || !fell_back) return api_fail(PARCEL_HIP_ERR_ARG, "batch_get_fallback: null argument");
  for (int s = 0; s < h->B.nsys; ++s) fell_back[s] = s < (int)h->B.fell_back.size() ? h->B.fell_back[s] : 0;
  return PARCEL_HIP_OK;
}

int parcel_hip_batch_get_stats(const parcel_hip_batch* h, parcel_hip_batch_stats* s) {
  if (!h || !s) return api_fail(PARCEL_HIP_ERR_ARG, "batch_stats: null argument");
  *s = h->B.stats;
  return PARCEL_HIP_OK;
}

int parcel_hip_batch_spmv(parcel_hip_batch* h, const double* x, double* y) {
  PARCEL_API_BEGIN
  if (!h || !x || !y) return api_fail(PARCEL_HIP_ERR_ARG, "batch_spmv: null argument");
  Batch& B = h->B;
  PARCEL_HIP_CHECK(hipSetDevice(B.device));
  const size_t nv = (size_t)B.nv();
  DevBuf<double> tx, ty;
  const double* dx = x;
  double* dy = y;
  if (!api_is_device_ptr(x)) { tx.alloc(nv); tx.upload(x, nv, B.stream); dx = tx.p; }
  const bool ydev = api_is_device_ptr(y);
  if (!ydev) { ty.alloc(nv); dy = ty.p; }
  const SellArgsT<double> A{B.A.sptr.p, B.A.sk.p, B.A.col.p, B.A.val.p};
  hipLaunchKernelGGL(k_spmv<double>, dim3(tiled_nb(B.n), B.nsys), dim3(kNT), 0, B.stream, B.n, A, dx, dy,
                     Bat{(long)B.n, B.A.entries});
  PARCEL_HIP_CHECK(hipGetLastError());
  if (!ydev) PARCEL_HIP_CHECK(hipMemcpyAsync(y, dy, sizeof(double) * nv, hipMemcpyDeviceToHost, B.stream));
  PARCEL_HIP_CHECK(hipStreamSynchronize(B.stream));
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

}  // extern "C"

namespace {

template <class T>
int precon_apply_any(parcel_hip_batch* h, const T* r, T* z) {
  if (!h || !r || !z) return api_fail(PARCEL_HIP_ERR_ARG, "batch_precon_apply: null argument");
  Batch& B = h->B;
  PARCEL_HIP_CHECK(hipSetDevice(B.device));
  if (B.precon == PARCEL_BATCH_FACTOR && !B.factor_src)
    return api_fail(PARCEL_HIP_ERR_ARG, "batch_precon_apply: no factor set");
  const size_t nv = (size_t)B.nv();
  DevBuf<T> tr, tz, tw;
  const T* dr = r;
  T* dz = z;
  if (!api_is_device_ptr(r)) { tr.alloc(nv); tr.upload(r, nv, B.stream); dr = tr.p; }
  const bool zdev = api_is_device_ptr(z);
  if (!zdev) { tz.alloc(nv); dz = tz.p; }
  BVecs<T> V = vecs_of<T>(B);
  if (V.F) {
    tw.alloc(nv);
    launch_lscan<T, T>(B, V.F, dr, tw.p, dz, nullptr, 0);
  } else {
    hipLaunchKernelGGL(k_b_scale<T>, dim3(tiled_nb(B.n), B.nsys), dim3(kNT), 0, B.stream, B.n, dr, V.dinv, dz,
                       Bat{(long)B.n, B.A.entries});
  }
  PARCEL_HIP_CHECK(hipGetLastError());
  if (!zdev) PARCEL_HIP_CHECK(hipMemcpyAsync(z, dz, sizeof(T) * nv, hipMemcpyDeviceToHost, B.stream));
  PARCEL_HIP_CHECK(hipStreamSynchronize(B.stream));
  return PARCEL_HIP_OK;
}

}  // namespace

extern "C" {

int parcel_hip_batch_precon_apply(parcel_hip_batch* h, const double* r, double* z) {
  PARCEL_API_BEGIN
  return precon_apply_any<double>(h, r, z);
  PARCEL_API_END
}

int parcel_hip_batch_precon_apply32(parcel_hip_batch* h, const float* r, float* z) {
  PARCEL_API_BEGIN
  return precon_apply_any<float>(h, r, z);
  PARCEL_API_END
}

int parcel_hip_batch_get_factor(const parcel_hip_batch* h, unsigned short* Ls, unsigned short* Lw,
                                unsigned short* Ld) {
  PARCEL_API_BEGIN
  if (!h || !Ls || !Lw || !Ld) return api_fail(PARCEL_HIP_ERR_ARG, "batch_get_factor: null argument");
  const Batch& B = h->B;
  if (B.factor_src != 2) return api_fail(PARCEL_HIP_ERR_ARG, "batch_get_factor: no BF16 factor set");
  PARCEL_HIP_CHECK(hipSetDevice(B.device));
  const long nv = B.nv();
  unsigned short* dst[3] = {Ls, Lw, Ld};
  for (int k = 0; k < 3; ++k)
    PARCEL_HIP_CHECK(hipMemcpyAsync(dst[k], B.Lbf.p + k * nv, sizeof(uint16_t) * nv, hipMemcpyDeviceToHost, B.stream));
  PARCEL_HIP_CHECK(hipStreamSynchronize(B.stream));
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

int parcel_hip_batch_destroy(parcel_hip_batch* h) {
  if (!h) return PARCEL_HIP_OK;
  (void)hipSetDevice(h->B.device);
  (void)hipStreamSynchronize(h->B.stream);
  hipStream_t s = h->B.stream;
  delete h;
  if (s) (void)hipStreamDestroy(s);
  return PARCEL_HIP_OK;
}

}  // extern "C"

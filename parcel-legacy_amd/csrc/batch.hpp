// batch.hpp -- the batched-solver object shared by batched.hip (solver) and
// mpnn.hip (the network that writes learned factors into a batch).
#pragma once

#include "internal.hpp"

namespace parcel {

// Outer (FP64 refinement) state of one system of a mixed batched solve.
struct OuterState {
  double bnorm;
  double rnorm;
  int done;
  int niter;
  int iret;
  int pad;
};

struct Batch {
  int device = 0;
  hipStream_t stream = nullptr;
  int nsys = 0, n = 0, nnz = 0;
  int nx = 0, ny = 0;  // five-point grid (0: the pattern is not one)
  int nstrip = 1;      // factor solves: independent strips of grid lines per system
  int precon = PARCEL_BATCH_NONE;
  int factor_src = 0;  // 0 none, 1 FP64 factor (IC(0)), 2 BF16 factor
  DevSELL A;           // shared pattern; A.val / A.val32 hold nsys*entries values
  std::vector<int> rowptr, col;  // 0-based host pattern
  DevBuf<int> diag_slot;         // SELL slot of each row's diagonal
  DevBuf<int4> nbr_slot;         // five-point grids: SELL slots of the s, w, e, n entries (-1: none)
  DevBuf<double> st5;     // five-point stencil c, e, n (3 * nsys * n) when the pattern is one and
  bool st5_ok = false;    // every system is exactly symmetric (the fused CG SpMV then reads these)
  DevBuf<double> dinv;           // nsys*n
  DevBuf<float> dinv32;
  DevBuf<double> L64;            // 3*nsys*n: Ls, Lw, Ld
  DevBuf<uint16_t> Lbf;          // 3*nsys*n bf16
  DevBuf<double> F64;            // 4*nsys*n: ls, 1/Ld, a_fwd, a_bwd
  DevBuf<double> sink;           // 1024: the factor apply's filler and padding stores (never read)
  DevBuf<float> F32;
  DevBuf<double> r, q, s, p0, p1, w, xb, bb;
  DevBuf<float> b32, d32, r32, q32, s32, p032, p132, w32;
  DevBuf<double> part;           // nsys*PART_N*kNBMax
  DevBuf<CGState> st;            // nsys
  DevBuf<OuterState> ost;        // nsys
  DevBuf<double> hist;
  int hist_cap = 0;
  PinnedBuf<CGState> hstate;     // host mirror of st, polled between chunks of iterations
  std::vector<long> hsptr;       // host copy of A.sptr (fallback: CSR values from SELL slots)
  // fallback of the learned path: systems whose solve ends with iret != 0
  // are solved again, from x = 0, by parcel_dcg's FP64 PCG with this
  // preconditioner (fb_itrmax <= 0: off)
  int fb_ipreflag = 2, fb_ilu = 0, fb_blocks = 32, fb_itrmax = 0;
  std::vector<int> fell_back;    // per system, last solve
  double inner_rtol = 1e-4;
  int inner_itrmax = 10000;
  bool profile = false;
  parcel_hip_batch_stats stats{};
  long vs() const { return n; }
  long nv() const { return (long)nsys * n; }
};

}  // namespace parcel

struct parcel_hip_batch {
  parcel::Batch B;
};

namespace parcel {
// Apply arrays F64/F32 from the batch's factor (L64 or Lbf); synchronises.
void derive_factor(Batch& B);
}  // namespace parcel

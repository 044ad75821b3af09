// internal.hpp -- plan objects and host-side entry points shared by the
// translation units of libparcel_hip.so (not part of the public C ABI).
#pragma once

#include <string>
#include <vector>

#include "common.hpp"
#include "parcel_hip.h"

namespace parcel {

// Owning device buffer.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    n = count;
    if (count) PARCEL_HIP_CHECK(hipMalloc(&p, count * sizeof(T)));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr; n = 0;
  }
  void upload(const T* h, size_t count, hipStream_t s) {
    if (count) PARCEL_HIP_CHECK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T>& h, hipStream_t s) {
    if (h.size() > n) alloc(h.size());
    upload(h.data(), h.size(), s);
  }
};

// Owning page-locked host buffer (device-to-host mirrors of small state).
template <class T>
struct PinnedBuf {
  T* p = nullptr;
  size_t n = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf() { release(); }
  void ensure(size_t count) {
    if (count <= n && p) return;
    release();
    PARCEL_HIP_CHECK(hipHostMalloc((void**)&p, count * sizeof(T), hipHostMallocDefault));
    n = count;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr; n = 0;
  }
};

// Sliced ELLPACK with slice height 64 (one wavefront), resident in HBM.
// Row i lives in slice i/64, lane i%64; its j-th stored entry (in the CSR
// order of the row) is at sptr[slice] + j*64 + lane; rows shorter than the
// slice width are padded with column -1.  A wavefront reads entry j of its
// 64 rows as one contiguous 512-byte (values) / 256-byte (columns) access,
// and every row still sums its entries in the reference's order.
struct DevSELL {
  int n = 0;
  long nnz = 0;          // real entries
  long entries = 0;      // stored entries (with padding)
  long nslices = 0;
  DevBuf<long> sptr;     // nslices+1
  DevBuf<int> sk;        // nslices: width of each slice
  DevBuf<int> col;       // entries, 0-based, -1 pads
  DevBuf<double> val;    // entries
  DevBuf<float> val32;   // FP32 copy (mixed-precision inner solve)
  DevBuf<double> val_lo; // lo parts of a double-double matrix (empty: none)
  // long rows: a slice is as wide as its rows after the few longest, whose
  // remaining entries follow the slices in col/val (entries counts them);
  // so[slice] = index of the slice's first overflow row (-1: none),
  // omask[slice] = its lanes holding one, optr = the rows' entry ranges.
  // Empty when no slice has one.
  long ovf_entries = 0;
  DevBuf<int> so;
  DevBuf<unsigned long long> omask;
  DevBuf<long> optr;
};

// 0-based CSR copy on the host (setup work: factorisations, level analysis).
struct HostCSR {
  int n = 0;
  std::vector<int> rowptr, col;
  std::vector<double> val;
  std::vector<double> val_lo;  // lo parts of a double-double matrix (empty: none)
};

// Level-scheduled triangular sweep of one block-Jacobi factor.
// Rows of every block are grouped into levels (a row's level is one more
// than the highest level of the rows it reads).  Level l of the whole
// preconditioner is described by meta[l] = {m, k, eo, ro}: m rows
// rows[ro .. ro+m) and an ELL slab of k entries per row stored entry-major
// (entry e of the t-th row at eo + e*m + t).  An entry's dependency is
// encoded in dep[]: >= 0 a global row id (read from the output vector),
// < 0 the LDS ring slot -(dep+1) holding a value computed `ring` levels or
// fewer before, kPad a padding entry.
constexpr int kPad = 0x7fffffff;
struct Sweep {
  int nlev = 0;
  DevBuf<int> blk_lev;   // nblocks+1: levels of block b are [blk_lev[b], blk_lev[b+1])
  DevBuf<int4> meta;     // nlev
  DevBuf<int> rows;      // global row ids in level order
  DevBuf<int> dep;       // encoded dependencies
  DevBuf<double> eval;   // factor values in accumulation order
  DevBuf<double> diag;   // per level-ordered row: scale (see mode)
  DevBuf<float> eval32;  // FP32 copies for the mixed-precision inner solve
  DevBuf<float> diag32;
  DevBuf<double> eval_lo;  // lo parts of double-double factors
  DevBuf<double> diag_lo;
  int max_width = 0;
  int max_k = 0;
  int ring = 0;          // levels kept in the LDS ring (0: none)
  int ring_stride = 0;   // doubles per ring level (>= max_width)
  long entries = 0;      // stored (padded) entries
  long real_entries = 0; // entries without padding
  long global_deps = 0;  // dependencies read from HBM instead of the ring
};

// Device-resident scalars of one CG solve (src/pcg.f90:240-255 keeps them in
// host variables and round-trips every one through PCIe; here they never
// leave HBM until the solve ends).  Iteration i reads slot (i-1)&1 and writes
// slot i&1, so no workgroup ever reads a value another workgroup of the same
// kernel is writing.
struct CGSlot {
  double rho0;   // rho0 of iteration `itr` (rho1 of itr-1)
  double alpha;  // alpha of iteration `itr` (written by its second kernel)
  double beta;   // beta used by iteration `itr`
  double rnorm;  // ||r|| entering iteration `itr`
  int itr;
  int pad[3];
};
struct CGState {
  double bnorm;
  int done_at;   // index of the K1 launch that ended the loop (0: running)
  int niter;
  int iret;
  int pad;
  CGSlot slot[2];
};

enum PreconKind { PRECON_NONE = 0, PRECON_PJ = 1, PRECON_BJ = 2, PRECON_RAS = 3 };
enum BJMode { BJ_ILU0 = 0, BJ_DILU = 1 };

// FP32 block-Jacobi apply as line scans of the block IC(0) factor (blocks of
// whole grid lines, symmetric couplings; precon.hip: lscan_setup): the apply
// arrays ls, 1/Ld, -Lw/Ld, -Lw(i+1)/Ld (each n floats), the forward values w
// and the sink line of the scans' filler stores.
struct LscanBJ {
  bool on = false;
  int nx = 0, ny = 0, nstrip = 0;
  DevBuf<float> F;
  DevBuf<float> w;
  DevBuf<double> sink;
};

// Line-wavefront form of the FP64 block-Jacobi apply: every block is whole
// grid lines of nx cells (whole planes of ny lines, or lines of one plane)
// whose factors couple a row only to the neighbouring cell of its line, the
// same cell of the neighbouring line and of the neighbouring plane.  The
// couplings (two per row, in the sweep's stored order) and the scales are
// kept per block in anti-diagonal order of the sweep (precon.hip,
// line_doff), so the rows one step of the wavefront visits are contiguous.
struct LineBJ {
  bool on = false;
  int nx = 0;
  int ny = 0;             // lines per plane when blocks hold several planes (0: one plane per block)
  int planes = 1;         // most planes in a block
  int waves = 1;          // wavefronts per workgroup: ceil(lines per plane / 64) x planes, rounded up to 2^k
  DevBuf<int> bstart;     // nblocks + 1 block starts
  DevBuf<double> fc, bc;  // line and cell couplings per row
  DevBuf<double> fp, bp;  // plane coupling per row (several planes per block)
  DevBuf<double> fd, bd;  // forward scale (D-ILU), backward scale
  DevBuf<double> sink;    // 64 doubles per block for the stores of idle lanes
};

struct Precon {
  int kind = PRECON_NONE;
  int mode = BJ_ILU0;
  int nblocks = 0;
  int nth = 256;        // workgroup size of the block-Jacobi apply kernel
  DevBuf<double> dinv;  // point Jacobi
  DevBuf<float> dinv32;
  Sweep fwd, bwd;       // block Jacobi
  LscanBJ lscan;        // FP32 line-scan form of the block factors (mixed solver)
  LineBJ line;          // line-wavefront form of the FP64 apply
  bool has32 = false;   // FP32 copies of the factors exist
  // double-double factors (precon_setup_dd): Dinv lo parts, apply precision
  DevBuf<double> dinv_lo;
  int dd_prec = 0;      // 0: not set up; 1: apply with hi parts; 2: double-double
  double setup_seconds = 0.0;
};

struct Plan {
  int device = 0;
  hipStream_t stream = nullptr;
  parcel_hip_param prm{};
  DevSELL A;
  HostCSR hA;  // kept only while setting up
  Precon M;
  // workspace
  DevBuf<double> r, q, s, p0, p1, x, b, t0;
  std::vector<DevBuf<double>> vec;  // extra work vectors of the other solvers
  DevBuf<double> basis;             // Krylov basis columns (GMRES family), n-strided
  DevBuf<double> basis2;            // second basis (orthonormal Q of CA-GMRES with MGS/CGS)
  DevBuf<double> wpart;             // partial-sum slots of the other solvers
  std::vector<DevBuf<float>> vec32; // FP32 work vectors (mixed-precision solve)
  std::vector<int> block_starts;    // explicit block-Jacobi blocks (nblocks+1 starts), empty: divide_region
  double inner_rtol = 1e-4;         // mixed-precision inner PCG tolerance
  int inner_itrmax = 10000;
  DevBuf<char> solver_state;        // device-resident scalars of the other solvers
  // restricted additive Schwarz (PRECON_RAS): the block D-ILU of the
  // expanded block-diagonal matrix (one block per extended box) lives in the
  // plan ras_plan; ras_src[p] is the row gathered into expanded row p,
  // ras_dst[p] the row it is scattered back to (-1: an overlap row, dropped)
  // (ras_src[p] = -1: the expanded row starts from 0).  With ras_add the
  // result is gathered instead: s[k] = (0 + z[ras_oa[k]]) + z[ras_ob[k]]
  // (the second term only where ras_ob[k] >= 0; CRS additive Schwarz with
  // iflagAS BASIC / INTERPOLATE)
  parcel_hip_plan* ras_owner = nullptr;
  Plan* ras_plan = nullptr;
  int ras_n = 0;
  DevBuf<int> ras_src, ras_dst;
  DevBuf<double> ras_r, ras_s;
  bool ras_add = false;
  DevBuf<int> ras_oa, ras_ob;
  // DIA storage of the matrix (plans created from diagonals): the fused CG
  // kernel reads these instead of the SELL slots (dia_nnd = 0: none)
  int dia_nnd = 0;
  DevBuf<double> dia_val;   // dia_nnd * n, diagonal-major
  DevBuf<int> dia_off;      // dia_nnd
  // symmetric stencil form of the matrix for the fused CG kernel (St7):
  // c, e, nn, t per row (FP64 and the FP32 copies of the mixed solver)
  bool st7 = false;
  int st7_nx = 0, st7_ny = 0;
  long st7_nxy = 0;
  DevBuf<double> st7_a[4];
  DevBuf<float> st7_a32[4];
  PinnedBuf<CGState> hstate;  // host mirror of cg, polled between chunks of iterations
  DevBuf<CGState> cg;       // device-resident solver scalars
  DevBuf<double> part;      // tiled partial sums (PART_N * kNBMax)
  DevBuf<double> hist;      // residual history
  int hist_cap = 0;
  parcel_hip_stats stats{};
};

// Partial-sum areas of P.part (each kNBMax doubles).
enum PartSlot { PART_BB = 0, PART_RR0 = 1, PART_RHO0 = 2, PART_GAMMA = 3, PART_RHO1 = 4, PART_RR = 5, PART_N = 8 };

// Fused residual update for the block-Jacobi apply inside CG iteration `it`:
// r -= alpha*q with alpha = slot.rho0 / gamma, gamma reduced from the
// partials of the preceding SpMV kernel.
template <class T>
struct CGFuseT {
  const T* q = nullptr;
  T* r = nullptr;
  CGState* st = nullptr;
  const double* part_gamma = nullptr;
  int nb_prod = 0;
  int it = 0;
};
using CGFuse = CGFuseT<double>;
using CGFuse32 = CGFuseT<float>;

// Vectors, matrix view and preconditioner of one CG run (T = double for the
// FP64 solver, float for the mixed-precision inner solve).
template <class T>
struct SellArgsT;
template <class T>
struct CGVecs {
  SellArgsT<T> A;
  const T* st7[4] = {nullptr, nullptr, nullptr, nullptr};  // the stencil arrays when the plan has them
  T *r = nullptr, *q = nullptr, *s = nullptr, *p0 = nullptr, *p1 = nullptr;
  const T* dinv = nullptr;   // point Jacobi scale, or null
  double* hist = nullptr;    // device residual history (hist_cap entries)
  int hist_cap = 0;
  void (*precon)(Plan&, const T*, T*, const CGFuseT<T>*) = nullptr;
  // the same without a fused update, returning at once when *skip != 0
  void (*precon_skip)(Plan&, const T*, T*, const int*) = nullptr;
};
template <class T>
int cg_run(Plan& P, const CGVecs<T>& V, T* x, const T* b, int itrmax, double rtol, double* hist_host, int* niter,
           int* iret);

// setup helpers (precon.hip)
void precon_setup(Plan& P);
// z = (L L^T)^-1 r in FP32 over nstrip strips of a grid of ny lines of nx
// cells (batched.hip, k_lscan with one system); w: forward values; a strip
// returns at once when st says the CG loop ended before iteration it
void lscan_apply32(const float* F, const float* r, float* w, float* z, int nx, int ny, int nstrip, long n,
                   const CGState* st, int it, double* sink, hipStream_t stream);
// divide_region (src/common.f90:125-182): 0-based rows [lo, hi) of block `rank` of `nproc`
void divide_region0(int rank, int nproc, int n, int* lo, int* hi);
// skip: if non-null, the kernels return at once when *skip != 0 (a solver
// whose loop has already ended on the device).
void precon_apply(Plan& P, const double* r, double* s, const CGFuse* fuse, const int* skip = nullptr);
// double-double preconditioner (factorisations in dd on the host; apply with
// the factors' lo parts when prec == 2); vectors as separate hi / lo arrays
void precon_setup_dd(Plan& P, int prec);
void precon_apply_dd(Plan& P, const double* rh, const double* rl, double* sh, double* sl, const int* skip = nullptr);
// FP32 apply with the factors rounded to float (precon_setup32 first)
void precon_setup32(Plan& P);
void precon_apply32(Plan& P, const float* r, float* s, const CGFuse32* fuse, const int* skip = nullptr);
void sell_build(Plan& P);
// the symmetric stencil arrays of a natural-order 5/7-point matrix (krylov.hip)
void stencil_build(Plan& P);

// solvers and kernels (krylov.hip)
void plan_workspace(Plan& P, int itrmax);
int cg_solve(Plan& P, double* x_dev, const double* b_dev, int itrmax, double rtol,
             double* hist_host, int* niter, int* iret);
int bicg_solve(Plan& P, double* x_dev, const double* b_dev, int itrmax, double rtol,
               double* hist_host, int* niter, int* iret);
int gmres_solve(Plan& P, double* x_dev, const double* b_dev, int itrmax, double rtol,
                double* hist_host, int* niter, int* iret);
int cagmres_solve(Plan& P, double* x_dev, const double* b_dev, int itrmax, double rtol,
                  double* hist_host, int* niter, int* iret);
// power_method_symMatrix in full (cbcg.hip): Emax of A M^-1, the eigenpair
// error ||A x - Emax x|| / ||Emax x|| and the eigenvector x (device, n)
void power_method_eigen(Plan& P, const double* x0, int itrmax, double* emax, double* errnorm, double* vec);
int cbcg_solve(Plan& P, double* x_dev, const double* b_dev, int itrmax, double rtol,
               double* hist_host, int* niter, int* iret);
int mixed_solve(Plan& P, double* x_dev, const double* b_dev, int itrmax, double rtol,
                double* hist_host, int* niter, int* iret);
int ddcg_solve(Plan& P, double* xh, double* xl, const double* bh, const double* bl, int itrmax, double rtol,
               double* hist_host, int* niter, int* iret);
void mixed_setup(Plan& P);
// n-vector work buffers 0..k-1 of P.vec
void ensure_vecs(Plan& P, int k);
// launch helpers shared by the solvers (krylov.hip)
void launch_residual(Plan& P, const double* x, const double* b, double* r, double* part);
void spmv(Plan& P, const double* x, double* y);
double dot_device(int n, const double* x, const double* y, hipStream_t st);
// allocation-free dot (krylov.hip): the partials live in a file-scope device array
double* noalloc_scratch(int dev);
void dot_noalloc(int dev, int n, const double* x, const double* y, double* out, bool out_dev, hipStream_t st);
double time_kernel(Plan& P, int which, int reps);

// C-API plumbing shared with the batched entry points (capi.hip)
int api_fail(int code, const std::string& msg);
int api_ensure_init();
int api_device();
bool api_is_device_ptr(const void* p);

}  // namespace parcel

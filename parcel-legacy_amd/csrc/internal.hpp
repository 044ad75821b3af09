// internal.hpp -- plan objects and host-side entry points shared by the
// translation units of libparcel_hip.so (not part of the public C ABI).
#pragma once

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

// 0-based CSR resident in HBM.
struct DevCSR {
  int n = 0;
  long nnz = 0;
  DevBuf<int> rowptr;   // n+1
  DevBuf<int> col;      // nnz, 0-based
  DevBuf<double> val;   // nnz
};

// 0-based CSR copy on the host (setup work: factorisations, level analysis).
struct HostCSR {
  int n = 0;
  std::vector<int> rowptr, col;
  std::vector<double> val;
};

// Level-scheduled triangular sweep of one block-Jacobi factor.
// Rows of every block are grouped into levels (a row's level is one more
// than the highest level of the rows it reads).  Level l of the whole
// preconditioner holds rows rows[row_off[l] .. row_off[l+1]) and an ELL slab
// of kmax[l] entries per row, stored entry-major: entry e of the t-th row is
// at ent_off[l] + e*m + t (m = rows in the level); a column id of -1 pads.
struct Sweep {
  int nlev = 0;
  DevBuf<int> blk_lev;   // nblocks+1: levels of block b are [blk_lev[b], blk_lev[b+1])
  DevBuf<int> row_off;   // nlev+1
  DevBuf<int> ent_off;   // nlev+1
  DevBuf<int> kmax;      // nlev
  DevBuf<int> rows;      // global row ids in level order
  DevBuf<int> ecol;      // global column ids (dependencies)
  DevBuf<double> eval;   // factor values in accumulation order
  DevBuf<double> diag;   // per level-ordered row: scale (see mode)
  int max_width = 0;
  long entries = 0;      // stored (padded) entries
  long real_entries = 0; // entries without padding
};

// Device-resident scalars of one CG solve (src/pcg.f90:240-255 keeps them in
// host variables and round-trips every one through PCIe; here they never
// leave HBM until the solve ends).
struct CGState {
  double alpha, beta, rho0, bnorm, rnorm, gamma, rho1, rr;
  int itr;   // iterations started
  int done;  // 1 once the loop has exited (converged, diverged or itrmax)
  int iret;  // PARCEL_RETURNCODE_*
  int pad;
};

enum PreconKind { PRECON_NONE = 0, PRECON_PJ = 1, PRECON_BJ = 2 };
enum BJMode { BJ_ILU0 = 0, BJ_DILU = 1 };

struct Precon {
  int kind = PRECON_NONE;
  int mode = BJ_ILU0;
  int nblocks = 0;
  DevBuf<double> dinv;  // point Jacobi
  Sweep fwd, bwd;       // block Jacobi
  double setup_seconds = 0.0;
};

struct Plan {
  int device = 0;
  hipStream_t stream = nullptr;
  parcel_hip_param prm{};
  DevCSR A;
  HostCSR hA;  // kept only while setting up
  Precon M;
  // workspace
  DevBuf<double> r, q, s, p0, p1, x, b, t0;
  DevBuf<CGState> cg;       // device-resident solver scalars
  DevBuf<double> part;      // tiled partial sums (4 * kNBMax)
  DevBuf<unsigned> counter; // last-workgroup counters (8)
  DevBuf<double> hist;      // residual history
  int hist_cap = 0;
  parcel_hip_stats stats{};
};

// setup helpers (precon.hip)
void precon_setup(Plan& P);
void precon_apply(Plan& P, const double* r, double* s, const int* skip);

// solvers and kernels (krylov.hip)
void plan_workspace(Plan& P, int itrmax);
int cg_solve(Plan& P, double* x_dev, const double* b_dev, int itrmax, double rtol,
             double* hist_host, int* niter, int* iret);
void spmv(Plan& P, const double* x, double* y);
double dot_device(int n, const double* x, const double* y, hipStream_t st);
double time_kernel(Plan& P, int which, int reps);

}  // namespace parcel

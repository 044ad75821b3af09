// capi.hip -- extern "C" entry points of libparcel_hip.so (include/parcel_hip.h).
//
// The one-shot entry points keep the argument lists of the reference Fortran
// wrappers (src/wrapper.f90); internally each one is plan_create +
// plan_solve + plan_destroy, the split KSP_CRS performs (src/krylov.f90:447-735).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>

#include "internal.hpp"

using namespace parcel;

struct parcel_hip_plan {
  Plan P;
};

namespace {

thread_local std::string g_last_error;
int g_device = -1;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

bool is_device_ptr(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t a;
  hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) { (void)hipGetLastError(); return false; }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

template <class T>
void fetch(const T* src, size_t count, std::vector<T>& dst) {
  dst.resize(count);
  if (!count) return;
  if (is_device_ptr(src))
    PARCEL_HIP_CHECK(hipMemcpy(dst.data(), src, count * sizeof(T), hipMemcpyDeviceToHost));
  else
    std::memcpy(dst.data(), src, count * sizeof(T));
}

int ensure_init() {
  if (g_device >= 0) return PARCEL_HIP_OK;
  return parcel_hip_init(0);
}

#define PARCEL_API_BEGIN try {
#define PARCEL_API_END                                               \
  }                                                                  \
  catch (const hip_error& e) {                                       \
    return fail(e.code < 0 ? e.code : PARCEL_HIP_ERR_HIP, e.what()); \
  }                                                                  \
  catch (const std::exception& e) {                                  \
    return fail(PARCEL_HIP_ERR_HIP, e.what());                       \
  }

void fill_defaults(parcel_hip_param* p) {
  // PARCEL_KSP_Default_Setting (src/type.f90:1007-1069)
  p->solver = PARCEL_SOLVER_CG;
  p->ipreflag = PARCEL_PRECON_NONE;
  p->ILU_method = PARCEL_ILUD;
  p->addL = 0;
  p->iflagAS = 4;
  p->iovlflag = 0;
  p->precon_thblock = -1;
  p->Neuman_itr = 1;
  p->independ_nvec = 1;
  p->nBlock = 2000;
  p->GMRES_m = 4;
  p->GMRES_GSflag = 1;
  p->CAGMRES_sstep = 4;
  p->CAGMRES_tstep = 1;
  p->CAGMRES_basis = 0;
  p->CAGMRES_QRflag = 4;
  p->CBCG_kstep = 4;
  p->CBCG_Eigenflag = 2;
  p->power_method_itrmax = 100;
  p->CAARNOLDI_sstep = 4;
  p->CAARNOLDI_tstep = 10;
  p->CAARNOLDI_basis = 1;
  p->CAARNOLDI_QRflag = 3;
  p->profile = 0;
}

// Merge user values over defaults (values < 0 keep the default), as
// SET_KSP_PARAM does with the -1 markers of INIT_PARCEL_INPUT.
parcel_hip_param merged(const parcel_hip_param* user) {
  parcel_hip_param d;
  fill_defaults(&d);
  if (!user) return d;
  parcel_hip_param r = *user;
  const int* u = reinterpret_cast<const int*>(user);
  int* o = reinterpret_cast<int*>(&r);
  const int* df = reinterpret_cast<const int*>(&d);
  for (size_t k = 0; k < sizeof(parcel_hip_param) / sizeof(int); ++k) o[k] = u[k] < 0 ? df[k] : u[k];
  // precon_thblock < 1 keeps its meaning (one block per compute unit)
  r.precon_thblock = user->precon_thblock;
  if (r.independ_nvec < 1) r.independ_nvec = 1;  // type.f90:1377
  return r;
}

// every field -1 ("unset, use the default"), as INIT_PARCEL_INPUT marks them
void unset_param(parcel_hip_param& prm) {
  for (size_t k = 0; k < sizeof(prm) / sizeof(int); ++k) reinterpret_cast<int*>(&prm)[k] = -1;
  prm.profile = 0;
}

}  // namespace

namespace parcel {
int api_fail(int code, const std::string& msg) { return fail(code, msg); }
int api_ensure_init() { return ensure_init(); }
int api_device() { return g_device; }
bool api_is_device_ptr(const void* p) { return is_device_ptr(p); }
}  // namespace parcel

extern "C" {

int parcel_hip_init(int myrank) {
  PARCEL_API_BEGIN
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count < 1) return fail(PARCEL_HIP_ERR_HIP, "no HIP device available");
  int dev = (myrank < 0 ? 0 : myrank) % count;
  PARCEL_HIP_CHECK(hipSetDevice(dev));
  noalloc_scratch(dev);  // the allocation-free dot's workspace on this device
  g_device = dev;
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

int parcel_hip_finalize(void) {
  g_device = -1;
  return PARCEL_HIP_OK;
}

const char* parcel_hip_last_error(void) { return g_last_error.c_str(); }

void parcel_hip_default_param(parcel_hip_param* prm) {
  if (prm) fill_defaults(prm);
}

}  // extern "C"

namespace {

int as_build(parcel_hip_plan* plan, const HostCSR& A);

// Plan set-up shared by the FP64 and the double-double entry points:
// dd_prec = 0 for FP64, else PRECISION_PRECONDITION of a dd plan; crsA_lo
// (dd only) may be null (PRECISION_A = 1).
int create_impl(parcel_hip_plan** out, int n, int gn, int nnz, int istart, const double* crsA,
                const double* crsA_lo, const int* crsRow_ptr, const int* crsCol, const parcel_hip_param* prm,
                int dd_prec, const std::vector<int>* block_starts = nullptr) {
  PARCEL_API_BEGIN
  if (!out) return fail(PARCEL_HIP_ERR_ARG, "plan pointer is NULL");
  *out = nullptr;
  if (n < 1 || nnz < 0 || !crsA || !crsRow_ptr || !crsCol)
    return fail(PARCEL_HIP_ERR_ARG, "invalid matrix arguments");
  if (gn != n || istart != 1)
    return fail(PARCEL_HIP_ERR_UNSUPPORTED,
                "one GPU holds one rank's whole system: expected gn == n and istart == 1");
  int rc = ensure_init();
  if (rc) return rc;
  auto t0 = std::chrono::steady_clock::now();
  auto* h = new parcel_hip_plan();
  Plan& P = h->P;
  P.device = g_device;
  P.prm = merged(prm);
  try {
    PARCEL_HIP_CHECK(hipStreamCreateWithFlags(&P.stream, hipStreamNonBlocking));
    HostCSR& A = P.hA;
    A.n = n;
    fetch(crsRow_ptr, (size_t)n + 1, A.rowptr);
    fetch(crsCol, (size_t)nnz, A.col);
    fetch(crsA, (size_t)nnz, A.val);
    if (crsA_lo) fetch(crsA_lo, (size_t)nnz, A.val_lo);
    if (A.rowptr[0] != 1 || A.rowptr[n] != nnz + 1)
      throw hip_error("crsRow_ptr must be 1-based with crsRow_ptr(n+1) = nnz+1", PARCEL_HIP_ERR_ARG);
    for (int i = 0; i <= n; ++i) A.rowptr[i] -= 1;
    for (int i = 0; i < n; ++i)
      if (A.rowptr[i + 1] < A.rowptr[i]) throw hip_error("crsRow_ptr is not monotone", PARCEL_HIP_ERR_ARG);
    for (long j = 0; j < nnz; ++j) {
      int c = A.col[j] - istart;
      if (c < 0 || c >= n)
        throw hip_error("column index outside this rank's rows (halo columns need >1 rank)",
                        PARCEL_HIP_ERR_UNSUPPORTED);
      A.col[j] = c;
    }
    sell_build(P);
    if (!dd_prec && (P.prm.solver == PARCEL_SOLVER_CG || P.prm.solver == PARCEL_SOLVER_CG_MIXED)) stencil_build(P);
    if (dd_prec && P.prm.solver != PARCEL_SOLVER_CG)
      throw hip_error("the double-double path has CG only", PARCEL_HIP_ERR_UNSUPPORTED);
    if (P.prm.solver != PARCEL_SOLVER_CG && P.prm.solver != PARCEL_SOLVER_BICG &&
        P.prm.solver != PARCEL_SOLVER_GMRES && P.prm.solver != PARCEL_SOLVER_CAGMRES &&
        P.prm.solver != PARCEL_SOLVER_CBCG && P.prm.solver != PARCEL_SOLVER_CG_MIXED)
      throw hip_error("solver not available on this path yet", PARCEL_HIP_ERR_UNSUPPORTED);
    if (P.prm.solver == PARCEL_SOLVER_GMRES && (P.prm.GMRES_m < 1 || P.prm.GMRES_m > 256))
      throw hip_error("GMRES_m must be in 1..256", PARCEL_HIP_ERR_ARG);
    if (P.prm.solver == PARCEL_SOLVER_CAGMRES) {
      if (P.prm.CAGMRES_sstep < 1 || P.prm.CAGMRES_sstep > 15)
        throw hip_error("CAGMRES_sstep must be in 1..15", PARCEL_HIP_ERR_ARG);
      if (P.prm.CAGMRES_tstep < 1 || P.prm.CAGMRES_sstep * P.prm.CAGMRES_tstep > 128)
        throw hip_error("CAGMRES_tstep must be >= 1 with s*t <= 128", PARCEL_HIP_ERR_ARG);
      const int q = P.prm.CAGMRES_QRflag;
      if (q < PARCEL_QR_MGS || q > PARCEL_QR_CHOLESKYQR2)
        throw hip_error("CAGMRES_QRflag must be in 1..5", PARCEL_HIP_ERR_ARG);
    }
    if (P.prm.solver == PARCEL_SOLVER_CBCG && (P.prm.CBCG_kstep < 1 || P.prm.CBCG_kstep > 8))
      throw hip_error("CBCG_kstep must be in 1..8", PARCEL_HIP_ERR_ARG);
    if (block_starts) P.block_starts = *block_starts;
    // additive Schwarz of the CRS path: the blocks' D-ILU runs in an
    // expanded plan (as_build); this plan's own set-up then has nothing to do
    const bool as = !dd_prec && !block_starts && P.prm.ipreflag == PARCEL_PRECON_ADDITIVESCHWARZ;
    if (as) P.prm.ipreflag = PARCEL_PRECON_NONE;
    if (dd_prec) precon_setup_dd(P, dd_prec);
    else precon_setup(P);
    if (as) {
      P.prm.ipreflag = PARCEL_PRECON_ADDITIVESCHWARZ;
      int arc = as_build(h, P.hA);
      if (arc != PARCEL_HIP_OK) throw hip_error(g_last_error, arc);
    }
    P.stats.precon_setup_s = P.M.setup_seconds;
    P.stats.bj_blocks = P.M.nblocks;
    P.stats.bj_levels_fwd = P.M.fwd.nlev;
    P.stats.bj_levels_bwd = P.M.bwd.nlev;
    P.stats.bj_entries = P.M.fwd.entries + P.M.bwd.entries;
    P.stats.bj_nnz_l = P.M.fwd.real_entries;
    P.stats.bj_nnz_u = P.M.bwd.real_entries;
    P.stats.bj_lscan = P.M.lscan.on ? 1 : 0;
    P.stats.spmv_stencil = P.st7 ? 1 : 0;
    P.stats.sell_entries = P.A.entries;
    P.stats.sell_overflow = P.A.ovf_entries;
    P.stats.bj_lines = P.M.line.on ? 1 : 0;
    P.hA = HostCSR();  // setup copy no longer needed
    plan_workspace(P, 1);
    PARCEL_HIP_CHECK(hipStreamSynchronize(P.stream));
  } catch (...) {
    if (P.stream) (void)hipStreamDestroy(P.stream);
    delete h;
    throw;
  }
  P.stats.setup_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  *out = h;
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

// Additive Schwarz of the CRS path at one rank (get_DILU_DInv_AS_crs,
// src/matrix.f90:3809-4375; AdditiveSchwarz_crs :4749-4858): the nblocks
// blocks of divide_region, each extended by addL rows on both sides within
// [1, n] (get_AS_region :3146-3218), D-ILU factorised on their own rows and
// columns (DILU_DInv_AS_crs :3428-3480: ILU_method 1 ILUD, any other value
// the row-sum form).  Expanded row q is row src[q] of its extended block;
// the expanded block-diagonal matrix's block D-ILU runs in a second plan.
// With addL = 0, n < addL or a block of fewer than 2 addL rows
// (check_AS_region :3220-3287) the reference falls back to block Jacobi
// with the D-ILU factor: this plan then becomes that block Jacobi.
// iflagAS 1 BASIC / 2 RESTRICT gather r on every expanded row, 3
// INTERPOLATE / 4 NONE only on owned rows (0 on the overlap); 2 / 4 keep
// the owned rows' results, 1 / 3 add the two results of an overlap row
// (AS_Local :5804-5904: s = 0, then lower overlaps, owned rows, upper
// overlaps, so a row's sum is (0 + first) + second).
int as_build(parcel_hip_plan* plan, const HostCSR& A) {
  Plan& P = plan->P;
  const int n = A.n, addL = P.prm.addL, flag = P.prm.iflagAS;
  if (addL < 0) return fail(PARCEL_HIP_ERR_ARG, "additive Schwarz: addL must be >= 0");
  if (flag < 1 || flag > 4) return fail(PARCEL_HIP_ERR_ARG, "additive Schwarz: iflagAS must be 1..4");
  int nblk = P.prm.precon_thblock;
  if (nblk < 1) {
    hipDeviceProp_t prop;
    PARCEL_HIP_CHECK(hipGetDeviceProperties(&prop, P.device));
    nblk = prop.multiProcessorCount;
  }
  nblk = std::max(1, std::min(nblk, n));
  const int dilu = P.prm.ILU_method == PARCEL_ILUD ? PARCEL_ILUD : PARCEL_MLUD;
  std::vector<int> lo(nblk), hi(nblk);
  bool ok = addL > 0 && n >= addL;
  for (int t = 0; t < nblk; ++t) {
    divide_region0(t, nblk, n, &lo[t], &hi[t]);
    if ((hi[t] - lo[t]) / 2 < addL) ok = false;
  }
  if (!ok) {
    // block Jacobi with the D-ILU factor the reference computes
    P.prm.ipreflag = PARCEL_PRECON_BLOCKJACOBI;
    P.prm.ILU_method = dilu;
    P.prm.precon_thblock = nblk;
    P.hA = A;
    precon_setup(P);
    return PARCEL_HIP_OK;
  }
  std::vector<int> src, starts, rp(1, 1), col, oa(n, -1), ob(n, -1);
  std::vector<unsigned char> own;
  std::vector<double> val;
  for (int t = 0; t < nblk; ++t) {
    const int a = std::max(lo[t] - addL, 0), b = std::min(hi[t] + addL, n);
    const int base = (int)src.size();
    starts.push_back(base);
    for (int k = a; k < b; ++k) {
      const int q = (int)src.size();
      const bool mine = lo[t] <= k && k < hi[t];
      src.push_back(k);
      own.push_back(mine ? 1 : 0);
      for (int j = A.rowptr[k]; j < A.rowptr[k + 1]; ++j) {
        const int c = A.col[j];
        if (c < a || c >= b) continue;
        col.push_back(c - a + base + 1);
        val.push_back(A.val[j]);
      }
      rp.push_back((int)col.size() + 1);
      // AS_Local's phases: a lower overlap row (k < lo[t]) is added first,
      // the owned row second, an upper overlap row (k >= hi[t]) last
      if (mine) {
        if (oa[k] < 0) oa[k] = q;
        else ob[k] = q;
      } else if (k < lo[t]) {
        // before the owner's contribution, which block t - 1 (processed
        // earlier) left in oa: it moves to the second term
        ob[k] = oa[k];
        oa[k] = q;
      } else {
        ob[k] = q;  // after it
      }
    }
  }
  const int ne = (int)src.size();
  starts.push_back(ne);
  std::vector<int> gsrc(ne), dst(ne);
  for (int q = 0; q < ne; ++q) {
    gsrc[q] = (own[q] || flag == 1 || flag == 2) ? src[q] : -1;
    dst[q] = own[q] ? src[q] : -1;
  }
  parcel_hip_param prm;
  unset_param(prm);
  prm.solver = PARCEL_SOLVER_CG;
  prm.ipreflag = PARCEL_PRECON_BLOCKJACOBI;
  prm.ILU_method = dilu;
  prm.independ_nvec = P.prm.independ_nvec;
  prm.precon_thblock = nblk;
  prm.profile = 0;
  parcel_hip_plan* aux = nullptr;
  int rc = create_impl(&aux, ne, ne, (int)col.size(), 1, val.data(), nullptr, rp.data(), col.data(), &prm, 0,
                       &starts);
  if (rc) return rc;
  (void)hipStreamSynchronize(aux->P.stream);
  (void)hipStreamDestroy(aux->P.stream);
  aux->P.stream = P.stream;
  P.ras_owner = aux;
  P.ras_plan = &aux->P;
  P.ras_n = ne;
  P.ras_src.upload(gsrc, P.stream);
  P.ras_dst.upload(dst, P.stream);
  if (flag == 1 || flag == 3) {
    P.ras_add = true;
    P.ras_oa.upload(oa, P.stream);
    P.ras_ob.upload(ob, P.stream);
  }
  P.ras_r.alloc((size_t)ne);
  P.ras_s.alloc((size_t)ne);
  P.M.kind = PRECON_RAS;
  P.M.nblocks = nblk;
  PARCEL_HIP_CHECK(hipStreamSynchronize(P.stream));
  return PARCEL_HIP_OK;
}

}  // namespace

extern "C" {

int parcel_hip_plan_create(parcel_hip_plan** out, int n, int gn, int nnz, int istart, const double* crsA,
                           const int* crsRow_ptr, const int* crsCol, const parcel_hip_param* prm) {
  return create_impl(out, n, gn, nnz, istart, crsA, nullptr, crsRow_ptr, crsCol, prm, 0);
}

int parcel_hip_plan_create_dd(parcel_hip_plan** out, int n, int gn, int nnz, int istart, const double* crsA_hi,
                              const double* crsA_lo, const int* crsRow_ptr, const int* crsCol,
                              const parcel_hip_param* prm, int PRECISION_A, int PRECISION_PRECONDITION) {
  if ((PRECISION_A != 1 && PRECISION_A != 2) || (PRECISION_PRECONDITION != 1 && PRECISION_PRECONDITION != 2))
    return fail(PARCEL_HIP_ERR_ARG, "PRECISION_A and PRECISION_PRECONDITION must be 1 or 2");
  if (PRECISION_A == 2 && !crsA_lo) return fail(PARCEL_HIP_ERR_ARG, "PRECISION_A = 2 needs crsA_lo");
  return create_impl(out, n, gn, nnz, istart, crsA_hi, PRECISION_A == 2 ? crsA_lo : nullptr, crsRow_ptr, crsCol,
                     prm, PRECISION_PRECONDITION);
}

int parcel_hip_plan_solve_dd(parcel_hip_plan* plan, double* vecx_hi, double* vecx_lo, const double* vecb_hi,
                             const double* vecb_lo, int* itrmax, double rtolmax, double* reshistory,
                             int PRECISION_b, int PRECISION_x, int* iret) {
  PARCEL_API_BEGIN
  if (!plan || !vecx_hi || !vecb_hi || !itrmax || !iret) return fail(PARCEL_HIP_ERR_ARG, "NULL argument");
  if ((PRECISION_b != 1 && PRECISION_b != 2) || (PRECISION_x != 1 && PRECISION_x != 2))
    return fail(PARCEL_HIP_ERR_ARG, "PRECISION_b and PRECISION_x must be 1 or 2");
  if ((PRECISION_b == 2 && !vecb_lo) || (PRECISION_x == 2 && !vecx_lo))
    return fail(PARCEL_HIP_ERR_ARG, "lo vectors required for precision 2");
  Plan& P = plan->P;
  if (!P.M.dd_prec) return fail(PARCEL_HIP_ERR_ARG, "plan was not created with parcel_hip_plan_create_dd");
  PARCEL_HIP_CHECK(hipSetDevice(P.device));
  const int n = P.A.n;
  // device copies: x (hi, lo) and b (hi, lo); lo parts zero at precision 1 (ddpcg.f90:330-345)
  DevBuf<double> d(4 * (size_t)n);
  double *xh = d.p, *xl = d.p + n, *bh = d.p + 2 * (size_t)n, *bl = d.p + 3 * (size_t)n;
  const size_t bytes = sizeof(double) * (size_t)n;
  auto put = [&](double* dst, const double* src) {
    PARCEL_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, is_device_ptr(src) ? hipMemcpyDeviceToDevice
                                                                         : hipMemcpyHostToDevice, P.stream));
  };
  put(xh, vecx_hi);
  put(bh, vecb_hi);
  if (PRECISION_x == 2) put(xl, vecx_lo);
  else PARCEL_HIP_CHECK(hipMemsetAsync(xl, 0, bytes, P.stream));
  if (PRECISION_b == 2) put(bl, vecb_lo);
  else PARCEL_HIP_CHECK(hipMemsetAsync(bl, 0, bytes, P.stream));
  int niter = 0, ret = 0;
  int rc = ddcg_solve(P, xh, xl, bh, bl, *itrmax, rtolmax, reshistory, &niter, &ret);
  if (rc) return rc;
  auto get = [&](double* dst, const double* src) {
    PARCEL_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, is_device_ptr(dst) ? hipMemcpyDeviceToDevice
                                                                         : hipMemcpyDeviceToHost, P.stream));
  };
  get(vecx_hi, xh);
  if (PRECISION_x == 2) get(vecx_lo, xl);
  PARCEL_HIP_CHECK(hipStreamSynchronize(P.stream));
  *itrmax = niter;
  *iret = ret;
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

int parcel_hip_ddcg(int icomm, double* vecx_hi, double* vecx_lo, const double* vecb_hi, const double* vecb_lo,
                    int n, int gn, int nnz, int istart, const double* crsA_hi, const double* crsA_lo,
                    const int* crsRow_ptr, const int* crsCol, int ipreflag, int ILU_method, int addL, int iflagAS,
                    int* itrmax, double rtolmax, double* reshistory, int iovlflag, int precon_thblock,
                    int independ_nvec, int nBlock, int PRECISION_A, int PRECISION_b, int PRECISION_x,
                    int PRECISION_PRECONDITION, int* iret) {
  (void)icomm;
  parcel_hip_param prm;
  unset_param(prm);
  prm.solver = PARCEL_SOLVER_CG;
  prm.ipreflag = ipreflag;
  prm.ILU_method = ILU_method;
  prm.addL = addL;
  prm.iflagAS = iflagAS;
  prm.iovlflag = iovlflag;
  prm.precon_thblock = precon_thblock;
  prm.independ_nvec = independ_nvec;
  prm.nBlock = nBlock;
  parcel_hip_plan* plan = nullptr;
  int rc = parcel_hip_plan_create_dd(&plan, n, gn, nnz, istart, crsA_hi, crsA_lo, crsRow_ptr, crsCol, &prm,
                                     PRECISION_A, PRECISION_PRECONDITION);
  if (rc) return rc;
  rc = parcel_hip_plan_solve_dd(plan, vecx_hi, vecx_lo, vecb_hi, vecb_lo, itrmax, rtolmax, reshistory, PRECISION_b,
                                PRECISION_x, iret);
  parcel_hip_plan_destroy(plan);
  return rc;
}

int parcel_hip_plan_destroy(parcel_hip_plan* plan) {
  PARCEL_API_BEGIN
  if (!plan) return PARCEL_HIP_OK;
  hipStream_t s = plan->P.stream;
  if (s) (void)hipStreamSynchronize(s);
  if (plan->P.ras_owner) {
    plan->P.ras_owner->P.stream = nullptr;  // shares this plan's stream
    (void)parcel_hip_plan_destroy(plan->P.ras_owner);
    plan->P.ras_owner = nullptr;
    plan->P.ras_plan = nullptr;
  }
  delete plan;
  if (s) (void)hipStreamDestroy(s);
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

int parcel_hip_plan_stats(const parcel_hip_plan* plan, parcel_hip_stats* stats) {
  if (!plan || !stats) return fail(PARCEL_HIP_ERR_ARG, "NULL argument");
  *stats = plan->P.stats;
  return PARCEL_HIP_OK;
}

int parcel_hip_plan_set_inner(parcel_hip_plan* plan, double inner_rtol, int inner_itrmax) {
  if (!plan) return fail(PARCEL_HIP_ERR_ARG, "NULL argument");
  if (!(inner_rtol > 0.0) || inner_itrmax < 1) return fail(PARCEL_HIP_ERR_ARG, "inner_rtol > 0 and inner_itrmax >= 1");
  plan->P.inner_rtol = inner_rtol;
  plan->P.inner_itrmax = inner_itrmax;
  return PARCEL_HIP_OK;
}

int parcel_hip_plan_set_profile(parcel_hip_plan* plan, int on) {
  if (!plan) return fail(PARCEL_HIP_ERR_ARG, "NULL argument");
  plan->P.prm.profile = on ? 1 : 0;
  return PARCEL_HIP_OK;
}

int parcel_hip_plan_solve(parcel_hip_plan* plan, double* vecx, const double* vecb, int* itrmax,
                          double rtolmax, double* reshistory, int* iret) {
  PARCEL_API_BEGIN
  if (!plan || !vecx || !vecb || !itrmax || !iret) return fail(PARCEL_HIP_ERR_ARG, "NULL argument");
  Plan& P = plan->P;
  PARCEL_HIP_CHECK(hipSetDevice(P.device));
  const int n = P.A.n;
  const int maxit = *itrmax;
  const bool xdev = is_device_ptr(vecx), bdev = is_device_ptr(vecb);
  double* x = vecx;
  const double* b = vecb;
  if (!xdev) {
    if (P.x.n < (size_t)n) P.x.alloc(n);
    P.x.upload(vecx, n, P.stream);
    x = P.x.p;
  }
  if (!bdev) {
    if (P.b.n < (size_t)n) P.b.alloc(n);
    P.b.upload(vecb, n, P.stream);
    b = P.b.p;
  }
  int niter = 0, ret = 0;
  int rc;
  switch (P.prm.solver) {
    case PARCEL_SOLVER_BICG: rc = bicg_solve(P, x, b, maxit, rtolmax, reshistory, &niter, &ret); break;
    case PARCEL_SOLVER_GMRES: rc = gmres_solve(P, x, b, maxit, rtolmax, reshistory, &niter, &ret); break;
    case PARCEL_SOLVER_CAGMRES: rc = cagmres_solve(P, x, b, maxit, rtolmax, reshistory, &niter, &ret); break;
    case PARCEL_SOLVER_CBCG: rc = cbcg_solve(P, x, b, maxit, rtolmax, reshistory, &niter, &ret); break;
    case PARCEL_SOLVER_CG_MIXED: rc = mixed_solve(P, x, b, maxit, rtolmax, reshistory, &niter, &ret); break;
    default: rc = cg_solve(P, x, b, maxit, rtolmax, reshistory, &niter, &ret); break;
  }
  if (rc) return rc;
  if (!xdev) {
    PARCEL_HIP_CHECK(hipMemcpyAsync(vecx, x, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, P.stream));
  }
  PARCEL_HIP_CHECK(hipStreamSynchronize(P.stream));
  *itrmax = niter;  // wrapper.f90:630 itrmax = niter
  *iret = ret;
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

// dia2crs (src/matrix.f90:7536-7576): row k takes diagonal j in stored
// order at column k + offset(j) (global, 1-based), dropping zero entries
// (minReal8 = 0, src/type.f90:445).  The plan keeps the diagonals too: its
// fused CG kernel reads them (MatVec_dia order, the same sums).
int parcel_hip_plan_create_dia(parcel_hip_plan** out, int n, int gn, int istart, const double* diaA, const int* offset,
                               int nnd, const parcel_hip_param* prm) {
  PARCEL_API_BEGIN
  if (!out) return fail(PARCEL_HIP_ERR_ARG, "plan pointer is NULL");
  *out = nullptr;
  if (n < 1 || nnd < 1 || !diaA || !offset) return fail(PARCEL_HIP_ERR_ARG, "invalid DIA matrix arguments");
  int rc = ensure_init();
  if (rc) return rc;
  std::vector<int> off;
  std::vector<double> dv;
  fetch(offset, (size_t)nnd, off);
  fetch(diaA, (size_t)nnd * n, dv);
  std::vector<int> rp(n + 1), col;
  std::vector<double> val;
  col.reserve((size_t)n * nnd);
  val.reserve((size_t)n * nnd);
  for (int k = 1; k <= n; ++k) {
    rp[k - 1] = (int)col.size() + 1;
    const long row = (long)k - 1 + istart;
    for (int j = 0; j < nnd; ++j) {
      const double a = dv[(size_t)j * n + (k - 1)];
      if (std::fabs(a) > 0.0) {
        col.push_back((int)(row + off[j]));
        val.push_back(a);
      }
    }
  }
  rp[n] = (int)col.size() + 1;
  rc = create_impl(out, n, gn, (int)col.size(), istart, val.data(), nullptr, rp.data(), col.data(), prm, 0);
  if (rc) return rc;
  Plan& P = (*out)->P;
  P.dia_nnd = nnd;
  P.dia_val.alloc((size_t)nnd * n);
  P.dia_val.upload(dv.data(), dv.size(), P.stream);
  P.dia_off.alloc((size_t)nnd);
  P.dia_off.upload(off.data(), off.size(), P.stream);
  PARCEL_HIP_CHECK(hipStreamSynchronize(P.stream));
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

static int one_shot_dia(parcel_hip_param& prm, double* vecx, const double* vecb, int n, int gn, int istart,
                        const double* diaA, const int* offset, int nnd, int* itrmax, double rtolmax,
                        double* reshistory, int* iret) {
  parcel_hip_plan* plan = nullptr;
  int rc = parcel_hip_plan_create_dia(&plan, n, gn, istart, diaA, offset, nnd, &prm);
  if (rc) return rc;
  rc = parcel_hip_plan_solve(plan, vecx, vecb, itrmax, rtolmax, reshistory, iret);
  parcel_hip_plan_destroy(plan);
  return rc;
}

static void common_param(parcel_hip_param& prm, int solver, int ipreflag, int ILU_method, int addL, int iflagAS,
                         int iovlflag, int precon_thblock, int Neuman_itr, int independ_nvec, int nBlock) {
  unset_param(prm);
  prm.solver = solver;
  prm.ipreflag = ipreflag;
  prm.ILU_method = ILU_method;
  prm.addL = addL;
  prm.iflagAS = iflagAS;
  prm.iovlflag = iovlflag;
  prm.precon_thblock = precon_thblock;
  prm.Neuman_itr = Neuman_itr;
  prm.independ_nvec = independ_nvec;
  prm.nBlock = nBlock;
}

int parcel_hip_dcg_dia(int icomm, double* vecx, const double* vecb, int n, int gn, int istart, const double* diaA,
                       const int* offset, int nnd, int ipreflag, int ILU_method, int addL, int iflagAS, int* itrmax,
                       double rtolmax, double* reshistory, int iovlflag, int precon_thblock, int Neuman_itr,
                       int independ_nvec, int nBlock, int* iret) {
  (void)icomm;
  parcel_hip_param prm;
  common_param(prm, PARCEL_SOLVER_CG, ipreflag, ILU_method, addL, iflagAS, iovlflag, precon_thblock, Neuman_itr,
               independ_nvec, nBlock);
  return one_shot_dia(prm, vecx, vecb, n, gn, istart, diaA, offset, nnd, itrmax, rtolmax, reshistory, iret);
}

int parcel_hip_dbicgstab_dia(int icomm, double* vecx, const double* vecb, int n, int gn, int istart,
                             const double* diaA, const int* offset, int nnd, int ipreflag, int ILU_method, int addL,
                             int iflagAS, int* itrmax, double rtolmax, double* reshistory, int iovlflag,
                             int precon_thblock, int Neuman_itr, int independ_nvec, int nBlock, int* iret) {
  (void)icomm;
  parcel_hip_param prm;
  common_param(prm, PARCEL_SOLVER_BICG, ipreflag, ILU_method, addL, iflagAS, iovlflag, precon_thblock, Neuman_itr,
               independ_nvec, nBlock);
  return one_shot_dia(prm, vecx, vecb, n, gn, istart, diaA, offset, nnd, itrmax, rtolmax, reshistory, iret);
}

int parcel_hip_dgmres_dia(int icomm, double* vecx, const double* vecb, int n, int gn, int istart, const double* diaA,
                          const int* offset, int nnd, int ipreflag, int ILU_method, int addL, int iflagAS,
                          int* itrmax, double rtolmax, double* reshistory, int iovlflag, int precon_thblock,
                          int Neuman_itr, int independ_nvec, int nBlock, int GMRES_m, int GMRES_GSflag, int* iret) {
  (void)icomm;
  parcel_hip_param prm;
  common_param(prm, PARCEL_SOLVER_GMRES, ipreflag, ILU_method, addL, iflagAS, iovlflag, precon_thblock, Neuman_itr,
               independ_nvec, nBlock);
  prm.GMRES_m = GMRES_m;
  prm.GMRES_GSflag = GMRES_GSflag;
  return one_shot_dia(prm, vecx, vecb, n, gn, istart, diaA, offset, nnd, itrmax, rtolmax, reshistory, iret);
}

int parcel_hip_dcagmres_dia(int icomm, double* vecx, const double* vecb, int n, int gn, int istart,
                            const double* diaA, const int* offset, int nnd, int ipreflag, int ILU_method, int addL,
                            int iflagAS, int* itrmax, double rtolmax, double* reshistory, int iovlflag,
                            int precon_thblock, int Neuman_itr, int independ_nvec, int nBlock, int CAGMRES_sstep,
                            int CAGMRES_tstep, int CAGMRES_basis, int CAGMRES_QRflag, int* iret) {
  (void)icomm;
  parcel_hip_param prm;
  common_param(prm, PARCEL_SOLVER_CAGMRES, ipreflag, ILU_method, addL, iflagAS, iovlflag, precon_thblock,
               Neuman_itr, independ_nvec, nBlock);
  prm.CAGMRES_sstep = CAGMRES_sstep;
  prm.CAGMRES_tstep = CAGMRES_tstep;
  prm.CAGMRES_basis = CAGMRES_basis;
  prm.CAGMRES_QRflag = CAGMRES_QRflag;
  return one_shot_dia(prm, vecx, vecb, n, gn, istart, diaA, offset, nnd, itrmax, rtolmax, reshistory, iret);
}

int parcel_hip_dcbcg_dia(int icomm, double* vecx, const double* vecb, int n, int gn, int istart, const double* diaA,
                         const int* offset, int nnd, int ipreflag, int ILU_method, int addL, int iflagAS, int* itrmax,
                         double rtolmax, double* reshistory, int iovlflag, int precon_thblock, int Neuman_itr,
                         int independ_nvec, int nBlock, int CBCG_kstep, int CBCG_Eigenflag, int power_method_itrmax,
                         int CAARNOLDI_sstep, int CAARNOLDI_tstep, int CAARNOLDI_basis, int CAARNOLDI_QRflag,
                         int* iret) {
  (void)icomm;
  parcel_hip_param prm;
  common_param(prm, PARCEL_SOLVER_CBCG, ipreflag, ILU_method, addL, iflagAS, iovlflag, precon_thblock, Neuman_itr,
               independ_nvec, nBlock);
  prm.CBCG_kstep = CBCG_kstep;
  prm.CBCG_Eigenflag = CBCG_Eigenflag;
  prm.power_method_itrmax = power_method_itrmax;
  prm.CAARNOLDI_sstep = CAARNOLDI_sstep;
  prm.CAARNOLDI_tstep = CAARNOLDI_tstep;
  prm.CAARNOLDI_basis = CAARNOLDI_basis;
  prm.CAARNOLDI_QRflag = CAARNOLDI_QRflag;
  return one_shot_dia(prm, vecx, vecb, n, gn, istart, diaA, offset, nnd, itrmax, rtolmax, reshistory, iret);
}

// power_method_symMatrix on a created plan (src/power_method.f90:188-437):
// max(1, itr) iterations from vecx (host or device; left unchanged), then the
// reference's eigen outputs: one eigenpair, EigenValueReal/Imag(1),
// EigenErrorNorm(1), MaxEigenID = MinEigenID = 1 (:400-425), and the first
// nvec components of the normalised eigenvector into evec (host or device).
static int eigen_run(parcel_hip_plan* plan, const double* vecx, int itr, double* ev_re, double* ev_im, double* evec,
                     int nvec, double* err, int* maxid, int* minid) {
  PARCEL_API_BEGIN
  if (!plan || !vecx) return fail(PARCEL_HIP_ERR_ARG, "eigen: null plan or start vector");
  Plan& P = plan->P;
  if (P.M.dd_prec) return fail(PARCEL_HIP_ERR_ARG, "eigen: the power method runs on FP64 plans");
  PARCEL_HIP_CHECK(hipSetDevice(P.device));
  const int n = P.A.n;
  DevBuf<double> d(2 * (size_t)n);
  const size_t bytes = sizeof(double) * (size_t)n;
  PARCEL_HIP_CHECK(hipMemcpyAsync(d.p, vecx, bytes, is_device_ptr(vecx) ? hipMemcpyDeviceToDevice
                                                                         : hipMemcpyHostToDevice, P.stream));
  double emax = 0.0, en = 0.0;
  // set_KSP_power_method_itrmax keeps at least one iteration (src/type.f90:1331-1336)
  power_method_eigen(P, d.p, std::max(1, itr), &emax, &en, d.p + n);
  if (evec && nvec > 0)
    PARCEL_HIP_CHECK(hipMemcpyAsync(evec, d.p + n, sizeof(double) * (size_t)std::min(nvec, n),
                                    is_device_ptr(evec) ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                    P.stream));
  PARCEL_HIP_CHECK(hipStreamSynchronize(P.stream));
  if (ev_re) *ev_re = emax;
  if (ev_im) *ev_im = 0.0;
  if (err) *err = en;
  if (maxid) *maxid = 1;
  if (minid) *minid = 1;
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

// parcel_dpower / parcel_dcaarnoldi through PARCEL_KSP_CRS / _DIA: the CRS
// driver copies size(EigenValueReal) = 1 component of the eigenvector
// (src/krylov.f90:360-369), the DIA driver all n (get_EigenPair_KSP,
// src/krylov.f90:665-673, src/EigenProblem.f90:18-46).  The reference never
// sets niter and iret for the power method (src/type.f90:805, :902 carry no
// initialiser); these return 0.
static int eigen_one_shot(parcel_hip_param& prm, const double* vecx, int n, int gn, int nnz, int istart,
                          const double* crsA, const int* crsRow_ptr, const int* crsCol, const double* diaA,
                          const int* offset, int nnd, int* itrmax, int itr, double* ev_re, double* ev_im,
                          double* evec, double* err, int* maxid, int* minid, int* iret) {
  if (!vecx || !iret) return fail(PARCEL_HIP_ERR_ARG, "eigen: null argument");
  parcel_hip_plan* plan = nullptr;
  int rc = diaA ? parcel_hip_plan_create_dia(&plan, n, gn, istart, diaA, offset, nnd, &prm)
                : parcel_hip_plan_create(&plan, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, &prm);
  if (rc) return rc;
  rc = eigen_run(plan, vecx, itr, ev_re, ev_im, evec, diaA ? n : 1, err, maxid, minid);
  parcel_hip_plan_destroy(plan);
  if (rc) return rc;
  if (itrmax) *itrmax = 0;
  *iret = 0;
  return PARCEL_HIP_OK;
}

static int one_shot(parcel_hip_param& prm, double* vecx, const double* vecb, int n, int gn, int nnz, int istart,
                    const double* crsA, const int* crsRow_ptr, const int* crsCol, int* itrmax, double rtolmax,
                    double* reshistory, int* iret) {
  parcel_hip_plan* plan = nullptr;
  int rc = parcel_hip_plan_create(&plan, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, &prm);
  if (rc) return rc;
  rc = parcel_hip_plan_solve(plan, vecx, vecb, itrmax, rtolmax, reshistory, iret);
  parcel_hip_plan_destroy(plan);
  return rc;
}


int parcel_hip_dbicgstab(int icomm, double* vecx, const double* vecb, int n, int gn, int nnz, int istart,
                         const double* crsA, const int* crsRow_ptr, const int* crsCol, int ipreflag,
                         int ILU_method, int addL, int iflagAS, int* itrmax, double rtolmax,
                         double* reshistory, int iovlflag, int precon_thblock, int Neuman_itr,
                         int independ_nvec, int nBlock, int* iret) {
  (void)icomm;
  parcel_hip_param prm;
  unset_param(prm);
  prm.solver = PARCEL_SOLVER_BICG;
  prm.ipreflag = ipreflag;
  prm.ILU_method = ILU_method;
  prm.addL = addL;
  prm.iflagAS = iflagAS;
  prm.iovlflag = iovlflag;
  prm.precon_thblock = precon_thblock;
  prm.Neuman_itr = Neuman_itr;
  prm.independ_nvec = independ_nvec;
  prm.nBlock = nBlock;
  return one_shot(prm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, itrmax, rtolmax, reshistory,
                  iret);
}

int parcel_hip_dcg(int icomm, double* vecx, const double* vecb, int n, int gn, int nnz, int istart,
                   const double* crsA, const int* crsRow_ptr, const int* crsCol, int ipreflag,
                   int ILU_method, int addL, int iflagAS, int* itrmax, double rtolmax,
                   double* reshistory, int iovlflag, int precon_thblock, int Neuman_itr,
                   int independ_nvec, int nBlock, int* iret) {
  (void)icomm;
  parcel_hip_param prm;
  for (size_t k = 0; k < sizeof(prm) / sizeof(int); ++k) reinterpret_cast<int*>(&prm)[k] = -1;
  prm.solver = PARCEL_SOLVER_CG;
  prm.ipreflag = ipreflag;
  prm.ILU_method = ILU_method;
  prm.addL = addL;
  prm.iflagAS = iflagAS;
  prm.iovlflag = iovlflag;
  prm.precon_thblock = precon_thblock;
  prm.Neuman_itr = Neuman_itr;
  prm.independ_nvec = independ_nvec;
  prm.nBlock = nBlock;
  prm.profile = 0;
  parcel_hip_plan* plan = nullptr;
  int rc = parcel_hip_plan_create(&plan, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, &prm);
  if (rc) return rc;
  rc = parcel_hip_plan_solve(plan, vecx, vecb, itrmax, rtolmax, reshistory, iret);
  parcel_hip_plan_destroy(plan);
  return rc;
}

int parcel_hip_dgmres(int icomm, double* vecx, const double* vecb, int n, int gn, int nnz, int istart,
                      const double* crsA, const int* crsRow_ptr, const int* crsCol, int ipreflag,
                      int ILU_method, int addL, int iflagAS, int* itrmax, double rtolmax,
                      double* reshistory, int iovlflag, int precon_thblock, int Neuman_itr,
                      int independ_nvec, int nBlock, int GMRES_m, int GMRES_GSflag, int* iret) {
  (void)icomm;
  parcel_hip_param prm;
  unset_param(prm);
  prm.solver = PARCEL_SOLVER_GMRES;
  prm.ipreflag = ipreflag;
  prm.ILU_method = ILU_method;
  prm.addL = addL;
  prm.iflagAS = iflagAS;
  prm.iovlflag = iovlflag;
  prm.precon_thblock = precon_thblock;
  prm.Neuman_itr = Neuman_itr;
  prm.independ_nvec = independ_nvec;
  prm.nBlock = nBlock;
  prm.GMRES_m = GMRES_m;
  prm.GMRES_GSflag = GMRES_GSflag;
  return one_shot(prm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, itrmax, rtolmax, reshistory,
                  iret);
}

int parcel_hip_dcagmres(int icomm, double* vecx, const double* vecb, int n, int gn, int nnz, int istart,
                        const double* crsA, const int* crsRow_ptr, const int* crsCol, int ipreflag,
                        int ILU_method, int addL, int iflagAS, int* itrmax, double rtolmax,
                        double* reshistory, int iovlflag, int precon_thblock, int Neuman_itr,
                        int independ_nvec, int nBlock, int CAGMRES_sstep, int CAGMRES_tstep,
                        int CAGMRES_basis, int CAGMRES_QRflag, int* iret) {
  (void)icomm;
  parcel_hip_param prm;
  unset_param(prm);
  prm.solver = PARCEL_SOLVER_CAGMRES;
  prm.ipreflag = ipreflag;
  prm.ILU_method = ILU_method;
  prm.addL = addL;
  prm.iflagAS = iflagAS;
  prm.iovlflag = iovlflag;
  prm.precon_thblock = precon_thblock;
  prm.Neuman_itr = Neuman_itr;
  prm.independ_nvec = independ_nvec;
  prm.nBlock = nBlock;
  prm.CAGMRES_sstep = CAGMRES_sstep;
  prm.CAGMRES_tstep = CAGMRES_tstep;
  prm.CAGMRES_basis = CAGMRES_basis;
  prm.CAGMRES_QRflag = CAGMRES_QRflag;
  return one_shot(prm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, itrmax, rtolmax, reshistory,
                  iret);
}

int parcel_hip_dcbcg(int icomm, double* vecx, const double* vecb, int n, int gn, int nnz, int istart,
                     const double* crsA, const int* crsRow_ptr, const int* crsCol, int ipreflag, int ILU_method,
                     int addL, int iflagAS, int* itrmax, double rtolmax, double* reshistory, int iovlflag,
                     int precon_thblock, int Neuman_itr, int independ_nvec, int nBlock, int CBCG_kstep,
                     int CBCG_Eigenflag, int power_method_itrmax, int CAARNOLDI_sstep, int CAARNOLDI_tstep,
                     int CAARNOLDI_basis, int CAARNOLDI_QRflag, int* iret) {
  (void)icomm;
  parcel_hip_param prm;
  unset_param(prm);
  prm.solver = PARCEL_SOLVER_CBCG;
  prm.ipreflag = ipreflag;
  prm.ILU_method = ILU_method;
  prm.addL = addL;
  prm.iflagAS = iflagAS;
  prm.iovlflag = iovlflag;
  prm.precon_thblock = precon_thblock;
  prm.Neuman_itr = Neuman_itr;
  prm.independ_nvec = independ_nvec;
  prm.nBlock = nBlock;
  prm.CBCG_kstep = CBCG_kstep;
  prm.CBCG_Eigenflag = CBCG_Eigenflag;
  prm.power_method_itrmax = power_method_itrmax;
  prm.CAARNOLDI_sstep = CAARNOLDI_sstep;
  prm.CAARNOLDI_tstep = CAARNOLDI_tstep;
  prm.CAARNOLDI_basis = CAARNOLDI_basis;
  prm.CAARNOLDI_QRflag = CAARNOLDI_QRflag;
  return one_shot(prm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, itrmax, rtolmax, reshistory,
                  iret);
}

int parcel_hip_dcg_mixed(int icomm, double* vecx, const double* vecb, int n, int gn, int nnz, int istart,
                         const double* crsA, const int* crsRow_ptr, const int* crsCol, int ipreflag,
                         int ILU_method, int addL, int iflagAS, int* itrmax, double rtolmax,
                         double* reshistory, int iovlflag, int precon_thblock, int Neuman_itr,
                         int independ_nvec, int nBlock, double inner_rtol, int inner_itrmax, int* iret) {
  (void)icomm;
  if (!(inner_rtol > 0.0) || inner_itrmax < 1) return fail(PARCEL_HIP_ERR_ARG, "inner_rtol > 0 and inner_itrmax >= 1");
  parcel_hip_param prm;
  unset_param(prm);
  prm.solver = PARCEL_SOLVER_CG_MIXED;
  prm.ipreflag = ipreflag;
  prm.ILU_method = ILU_method;
  prm.addL = addL;
  prm.iflagAS = iflagAS;
  prm.iovlflag = iovlflag;
  prm.precon_thblock = precon_thblock;
  prm.Neuman_itr = Neuman_itr;
  prm.independ_nvec = independ_nvec;
  prm.nBlock = nBlock;
  parcel_hip_plan* plan = nullptr;
  int rc = parcel_hip_plan_create(&plan, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, &prm);
  if (rc) return rc;
  rc = parcel_hip_plan_set_inner(plan, inner_rtol, inner_itrmax);
  if (!rc) rc = parcel_hip_plan_solve(plan, vecx, vecb, itrmax, rtolmax, reshistory, iret);
  parcel_hip_plan_destroy(plan);
  return rc;
}

int parcel_hip_plan_spmv(parcel_hip_plan* plan, const double* x, double* y) {
  PARCEL_API_BEGIN
  if (!plan || !x || !y) return fail(PARCEL_HIP_ERR_ARG, "NULL argument");
  Plan& P = plan->P;
  const int n = P.A.n;
  DevBuf<double> tx, ty;
  const double* dx = x;
  double* dy = y;
  if (!is_device_ptr(x)) { tx.alloc(n); tx.upload(x, n, P.stream); dx = tx.p; }
  bool ydev = is_device_ptr(y);
  if (!ydev) { ty.alloc(n); dy = ty.p; }
  spmv(P, dx, dy);
  if (!ydev) PARCEL_HIP_CHECK(hipMemcpyAsync(y, dy, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, P.stream));
  PARCEL_HIP_CHECK(hipStreamSynchronize(P.stream));
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

int parcel_hip_plan_precon(parcel_hip_plan* plan, const double* r, double* s) {
  PARCEL_API_BEGIN
  if (!plan || !r || !s) return fail(PARCEL_HIP_ERR_ARG, "NULL argument");
  Plan& P = plan->P;
  const int n = P.A.n;
  DevBuf<double> tr, ts;
  const double* dr = r;
  double* ds = s;
  if (!is_device_ptr(r)) { tr.alloc(n); tr.upload(r, n, P.stream); dr = tr.p; }
  bool sdev = is_device_ptr(s);
  if (!sdev) { ts.alloc(n); ds = ts.p; }
  precon_apply(P, dr, ds, nullptr);
  if (!sdev) PARCEL_HIP_CHECK(hipMemcpyAsync(s, ds, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, P.stream));
  PARCEL_HIP_CHECK(hipStreamSynchronize(P.stream));
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

int parcel_hip_dot(int n, const double* x, const double* y, double* result) {
  PARCEL_API_BEGIN
  if (n < 0 || !x || !y || !result) return fail(PARCEL_HIP_ERR_ARG, "invalid argument");
  int rc = ensure_init();
  if (rc) return rc;
  DevBuf<double> tx, ty;
  const double* dx = x;
  const double* dy = y;
  if (!is_device_ptr(x)) { tx.alloc(n); tx.upload(x, n, nullptr); dx = tx.p; }
  if (!is_device_ptr(y)) { ty.alloc(n); ty.upload(y, n, nullptr); dy = ty.p; }
  *result = dot_device(n, dx, dy, nullptr);
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

int parcel_hip_dot_noalloc(int n, const double* x, const double* y, double* result) {
  PARCEL_API_BEGIN
  if (n < 0 || !x || !y || !result) return fail(PARCEL_HIP_ERR_ARG, "invalid argument");
  int rc = ensure_init();
  if (rc) return rc;
  if (n > 0 && (!is_device_ptr(x) || !is_device_ptr(y)))
    return fail(PARCEL_HIP_ERR_ARG, "dot_noalloc: x and y must be device pointers (staging host arrays would allocate)");
  int dev = 0;
  PARCEL_HIP_CHECK(hipGetDevice(&dev));
  dot_noalloc(dev, n, x, y, result, is_device_ptr(result), nullptr);
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

int parcel_hip_plan_time_kernel(parcel_hip_plan* plan, int which, int reps, double* mean_ms) {
  PARCEL_API_BEGIN
  if (!plan || !mean_ms || reps < 1 || (which != 1 && which != 2))
    return fail(PARCEL_HIP_ERR_ARG, "invalid argument");
  *mean_ms = time_kernel(plan->P, which, reps);
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

int parcel_hip_malloc(void** ptr, unsigned long bytes) {
  PARCEL_API_BEGIN
  if (!ptr) return fail(PARCEL_HIP_ERR_ARG, "NULL argument");
  int rc = ensure_init();
  if (rc) return rc;
  *ptr = nullptr;
  PARCEL_HIP_CHECK(hipMalloc(ptr, bytes ? bytes : 1));
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

int parcel_hip_free(void* ptr) {
  PARCEL_API_BEGIN
  if (ptr) PARCEL_HIP_CHECK(hipFree(ptr));
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

int parcel_hip_memcpy(void* dst, const void* src, unsigned long bytes, int kind) {
  PARCEL_API_BEGIN
  if ((!dst || !src) && bytes) return fail(PARCEL_HIP_ERR_ARG, "NULL argument");
  hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice
                  : kind == 2 ? hipMemcpyDeviceToHost
                  : kind == 3 ? hipMemcpyDeviceToDevice : hipMemcpyDefault;
  if (kind < 1 || kind > 3) return fail(PARCEL_HIP_ERR_ARG, "kind must be 1, 2 or 3");
  if (bytes) PARCEL_HIP_CHECK(hipMemcpy(dst, src, bytes, k));
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

int parcel_hip_memset(void* ptr, int value, unsigned long bytes) {
  PARCEL_API_BEGIN
  if (!ptr && bytes) return fail(PARCEL_HIP_ERR_ARG, "NULL argument");
  if (bytes) {
    PARCEL_HIP_CHECK(hipMemset(ptr, value, bytes));
    PARCEL_HIP_CHECK(hipDeviceSynchronize());
  }
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

int parcel_hip_device_synchronize(void) {
  PARCEL_API_BEGIN
  PARCEL_HIP_CHECK(hipDeviceSynchronize());
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

}  // extern "C"

// --- stencil (DDM) entry point -------------------------------------------------
//
// parcel_dcg_ddm (src/ddm_kernel.f90:2893-2986): the matrix comes as nnd
// diagonals over the rank's ghost-padded grid (margin cells on every side),
// vectors in the same padded layout.  With the block-Jacobi preconditioner the
// GPU build of the reference solves D-ILU on boxes (init_DILU_DDM with
// nx/4, ny/4, nz/4 boxes per direction, src/src_DDM/ddm_kernel.f90:1009-1016
// and src/src_DDM/precon_ddm_GPU.f90:370-470; get_DILU_DDM :2646-2922).  Here
// the interior cells are renumbered box by box (boxes x fastest, cells x
// fastest inside a box), the diagonals become a CRS matrix in that order, and
// the boxes become contiguous block-Jacobi blocks of the CRS path, whose
// D-ILU (ILU_method 1) is the box D-ILU above.

namespace {

struct DdmGrid {
  int dim = 0;
  int nx = 1, ny = 1, nz = 1;        // interior cells per direction
  int mx = 1, my = 1, mz = 1;        // padded
  int margin = 0;
};

// Box split of one direction exactly as init_DILU_DDM: div = max(n/4, 1)
// boxes of n_local = ceil(n/div) cells, the last ones cut at n.
std::vector<int> box_edges(int n) {
  int div = std::max(n / 4, 1);
  const int loc = (n % div) ? n / div + 1 : n / div;
  std::vector<int> e;
  for (int i = 0; i < div; ++i) {
    const int s0 = loc * i, s1 = std::min(n, loc * (i + 1));
    if (s1 > s0) e.push_back(s0);
  }
  e.push_back(n);
  return e;
}

}  // namespace

extern "C" {

// Restricted additive Schwarz on the box-major system: expanded row p of box
// b is a cell of b's extended box (lexicographic z, y, x order inside the
// box); the expanded matrix keeps the couplings inside the extended box; its
// block D-ILU (ILUD, one block per box) runs in a second plan sharing this
// plan's stream.
static int ras_build(parcel_hip_plan* plan, const DdmGrid& G, int sx, int sy, int sz, const std::vector<int>& ex,
                     const std::vector<int>& ey, const std::vector<int>& ez, const std::vector<int>& pos,
                     const std::vector<int>& ioff, const std::vector<double>& vd, long m, int nnd, int addL) {
  Plan& P = plan->P;
  const int nbx = (int)ex.size() - 1, nby = (int)ey.size() - 1, nbz = (int)ez.size() - 1;
  auto lo = [&](const std::vector<int>& e, int k) { return k > 0 ? e[k] - addL : e[k]; };
  auto hi = [&](const std::vector<int>& e, int k, int nb) { return k < nb - 1 ? e[k + 1] + addL : e[k + 1]; };
  for (int k = 1; k < nbx; ++k)
    if (ex[k] - addL < 0 || ex[k] + addL > G.nx) return fail(PARCEL_HIP_ERR_UNSUPPORTED, "dcg_ddm: addL wider than the grid");
  for (int k = 1; k < nby; ++k)
    if (ey[k] - addL < 0 || ey[k] + addL > G.ny) return fail(PARCEL_HIP_ERR_UNSUPPORTED, "dcg_ddm: addL wider than the grid");
  for (int k = 1; k < nbz; ++k)
    if (ez[k] - addL < 0 || ez[k] + addL > G.nz) return fail(PARCEL_HIP_ERR_UNSUPPORTED, "dcg_ddm: addL wider than the grid");
  std::vector<int> src, dst, starts, rp(1, 1), col;
  std::vector<double> val;
  std::vector<long> cell;  // padded index of each expanded row of the current box
  std::vector<int> local;  // padded index -> expanded position (current box), -1 outside
  std::vector<std::pair<int, double>> row;
  for (int bz = 0; bz < nbz; ++bz)
    for (int by = 0; by < nby; ++by)
      for (int bx = 0; bx < nbx; ++bx) {
        const int x0 = lo(ex, bx), x1 = hi(ex, bx, nbx), y0 = lo(ey, by), y1 = hi(ey, by, nby);
        const int z0 = lo(ez, bz), z1 = hi(ez, bz, nbz);
        const int base = (int)src.size();
        starts.push_back(base);
        cell.clear();
        for (int jz = z0; jz < z1; ++jz)
          for (int jy = y0; jy < y1; ++jy)
            for (int jx = x0; jx < x1; ++jx) {
              const long j = (jx + sx) + (long)G.mx * (jy + sy) + (long)G.mx * G.my * (jz + sz);
              const bool own = jx >= ex[bx] && jx < ex[bx + 1] && jy >= ey[by] && jy < ey[by + 1] &&
                               jz >= ez[bz] && jz < ez[bz + 1];
              src.push_back(pos[j]);
              dst.push_back(own ? pos[j] : -1);
              cell.push_back(j);
            }
        if (local.empty()) local.assign((size_t)m, -1);
        for (size_t q = 0; q < cell.size(); ++q) local[cell[q]] = base + (int)q;
        for (size_t q = 0; q < cell.size(); ++q) {
          const long j = cell[q];
          row.clear();
          for (int ip = 0; ip < nnd; ++ip) {
            const long jn = j + ioff[ip];
            if (jn < 0 || jn >= m || local[jn] < 0) continue;  // outside the extended box
            const double v = vd[(size_t)ip * m + j];
            if (v == 0.0 && jn != j) continue;
            row.emplace_back(local[jn], v);
          }
          std::sort(row.begin(), row.end(), [](const std::pair<int, double>& a, const std::pair<int, double>& c) {
            return a.first < c.first;
          });
          for (auto& e : row) { col.push_back(e.first + 1); val.push_back(e.second); }
          rp.push_back((int)col.size() + 1);
        }
        for (size_t q = 0; q < cell.size(); ++q) local[cell[q]] = -1;
      }
  const int ne = (int)src.size();
  starts.push_back(ne);
  parcel_hip_param prm;
  unset_param(prm);
  prm.solver = PARCEL_SOLVER_CG;
  prm.ipreflag = PARCEL_PRECON_BLOCKJACOBI;
  prm.ILU_method = PARCEL_ILUD;
  prm.independ_nvec = 1;
  prm.precon_thblock = (int)starts.size() - 1;
  prm.profile = 0;
  parcel_hip_plan* aux = nullptr;
  int rc = create_impl(&aux, ne, ne, (int)col.size(), 1, val.data(), nullptr, rp.data(), col.data(), &prm, 0,
                       &starts);
  if (rc) return rc;
  // the block solves run on the outer plan's stream
  (void)hipStreamSynchronize(aux->P.stream);
  (void)hipStreamDestroy(aux->P.stream);
  aux->P.stream = P.stream;
  P.ras_owner = aux;
  P.ras_plan = &aux->P;
  P.ras_n = ne;
  P.ras_src.upload(src, P.stream);
  P.ras_dst.upload(dst, P.stream);
  P.ras_r.alloc((size_t)ne);
  P.ras_s.alloc((size_t)ne);
  P.M.kind = PRECON_RAS;
  PARCEL_HIP_CHECK(hipStreamSynchronize(P.stream));
  return PARCEL_HIP_OK;
}

int parcel_hip_dcg_ddm(int icomm, double* x, const double* b, int nrow, int nnd, const int* ioff_dia,
                       const double* val_dia, int num_neighbor_ranks, const int* neighbor_ranks, int margin,
                       int ncomp_grids, const int* ndiv_grids, const int* num_grids, const int* ioff_grids,
                       int div_direc_th, int ipreflag, int addL, int iflagAS, int* itrmax, double rtolmax,
                       double* reshistory, int iovlflag, int precon_thblock, int Neuman_itr, int independ_nvec,
                       int nBlock, int* iret) {
  PARCEL_API_BEGIN
  (void)icomm; (void)neighbor_ranks; (void)ndiv_grids; (void)div_direc_th; (void)iflagAS;
  (void)iovlflag; (void)precon_thblock; (void)Neuman_itr; (void)nBlock;
  if (!x || !b || !ioff_dia || !val_dia || !num_grids || !ioff_grids || !itrmax || !iret)
    return fail(PARCEL_HIP_ERR_ARG, "dcg_ddm: null argument");
  if (num_neighbor_ranks != 0)
    return fail(PARCEL_HIP_ERR_UNSUPPORTED, "dcg_ddm: one GPU holds one rank's whole grid (no neighbour ranks)");
  if (ncomp_grids < 1 || ncomp_grids > 3 || margin < 0 || nnd < 1)
    return fail(PARCEL_HIP_ERR_ARG, "dcg_ddm: bad grid description");
  if (ipreflag != PARCEL_PRECON_NONE && ipreflag != PARCEL_PRECON_POINTJACOBI &&
      ipreflag != PARCEL_PRECON_BLOCKJACOBI && ipreflag != PARCEL_PRECON_ADDITIVESCHWARZ)
    return fail(PARCEL_HIP_ERR_UNSUPPORTED, "dcg_ddm: preconditioner not available on the stencil path");
  const bool as = ipreflag == PARCEL_PRECON_ADDITIVESCHWARZ;
  if (as && addL < 0) return fail(PARCEL_HIP_ERR_ARG, "dcg_ddm: addL must be >= 0");
  DdmGrid G;
  G.dim = ncomp_grids;
  G.margin = margin;
  int* nn[3] = {&G.nx, &G.ny, &G.nz};
  for (int d = 0; d < ncomp_grids; ++d) *nn[d] = num_grids[d];
  if (G.nx < 1 || G.ny < 1 || G.nz < 1) return fail(PARCEL_HIP_ERR_ARG, "dcg_ddm: empty grid");
  G.mx = G.nx + 2 * margin;
  G.my = ncomp_grids >= 2 ? G.ny + 2 * margin : 1;
  G.mz = ncomp_grids >= 3 ? G.nz + 2 * margin : 1;
  const long m = (long)G.mx * G.my * G.mz;
  if (m != nrow) return fail(PARCEL_HIP_ERR_ARG, "dcg_ddm: nrow differs from the padded grid size");
  const int sx = margin, sy = ncomp_grids >= 2 ? margin : 0, sz = ncomp_grids >= 3 ? margin : 0;
  const long n = (long)G.nx * G.ny * G.nz;
  // box-major numbering of the interior cells
  const std::vector<int> ex = box_edges(G.nx), ey = box_edges(G.ny), ez = box_edges(G.nz);
  std::vector<int> perm(n), pos(m, -1), starts;
  long p = 0;
  for (size_t bz = 0; bz + 1 < ez.size(); ++bz)
    for (size_t by = 0; by + 1 < ey.size(); ++by)
      for (size_t bx = 0; bx + 1 < ex.size(); ++bx) {
        starts.push_back((int)p);
        for (int jz = ez[bz]; jz < ez[bz + 1]; ++jz)
          for (int jy = ey[by]; jy < ey[by + 1]; ++jy)
            for (int jx = ex[bx]; jx < ex[bx + 1]; ++jx) {
              const long j = (jx + sx) + (long)G.mx * (jy + sy) + (long)G.mx * G.my * (jz + sz);
              perm[p] = (int)j;
              pos[j] = (int)p;
              ++p;
            }
      }
  starts.push_back((int)p);
  std::vector<int> ioff(nnd);
  std::vector<double> vd;
  fetch(ioff_dia, (size_t)nnd, ioff);
  fetch(val_dia, (size_t)m * nnd, vd);
  // CRS in the new numbering, columns ascending; entries coupling to ghost
  // cells must be zero (a single rank has no halo values)
  std::vector<int> rp(n + 1), col;
  std::vector<double> val;
  col.reserve((size_t)n * nnd);
  val.reserve((size_t)n * nnd);
  std::vector<std::pair<int, double>> row;
  for (long r = 0; r < n; ++r) {
    const long j = perm[r];
    row.clear();
    for (int ip = 0; ip < nnd; ++ip) {
      const long jn = j + ioff[ip];
      const double v = vd[(size_t)ip * m + j];
      const bool inside = jn >= 0 && jn < m && pos[jn] >= 0;
      if (!inside) {
        if (v != 0.0) throw hip_error("dcg_ddm: coupling to a ghost cell (halo values need >1 rank)",
                                      PARCEL_HIP_ERR_UNSUPPORTED);
        continue;
      }
      if (v == 0.0 && jn != j) continue;
      row.emplace_back(pos[jn], v);
    }
    std::sort(row.begin(), row.end(), [](const std::pair<int, double>& a, const std::pair<int, double>& c) {
      return a.first < c.first;
    });
    rp[r] = (int)col.size() + 1;
    for (auto& e : row) { col.push_back(e.first + 1); val.push_back(e.second); }
  }
  rp[n] = (int)col.size() + 1;
  // vectors in the new numbering
  std::vector<double> xb, bb, xp(n), bp(n);
  fetch(x, (size_t)m, xb);
  fetch(b, (size_t)m, bb);
  for (long r = 0; r < n; ++r) { xp[r] = xb[perm[r]]; bp[r] = bb[perm[r]]; }
  parcel_hip_param prm;
  unset_param(prm);
  prm.solver = PARCEL_SOLVER_CG;
  prm.ipreflag = as ? PARCEL_PRECON_NONE : ipreflag;
  prm.ILU_method = PARCEL_ILUD;
  prm.independ_nvec = independ_nvec;
  prm.precon_thblock = (int)starts.size() - 1;
  prm.profile = 0;
  parcel_hip_plan* plan = nullptr;
  int rc = create_impl(&plan, (int)n, (int)n, (int)col.size(), 1, val.data(), nullptr, rp.data(), col.data(), &prm,
                       0, ipreflag == PARCEL_PRECON_BLOCKJACOBI ? &starts : nullptr);
  if (rc) return rc;
  if (as) {
    // solve_DILU_AS_DDM_GPU (precon_ddm_GPU.f90:1573-1868) at one rank: the
    // boxes of init_AS_DDM (:370-643) extended by addL cells on every side
    // facing another box, D-ILU on each extended box, the box's own cells
    // kept (the overlap sums of the update kernel only reach extension
    // cells, which are never copied out)
    rc = ras_build(plan, G, sx, sy, sz, ex, ey, ez, pos, ioff, vd, m, nnd, addL);
    if (rc) {
      parcel_hip_plan_destroy(plan);
      return rc;
    }
  }
  rc = parcel_hip_plan_solve(plan, xp.data(), bp.data(), itrmax, rtolmax, reshistory, iret);
  parcel_hip_plan_destroy(plan);
  if (rc) return rc;
  for (long r = 0; r < n; ++r) xb[perm[r]] = xp[r];
  if (is_device_ptr(x))
    PARCEL_HIP_CHECK(hipMemcpy(x, xb.data(), sizeof(double) * m, hipMemcpyHostToDevice));
  else
    std::memcpy(x, xb.data(), sizeof(double) * m);
  return PARCEL_HIP_OK;
  PARCEL_API_END
}

int parcel_hip_dpower(int icomm, double* vecx, const double* vecb, int n, int gn, int nnz, int istart,
                      const double* crsA, const int* crsRow_ptr, const int* crsCol, int ipreflag, int ILU_method,
                      int addL, int iflagAS, int* itrmax, int iovlflag, int precon_thblock, int Neuman_itr,
                      int independ_nvec, int nBlock, int power_method_itrmax, double* EigenValueReal,
                      double* EigenValueImag, double* EigenVector, double* EigenErrorNorm, int* MaxEigenID,
                      int* MinEigenID, int* iret) {
  (void)icomm; (void)vecb;
  parcel_hip_param prm;
  common_param(prm, PARCEL_SOLVER_CG, ipreflag, ILU_method, addL, iflagAS, iovlflag, precon_thblock, Neuman_itr,
               independ_nvec, nBlock);
  prm.power_method_itrmax = power_method_itrmax;
  return eigen_one_shot(prm, vecx, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, nullptr, nullptr, 0, itrmax,
                        power_method_itrmax,
                        EigenValueReal, EigenValueImag, EigenVector, EigenErrorNorm, MaxEigenID, MinEigenID, iret);
}

int parcel_hip_dpower_dia(int icomm, double* vecx, const double* vecb, int n, int gn, int istart, const double* diaA,
                          const int* offset, int nnd, int ipreflag, int ILU_method, int addL, int iflagAS,
                          int* itrmax, int iovlflag, int precon_thblock, int Neuman_itr, int independ_nvec,
                          int nBlock, int power_method_itrmax, double* EigenValueReal, double* EigenValueImag,
                          double* EigenVector, double* EigenErrorNorm, int* MaxEigenID, int* MinEigenID, int* iret) {
  (void)icomm; (void)vecb;
  parcel_hip_param prm;
  common_param(prm, PARCEL_SOLVER_CG, ipreflag, ILU_method, addL, iflagAS, iovlflag, precon_thblock, Neuman_itr,
               independ_nvec, nBlock);
  prm.power_method_itrmax = power_method_itrmax;
  return eigen_one_shot(prm, vecx, n, gn, 0, istart, nullptr, nullptr, nullptr, diaA, offset, nnd, itrmax,
                        power_method_itrmax,
                        EigenValueReal, EigenValueImag, EigenVector, EigenErrorNorm, MaxEigenID, MinEigenID, iret);
}

// KSP_pCAARNOLDI runs power_method_symMatrix (src/pcagmres.f90:2647-2651).
// parcel_dcaarnoldi leaves power_method_itrmax at INIT_PARCEL_INPUT's -1
// (src/wrapper.f90:24, :1561-1680), which SET_KSP_PARAM turns into
// max(1, -1) = 1 iteration (src/krylov.f90:32, src/type.f90:1331-1336); the
// CA-Arnoldi parameters are kept in the plan's parameters, unused.
int parcel_hip_dcaarnoldi(int icomm, double* vecx, const double* vecb, int n, int gn, int nnz, int istart,
                          const double* crsA, const int* crsRow_ptr, const int* crsCol, int ipreflag,
                          int ILU_method, int addL, int iflagAS, int iovlflag, int precon_thblock, int Neuman_itr,
                          int independ_nvec, int nBlock, int CAARNOLDI_sstep, int CAARNOLDI_tstep,
                          int CAARNOLDI_basis, int CAARNOLDI_QRflag, double* EigenValueReal,
                          double* EigenValueImag, double* EigenVector, double* EigenErrorNorm, int* MaxEigenID,
                          int* MinEigenID, int* iret) {
  (void)icomm; (void)vecb;
  parcel_hip_param prm;
  common_param(prm, PARCEL_SOLVER_CG, ipreflag, ILU_method, addL, iflagAS, iovlflag, precon_thblock, Neuman_itr,
               independ_nvec, nBlock);
  prm.CAARNOLDI_sstep = CAARNOLDI_sstep;
  prm.CAARNOLDI_tstep = CAARNOLDI_tstep;
  prm.CAARNOLDI_basis = CAARNOLDI_basis;
  prm.CAARNOLDI_QRflag = CAARNOLDI_QRflag;
  return eigen_one_shot(prm, vecx, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, nullptr, nullptr, 0, nullptr, 1,
                        EigenValueReal, EigenValueImag, EigenVector, EigenErrorNorm, MaxEigenID, MinEigenID, iret);
}

int parcel_hip_dcaarnoldi_dia(int icomm, double* vecx, const double* vecb, int n, int gn, int istart,
                              const double* diaA, const int* offset, int nnd, int ipreflag, int ILU_method, int addL,
                              int iflagAS, int iovlflag, int precon_thblock, int Neuman_itr, int independ_nvec,
                              int nBlock, int CAARNOLDI_sstep, int CAARNOLDI_tstep, int CAARNOLDI_basis,
                              int CAARNOLDI_QRflag, double* EigenValueReal, double* EigenValueImag,
                              double* EigenVector, double* EigenErrorNorm, int* MaxEigenID, int* MinEigenID,
                              int* iret) {
  (void)icomm; (void)vecb;
  parcel_hip_param prm;
  common_param(prm, PARCEL_SOLVER_CG, ipreflag, ILU_method, addL, iflagAS, iovlflag, precon_thblock, Neuman_itr,
               independ_nvec, nBlock);
  prm.CAARNOLDI_sstep = CAARNOLDI_sstep;
  prm.CAARNOLDI_tstep = CAARNOLDI_tstep;
  prm.CAARNOLDI_basis = CAARNOLDI_basis;
  prm.CAARNOLDI_QRflag = CAARNOLDI_QRflag;
  return eigen_one_shot(prm, vecx, n, gn, 0, istart, nullptr, nullptr, nullptr, diaA, offset, nnd, nullptr, 1,
                        EigenValueReal, EigenValueImag, EigenVector, EigenErrorNorm, MaxEigenID, MinEigenID, iret);
}

int parcel_hip_plan_eigen(parcel_hip_plan* plan, const double* vecx, int power_method_itrmax,
                          double* EigenValueReal, double* EigenValueImag, double* EigenVector,
                          double* EigenErrorNorm, int* MaxEigenID, int* MinEigenID) {
  return eigen_run(plan, vecx, power_method_itrmax, EigenValueReal, EigenValueImag, EigenVector,
                   plan ? plan->P.A.n : 0, EigenErrorNorm, MaxEigenID, MinEigenID);
}

}  // extern "C"

/*
 * parcel_hip.h -- C ABI of the MI355X-native PARCEL Krylov library
 * (libparcel_hip.so).  Plain C types only; every entry point returns an
 * error code (PARCEL_HIP_OK = 0) instead of stopping the process.
 *
 * Array conventions follow the reference API (src/wrapper.f90:1995-2010):
 * CRS arrays are 1-based (crsRow_ptr(n+1) offsets, crsCol global column ids),
 * istart is the 1-based first global row of this rank; with one GPU holding
 * one rank's whole subdomain, istart = 1 and gn = n.  Vector and matrix
 * pointers may be host or device (HBM) pointers: the library inspects each
 * one (the reference GPU build takes device pointers, src/wrapper.f90:569-573,
 * the CPU build host arrays).  reshistory is always a host array of itrmax
 * entries, as in the reference.
 */
#ifndef PARCEL_HIP_H
#define PARCEL_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define PARCEL_HIP_OK 0
#define PARCEL_HIP_ERR_ARG (-1)        /* invalid argument */
#define PARCEL_HIP_ERR_HIP (-2)        /* HIP runtime error (no device, OOM ...) */
#define PARCEL_HIP_ERR_ZERO_DIAG (-3)  /* missing diagonal in a factorised block */
#define PARCEL_HIP_ERR_UNSUPPORTED (-4)
#define PARCEL_HIP_ERR_NOT_INIT (-5)

/* solver ids, src/define.f90:14-22 */
#define PARCEL_SOLVER_CG 1
#define PARCEL_SOLVER_BICG 2
#define PARCEL_SOLVER_GMRES 3
#define PARCEL_SOLVER_CAGMRES 4
#define PARCEL_SOLVER_CBCG 5
/* FP64 iterative refinement around an FP32 PCG (no reference counterpart) */
#define PARCEL_SOLVER_CG_MIXED 11

/* preconditioner ids, src/define.f90:30-35 */
#define PARCEL_PRECON_NONE 0
#define PARCEL_PRECON_POINTJACOBI 1
#define PARCEL_PRECON_BLOCKJACOBI 2
#define PARCEL_PRECON_ADDITIVESCHWARZ 3

/* ILU_method, src/define.f90:45-47 */
#define PARCEL_ILU0 0
#define PARCEL_ILUD 1
#define PARCEL_MLUD 2

/* Gram-Schmidt / QR variants, src/define.f90:56-68 (GMRES_GSflag: 1 MGS,
 * any other value CGS, as src/pgmres.f90:336 tests only for MGS) */
#define PARCEL_QR_MGS 1
#define PARCEL_QR_CGS 2
#define PARCEL_QR_TSQR 3
#define PARCEL_QR_CHOLESKYQR 4
#define PARCEL_QR_CHOLESKYQR2 5

/* return codes of the solvers (iret), src/define.f90:101-102 */
#define PARCEL_RETURNCODE_NORMAL 0
#define PARCEL_RETURNCODE_MAXITERATION 1

/*
 * Solver parameters: the fields of KSP_INPUT_PARAM (src/type.f90:907-946)
 * that the solver wrappers fill (src/wrapper.f90:2040-2070).  A value < 0
 * means "unset, use the default" exactly as INIT_PARCEL_INPUT
 * (src/wrapper.f90:5-44) marks unset fields with -1.
 */
typedef struct parcel_hip_param {
  int solver;
  int ipreflag;
  int ILU_method;
  int addL;
  int iflagAS;
  int iovlflag;
  int precon_thblock; /* block-Jacobi blocks; < 1: one per compute unit */
  int Neuman_itr;
  int independ_nvec;
  int nBlock;
  int GMRES_m;
  int GMRES_GSflag;
  int CAGMRES_sstep;
  int CAGMRES_tstep;
  int CAGMRES_basis;
  int CAGMRES_QRflag;
  int CBCG_kstep;
  int CBCG_Eigenflag;
  int power_method_itrmax;
  int CAARNOLDI_sstep;
  int CAARNOLDI_tstep;
  int CAARNOLDI_basis;
  int CAARNOLDI_QRflag;
  int profile; /* 1: time kernel classes with HIP events during solves */
} parcel_hip_param;

/* Per-plan measurements (filled by setup and by every solve). */
typedef struct parcel_hip_stats {
  double setup_s;          /* host wall time of plan setup (upload + factorisation) */
  double precon_setup_s;   /* part of setup_s spent on the preconditioner */
  double solve_s;          /* host wall time of the last solve (inputs already in HBM) */
  double solve_device_ms;  /* HIP-event time of the last solve on the plan's stream */
  int niter;
  int launches;            /* kernel launches issued by the last solve */
  /* with param.profile = 1: summed HIP-event times of kernel classes */
  double spmv_ms;          /* fused SpMV + CG update kernel */
  int spmv_calls;
  double precon_ms;        /* preconditioner apply */
  int precon_calls;
  double vec_ms;           /* vector update + reduction kernels */
  int vec_calls;
  /* preconditioner structure */
  int bj_blocks;
  int bj_levels_fwd;
  int bj_levels_bwd;
  long bj_entries;         /* stored factor entries (both sweeps, with padding) */
  long bj_nnz_l;           /* off-diagonal entries read by the forward sweep */
  long bj_nnz_u;           /* off-diagonal entries read by the backward sweep */
  /* mixed-precision refinement (PARCEL_SOLVER_CG_MIXED) */
  int inner_niter;         /* FP32 PCG iterations summed over the corrections */
  double final_rel_res;    /* ||b - A x|| / ||b|| after the last correction (FP64) */
  int bj_lscan;            /* 1: the mixed solver's FP32 block-Jacobi apply runs block IC(0) line scans */
  int spmv_stencil;        /* 1: the CG kernel reads the matrix as a symmetric 5/7-point stencil (4 arrays) */
  long sell_entries;       /* stored SELL-64 entries (padding and the long rows' overflow included) */
  long sell_overflow;      /* entries of long rows kept after the slices (0: none) */
  int bj_lines;            /* 1: the FP64 block-Jacobi apply runs as line wavefronts (blocks of whole
                              grid lines with 5-point couplings inside the block) */
} parcel_hip_stats;

/* --- library life cycle: PARCEL_Init / PARCEL_Free (src/c_krylov_wrapper.f90:6-22) */
/* Selects device myrank % ndevices (src/CUDA/src/cu_common.cu:489-518). */
int parcel_hip_init(int myrank);
int parcel_hip_finalize(void);
const char* parcel_hip_last_error(void);
void parcel_hip_default_param(parcel_hip_param* prm);

/* --- one-shot solver entry points, same argument lists and meaning as the
 * reference wrappers minus the MPI communicator handle's semantics (icomm is
 * accepted and ignored: one GPU = one rank).  itrmax is in: maximum
 * iterations, out: iterations done.  Returns an error code; the solver's own
 * status goes to *iret. */

/* parcel_dcg (src/wrapper.f90:544-632 GPU build, :1995-2083 CPU build) */
int parcel_hip_dcg(int icomm, double* vecx, const double* vecb, int n, int gn,
                   int nnz, int istart, const double* crsA, const int* crsRow_ptr,
                   const int* crsCol, int ipreflag, int ILU_method, int addL,
                   int iflagAS, int* itrmax, double rtolmax, double* reshistory,
                   int iovlflag, int precon_thblock, int Neuman_itr,
                   int independ_nvec, int nBlock, int* iret);

/* parcel_dbicgstab (src/wrapper.f90:362 GPU build, :1813 CPU build): same
 * arguments as parcel_dcg; BiCGStab of src/pbicg.f90:280-698. */
int parcel_hip_dbicgstab(int icomm, double* vecx, const double* vecb, int n, int gn,
                         int nnz, int istart, const double* crsA, const int* crsRow_ptr,
                         const int* crsCol, int ipreflag, int ILU_method, int addL,
                         int iflagAS, int* itrmax, double rtolmax, double* reshistory,
                         int iovlflag, int precon_thblock, int Neuman_itr,
                         int independ_nvec, int nBlock, int* iret);

/* parcel_dgmres (src/wrapper.f90:727-818 GPU build, :2179-2270 CPU build):
 * parcel_dcg's arguments plus the restart length GMRES_m and the
 * orthogonalisation GMRES_GSflag; restarted GMRES of src/pgmres.f90:62-525
 * with right preconditioning.  The optional machineflag argument of the
 * reference selects CPU or GPU and has no counterpart here. */
int parcel_hip_dgmres(int icomm, double* vecx, const double* vecb, int n, int gn,
                      int nnz, int istart, const double* crsA, const int* crsRow_ptr,
                      const int* crsCol, int ipreflag, int ILU_method, int addL,
                      int iflagAS, int* itrmax, double rtolmax, double* reshistory,
                      int iovlflag, int precon_thblock, int Neuman_itr,
                      int independ_nvec, int nBlock, int GMRES_m, int GMRES_GSflag,
                      int* iret);

/* parcel_dcagmres (src/wrapper.f90:1355-1454 GPU build, :2807-2906 CPU build):
 * parcel_dcg's arguments plus CAGMRES_sstep (s, 1..15), CAGMRES_tstep (t >= 1,
 * s*t <= 128), CAGMRES_basis (0 monomial, 1 Newton with Leja-ordered Ritz
 * shifts, 2 the same Newton basis with the conjugate-pair step of the
 * Hessenberg's own recurrence) and CAGMRES_QRflag (PARCEL_QR_MGS, _CGS, _TSQR,
 * _CHOLESKYQR, _CHOLESKYQR2).  CA-GMRES of src/pcagmres.f90:1557-2517.
 * Basis 1 is the reference's: for a complex shift pair a +- ib its
 * calc_NewtonBasisVec (:1519-1530) forms V_i = (A - a) V_{i-1} - b^2 V_{i-2}
 * while the Hessenberg (calc_NewtonBasis, :586-611) describes + b^2, so with
 * complex Ritz values the residual estimate drifts from the true residual.
 * Basis 2 (no counterpart in the reference) forms + b^2. */
int parcel_hip_dcagmres(int icomm, double* vecx, const double* vecb, int n, int gn,
                        int nnz, int istart, const double* crsA, const int* crsRow_ptr,
                        const int* crsCol, int ipreflag, int ILU_method, int addL,
                        int iflagAS, int* itrmax, double rtolmax, double* reshistory,
                        int iovlflag, int precon_thblock, int Neuman_itr,
                        int independ_nvec, int nBlock, int CAGMRES_sstep, int CAGMRES_tstep,
                        int CAGMRES_basis, int CAGMRES_QRflag, int* iret);

/* parcel_dcbcg (src/wrapper.f90:1139-1233 CPU build; the GPU build's
 * counterpart has the same arguments): parcel_dcg's arguments plus
 * CBCG_kstep (k, 1..8), CBCG_Eigenflag (1 power method; 2 CA-Arnoldi, which
 * the reference's GPU build runs as the power method, src/pcagmres.f90:
 * 2647-2651, and so does this library; any other value uses Emax = 1,
 * Emin = 0), power_method_itrmax and the
 * CA-Arnoldi parameters CAARNOLDI_sstep/_tstep/_basis/_QRflag.  Chebyshev-
 * basis CG of src/pcbcg.f90:1182-2025: reshistory(j) = ||r|| / ||b|| at the
 * start of outer step j (k iterations each); niter = min(itrmax, j) of the
 * last step (itrmax / k + 1 when no step converged, its history entry 0). */
int parcel_hip_dcbcg(int icomm, double* vecx, const double* vecb, int n, int gn, int nnz,
                     int istart, const double* crsA, const int* crsRow_ptr, const int* crsCol,
                     int ipreflag, int ILU_method, int addL, int iflagAS, int* itrmax,
                     double rtolmax, double* reshistory, int iovlflag, int precon_thblock,
                     int Neuman_itr, int independ_nvec, int nBlock, int CBCG_kstep,
                     int CBCG_Eigenflag, int power_method_itrmax, int CAARNOLDI_sstep,
                     int CAARNOLDI_tstep, int CAARNOLDI_basis, int CAARNOLDI_QRflag, int* iret);

/* parcel_dpower (src/wrapper.f90:913-1022) and parcel_dcaarnoldi (:1561-1680),
 * with their DIA forms (:1025, :1683): power_method_symMatrix
 * (src/power_method.f90:188-437) from the start vector vecx (all ones when it
 * is 0), max(1, power_method_itrmax) iterations of x' = y/||y||,
 * y = A M^-1 x', Emax = <x', y>; parcel_dcaarnoldi runs the same method
 * (KSP_pCAARNOLDI compiles to power_method_symMatrix,
 * src/pcagmres.f90:2647-2651) with one iteration, its power_method_itrmax
 * being left at -1 (src/wrapper.f90:24, src/krylov.f90:32,
 * src/type.f90:1331-1336).  Outputs one eigenpair: EigenValueReal[0] = Emax,
 * EigenValueImag[0] = 0, EigenErrorNorm[0] = ||A x - Emax x|| / ||Emax x||,
 * MaxEigenID = MinEigenID = 1, and the normalised last iterate x: its first
 * component only for the CRS forms (src/krylov.f90:360-369 copies
 * size(EigenValueReal) = 1 entries), all n for the DIA forms
 * (get_EigenPair_KSP, src/EigenProblem.f90:18-46); EigenVector is host or
 * device memory.  vecx and vecb are left unchanged; itrmax (parcel_dpower)
 * and iret, which the reference leaves unset, are returned as 0. */
int parcel_hip_dpower(int icomm, double* vecx, const double* vecb, int n, int gn, int nnz, int istart,
                      const double* crsA, const int* crsRow_ptr, const int* crsCol, int ipreflag, int ILU_method,
                      int addL, int iflagAS, int* itrmax, int iovlflag, int precon_thblock, int Neuman_itr,
                      int independ_nvec, int nBlock, int power_method_itrmax, double* EigenValueReal,
                      double* EigenValueImag, double* EigenVector, double* EigenErrorNorm, int* MaxEigenID,
                      int* MinEigenID, int* iret);
int parcel_hip_dpower_dia(int icomm, double* vecx, const double* vecb, int n, int gn, int istart, const double* diaA,
                          const int* offset, int nnd, int ipreflag, int ILU_method, int addL, int iflagAS,
                          int* itrmax, int iovlflag, int precon_thblock, int Neuman_itr, int independ_nvec,
                          int nBlock, int power_method_itrmax, double* EigenValueReal, double* EigenValueImag,
                          double* EigenVector, double* EigenErrorNorm, int* MaxEigenID, int* MinEigenID, int* iret);
int parcel_hip_dcaarnoldi(int icomm, double* vecx, const double* vecb, int n, int gn, int nnz, int istart,
                          const double* crsA, const int* crsRow_ptr, const int* crsCol, int ipreflag,
                          int ILU_method, int addL, int iflagAS, int iovlflag, int precon_thblock, int Neuman_itr,
                          int independ_nvec, int nBlock, int CAARNOLDI_sstep, int CAARNOLDI_tstep,
                          int CAARNOLDI_basis, int CAARNOLDI_QRflag, double* EigenValueReal,
                          double* EigenValueImag, double* EigenVector, double* EigenErrorNorm, int* MaxEigenID,
                          int* MinEigenID, int* iret);
int parcel_hip_dcaarnoldi_dia(int icomm, double* vecx, const double* vecb, int n, int gn, int istart,
                              const double* diaA, const int* offset, int nnd, int ipreflag, int ILU_method, int addL,
                              int iflagAS, int iovlflag, int precon_thblock, int Neuman_itr, int independ_nvec,
                              int nBlock, int CAARNOLDI_sstep, int CAARNOLDI_tstep, int CAARNOLDI_basis,
                              int CAARNOLDI_QRflag, double* EigenValueReal, double* EigenValueImag,
                              double* EigenVector, double* EigenErrorNorm, int* MaxEigenID, int* MinEigenID,
                              int* iret);

/* The R factor of TSQR_Ronly_ (src/QRDecomposition.f90:887-1002, the
 * CAGMRES_QRflag = 3 path of QRDecomposition_) for the n x nc column-major
 * matrix V (ld ldv >= n; host or device memory; nc <= 16): nc x nc upper
 * triangular with a non-negative diagonal, written to host memory R. */
int parcel_hip_tsqr_r(int n, int nc, const double* V, int ldv, double* R);

/* --- DIA format entry points: the matrix as nnd diagonals of n values
 * (diaA(k + (j-1)*n) is row k of diagonal j, coupling it to global column
 * k + istart - 1 + offset(j)), src/wrapper.f90: parcel_dcg_dia :635-724 (GPU
 * build) / :2086-2176 (CPU build), parcel_dbicgstab_dia :454 / :1905,
 * parcel_dgmres_dia :821 / :2273, parcel_dcbcg_dia :1246 / :2698,
 * parcel_dcagmres_dia :1457 / :2909; the arguments after nnd are the CRS
 * entry points'.  The diagonals are converted to CRS as dia2crs does
 * (src/matrix.f90:7536-7576, zero entries dropped) for the preconditioner
 * set-up, and the fused CG kernel reads the diagonals themselves (MatVec_dia,
 * :2309-2379, diagonals in stored order: the same row sums). */
int parcel_hip_dcg_dia(int icomm, double* vecx, const double* vecb, int n, int gn, int istart, const double* diaA,
                       const int* offset, int nnd, int ipreflag, int ILU_method, int addL, int iflagAS, int* itrmax,
                       double rtolmax, double* reshistory, int iovlflag, int precon_thblock, int Neuman_itr,
                       int independ_nvec, int nBlock, int* iret);
int parcel_hip_dbicgstab_dia(int icomm, double* vecx, const double* vecb, int n, int gn, int istart,
                             const double* diaA, const int* offset, int nnd, int ipreflag, int ILU_method, int addL,
                             int iflagAS, int* itrmax, double rtolmax, double* reshistory, int iovlflag,
                             int precon_thblock, int Neuman_itr, int independ_nvec, int nBlock, int* iret);
int parcel_hip_dgmres_dia(int icomm, double* vecx, const double* vecb, int n, int gn, int istart, const double* diaA,
                          const int* offset, int nnd, int ipreflag, int ILU_method, int addL, int iflagAS,
                          int* itrmax, double rtolmax, double* reshistory, int iovlflag, int precon_thblock,
                          int Neuman_itr, int independ_nvec, int nBlock, int GMRES_m, int GMRES_GSflag, int* iret);
int parcel_hip_dcagmres_dia(int icomm, double* vecx, const double* vecb, int n, int gn, int istart,
                            const double* diaA, const int* offset, int nnd, int ipreflag, int ILU_method, int addL,
                            int iflagAS, int* itrmax, double rtolmax, double* reshistory, int iovlflag,
                            int precon_thblock, int Neuman_itr, int independ_nvec, int nBlock, int CAGMRES_sstep,
                            int CAGMRES_tstep, int CAGMRES_basis, int CAGMRES_QRflag, int* iret);
int parcel_hip_dcbcg_dia(int icomm, double* vecx, const double* vecb, int n, int gn, int istart, const double* diaA,
                         const int* offset, int nnd, int ipreflag, int ILU_method, int addL, int iflagAS, int* itrmax,
                         double rtolmax, double* reshistory, int iovlflag, int precon_thblock, int Neuman_itr,
                         int independ_nvec, int nBlock, int CBCG_kstep, int CBCG_Eigenflag, int power_method_itrmax,
                         int CAARNOLDI_sstep, int CAARNOLDI_tstep, int CAARNOLDI_basis, int CAARNOLDI_QRflag,
                         int* iret);

/* Mixed-precision PCG (no counterpart in the reference; its FP64 PCG is
 * KSP_pCG, src/pcg.f90:224-892): parcel_dcg's arguments plus the inner
 * tolerance and iteration limit.  FP64 refinement x += d, r = b - A x, each
 * d from an FP32 PCG on A d = r (FP32 matrix, vectors and preconditioner
 * factors; dots in double) to inner_rtol.  reshistory(k) = ||b - A x|| / ||b||
 * before correction k; itrmax/niter count corrections. */
int parcel_hip_dcg_mixed(int icomm, double* vecx, const double* vecb, int n, int gn,
                         int nnz, int istart, const double* crsA, const int* crsRow_ptr,
                         const int* crsCol, int ipreflag, int ILU_method, int addL,
                         int iflagAS, int* itrmax, double rtolmax, double* reshistory,
                         int iovlflag, int precon_thblock, int Neuman_itr,
                         int independ_nvec, int nBlock, double inner_rtol, int inner_itrmax,
                         int* iret);

/* parcel_ddcg (src/ddwrapper.f90:184-282): double-double PCG of
 * src/ddpcg.f90:247-563.  Vectors and the matrix as hi / lo pairs;
 * PRECISION_A / _b / _x / _PRECONDITION: 1 double (the lo array is ignored
 * and may be NULL), 2 double-double.  The preconditioner factors are
 * always computed in double-double; precision 1 applies their hi parts.
 * ipreflag 0, 1 or 2 (ILU_method 0, 1, 2). */
int parcel_hip_ddcg(int icomm, double* vecx_hi, double* vecx_lo, const double* vecb_hi,
                    const double* vecb_lo, int n, int gn, int nnz, int istart,
                    const double* crsA_hi, const double* crsA_lo, const int* crsRow_ptr,
                    const int* crsCol, int ipreflag, int ILU_method, int addL, int iflagAS,
                    int* itrmax, double rtolmax, double* reshistory, int iovlflag,
                    int precon_thblock, int independ_nvec, int nBlock, int PRECISION_A,
                    int PRECISION_b, int PRECISION_x, int PRECISION_PRECONDITION, int* iret);

/* parcel_dcg_ddm (src/ddm_kernel.f90:2893-2986 GPU build, :2266-2360 CPU
 * build): PCG on a stencil matrix given as nnd diagonals over the rank's
 * ghost-padded grid (num_grids interior cells per direction, margin ghost
 * cells on each side, nrow = the padded size; ioff_dia the linear offsets,
 * ioff_grids(nnd, ncomp_grids) the per-direction offsets); x, b padded too.
 * ipreflag 0 none, 1 point Jacobi, 2 block Jacobi = D-ILU on boxes of the
 * GPU build (nx/4 x ny/4 x nz/4 boxes, init_DILU_DDM), 3 additive Schwarz
 * (solve_DILU_AS_DDM_GPU, src/src_DDM/precon_ddm_GPU.f90:1573-1868): the same
 * boxes extended by addL >= 0 cells on every side facing another box, D-ILU
 * per extended box, each box's own cells kept (restricted additive Schwarz,
 * what the reference computes at one rank; iflagAS only selects how rank
 * sleeves are averaged).  One GPU holds one rank's whole grid:
 * num_neighbor_ranks must be 0. */
int parcel_hip_dcg_ddm(int icomm, double* x, const double* b, int nrow, int nnd, const int* ioff_dia,
                       const double* val_dia, int num_neighbor_ranks, const int* neighbor_ranks, int margin,
                       int ncomp_grids, const int* ndiv_grids, const int* num_grids, const int* ioff_grids,
                       int div_direc_th, int ipreflag, int addL, int iflagAS, int* itrmax, double rtolmax,
                       double* reshistory, int iovlflag, int precon_thblock, int Neuman_itr, int independ_nvec,
                       int nBlock, int* iret);

/* --- plan API: setup once (upload, factorise), solve many times.  This is the
 * split that KSP_CRS performs internally (src/krylov.f90:447-735: set_KSP_CRS,
 * then START_KSP_SOLVER); exposing it lets a caller keep the matrix resident
 * in HBM across right-hand sides. */
typedef struct parcel_hip_plan parcel_hip_plan;

int parcel_hip_plan_create(parcel_hip_plan** plan, int n, int gn, int nnz, int istart,
                           const double* crsA, const int* crsRow_ptr,
                           const int* crsCol, const parcel_hip_param* prm);
int parcel_hip_plan_solve(parcel_hip_plan* plan, double* vecx, const double* vecb,
                          int* itrmax, double rtolmax, double* reshistory, int* iret);
int parcel_hip_plan_stats(const parcel_hip_plan* plan, parcel_hip_stats* stats);
/* a plan from DIA diagonals (the set-up split of PARCEL_KSP_DIA,
 * src/krylov.f90); arguments as parcel_hip_dcg_dia's */
int parcel_hip_plan_create_dia(parcel_hip_plan** plan, int n, int gn, int istart, const double* diaA,
                               const int* offset, int nnd, const parcel_hip_param* prm);
/* double-double plans (the split of parcel_ddcg): the solver must be CG */
int parcel_hip_plan_create_dd(parcel_hip_plan** plan, int n, int gn, int nnz, int istart,
                              const double* crsA_hi, const double* crsA_lo,
                              const int* crsRow_ptr, const int* crsCol,
                              const parcel_hip_param* prm, int PRECISION_A,
                              int PRECISION_PRECONDITION);
int parcel_hip_plan_solve_dd(parcel_hip_plan* plan, double* vecx_hi, double* vecx_lo,
                             const double* vecb_hi, const double* vecb_lo, int* itrmax,
                             double rtolmax, double* reshistory, int PRECISION_b,
                             int PRECISION_x, int* iret);
/* Inner tolerance and iteration limit of PARCEL_SOLVER_CG_MIXED plans
 * (defaults 1e-4 and 10000). */
int parcel_hip_plan_set_inner(parcel_hip_plan* plan, double inner_rtol, int inner_itrmax);
/* Switch per-kernel HIP-event timing of later solves on (1) or off (0). */
int parcel_hip_plan_set_profile(parcel_hip_plan* plan, int on);
int parcel_hip_plan_destroy(parcel_hip_plan* plan);

/* The power method of parcel_dpower on a created FP64 plan (its matrix and
 * preconditioner), with the whole eigenvector: EigenVector receives all n
 * components (host or device); vecx is host or device.  No counterpart in
 * the reference's API: its CRS driver returns one component. */
int parcel_hip_plan_eigen(parcel_hip_plan* plan, const double* vecx, int power_method_itrmax,
                          double* EigenValueReal, double* EigenValueImag, double* EigenVector,
                          double* EigenErrorNorm, int* MaxEigenID, int* MinEigenID);

/* --- kernels exposed for unit parity tests and measurement --------------- */
/* y = A x with the plan's matrix (src/matrix.f90:9051 MatVec_crs_FP64_CPU;
 * GPU reference: cusparseSpMV, src/CUDA/src/cusparse.cu:632).  Device or host
 * pointers of length n. */
int parcel_hip_plan_spmv(parcel_hip_plan* plan, const double* x, double* y);
/* s = M^-1 r with the plan's preconditioner (src/krylov_h.f90:1684 precon_crs) */
int parcel_hip_plan_precon(parcel_hip_plan* plan, const double* r, double* s);
/* <x,y> in the library's fixed reduction order (src/vec.f90:1027
 * calc_dot_FP64_CPU; GPU reference src/CUDA/src/cu_common.cu:412 cuda_dot). */
int parcel_hip_dot(int n, const double* x, const double* y, double* result);
/* parcel_hip_dot without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  x and y must be
 * device pointers (n > 0).  The kernels keep their partial sums in a fixed
 * device array of the library (2049 doubles, 16 KiB per GPU, held until the
 * process ends), which bounds the partials to 2048 workgroups: every n >= 0
 * fits.  result: a device pointer receives the value without a wait; a host
 * pointer receives it after a synchronisation.  Same bits as parcel_hip_dot.
 * All calls on one GPU share that array: calls that overlap on one GPU (from
 * several host threads, or streams that run concurrently) race. */
int parcel_hip_dot_noalloc(int n, const double* x, const double* y, double* result);
/* Time `reps` launches of the plan's dominant kernel class (1: fused CG SpMV,
 * 2: preconditioner apply) with HIP events on the plan's stream; returns the
 * mean duration in milliseconds. */
int parcel_hip_plan_time_kernel(parcel_hip_plan* plan, int which, int reps, double* mean_ms);

/* --- device memory helpers: the counterpart of cudaMalloc_FP64 /
 * cudaMemcpy_h2d_FP64 / cudaFree of the reference drivers
 * (src/CUDA/src/cu_common.cu:573,850; src/c_common.f90:206), which put the
 * matrix and vectors in device memory before calling the GPU-build wrappers.
 * kind: 1 host->device, 2 device->host, 3 device->device. */
int parcel_hip_malloc(void** ptr, unsigned long bytes);
int parcel_hip_free(void* ptr);
int parcel_hip_memcpy(void* dst, const void* src, unsigned long bytes, int kind);
int parcel_hip_memset(void* ptr, int value, unsigned long bytes);
int parcel_hip_device_synchronize(void);

/* --- batched systems: many matrices with ONE sparsity pattern (no
 * counterpart in the reference, whose solvers take one system per call).
 * Every system runs the arithmetic of the single-system solver
 * (KSP_pCG, src/pcg.f90:224-892) on its own vectors; a launch covers all
 * systems.  Preconditioners:
 *   NONE, JACOBI      as ipreflag 0 / 1 of parcel_dcg;
 *   IC0               incomplete Cholesky on the lower five-point pattern;
 *   FACTOR            M = L L^T from a caller-supplied (or network-predicted)
 *                     lower factor on the same pattern, BF16 values.
 * IC0 and FACTOR need the pattern of a 2D five-point grid in natural order
 * (rows x fastest, grid lines of at most 1024 cells, n < 2^29 per system);
 * the factor solves are line recurrences evaluated as 64-lane scans, a line
 * split over up to four wavefronts. */
#define PARCEL_BATCH_NONE 0
#define PARCEL_BATCH_JACOBI 1
#define PARCEL_BATCH_IC0 2
#define PARCEL_BATCH_FACTOR 3
/* precision of parcel_hip_batch_solve */
#define PARCEL_BATCH_FP64 64       /* FP64 PCG (parcel_dcg's iteration) */
#define PARCEL_BATCH_REFINE32 32   /* FP64 refinement around FP32 PCG */
#define PARCEL_BATCH_APPLY32 48    /* FP64 PCG with the factor apply in FP32 */

typedef struct parcel_hip_batch parcel_hip_batch;

typedef struct {
  double setup_s;          /* create + preconditioner setup (host clock) */
  double solve_s;          /* last solve (host clock) */
  double solve_device_ms;  /* last solve, HIP events on the batch stream */
  int max_niter;           /* largest (outer) iteration count of the last solve */
  int outer_steps;         /* mixed solves: batched FP64 corrections launched */
  long inner_total;        /* mixed solves: FP32 iterations summed over systems */
  int inner_steps;         /* mixed solves: batched FP32 iterations launched */
  int launches;
  double apply_ms;         /* factor-apply kernel time (profile on) */
  int apply_calls;
  double spmv_ms;          /* fused SpMV kernel time (profile on) */
  int spmv_calls;
  int nx, ny;              /* detected five-point grid (0 if none) */
  int stencil5;            /* 1: every system exactly symmetric, the CG SpMV reads c/e/n arrays */
  int fallback_systems;    /* systems of the last solve re-solved by the fallback */
} parcel_hip_batch_stats;

/* nsys systems of order n; rowptr/col: 1-based CSR (host), shared; vals:
 * nsys*nnz values in CSR order, system s at vals + s*nnz (host or device). */
int parcel_hip_batch_create(parcel_hip_batch** batch, int nsys, int n, int nnz, const int* rowptr,
                            const int* col, const double* vals, int precon);
/* Factor of every system for PARCEL_BATCH_FACTOR: bf16 bit patterns, each
 * array nsys*n (host or device): Ls[i] = L(i, i-nx), Lw[i] = L(i, i-1),
 * Ld[i] = L(i, i) > 0.  Entries outside the pattern must be 0. */
int parcel_hip_batch_set_factor(parcel_hip_batch* batch, const unsigned short* Ls, const unsigned short* Lw,
                                const unsigned short* Ld);
/* Solve all systems to ||b - A x|| / ||b|| < rtol.  precision 64: FP64 PCG
 * (per system exactly parcel_dcg's iteration); precision 32: FP64
 * refinement around FP32 PCG with the FP32 preconditioner (as
 * PARCEL_SOLVER_CG_MIXED; itrmax then counts FP64 corrections); precision
 * 48: FP64 PCG whose factor apply (IC0 / FACTOR) runs in FP32 (r rounded to
 * FP32, z widened back).  x, b:
 * nsys*n (host or device), x holds the initial guess.  hist (nsys*itrmax,
 * may be NULL), niter and iret (nsys each, may be NULL) are host arrays. */
int parcel_hip_batch_solve(parcel_hip_batch* batch, double* x, const double* b, int itrmax, double rtol,
                           int precision, double* hist, int* niter, int* iret);
int parcel_hip_batch_set_inner(parcel_hip_batch* batch, double inner_rtol, int inner_itrmax);
/* Factor solves on nstrip independent strips of grid lines per system (lines
 * [ny k / nstrip, ny (k+1) / nstrip) for strip k): M = diag(L_k L_k^T), the
 * factor's couplings between strips dropped; nstrip wavefronts per system.
 * Default 1 (M = L L^T exactly). */
int parcel_hip_batch_set_strips(parcel_hip_batch* batch, int nstrip);
/* Block-Jacobi IC(0) for a PARCEL_BATCH_IC0 batch: every system's factor is
 * recomputed with each of nblocks strips of grid lines (lines
 * [ny k / nblocks, ny (k+1) / nblocks)) factorised on its own, its
 * couplings to the strip below dropped, and the factor solves run per strip
 * (set_strips(nblocks)).  On these symmetric systems this is block-Jacobi
 * ILU(0) (parcel_dcg ipreflag 2, ILU_method 0, src/matrix.f90:8249-8392) in
 * exact arithmetic when nblocks divides ny (the strips are then
 * divide_region's row blocks); the rounding is the factor solves' own. */
int parcel_hip_batch_set_ic0_blocks(parcel_hip_batch* batch, int nblocks);
/* Fallback of the learned path: after a batched solve, every system whose
 * solve ended with iret != 0 is solved again, from x = 0, by parcel_dcg's
 * FP64 PCG (KSP_pCG, src/pcg.f90:224-892) with preconditioner ipreflag
 * (0..2) / ILU_method / precon_thblock and at most itrmax iterations; that
 * solve's x, niter, iret and history (its first itrmax entries of the
 * batch's row) replace the system's.  itrmax <= 0 switches it off.  FACTOR
 * batches start with (2, ILU(0), 32 blocks, 20000), other batches off.  No
 * counterpart in the reference. */
int parcel_hip_batch_set_fallback(parcel_hip_batch* batch, int ipreflag, int ILU_method, int precon_thblock,
                                  int itrmax);
/* per system (nsys ints, host): 1 if the last solve fell back */
int parcel_hip_batch_get_fallback(const parcel_hip_batch* batch, int* fell_back);
int parcel_hip_batch_set_profile(parcel_hip_batch* batch, int on);
int parcel_hip_batch_get_stats(const parcel_hip_batch* batch, parcel_hip_batch_stats* stats);
/* y = A x for every system (x, y: nsys*n, host or device) */
int parcel_hip_batch_spmv(parcel_hip_batch* batch, const double* x, double* y);
/* z = M^-1 r with the preconditioner of the batch (FP64 arrays) */
int parcel_hip_batch_precon_apply(parcel_hip_batch* batch, const double* r, double* z);
/* FP32 apply of the same preconditioner (r, z: nsys*n floats) */
int parcel_hip_batch_precon_apply32(parcel_hip_batch* batch, const float* r, float* z);
/* The BF16 factor of every system (FACTOR batches; nsys*n each, host) */
int parcel_hip_batch_get_factor(const parcel_hip_batch* batch, unsigned short* Ls, unsigned short* Lw,
                                unsigned short* Ld);
int parcel_hip_batch_destroy(parcel_hip_batch* batch);

/* --- learned factor: a message-passing network over the graph of a 2D
 * five-point matrix predicts L on the matrix's lower pattern (M = L L^T,
 * BF16 factor); its layers run on BF16 MFMA with FP32 accumulation.  No
 * counterpart in the reference.  Parameters: FP32, parcel_hip_mpnn_nparams()
 * of them (6 layers of W 32x64 + b 32, head Wo 8x32 + bo 8, row-major). */
typedef struct parcel_hip_mpnn parcel_hip_mpnn;
int parcel_hip_mpnn_nparams(void);
int parcel_hip_mpnn_create(parcel_hip_mpnn** net, const float* params, int nparams);
/* Predict and install the factor of every system of a PARCEL_BATCH_FACTOR batch */
int parcel_hip_mpnn_factor(parcel_hip_mpnn* net, parcel_hip_batch* batch);
/* Network internals for a batch (host arrays, each may be NULL): head
 * outputs o (nsys*n*8), last activation (nsys*n*32 BF16 bits), scaled
 * couplings ahat (nsys*n*4: s, w, e, n) and sqrt(diag) sd (nsys*n). */
int parcel_hip_mpnn_inspect(parcel_hip_mpnn* net, parcel_hip_batch* batch, float* o, unsigned short* hlast,
                            float* ahat, float* sd);
int parcel_hip_mpnn_destroy(parcel_hip_mpnn* net);

/* Training on the device: a pool of `capacity` systems of one grid (n cells,
 * lines of nx), minibatch gradients by the network's backward kernels, Adam. */
typedef struct parcel_hip_mpnn_trainer parcel_hip_mpnn_trainer;
int parcel_hip_mpnn_trainer_create(parcel_hip_mpnn_trainer** trainer, int n, int nx, int capacity,
                                   const float* params);
/* add a batch's systems with solutions x and right-hand sides b (nsys*n) */
int parcel_hip_mpnn_trainer_add(parcel_hip_mpnn_trainer* trainer, parcel_hip_batch* batch, const double* x,
                                const double* b);
/* loss and gradient (nparams floats, host; may be NULL) of pool systems idx[0..S) */
int parcel_hip_mpnn_trainer_grad(parcel_hip_mpnn_trainer* trainer, const int* idx, int S, float* grad,
                                 double* loss);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8) on the minibatch; loss before the step */
int parcel_hip_mpnn_trainer_step(parcel_hip_mpnn_trainer* trainer, const int* idx, int S, float lr, double* loss);
int parcel_hip_mpnn_trainer_params(parcel_hip_mpnn_trainer* trainer, float* params);
int parcel_hip_mpnn_trainer_destroy(parcel_hip_mpnn_trainer* trainer);

/* --- synthetic problem generators (host code): the counterpart of the
 * reference's example generators (src/example/problem.f90) plus the build's
 * own seeded generators.  Output CSR is 1-based with sorted columns. ------ */
/* make_poisson3D_CRS (src/example/problem.f90:448-618): diag -6, off +1,
 * whole domain (one rank). nnz of the result is returned in *nnz; pass NULL
 * arrays to query the size. */
int parcel_gen_poisson3d_parcel(int gnx, int gny, int gnz, int* nnz, int* rowptr,
                                int* col, double* val);
/* make_poisson3D_right_vector (src/example/problem.f90:772-875) */
int parcel_gen_poisson3d_rhs_parcel(int gnx, int gny, int gnz, double* vecb, double* vecx_exact);
/* make_poisson2D_CRS (src/example/problem.f90:210-340) scaled by `scale`
 * (scale -1 gives the SPD 4/-1 operator) */
int parcel_gen_poisson2d(int gnx, int gny, double scale, int* nnz, int* rowptr, int* col,
                         double* val);
/* 3D 7-point constant Poisson, diag 6 off -1 (SPD) */
int parcel_gen_poisson3d_spd(int gnx, int gny, int gnz, int* nnz, int* rowptr, int* col,
                             double* val);
/* make_poisson_3D_DDM (src/example/ddm_setup.f90:133-349) for one rank:
 * 7 diagonals over the (gnx+2m)(gny+2m)(gnz+2m) padded grid (val_dia of
 * 7*padded size, order -z -y -x 0 +x +y +z), ioff_dia (7), ioff_grids (7x3,
 * column-major), b and the exact solution in the padded layout (ghosts 0). */
int parcel_gen_poisson3d_ddm(int gnx, int gny, int gnz, int margin, double* val_dia, int* ioff_dia,
                             int* ioff_grids, double* vecb, double* vecx_exact);
/* x ~ U[-1,1] from the counter-based stream `seed` (SplitMix64, 53-bit) */
int parcel_gen_uniform(long n, unsigned long long seed, double lo, double hi, double* x);
/* 3D 7-point first-order upwind convection-diffusion (finite volumes on cubic
 * cells, divided by D*h: a diffusion face is -1).  The velocity direction
 * (vx, vy, vz) is scaled to cell Peclet number max|v_d| h / (2D) = peclet;
 * with c_d = 2 peclet v_d / max|v| the upwind neighbour gets -(1 + |c_d|),
 * the downwind one -1, the diagonal 6 + sum |c_d| (Dirichlet walls).  Entry
 * order b, s, w, c, e, n, t as make_poisson3D_CRS. */
int parcel_gen_convdiff3d(int gnx, int gny, int gnz, double peclet, double vx, double vy, double vz, int* nnz,
                          int* rowptr, int* col, double* val);
/* kappa ~ log-uniform in [kmin, kmax) from the SplitMix64 stream `seed` */
int parcel_gen_loguniform(long n, unsigned long long seed, double kmin, double kmax, double* kappa);
/* make_poisson3D_adv_CRS (src/example/problem.f90:620-770) with per-cell
 * conductivity kappa (x fastest), one rank holding the whole grid; a ghost
 * cell carries the boundary cell's conductivity.  Pass NULL arrays to query
 * nnz. */
int parcel_gen_diffusion3d(int gnx, int gny, int gnz, const double* kappa, int* nnz, int* rowptr, int* col,
                           double* val);
/* make_poisson2D_matrix_DIA_get_nnd / _from_CRS (src/example/problem.f90:
 * 342-446) for one rank holding all rows: the distinct column offsets of
 * the CSR matrix in ascending order (*nnd of them; pass offset = NULL to
 * query nnd) and diaA (nnd*n, diagonal-major, zero where a row has no entry
 * on the diagonal). */
int parcel_gen_crs_to_dia(int n, const int* rowptr, const int* col, const double* val, int* nnd, int* offset,
                          double* diaA);
/* b = A x in double-double (exact products, dd_add in CRS order) */
int parcel_gen_spmv_dd(int n, const int* rowptr, const int* col, const double* val, const double* x, double* b_hi,
                       double* b_lo);
/* x ~ N(0,1): normal k of stream `seed` is Box-Muller on uniforms 2*(k/2)
 * and 2*(k/2)+1 (cosine for even k, sine for odd k) */
int parcel_gen_normal(long n, unsigned long long seed, double* x);
/* Gaussian random field on an nx*ny torus (powers of two), squared-
 * exponential covariance sigma^2 exp(-d^2 / (2 ell^2)), by circulant
 * embedding in FP64 (white noise: parcel_gen_normal of `seed`) */
int parcel_gen_grf2d(int nx, int ny, double ell, double sigma, unsigned long long seed, double* g);
/* 2D five-point diffusion with per-cell conductivity kappa (nx*ny, x
 * fastest), harmonic-mean faces (make_poisson3D_adv_CRS,
 * src/example/problem.f90:620-770, in 2D with dx = dy), ghost cells carrying
 * the boundary cell's kappa; SPD: off-diagonals -face, diagonal the sum. */
int parcel_gen_diffusion2d(int nx, int ny, const double* kappa, int* nnz, int* rowptr, int* col,
                           double* val);
/* nsys systems kappa = exp(grf(seed0 + s)); shared pattern (rowptr/col may be
 * NULL), values of system s at vals + s*nnz; g_out (nsys*nx*ny) optional.
 * Pass vals = NULL to query nnz. */
int parcel_gen_grf_diffusion2d(int nx, int ny, double ell, double sigma, unsigned long long seed0, int nsys,
                               int* nnz, int* rowptr, int* col, double* vals, double* g_out);

#ifdef __cplusplus
}
#endif
#endif

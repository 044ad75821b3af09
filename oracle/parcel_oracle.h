/*
 * parcel_oracle.h -- CPU oracle for the PARCEL Krylov hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, as the checker and as the CPU baseline.
 *
 * Every routine restates the reference algorithm in the reference's own
 * operation order; each one cites the reference file:line it follows
 * (paths relative to the PARCEL 1.2 source tree).
 *
 * Conventions shared with the reference (src/wrapper.f90:1995-2010):
 *   - CSR is 1-based: rowptr[0..n] holds 1-based offsets, col[] holds 1-based
 *     global column ids; the oracle is single-rank, so istart = 1, gn = n.
 *   - every `a*b + c` of the reference is evaluated as one fused multiply-add
 *     fma(a,b,c); the reference is built with -O3 -march=native
 *     (arch/make_config:17), which contracts these loops the same way.
 *     The oracle itself is compiled with -ffp-contract=off so that no other
 *     contraction happens.
 *
 * Reduction order.  The reference's dots are OpenMP/MPI reductions whose
 * order depends on the thread and rank count (src/vec.f90:1008-1045).  The
 * oracle commits to one of two documented orders, chosen per call:
 *   ORACLE_RED_SEQ   : left-to-right fma accumulation over i = 0..n-1;
 *   ORACLE_RED_TILED : the tiled tree order the HIP kernels use (see
 *                      oracle_dot() for the exact definition), so GPU and
 *                      oracle agree bit for bit.
 */
#ifndef PARCEL_ORACLE_H
#define PARCEL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORACLE_RED_SEQ = 0, ORACLE_RED_TILED = 1 };

/* reduction layout: mode, threads per tile (nt), max number of tiles owners
 * (nbmax). For ORACLE_RED_TILED the number of owner blocks is
 * nb = min(ceil(n/nt), nbmax). */
typedef struct {
  int mode;
  int nt;
  int nbmax;
} oracle_red;

/* --- reductions --------------------------------------------------------- */
double oracle_dot(const oracle_red* red, int n, const double* x, const double* y);
/* Reduce nb partial sums in the tiled order used by the consumers. */
double oracle_reduce_partials(const double* part, int nb);

/* --- block layout (src/common.f90:125-182) ------------------------------ */
/* start/end (1-based, inclusive) of block `rank` of `nproc` over [1, n] */
void oracle_divide_region(int rank, int nproc, int n, int* lo, int* hi);

/* --- SpMV (src/matrix.f90:9051-9110) ------------------------------------ */
void oracle_spmv(int n, const int* rowptr, const int* col, const double* val,
                 const double* x, double* y);

/* --- preconditioners (src/krylov.f90:1278-1350, src/krylov_h.f90:1684) --- */
typedef struct oracle_precon oracle_precon;
/* ipreflag: 0 none, 1 point Jacobi, 2 block Jacobi.
 * ilu_method: 0 ILU(0), 1 D-ILU (ILUD), 2 D-ILU (MLUD).
 * nblocks: number of contiguous row blocks (divide_region), >= 1.
 * Returns NULL on error (zero diagonal inside a block), with *err set. */
oracle_precon* oracle_precon_create(int ipreflag, int ilu_method, int nblocks,
                                    int independ_nvec, int n, const int* rowptr,
                                    const int* col, const double* val, int* err);
/* the same with explicit blocks: block b is 0-based rows [starts[b], starts[b+1]) */
oracle_precon* oracle_precon_create_blocks(int ipreflag, int ilu_method, int nblocks, const int* starts,
                                           int independ_nvec, int n, const int* rowptr,
                                           const int* col, const double* val, int* err);
void oracle_precon_apply(const oracle_precon* p, const double* r, double* s);
void oracle_precon_free(oracle_precon* p);
/* The same preconditioner in FP32: sweeps in float with the factors rounded
 * to float where they are read. */
void oracle_precon_apply32(const oracle_precon* p, const float* r, float* s);
/* Inspect the factor data (for tests). */
int oracle_precon_nblocks(const oracle_precon* p);
const double* oracle_precon_dinv(const oracle_precon* p);

/* --- solvers ----------------------------------------------------------- */
/* KSP_pCG (src/pcg.f90:224-892).  x is the initial guess in, solution out.
 * hist must hold itrmax entries.  Returns 0, or a negative error code. */
int oracle_pcg(int n, const int* rowptr, const int* col, const double* val,
               const double* b, double* x, int ipreflag, int ilu_method,
               int nblocks, int independ_nvec, int itrmax, double rtol,
               const oracle_red* red, double* hist, int* niter, int* iret);
/* KSP_pBiCG (src/pbicg.f90:280-698), BiCGStab with right preconditioning.
 * The 3-term updates of axpy_vec3 are evaluated as
 *   x = fma(c3, pe, fma(alpha, pp, x)),  p = fma(beta, p, fma(-(beta*c3), q, r)),
 *   e = fma(-alpha, q, r),  r = fma(-c3, v, e). */
int oracle_bicgstab(int n, const int* rowptr, const int* col, const double* val,
                    const double* b, double* x, int ipreflag, int ilu_method,
                    int nblocks, int independ_nvec, int itrmax, double rtol,
                    const oracle_red* red, double* hist, int* niter, int* iret);

/* KSP_pGMRES (src/pgmres.f90:62-525), restarted GMRES(m) with right
 * preconditioning.  gsflag 1: modified Gram-Schmidt; any other value:
 * classical Gram-Schmidt (CGS_Orth_FP64_CPU, src/pcagmres.f90:5-51).
 * Givens rotations: src/pcagmres.f90:190-242; solution update: update_xVec,
 * src/pcagmres.f90:1136-1236.  hist must hold max(itrmax, m+1) entries (the
 * reference writes reshistory(itr+inner) past itrmax when itrmax < m+1). */
int oracle_gmres(int n, const int* rowptr, const int* col, const double* val,
                 const double* b, double* x, int ipreflag, int ilu_method,
                 int nblocks, int independ_nvec, int itrmax, double rtol, int m,
                 int gsflag, const oracle_red* red, double* hist, int* niter, int* iret);

/* KSP_pCAGMRES (src/pcagmres.f90:1557-2517), CA-GMRES(s,t) with right
 * preconditioning.  basis: 0 monomial, 1 Newton with Leja-ordered
 * Ritz shifts (the reference's conjugate-pair step, + B1(i-1,i) V_{i-2}),
 * 2 the same with - B1(i-1,i) V_{i-2} (consistent with the Hessenberg).  qrflag: 1 MGS, 2 CGS, 3 TSQR (CholeskyQR2 when n <= max(s+1, 64)),
 * 4 CholeskyQR, 5 CholeskyQR2.  hist must
 * hold itrmax + s entries (the reference writes reshistory(itr+i) up to
 * itrmax + s when a cycle starts at itr = itrmax).  Returns -4 for t != 1
 * or another qrflag. */
int oracle_cagmres(int n, const int* rowptr, const int* col, const double* val,
                   const double* b, double* x, int ipreflag, int ilu_method,
                   int nblocks, int independ_nvec, int itrmax, double rtol, int s,
                   int t, int basis, int qrflag, const oracle_red* red, double* hist,
                   int* niter, int* iret);

/* KSP_pCBCG (src/pcbcg.f90:1182-2025), k-step Chebyshev-basis CG, k in 1..8.
 * eigflag 1 and 2: Emax from the power method (pm_itrmax steps from x0),
 * Emin = 0 (2 is CA-Arnoldi, whose compiled body is the power method);
 * other: Emax = 1, Emin = 0.  hist must hold
 * itrmax / k + 1 entries; hist[j-1] = ||r|| / ||b|| at the start of outer
 * step j; niter = min(itrmax, j_last) (entries never written are 0). */
int oracle_cbcg(int n, const int* rowptr, const int* col, const double* val, const double* b, double* x,
                int ipreflag, int ilu_method, int nblocks, int independ_nvec, int itrmax, double rtol, int k,
                int eigflag, int pm_itrmax, const oracle_red* red, double* hist, int* niter, int* iret,
                double* emax_out, double* emin_out);
/* the same with a prebuilt preconditioner and given bounds */
int oracle_cbcg_with(const oracle_precon* M, int n, const int* rowptr, const int* col, const double* val,
                     const double* b, double* x, int itrmax, double rtol, int k, double Emax, double Emin,
                     const oracle_red* red, double* hist, int* niter, int* iret);
double oracle_power_emax(const oracle_precon* M, int n, const int* rowptr, const int* col, const double* val,
                         const double* x0, int itrmax, const oracle_red* red);
/* power_method_symMatrix in full (src/power_method.f90:188-437), the body of
 * parcel_dpower and parcel_dcaarnoldi: Emax, the eigenpair error
 * ||A x - Emax x|| / ||Emax x|| and the normalised eigenvector x (vec, n). */
void oracle_power_eigen(const oracle_precon* M, int n, const int* rowptr, const int* col, const double* val,
                        const double* x0, int itrmax, const oracle_red* red, double* emax_out, double* errnorm_out,
                        double* vec);
/* TSQR R factor (tsqr.c): R (nc x nc, ld nc, nc <= 16) of V (n x nc, ld ldv)
 * with a non-negative diagonal; oracle_house_r is the second level
 * (House_local_Ronly_omp) on its own; oracle_tsqr_waves the number of
 * first-level factors. */
int oracle_tsqr_r(int n, int nc, const double* V, int ldv, double* R);
void oracle_house_r(int m, int s, double* A, int lda, double* R);
int oracle_tsqr_waves(int n);
/* pseudo-inverse of a symmetric n x n matrix (upper triangle of A read) */
void oracle_sym_pinv(int n, const double* A, int lda, double* P, int ldp);

/* Dense kernels used by CA-GMRES (column-major, 0-based); exposed for tests. */
void oracle_gemm_nn(int m, int n, int k, const double* A, int lda, const double* B, int ldb,
                    double* C, int ldc);
void oracle_trmm_lunn(int n, const double* A, int lda, double* B, int ldb);
void oracle_trinv_upper(int n, double* A, int lda);
int oracle_chol_upper(int n, double* A, int lda);
int oracle_hqr_eigs(int n, double* H, int ldh, double* wr, double* wi);
void oracle_leja(int n, double* re, double* im);

/* Mixed-precision iterative refinement (no reference counterpart): FP64
 * outer loop (PCG-style history: hist[k-1] = ||b - A x_{k-1}|| / ||b|| before
 * correction k; exit after the correction whose new residual is below
 * rtol; niter = corrections, iret = 1 at the limit), each correction an
 * FP32 PCG from 0 on A d = r to inner_rtol with the FP32 preconditioner
 * (dots and scalars in double, in the tiled order).  *inner_total = FP32
 * iterations summed over the corrections. */
int oracle_pcg_mixed(int n, const int* rowptr, const int* col, const double* val, const double* b,
                     double* x, int ipreflag, int ilu_method, int nblocks, int independ_nvec,
                     int itrmax, double rtol, int inner_itrmax, double inner_rtol,
                     const oracle_red* red, double* hist, int* niter, int* iret, int* inner_total);

/* --- batched solver's factor preconditioners (batched.c; no reference
 * counterpart).  M = L L^T on a 2D five-point grid in natural order. ------ */
#define ORACLE_PRECON_FACTOR 9
/* restricted additive Schwarz over overlapping blocks (see oracle_precon_create_ras) */
#define ORACLE_PRECON_RAS 10
/* additive Schwarz of the CRS path (AdditiveSchwarz_crs at one rank, as.c):
 * divide_region blocks extended by addL rows, D-ILU (ilu_method 1 ILUD,
 * else the row-sum form) per extended block, iflagAS 1 BASIC, 2 RESTRICT,
 * 3 INTERPOLATE, 4 NONE; block Jacobi D-ILU when addL == 0 or a block has
 * fewer than 2 addL rows */
#define ORACLE_PRECON_AS 11
oracle_precon* oracle_precon_create_as(int ilu_method, int nblocks, int addL, int iflagAS, int independ_nvec, int n,
                                       const int* rowptr, const int* col, const double* val, int* err);
oracle_precon* oracle_precon_create_ras(int n, int ne, const int* src, const int* dst, const int* rp_e,
                                        const int* col_e, const double* val_e, int nblocks, const int* starts,
                                        int* err);
/* PCG / mixed refinement with a prebuilt preconditioner (any kind) */
int oracle_pcg_with(const oracle_precon* M, int n, const int* rowptr, const int* col, const double* val,
                    const double* b, double* x, int itrmax, double rtol, const oracle_red* red,
                    double* hist, int* niter, int* iret);
int oracle_pcg_mixed_with(const oracle_precon* M, int n, const int* rowptr, const int* col, const double* val,
                          const double* b, double* x, int itrmax, double rtol, int inner_itrmax, double inner_rtol,
                          const oracle_red* red, double* hist, int* niter, int* iret, int* inner_total);
/* IC(0) factor (Ls, Lw, Ld; n each) of a five-point matrix, 1-based CSR */
int oracle_ic0_5pt(int n, int nx, const int* rowptr, const double* val, double* L);
/* block-Jacobi IC(0): each of nstrip strips of grid lines factorised alone */
int oracle_ic0_5pt_blocks(int n, int nx, int nstrip, const int* rowptr, const double* val, double* L);
int oracle_ic0_grid_blocks(int n, int nx, int nstrip, const int* rowptr, const int* col, const double* val,
                           double* L);
/* apply arrays (ls, 1/Ld, a_fwd, a_bwd) from a double (src 0) or bf16
 * (src 1) factor; F64 / F32 may be NULL */
void oracle_derive(int n, int nx, int src, const void* L, double* F64, float* F32);
oracle_precon* oracle_precon_create_factor(int n, int nx, const double* F64, const float* F32);
/* on != 0: FP64 vectors go through the FP32 apply (r rounded, z widened) */
void oracle_precon_factor_apply32(oracle_precon* p, int on);
/* factor solves on nstrip independent strips of grid lines (default 1) */
void oracle_precon_factor_strips(oracle_precon* p, int nstrip);
void oracle_lscan_apply(int nx, int ny, const double* F, const double* r, double* z);
void oracle_lscan_apply32(int nx, int ny, const float* F, const float* r, float* z);
float oracle_bf16_to_f32(unsigned short h);
unsigned short oracle_f32_to_bf16(float f);

/* --- double-double (src/dd_func.f90, src/ddpcg.f90, src/ddmatrix.f90) ---- */
/* op: 0 add, 1 sub, 2 mul, 3 div, 4 sqrt (b unused), elementwise over n */
void oracle_dd_op(int op, int n, const double* ah, const double* al, const double* bh,
                  const double* bl, double* ch, double* cl);
void oracle_dd_dot(const oracle_red* red, int n, const double* xh, const double* xl,
                   const double* yh, const double* yl, double* rh, double* rl);
/* vl may be NULL (matrix without lo part, PRECISION_A = 1) */
void oracle_dd_spmv(int n, const int* rowptr, const int* col, const double* vh,
                    const double* vl, const double* xh, const double* xl, double* yh,
                    double* yl);
typedef struct oracle_ddprecon oracle_ddprecon;
/* prec: PRECISION_PRECONDITION (1: apply with the hi parts only, 2: dd) */
oracle_ddprecon* oracle_ddprecon_create(int ipreflag, int ilu_method, int nblocks,
                                        int independ_nvec, int prec, int n, const int* rowptr,
                                        const int* col, const double* vh, const double* vl,
                                        int* err);
void oracle_ddprecon_apply(const oracle_ddprecon* p, const double* rh, const double* rl,
                           double* sh, double* sl);
void oracle_ddprecon_free(oracle_ddprecon* p);
/* ddKSP_pCG (src/ddpcg.f90:247-563); x (hi, lo) in/out; hist[i] = rnorm_hi/bnorm_hi */
int oracle_ddcg(int n, const int* rowptr, const int* col, const double* vh, const double* vl,
                const double* bh, const double* bl, double* xh, double* xl, int ipreflag,
                int ilu_method, int nblocks, int independ_nvec, int prec, int itrmax,
                double rtol, const oracle_red* red, double* hist, int* niter, int* iret);

/* Wall seconds of the last oracle_pcg call: [0] preconditioner setup,
 * [1] iterations (everything after setup). */
void oracle_last_times(double* t);

#ifdef __cplusplus
}
#endif
#endif

/*
 * parcel_fortran.h -- by-reference entry points for C and Fortran callers of
 * libparcel_hip.so.
 *
 * These keep the names and argument lists of the reference's external
 * Fortran subroutines that C programs call with every argument by reference
 * (src/example/c2fortran_kernel.f90: parcel_init_c, parcel_free_c,
 * parcel_dcg_c :334-432, parcel_dbicgstab_c, parcel_dgmres_c :523-600,
 * parcel_ddcg_c :620, parcel_dcagmres_c :901, parcel_dcbcg_c :1006,
 * parcel_dcg_ddm_c :2964; called as parcel_dcg_c_(...) from
 * src/example/main_direct_c.c:102,233,272,311,514) in the form CPU-build
 * drivers use: vectors and matrices are arrays.  GPU-build drivers pass
 * devicemem handles instead: include/parcel_devicemem.h and
 * libparcel_devicemem.so export the same names for them.
 * `niter` is the maximum iteration count in and the iterations done out.
 * Vectors may be host or device pointers; the CRS arrays may be host or
 * device pointers (they are read once, for the set-up).  Errors that the
 * reference turns into `stop` set *iret to the negative library error code
 * (parcel_hip_last_error() has the message).
 */
#ifndef PARCEL_FORTRAN_H
#define PARCEL_FORTRAN_H

#ifdef __cplusplus
extern "C" {
#endif

/* parcel_init_c / parcel_free_c: PARCEL_Init(myrank) / PARCEL_Free()
 * (src/c_krylov_wrapper.f90:6-22) */
void parcel_init_c_(const int* myrank);
void parcel_free_c_(void);

/* parcel_dcg_c (src/example/c2fortran_kernel.f90:334-432) */
void parcel_dcg_c_(const int* icomm, double* vecx, const double* vecb, const int* n,
                   const int* gn, const int* nnz, const int* istart, const double* crsa,
                   const int* crsrow_ptr, const int* crscol, const int* ipreflag,
                   const int* ilu_method, const int* addl, const int* iflagas, int* niter,
                   const double* rtolmax, double* reshistory, const int* iovlflag,
                   const int* precon_thblock, const int* Neuman_itr,
                   const int* independ_nvec, const int* nBlock, int* iret);

/* parcel_dbicgstab_c (src/example/c2fortran_kernel.f90:434-521) */
void parcel_dbicgstab_c_(const int* icomm, double* vecx, const double* vecb, const int* n,
                         const int* gn, const int* nnz, const int* istart, const double* crsa,
                         const int* crsrow_ptr, const int* crscol, const int* ipreflag,
                         const int* ilu_method, const int* addl, const int* iflagas,
                         int* niter, const double* rtolmax, double* reshistory,
                         const int* iovlflag, const int* precon_thblock,
                         const int* Neuman_itr, const int* independ_nvec,
                         const int* nBlock, int* iret);

/* parcel_dgmres_c (src/example/c2fortran_kernel.f90:523-600): adds the
 * restart length and the Gram-Schmidt flag before iret */
void parcel_dgmres_c_(const int* icomm, double* vecx, const double* vecb, const int* n,
                      const int* gn, const int* nnz, const int* istart, const double* crsa,
                      const int* crsrow_ptr, const int* crscol, const int* ipreflag,
                      const int* ilu_method, const int* addl, const int* iflagas, int* niter,
                      const double* rtolmax, double* reshistory, const int* iovlflag,
                      const int* precon_thblock, const int* Neuman_itr,
                      const int* independ_nvec, const int* nBlock, const int* restart,
                      const int* gmres_gsflag, int* iret);

/* parcel_dcagmres_c (src/example/c2fortran_kernel.f90:901-1004): adds the
 * CA-GMRES s, t, basis and QR flag before iret */
void parcel_dcagmres_c_(const int* icomm, double* vecx, const double* vecb, const int* n,
                        const int* gn, const int* nnz, const int* istart, const double* crsa,
                        const int* crsrow_ptr, const int* crscol, const int* ipreflag,
                        const int* ilu_method, const int* addl, const int* iflagas, int* niter,
                        const double* rtolmax, double* reshistory, const int* iovlflag,
                        const int* precon_thblock, const int* Neuman_itr,
                        const int* independ_nvec, const int* nBlock, const int* cagmres_sstep,
                        const int* cagmres_tstep, const int* cagmres_basis,
                        const int* cagmres_qrflag, int* iret);

/* parcel_dcbcg_c (src/example/c2fortran_kernel.f90:1006-1117): adds the
 * CBCG k, eigenvalue-bound flag, power-method limit and CA-Arnoldi
 * parameters before iret */
void parcel_dcbcg_c_(const int* icomm, double* vecx, const double* vecb, const int* n,
                     const int* gn, const int* nnz, const int* istart, const double* crsa,
                     const int* crsrow_ptr, const int* crscol, const int* ipreflag,
                     const int* ilu_method, const int* addl, const int* iflagas, int* niter,
                     const double* rtolmax, double* reshistory, const int* iovlflag,
                     const int* precon_thblock, const int* Neuman_itr,
                     const int* independ_nvec, const int* nBlock, const int* cbcg_kstep,
                     const int* cbcg_eigenflag, const int* power_method_itrmax,
                     const int* caarnoldi_sstep, const int* caarnoldi_tstep,
                     const int* caarnoldi_basis, const int* caarnoldi_qrflag, int* iret);

/* parcel_dcaarnoldi_c (src/example/c2fortran_kernel.f90:1119-1226): the
 * eigen outputs of parcel_hip_dcaarnoldi (one eigenpair, the first
 * component of the eigenvector) before iret */
void parcel_dcaarnoldi_c_(const int* icomm, double* vecx, const double* vecb, const int* n,
                          const int* gn, const int* nnz, const int* istart, const double* crsa,
                          const int* crsrow_ptr, const int* crscol, const int* ipreflag,
                          const int* ilu_method, const int* addl, const int* iflagas,
                          const int* iovlflag, const int* precon_thblock, const int* Neuman_itr,
                          const int* independ_nvec, const int* nBlock, const int* caarnoldi_sstep,
                          const int* caarnoldi_tstep, const int* caarnoldi_basis,
                          const int* caarnoldi_qrflag, double* EigenValueReal,
                          double* EigenValueImag, double* EigenVector, double* EigenErrorNorm,
                          int* EmaxID, int* EminID, int* iret);

/* parcel_ddcg_c (src/example/c2fortran_kernel.f90:620-708): double-double
 * PCG, hi / lo arrays; Neuman_itr is in its list (not in parcel_ddcg's) and
 * unused */
void parcel_ddcg_c_(const int* icomm, double* vecx, double* vecx_lo, const double* vecb,
                    const double* vecb_lo, const int* n, const int* gn, const int* nnz,
                    const int* istart, const double* crsa, const double* crsa_lo,
                    const int* crsrow_ptr, const int* crscol, const int* ipreflag,
                    const int* ilu_method, const int* addl, const int* iflagas, int* niter,
                    const double* rtolmax, double* reshistory, const int* iovlflag,
                    const int* precon_thblock, const int* Neuman_itr,
                    const int* independ_nvec, const int* nBlock, const int* precision_a,
                    const int* precision_b, const int* precision_x,
                    const int* precision_precondition, int* iret);

/* parcel_dcg_ddm_c (src/example/c2fortran_kernel.f90:2964-3064): the
 * stencil (DIA over a ghost-padded grid) entry point parcel_dcg_ddm */
void parcel_dcg_ddm_c_(const int* icomm, double* vecx, const double* vecb, const int* n,
                       const int* nnd, const int* offset, const double* diaA,
                       const int* num_neighbor_ranks, const int* neighbor_ranks,
                       const int* margin, const int* ndim, const int* npe_dim,
                       const int* n_grids, const int* offset_dim, const int* div_direc_th,
                       const int* ipreflag, const int* addl, const int* iflagas, int* niter,
                       const double* rtolmax, double* reshistory, const int* iovlflag,
                       const int* precon_thblock, const int* Neuman_itr,
                       const int* independ_nvec, const int* nBlock, int* iret);

/* DIA forms (src/example/c2fortran_kernel.f90:1228-2091): the CRS forms'
 * arguments with (n, gn, istart, diaa, offset, nnd) in place of
 * (n, gn, nnz, istart, crsa, crsrow_ptr, crscol); parcel_dcaarnoldi_dia_c
 * has no Neuman_itr. */
void parcel_dcg_dia_c_(const int* icomm, double* vecx, const double* vecb, const int* n, const int* gn,
                       const int* istart, const double* diaa, const int* offset, const int* nnd,
                       const int* ipreflag, const int* ilu_method, const int* addl, const int* iflagas, int* niter,
                       const double* rtolmax, double* reshistory, const int* iovlflag, const int* precon_thblock,
                       const int* Neuman_itr, const int* independ_nvec, const int* nBlock, int* iret);
void parcel_dbicgstab_dia_c_(const int* icomm, double* vecx, const double* vecb, const int* n, const int* gn,
                             const int* istart, const double* diaa, const int* offset, const int* nnd,
                             const int* ipreflag, const int* ilu_method, const int* addl, const int* iflagas,
                             int* niter, const double* rtolmax, double* reshistory, const int* iovlflag,
                             const int* precon_thblock, const int* Neuman_itr, const int* independ_nvec,
                             const int* nBlock, int* iret);
void parcel_dgmres_dia_c_(const int* icomm, double* vecx, const double* vecb, const int* n, const int* gn,
                          const int* istart, const double* diaa, const int* offset, const int* nnd,
                          const int* ipreflag, const int* ilu_method, const int* addl, const int* iflagas,
                          int* niter, const double* rtolmax, double* reshistory, const int* iovlflag,
                          const int* precon_thblock, const int* Neuman_itr, const int* independ_nvec,
                          const int* nBlock, const int* restart, const int* gmres_gsflag, int* iret);
void parcel_dcagmres_dia_c_(const int* icomm, double* vecx, const double* vecb, const int* n, const int* gn,
                            const int* istart, const double* diaa, const int* offset, const int* nnd,
                            const int* ipreflag, const int* ilu_method, const int* addl, const int* iflagas,
                            int* niter, const double* rtolmax, double* reshistory, const int* iovlflag,
                            const int* precon_thblock, const int* Neuman_itr, const int* independ_nvec,
                            const int* nBlock, const int* cagmres_sstep, const int* cagmres_tstep,
                            const int* cagmres_basis, const int* cagmres_qrflag, int* iret);
void parcel_dcbcg_dia_c_(const int* icomm, double* vecx, const double* vecb, const int* n, const int* gn,
                         const int* istart, const double* diaa, const int* offset, const int* nnd,
                         const int* ipreflag, const int* ilu_method, const int* addl, const int* iflagas,
                         int* niter, const double* rtolmax, double* reshistory, const int* iovlflag,
                         const int* precon_thblock, const int* Neuman_itr, const int* independ_nvec,
                         const int* nBlock, const int* cbcg_kstep, const int* cbcg_eigenflag,
                         const int* power_method_itrmax, const int* caarnoldi_sstep, const int* caarnoldi_tstep,
                         const int* caarnoldi_basis, const int* caarnoldi_qrflag, int* iret);
void parcel_dcaarnoldi_dia_c_(const int* icomm, double* vecx, const double* vecb, const int* n, const int* gn,
                              const int* istart, const double* diaa, const int* offset, const int* nnd,
                              const int* ipreflag, const int* ilu_method, const int* addl, const int* iflagas,
                              const int* iovlflag, const int* precon_thblock, const int* independ_nvec,
                              const int* nBlock, const int* caarnoldi_sstep, const int* caarnoldi_tstep,
                              const int* caarnoldi_basis, const int* caarnoldi_qrflag, double* EigenValueReal,
                              double* EigenValueImag, double* EigenVector, double* EigenErrorNorm, int* EmaxID,
                              int* EminID, int* iret);

#ifdef __cplusplus
}
#endif
#endif

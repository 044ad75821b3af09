/*
 * parcel_fortran.h -- by-reference entry points for C and Fortran callers of
 * libparcel_hip.so.
 *
 * These keep the names and argument lists of the reference's external
 * Fortran subroutines that C programs call with every argument by reference
 * (src/example/c2fortran_kernel.f90: parcel_init_c, parcel_free_c,
 * parcel_dcg_c :334-432, parcel_dbicgstab_c, parcel_dgmres_c :523-600;
 * called as parcel_dcg_c_(...) from src/example/main_direct_c.c:102,233,272,311,514).
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

#ifdef __cplusplus
}
#endif
#endif

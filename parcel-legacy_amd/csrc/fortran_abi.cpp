// fortran_abi.cpp -- by-reference entry points (include/parcel_fortran.h):
// the argument-by-reference convention of the reference's external Fortran
// subroutines called from C (src/example/c2fortran_kernel.f90), forwarded to
// the by-value C ABI.
#include "parcel_fortran.h"
#include "parcel_hip.h"

extern "C" {

void parcel_init_c_(const int* myrank) { (void)parcel_hip_init(myrank ? *myrank : 0); }

void parcel_free_c_(void) { (void)parcel_hip_finalize(); }

void parcel_dcg_c_(const int* icomm, double* vecx, const double* vecb, const int* n, const int* gn,
                   const int* nnz, const int* istart, const double* crsa, const int* crsrow_ptr,
                   const int* crscol, const int* ipreflag, const int* ilu_method, const int* addl,
                   const int* iflagas, int* niter, const double* rtolmax, double* reshistory,
                   const int* iovlflag, const int* precon_thblock, const int* Neuman_itr,
                   const int* independ_nvec, const int* nBlock, int* iret) {
  int rc = parcel_hip_dcg(*icomm, vecx, vecb, *n, *gn, *nnz, *istart, crsa, crsrow_ptr, crscol, *ipreflag,
                          *ilu_method, *addl, *iflagas, niter, *rtolmax, reshistory, *iovlflag, *precon_thblock,
                          *Neuman_itr, *independ_nvec, *nBlock, iret);
  if (rc) *iret = rc;
}

void parcel_dbicgstab_c_(const int* icomm, double* vecx, const double* vecb, const int* n, const int* gn,
                         const int* nnz, const int* istart, const double* crsa, const int* crsrow_ptr,
                         const int* crscol, const int* ipreflag, const int* ilu_method, const int* addl,
                         const int* iflagas, int* niter, const double* rtolmax, double* reshistory,
                         const int* iovlflag, const int* precon_thblock, const int* Neuman_itr,
                         const int* independ_nvec, const int* nBlock, int* iret) {
  int rc = parcel_hip_dbicgstab(*icomm, vecx, vecb, *n, *gn, *nnz, *istart, crsa, crsrow_ptr, crscol, *ipreflag,
                                *ilu_method, *addl, *iflagas, niter, *rtolmax, reshistory, *iovlflag,
                                *precon_thblock, *Neuman_itr, *independ_nvec, *nBlock, iret);
  if (rc) *iret = rc;
}

void parcel_dgmres_c_(const int* icomm, double* vecx, const double* vecb, const int* n, const int* gn,
                      const int* nnz, const int* istart, const double* crsa, const int* crsrow_ptr,
                      const int* crscol, const int* ipreflag, const int* ilu_method, const int* addl,
                      const int* iflagas, int* niter, const double* rtolmax, double* reshistory,
                      const int* iovlflag, const int* precon_thblock, const int* Neuman_itr,
                      const int* independ_nvec, const int* nBlock, const int* restart,
                      const int* gmres_gsflag, int* iret) {
  int rc = parcel_hip_dgmres(*icomm, vecx, vecb, *n, *gn, *nnz, *istart, crsa, crsrow_ptr, crscol, *ipreflag,
                             *ilu_method, *addl, *iflagas, niter, *rtolmax, reshistory, *iovlflag,
                             *precon_thblock, *Neuman_itr, *independ_nvec, *nBlock, *restart, *gmres_gsflag,
                             iret);
  if (rc) *iret = rc;
}

}  // extern "C"

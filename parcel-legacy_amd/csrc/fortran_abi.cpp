// fortran_abi.cpp -- by-reference entry points (include/parcel_fortran.h):
// the argument-by-reference convention of the reference's external Fortran
// subroutines called from C (src/example/c2fortran_kernel.f90) as CPU-build
// drivers use them (vectors and matrices passed as arrays), forwarded to the
// by-value C ABI.  devicemem_abi.cpp has the GPU-build convention.
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

void parcel_dcagmres_c_(const int* icomm, double* vecx, const double* vecb, const int* n, const int* gn,
                        const int* nnz, const int* istart, const double* crsa, const int* crsrow_ptr,
                        const int* crscol, const int* ipreflag, const int* ilu_method, const int* addl,
                        const int* iflagas, int* niter, const double* rtolmax, double* reshistory,
                        const int* iovlflag, const int* precon_thblock, const int* Neuman_itr,
                        const int* independ_nvec, const int* nBlock, const int* cagmres_sstep,
                        const int* cagmres_tstep, const int* cagmres_basis, const int* cagmres_qrflag, int* iret) {
  int rc = parcel_hip_dcagmres(*icomm, vecx, vecb, *n, *gn, *nnz, *istart, crsa, crsrow_ptr, crscol, *ipreflag,
                               *ilu_method, *addl, *iflagas, niter, *rtolmax, reshistory, *iovlflag,
                               *precon_thblock, *Neuman_itr, *independ_nvec, *nBlock, *cagmres_sstep,
                               *cagmres_tstep, *cagmres_basis, *cagmres_qrflag, iret);
  if (rc) *iret = rc;
}

void parcel_dcbcg_c_(const int* icomm, double* vecx, const double* vecb, const int* n, const int* gn,
                     const int* nnz, const int* istart, const double* crsa, const int* crsrow_ptr,
                     const int* crscol, const int* ipreflag, const int* ilu_method, const int* addl,
                     const int* iflagas, int* niter, const double* rtolmax, double* reshistory,
                     const int* iovlflag, const int* precon_thblock, const int* Neuman_itr,
                     const int* independ_nvec, const int* nBlock, const int* cbcg_kstep,
                     const int* cbcg_eigenflag, const int* power_method_itrmax, const int* caarnoldi_sstep,
                     const int* caarnoldi_tstep, const int* caarnoldi_basis, const int* caarnoldi_qrflag,
                     int* iret) {
  int rc = parcel_hip_dcbcg(*icomm, vecx, vecb, *n, *gn, *nnz, *istart, crsa, crsrow_ptr, crscol, *ipreflag,
                            *ilu_method, *addl, *iflagas, niter, *rtolmax, reshistory, *iovlflag, *precon_thblock,
                            *Neuman_itr, *independ_nvec, *nBlock, *cbcg_kstep, *cbcg_eigenflag,
                            *power_method_itrmax, *caarnoldi_sstep, *caarnoldi_tstep, *caarnoldi_basis,
                            *caarnoldi_qrflag, iret);
  if (rc) *iret = rc;
}

// parcel_dcaarnoldi_c (src/example/c2fortran_kernel.f90:1119-1226)
void parcel_dcaarnoldi_c_(const int* icomm, double* vecx, const double* vecb, const int* n, const int* gn,
                          const int* nnz, const int* istart, const double* crsa, const int* crsrow_ptr,
                          const int* crscol, const int* ipreflag, const int* ilu_method, const int* addl,
                          const int* iflagas, const int* iovlflag, const int* precon_thblock,
                          const int* Neuman_itr, const int* independ_nvec, const int* nBlock,
                          const int* caarnoldi_sstep, const int* caarnoldi_tstep, const int* caarnoldi_basis,
                          const int* caarnoldi_qrflag, double* EigenValueReal, double* EigenValueImag,
                          double* EigenVector, double* EigenErrorNorm, int* EmaxID, int* EminID, int* iret) {
  int rc = parcel_hip_dcaarnoldi(*icomm, vecx, vecb, *n, *gn, *nnz, *istart, crsa, crsrow_ptr, crscol, *ipreflag,
                                 *ilu_method, *addl, *iflagas, *iovlflag, *precon_thblock, *Neuman_itr,
                                 *independ_nvec, *nBlock, *caarnoldi_sstep, *caarnoldi_tstep, *caarnoldi_basis,
                                 *caarnoldi_qrflag, EigenValueReal, EigenValueImag, EigenVector, EigenErrorNorm,
                                 EmaxID, EminID, iret);
  if (rc) *iret = rc;
}

// parcel_ddcg_c takes Neuman_itr after precon_thblock although parcel_ddcg
// has no such argument; it is accepted and unused.
void parcel_ddcg_c_(const int* icomm, double* vecx, double* vecx_lo, const double* vecb, const double* vecb_lo,
                    const int* n, const int* gn, const int* nnz, const int* istart, const double* crsa,
                    const double* crsa_lo, const int* crsrow_ptr, const int* crscol, const int* ipreflag,
                    const int* ilu_method, const int* addl, const int* iflagas, int* niter, const double* rtolmax,
                    double* reshistory, const int* iovlflag, const int* precon_thblock, const int* Neuman_itr,
                    const int* independ_nvec, const int* nBlock, const int* precision_a, const int* precision_b,
                    const int* precision_x, const int* precision_precondition, int* iret) {
  (void)Neuman_itr;
  int rc = parcel_hip_ddcg(*icomm, vecx, vecx_lo, vecb, vecb_lo, *n, *gn, *nnz, *istart, crsa, crsa_lo, crsrow_ptr,
                           crscol, *ipreflag, *ilu_method, *addl, *iflagas, niter, *rtolmax, reshistory, *iovlflag,
                           *precon_thblock, *independ_nvec, *nBlock, *precision_a, *precision_b, *precision_x,
                           *precision_precondition, iret);
  if (rc) *iret = rc;
}

void parcel_dcg_ddm_c_(const int* icomm, double* vecx, const double* vecb, const int* n, const int* nnd,
                       const int* offset, const double* diaA, const int* num_neighbor_ranks, const int* neighbor_ranks,
                       const int* margin, const int* ndim, const int* npe_dim, const int* n_grids,
                       const int* offset_dim, const int* div_direc_th, const int* ipreflag, const int* addl,
                       const int* iflagas, int* niter, const double* rtolmax, double* reshistory,
                       const int* iovlflag, const int* precon_thblock, const int* Neuman_itr,
                       const int* independ_nvec, const int* nBlock, int* iret) {
  int rc = parcel_hip_dcg_ddm(*icomm, vecx, vecb, *n, *nnd, offset, diaA, *num_neighbor_ranks, neighbor_ranks,
                              *margin, *ndim, npe_dim, n_grids, offset_dim, *div_direc_th, *ipreflag, *addl, *iflagas,
                              niter, *rtolmax, reshistory, *iovlflag, *precon_thblock, *Neuman_itr, *independ_nvec,
                              *nBlock, iret);
  if (rc) *iret = rc;
}

// ---- DIA forms (src/example/c2fortran_kernel.f90:1228-2091): the CRS
// entry points' arguments with (n, gn, istart, diaA, offset, nnd) in place of
// (n, gn, nnz, istart, crsa, crsrow_ptr, crscol)

void parcel_dcg_dia_c_(const int* icomm, double* vecx, const double* vecb, const int* n, const int* gn,
                       const int* istart, const double* diaa, const int* offset, const int* nnd,
                       const int* ipreflag, const int* ilu_method, const int* addl, const int* iflagas, int* niter,
                       const double* rtolmax, double* reshistory, const int* iovlflag, const int* precon_thblock,
                       const int* Neuman_itr, const int* independ_nvec, const int* nBlock, int* iret) {
  int rc = parcel_hip_dcg_dia(*icomm, vecx, vecb, *n, *gn, *istart, diaa, offset, *nnd, *ipreflag, *ilu_method,
                              *addl, *iflagas, niter, *rtolmax, reshistory, *iovlflag, *precon_thblock, *Neuman_itr,
                              *independ_nvec, *nBlock, iret);
  if (rc) *iret = rc;
}

void parcel_dbicgstab_dia_c_(const int* icomm, double* vecx, const double* vecb, const int* n, const int* gn,
                             const int* istart, const double* diaa, const int* offset, const int* nnd,
                             const int* ipreflag, const int* ilu_method, const int* addl, const int* iflagas,
                             int* niter, const double* rtolmax, double* reshistory, const int* iovlflag,
                             const int* precon_thblock, const int* Neuman_itr, const int* independ_nvec,
                             const int* nBlock, int* iret) {
  int rc = parcel_hip_dbicgstab_dia(*icomm, vecx, vecb, *n, *gn, *istart, diaa, offset, *nnd, *ipreflag,
                                    *ilu_method, *addl, *iflagas, niter, *rtolmax, reshistory, *iovlflag,
                                    *precon_thblock, *Neuman_itr, *independ_nvec, *nBlock, iret);
  if (rc) *iret = rc;
}

void parcel_dgmres_dia_c_(const int* icomm, double* vecx, const double* vecb, const int* n, const int* gn,
                          const int* istart, const double* diaa, const int* offset, const int* nnd,
                          const int* ipreflag, const int* ilu_method, const int* addl, const int* iflagas,
                          int* niter, const double* rtolmax, double* reshistory, const int* iovlflag,
                          const int* precon_thblock, const int* Neuman_itr, const int* independ_nvec,
                          const int* nBlock, const int* restart, const int* gmres_gsflag, int* iret) {
  int rc = parcel_hip_dgmres_dia(*icomm, vecx, vecb, *n, *gn, *istart, diaa, offset, *nnd, *ipreflag, *ilu_method,
                                 *addl, *iflagas, niter, *rtolmax, reshistory, *iovlflag, *precon_thblock,
                                 *Neuman_itr, *independ_nvec, *nBlock, *restart, *gmres_gsflag, iret);
  if (rc) *iret = rc;
}

void parcel_dcagmres_dia_c_(const int* icomm, double* vecx, const double* vecb, const int* n, const int* gn,
                            const int* istart, const double* diaa, const int* offset, const int* nnd,
                            const int* ipreflag, const int* ilu_method, const int* addl, const int* iflagas,
                            int* niter, const double* rtolmax, double* reshistory, const int* iovlflag,
                            const int* precon_thblock, const int* Neuman_itr, const int* independ_nvec,
                            const int* nBlock, const int* cagmres_sstep, const int* cagmres_tstep,
                            const int* cagmres_basis, const int* cagmres_qrflag, int* iret) {
  int rc = parcel_hip_dcagmres_dia(*icomm, vecx, vecb, *n, *gn, *istart, diaa, offset, *nnd, *ipreflag,
                                   *ilu_method, *addl, *iflagas, niter, *rtolmax, reshistory, *iovlflag,
                                   *precon_thblock, *Neuman_itr, *independ_nvec, *nBlock, *cagmres_sstep,
                                   *cagmres_tstep, *cagmres_basis, *cagmres_qrflag, iret);
  if (rc) *iret = rc;
}

void parcel_dcbcg_dia_c_(const int* icomm, double* vecx, const double* vecb, const int* n, const int* gn,
                         const int* istart, const double* diaa, const int* offset, const int* nnd,
                         const int* ipreflag, const int* ilu_method, const int* addl, const int* iflagas,
                         int* niter, const double* rtolmax, double* reshistory, const int* iovlflag,
                         const int* precon_thblock, const int* Neuman_itr, const int* independ_nvec,
                         const int* nBlock, const int* cbcg_kstep, const int* cbcg_eigenflag,
                         const int* power_method_itrmax, const int* caarnoldi_sstep, const int* caarnoldi_tstep,
                         const int* caarnoldi_basis, const int* caarnoldi_qrflag, int* iret) {
  int rc = parcel_hip_dcbcg_dia(*icomm, vecx, vecb, *n, *gn, *istart, diaa, offset, *nnd, *ipreflag, *ilu_method,
                                *addl, *iflagas, niter, *rtolmax, reshistory, *iovlflag, *precon_thblock,
                                *Neuman_itr, *independ_nvec, *nBlock, *cbcg_kstep, *cbcg_eigenflag,
                                *power_method_itrmax, *caarnoldi_sstep, *caarnoldi_tstep, *caarnoldi_basis,
                                *caarnoldi_qrflag, iret);
  if (rc) *iret = rc;
}

// parcel_dcaarnoldi_dia_c has no Neuman_itr argument (it forwards an
// undeclared local to parcel_dcaarnoldi_dia); the power method that runs
// never reads it, so 1 is passed.
void parcel_dcaarnoldi_dia_c_(const int* icomm, double* vecx, const double* vecb, const int* n, const int* gn,
                              const int* istart, const double* diaa, const int* offset, const int* nnd,
                              const int* ipreflag, const int* ilu_method, const int* addl, const int* iflagas,
                              const int* iovlflag, const int* precon_thblock, const int* independ_nvec,
                              const int* nBlock, const int* caarnoldi_sstep, const int* caarnoldi_tstep,
                              const int* caarnoldi_basis, const int* caarnoldi_qrflag, double* EigenValueReal,
                              double* EigenValueImag, double* EigenVector, double* EigenErrorNorm, int* EmaxID,
                              int* EminID, int* iret) {
  int rc = parcel_hip_dcaarnoldi_dia(*icomm, vecx, vecb, *n, *gn, *istart, diaa, offset, *nnd, *ipreflag,
                                     *ilu_method, *addl, *iflagas, *iovlflag, *precon_thblock, 1, *independ_nvec,
                                     *nBlock, *caarnoldi_sstep, *caarnoldi_tstep, *caarnoldi_basis,
                                     *caarnoldi_qrflag, EigenValueReal, EigenValueImag, EigenVector,
                                     EigenErrorNorm, EmaxID, EminID, iret);
  if (rc) *iret = rc;
}

}  // extern "C"

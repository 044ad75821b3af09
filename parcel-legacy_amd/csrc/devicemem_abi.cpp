// devicemem_abi.cpp -- libparcel_devicemem.so: the by-reference entry points
// (include/parcel_devicemem.h) in the convention of the reference's
// GPU-build C drivers, where every vector and matrix argument is the address
// of a devicemem handle whose first member is the device array
// (src/c_common.f90:7-13, src/example/main_direct_c.c:206-243).  Each
// function reads the arrays out of the handles and forwards to the by-value
// C ABI of libparcel_hip.so.
#include "parcel_devicemem.h"
#include "parcel_hip.h"

namespace {

inline double* dptr(parcel_devicemem* h) { return h ? static_cast<double*>(h->mem) : nullptr; }
inline const double* dptr(const parcel_devicemem* h) { return h ? static_cast<const double*>(h->mem) : nullptr; }
inline const int* iptr(const parcel_devicemem* h) { return h ? static_cast<const int*>(h->mem) : nullptr; }

}  // namespace

extern "C" {

void parcel_dcg_c_(const int* icomm, parcel_devicemem* vecx, const parcel_devicemem* vecb, const int* n,
                   const int* gn, const int* nnz, const int* istart, const parcel_devicemem* crsa,
                   const parcel_devicemem* crsrow_ptr, const parcel_devicemem* crscol, const int* ipreflag,
                   const int* ilu_method, const int* addl, const int* iflagas, int* niter, const double* rtolmax,
                   double* reshistory, const int* iovlflag, const int* precon_thblock, const int* Neuman_itr,
                   const int* independ_nvec, const int* nBlock, int* iret) {
  int rc = parcel_hip_dcg(*icomm, dptr(vecx), dptr(vecb), *n, *gn, *nnz, *istart, dptr(crsa), iptr(crsrow_ptr),
                          iptr(crscol), *ipreflag, *ilu_method, *addl, *iflagas, niter, *rtolmax, reshistory,
                          *iovlflag, *precon_thblock, *Neuman_itr, *independ_nvec, *nBlock, iret);
  if (rc) *iret = rc;
}

void parcel_dbicgstab_c_(const int* icomm, parcel_devicemem* vecx, const parcel_devicemem* vecb, const int* n,
                         const int* gn, const int* nnz, const int* istart, const parcel_devicemem* crsa,
                         const parcel_devicemem* crsrow_ptr, const parcel_devicemem* crscol, const int* ipreflag,
                         const int* ilu_method, const int* addl, const int* iflagas, int* niter,
                         const double* rtolmax, double* reshistory, const int* iovlflag, const int* precon_thblock,
                         const int* Neuman_itr, const int* independ_nvec, const int* nBlock, int* iret) {
  int rc = parcel_hip_dbicgstab(*icomm, dptr(vecx), dptr(vecb), *n, *gn, *nnz, *istart, dptr(crsa),
                                iptr(crsrow_ptr), iptr(crscol), *ipreflag, *ilu_method, *addl, *iflagas, niter,
                                *rtolmax, reshistory, *iovlflag, *precon_thblock, *Neuman_itr, *independ_nvec,
                                *nBlock, iret);
  if (rc) *iret = rc;
}

void parcel_dgmres_c_(const int* icomm, parcel_devicemem* vecx, const parcel_devicemem* vecb, const int* n,
                      const int* gn, const int* nnz, const int* istart, const parcel_devicemem* crsa,
                      const parcel_devicemem* crsrow_ptr, const parcel_devicemem* crscol, const int* ipreflag,
                      const int* ilu_method, const int* addl, const int* iflagas, int* niter,
                      const double* rtolmax, double* reshistory, const int* iovlflag, const int* precon_thblock,
                      const int* Neuman_itr, const int* independ_nvec, const int* nBlock, const int* restart,
                      const int* gmres_gsflag, int* iret) {
  int rc = parcel_hip_dgmres(*icomm, dptr(vecx), dptr(vecb), *n, *gn, *nnz, *istart, dptr(crsa), iptr(crsrow_ptr),
                             iptr(crscol), *ipreflag, *ilu_method, *addl, *iflagas, niter, *rtolmax, reshistory,
                             *iovlflag, *precon_thblock, *Neuman_itr, *independ_nvec, *nBlock, *restart,
                             *gmres_gsflag, iret);
  if (rc) *iret = rc;
}

void parcel_dcagmres_c_(const int* icomm, parcel_devicemem* vecx, const parcel_devicemem* vecb, const int* n,
                        const int* gn, const int* nnz, const int* istart, const parcel_devicemem* crsa,
                        const parcel_devicemem* crsrow_ptr, const parcel_devicemem* crscol, const int* ipreflag,
                        const int* ilu_method, const int* addl, const int* iflagas, int* niter,
                        const double* rtolmax, double* reshistory, const int* iovlflag, const int* precon_thblock,
                        const int* Neuman_itr, const int* independ_nvec, const int* nBlock,
                        const int* cagmres_sstep, const int* cagmres_tstep, const int* cagmres_basis,
                        const int* cagmres_qrflag, int* iret) {
  int rc = parcel_hip_dcagmres(*icomm, dptr(vecx), dptr(vecb), *n, *gn, *nnz, *istart, dptr(crsa),
                               iptr(crsrow_ptr), iptr(crscol), *ipreflag, *ilu_method, *addl, *iflagas, niter,
                               *rtolmax, reshistory, *iovlflag, *precon_thblock, *Neuman_itr, *independ_nvec,
                               *nBlock, *cagmres_sstep, *cagmres_tstep, *cagmres_basis, *cagmres_qrflag, iret);
  if (rc) *iret = rc;
}

void parcel_dcbcg_c_(const int* icomm, parcel_devicemem* vecx, const parcel_devicemem* vecb, const int* n,
                     const int* gn, const int* nnz, const int* istart, const parcel_devicemem* crsa,
                     const parcel_devicemem* crsrow_ptr, const parcel_devicemem* crscol, const int* ipreflag,
                     const int* ilu_method, const int* addl, const int* iflagas, int* niter, const double* rtolmax,
                     double* reshistory, const int* iovlflag, const int* precon_thblock, const int* Neuman_itr,
                     const int* independ_nvec, const int* nBlock, const int* cbcg_kstep,
                     const int* cbcg_eigenflag, const int* power_method_itrmax, const int* caarnoldi_sstep,
                     const int* caarnoldi_tstep, const int* caarnoldi_basis, const int* caarnoldi_qrflag,
                     int* iret) {
  int rc = parcel_hip_dcbcg(*icomm, dptr(vecx), dptr(vecb), *n, *gn, *nnz, *istart, dptr(crsa), iptr(crsrow_ptr),
                            iptr(crscol), *ipreflag, *ilu_method, *addl, *iflagas, niter, *rtolmax, reshistory,
                            *iovlflag, *precon_thblock, *Neuman_itr, *independ_nvec, *nBlock, *cbcg_kstep,
                            *cbcg_eigenflag, *power_method_itrmax, *caarnoldi_sstep, *caarnoldi_tstep,
                            *caarnoldi_basis, *caarnoldi_qrflag, iret);
  if (rc) *iret = rc;
}

void parcel_ddcg_c_(const int* icomm, parcel_devicemem* vecx, parcel_devicemem* vecx_lo,
                    const parcel_devicemem* vecb, const parcel_devicemem* vecb_lo, const int* n, const int* gn,
                    const int* nnz, const int* istart, const parcel_devicemem* crsa, const parcel_devicemem* crsa_lo,
                    const parcel_devicemem* crsrow_ptr, const parcel_devicemem* crscol, const int* ipreflag,
                    const int* ilu_method, const int* addl, const int* iflagas, int* niter, const double* rtolmax,
                    double* reshistory, const int* iovlflag, const int* precon_thblock, const int* Neuman_itr,
                    const int* independ_nvec, const int* nBlock, const int* precision_a, const int* precision_b,
                    const int* precision_x, const int* precision_precondition, int* iret) {
  (void)Neuman_itr;
  int rc = parcel_hip_ddcg(*icomm, dptr(vecx), dptr(vecx_lo), dptr(vecb), dptr(vecb_lo), *n, *gn, *nnz, *istart,
                           dptr(crsa), dptr(crsa_lo), iptr(crsrow_ptr), iptr(crscol), *ipreflag, *ilu_method, *addl,
                           *iflagas, niter, *rtolmax, reshistory, *iovlflag, *precon_thblock, *independ_nvec,
                           *nBlock, *precision_a, *precision_b, *precision_x, *precision_precondition, iret);
  if (rc) *iret = rc;
}

void parcel_dcg_ddm_c_(const int* icomm, parcel_devicemem* vecx, const parcel_devicemem* vecb, const int* n,
                       const int* nnd, const parcel_devicemem* offset, const parcel_devicemem* diaA,
                       const int* num_neighbor_ranks, const int* neighbor_ranks, const int* margin, const int* ndim,
                       const int* npe_dim, const int* n_grids, const int* offset_dim, const int* div_direc_th,
                       const int* ipreflag, const int* addl, const int* iflagas, int* niter, const double* rtolmax,
                       double* reshistory, const int* iovlflag, const int* precon_thblock, const int* Neuman_itr,
                       const int* independ_nvec, const int* nBlock, int* iret) {
  int rc = parcel_hip_dcg_ddm(*icomm, dptr(vecx), dptr(vecb), *n, *nnd, iptr(offset), dptr(diaA),
                              *num_neighbor_ranks, neighbor_ranks, *margin, *ndim, npe_dim, n_grids, offset_dim,
                              *div_direc_th, *ipreflag, *addl, *iflagas, niter, *rtolmax, reshistory, *iovlflag,
                              *precon_thblock, *Neuman_itr, *independ_nvec, *nBlock, iret);
  if (rc) *iret = rc;
}

}  // extern "C"

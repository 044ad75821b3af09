/*
 * parcel_devicemem.h -- by-reference entry points in the convention of the
 * reference's GPU-build C drivers (libparcel_devicemem.so).
 *
 * A GPU-build driver allocates each vector and matrix array on the device
 * and passes the address of a `devicemem` handle whose first member is the
 * device pointer (src/c_common.f90:7-13; src/example/main_direct_c.c:206-243
 * fills cu_vecx.a_d with cudaMalloc and passes &cu_vecx).  These functions
 * have the names and argument lists of include/parcel_fortran.h with every
 * vector / matrix argument such a handle; they read `mem` and forward to
 * the C ABI of parcel_hip.h.  Link libparcel_devicemem.so before
 * libparcel_hip.so (which exports the array form under the same names);
 * parcel_init_c_ / parcel_free_c_ come from libparcel_hip.so.  Device memory
 * comes from parcel_hip_malloc.  Errors set *iret to the negative library
 * error code.  Do not include this header together with parcel_fortran.h.
 */
#ifndef PARCEL_DEVICEMEM_H
#define PARCEL_DEVICEMEM_H

#ifdef __cplusplus
extern "C" {
#endif

/* type,bind(c) :: devicemem (src/c_common.f90:7-13): the device array and
 * two descriptor slots of the reference's cuSPARSE build (unused here) */
typedef struct parcel_devicemem {
  void* mem;
  void* descr;
  void* descr_add;
} parcel_devicemem;

/* parcel_init_c / parcel_free_c (src/example/c2fortran_kernel.f90:1-10),
 * exported by libparcel_hip.so */
void parcel_init_c_(const int* myrank);
void parcel_free_c_(void);

/* parcel_dcg_c (src/example/c2fortran_kernel.f90:334-432, CUDA_SOLVER form) */
void parcel_dcg_c_(const int* icomm, parcel_devicemem* vecx, const parcel_devicemem* vecb,
                   const int* n, const int* gn, const int* nnz, const int* istart,
                   const parcel_devicemem* crsa, const parcel_devicemem* crsrow_ptr,
                   const parcel_devicemem* crscol, const int* ipreflag, const int* ilu_method,
                   const int* addl, const int* iflagas, int* niter, const double* rtolmax,
                   double* reshistory, const int* iovlflag, const int* precon_thblock,
                   const int* Neuman_itr, const int* independ_nvec, const int* nBlock, int* iret);

/* parcel_dbicgstab_c (:434-521, CUDA_SOLVER form) */
void parcel_dbicgstab_c_(const int* icomm, parcel_devicemem* vecx, const parcel_devicemem* vecb,
                         const int* n, const int* gn, const int* nnz, const int* istart,
                         const parcel_devicemem* crsa, const parcel_devicemem* crsrow_ptr,
                         const parcel_devicemem* crscol, const int* ipreflag,
                         const int* ilu_method, const int* addl, const int* iflagas, int* niter,
                         const double* rtolmax, double* reshistory, const int* iovlflag,
                         const int* precon_thblock, const int* Neuman_itr,
                         const int* independ_nvec, const int* nBlock, int* iret);

/* parcel_dgmres_c (:523-618, CUDA_SOLVER form) */
void parcel_dgmres_c_(const int* icomm, parcel_devicemem* vecx, const parcel_devicemem* vecb,
                      const int* n, const int* gn, const int* nnz, const int* istart,
                      const parcel_devicemem* crsa, const parcel_devicemem* crsrow_ptr,
                      const parcel_devicemem* crscol, const int* ipreflag, const int* ilu_method,
                      const int* addl, const int* iflagas, int* niter, const double* rtolmax,
                      double* reshistory, const int* iovlflag, const int* precon_thblock,
                      const int* Neuman_itr, const int* independ_nvec, const int* nBlock,
                      const int* restart, const int* gmres_gsflag, int* iret);

/* parcel_dcagmres_c (:901-1004, CUDA_SOLVER form: the driver passes
 * &devicemem where the shim reads a c_ptr, i.e. the handle's first member) */
void parcel_dcagmres_c_(const int* icomm, parcel_devicemem* vecx, const parcel_devicemem* vecb,
                        const int* n, const int* gn, const int* nnz, const int* istart,
                        const parcel_devicemem* crsa, const parcel_devicemem* crsrow_ptr,
                        const parcel_devicemem* crscol, const int* ipreflag,
                        const int* ilu_method, const int* addl, const int* iflagas, int* niter,
                        const double* rtolmax, double* reshistory, const int* iovlflag,
                        const int* precon_thblock, const int* Neuman_itr,
                        const int* independ_nvec, const int* nBlock, const int* cagmres_sstep,
                        const int* cagmres_tstep, const int* cagmres_basis,
                        const int* cagmres_qrflag, int* iret);

/* parcel_dcbcg_c (:1006-1117, CUDA_SOLVER form) */
void parcel_dcbcg_c_(const int* icomm, parcel_devicemem* vecx, const parcel_devicemem* vecb,
                     const int* n, const int* gn, const int* nnz, const int* istart,
                     const parcel_devicemem* crsa, const parcel_devicemem* crsrow_ptr,
                     const parcel_devicemem* crscol, const int* ipreflag, const int* ilu_method,
                     const int* addl, const int* iflagas, int* niter, const double* rtolmax,
                     double* reshistory, const int* iovlflag, const int* precon_thblock,
                     const int* Neuman_itr, const int* independ_nvec, const int* nBlock,
                     const int* cbcg_kstep, const int* cbcg_eigenflag,
                     const int* power_method_itrmax, const int* caarnoldi_sstep,
                     const int* caarnoldi_tstep, const int* caarnoldi_basis,
                     const int* caarnoldi_qrflag, int* iret);

/* parcel_ddcg_c (:620-708): the reference builds it for the CPU only (its
 * dd path forces the CPU, src/ddwrapper.f90:268); here the hi / lo arrays
 * may be device handles too (a NULL handle or mem: no lo part) */
void parcel_ddcg_c_(const int* icomm, parcel_devicemem* vecx, parcel_devicemem* vecx_lo,
                    const parcel_devicemem* vecb, const parcel_devicemem* vecb_lo, const int* n,
                    const int* gn, const int* nnz, const int* istart, const parcel_devicemem* crsa,
                    const parcel_devicemem* crsa_lo, const parcel_devicemem* crsrow_ptr,
                    const parcel_devicemem* crscol, const int* ipreflag, const int* ilu_method,
                    const int* addl, const int* iflagas, int* niter, const double* rtolmax,
                    double* reshistory, const int* iovlflag, const int* precon_thblock,
                    const int* Neuman_itr, const int* independ_nvec, const int* nBlock,
                    const int* precision_a, const int* precision_b, const int* precision_x,
                    const int* precision_precondition, int* iret);

/* parcel_dcg_ddm_c (:2964-3064, CUDA_SOLVER form): vecx, vecb, offset and
 * diaA as handles; the grid description stays host arrays */
void parcel_dcg_ddm_c_(const int* icomm, parcel_devicemem* vecx, const parcel_devicemem* vecb,
                       const int* n, const int* nnd, const parcel_devicemem* offset,
                       const parcel_devicemem* diaA, const int* num_neighbor_ranks,
                       const int* neighbor_ranks, const int* margin, const int* ndim,
                       const int* npe_dim, const int* n_grids, const int* offset_dim,
                       const int* div_direc_th, const int* ipreflag, const int* addl,
                       const int* iflagas, int* niter, const double* rtolmax, double* reshistory,
                       const int* iovlflag, const int* precon_thblock, const int* Neuman_itr,
                       const int* independ_nvec, const int* nBlock, int* iret);

#ifdef __cplusplus
}
#endif
#endif

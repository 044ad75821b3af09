! parcel_wrapper.f90 -- Fortran module PARCEL_wrapper over libparcel_hip.so.
!
! Keeps the procedure names and argument lists of the reference modules
! (src/wrapper.f90): parcel_dcg (:544-632 GPU build, :1995-2083 CPU build),
! parcel_dbicgstab (:362, :1813), parcel_dgmres (:727, :2179),
! parcel_dcagmres (:1355, :2807), parcel_dcbcg (:1139), parcel_dpower (:913)
! and parcel_dcaarnoldi (:1561), the DIA forms of the first five (:635, :454,
! :821, :1457, :1246); parcel_ddcg
! (src/ddwrapper.f90:184); parcel_dcg_ddm (src/src_DDM/ddm_kernel.f90:2266);
! and PARCEL_Init / PARCEL_Free (src/c_krylov_wrapper.f90:6-22).  Each generic
! name takes either host arrays (the CPU build's real*8 / integer arrays) or
! device pointers (the GPU build's type(c_ptr)); both forward to the C ABI of
! include/parcel_hip.h.  A library error stops the program with its message,
! as the reference stops on set-up errors (src/krylov.f90:551-560).
module PARCEL_wrapper
  use iso_c_binding
  implicit none
  private
  public :: PARCEL_Init, PARCEL_Free, parcel_dcg, parcel_dbicgstab, parcel_dgmres, parcel_dcagmres, &
       parcel_dcbcg, parcel_ddcg, parcel_dcg_ddm, parcel_dpower, parcel_dcaarnoldi, &
       parcel_dcg_dia, parcel_dbicgstab_dia, parcel_dgmres_dia, parcel_dcagmres_dia, parcel_dcbcg_dia

  interface
    integer(c_int) function c_init(myrank) bind(c, name="parcel_hip_init")
      import :: c_int
      integer(c_int), value :: myrank
    end function
    integer(c_int) function c_finalize() bind(c, name="parcel_hip_finalize")
      import :: c_int
    end function
    type(c_ptr) function c_last_error() bind(c, name="parcel_hip_last_error")
      import :: c_ptr
    end function
    ! parcel_hip_dcg and parcel_hip_dbicgstab share one argument list
    integer(c_int) function c_krylov(icomm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, &
         ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, precon_thblock, &
         Neuman_itr, independ_nvec, nBlock, iret) bind(c)
      import :: c_int, c_ptr, c_double
      integer(c_int), value :: icomm, n, gn, nnz, istart, ipreflag, ILU_method, addL, iflagAS
      type(c_ptr), value :: vecx, vecb, crsA, crsRow_ptr, crsCol, reshistory
      integer(c_int) :: itrmax, iret
      real(c_double), value :: rtolmax
      integer(c_int), value :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
    end function
    integer(c_int) function c_dgmres(icomm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, &
         ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, precon_thblock, &
         Neuman_itr, independ_nvec, nBlock, GMRES_m, GMRES_GSflag, iret) bind(c, name="parcel_hip_dgmres")
      import :: c_int, c_ptr, c_double
      integer(c_int), value :: icomm, n, gn, nnz, istart, ipreflag, ILU_method, addL, iflagAS
      type(c_ptr), value :: vecx, vecb, crsA, crsRow_ptr, crsCol, reshistory
      integer(c_int) :: itrmax, iret
      real(c_double), value :: rtolmax
      integer(c_int), value :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
      integer(c_int), value :: GMRES_m, GMRES_GSflag
    end function
    integer(c_int) function c_dcagmres(icomm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, &
         ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, precon_thblock, &
         Neuman_itr, independ_nvec, nBlock, sstep, tstep, basis, qrflag, iret) bind(c, name="parcel_hip_dcagmres")
      import :: c_int, c_ptr, c_double
      integer(c_int), value :: icomm, n, gn, nnz, istart, ipreflag, ILU_method, addL, iflagAS
      type(c_ptr), value :: vecx, vecb, crsA, crsRow_ptr, crsCol, reshistory
      integer(c_int) :: itrmax, iret
      real(c_double), value :: rtolmax
      integer(c_int), value :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
      integer(c_int), value :: sstep, tstep, basis, qrflag
    end function
    integer(c_int) function c_dcbcg(icomm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, &
         ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, precon_thblock, &
         Neuman_itr, independ_nvec, nBlock, kstep, eigenflag, pm_itrmax, ca_s, ca_t, ca_basis, ca_qr, iret) &
         bind(c, name="parcel_hip_dcbcg")
      import :: c_int, c_ptr, c_double
      integer(c_int), value :: icomm, n, gn, nnz, istart, ipreflag, ILU_method, addL, iflagAS
      type(c_ptr), value :: vecx, vecb, crsA, crsRow_ptr, crsCol, reshistory
      integer(c_int) :: itrmax, iret
      real(c_double), value :: rtolmax
      integer(c_int), value :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
      integer(c_int), value :: kstep, eigenflag, pm_itrmax, ca_s, ca_t, ca_basis, ca_qr
    end function
    integer(c_int) function c_dcg_dia(icomm, vecx, vecb, n, gn, istart, diaA, offset, nnd, &
         ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, &
         precon_thblock, Neuman_itr, independ_nvec, nBlock, iret) &
         bind(c, name="parcel_hip_dcg_dia")
      import :: c_int, c_ptr, c_double
      integer(c_int), value :: icomm, n, gn, istart, nnd, ipreflag, ILU_method, addL, iflagAS
      type(c_ptr), value :: vecx, vecb, diaA, offset, reshistory
      integer(c_int) :: itrmax, iret
      real(c_double), value :: rtolmax
      integer(c_int), value :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
    end function
    integer(c_int) function c_dbicgstab_dia(icomm, vecx, vecb, n, gn, istart, diaA, offset, nnd, &
         ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, &
         precon_thblock, Neuman_itr, independ_nvec, nBlock, iret) &
         bind(c, name="parcel_hip_dbicgstab_dia")
      import :: c_int, c_ptr, c_double
      integer(c_int), value :: icomm, n, gn, istart, nnd, ipreflag, ILU_method, addL, iflagAS
      type(c_ptr), value :: vecx, vecb, diaA, offset, reshistory
      integer(c_int) :: itrmax, iret
      real(c_double), value :: rtolmax
      integer(c_int), value :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
    end function
    integer(c_int) function c_dgmres_dia(icomm, vecx, vecb, n, gn, istart, diaA, offset, nnd, &
         ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, &
         precon_thblock, Neuman_itr, independ_nvec, nBlock, GMRES_m, GMRES_GSflag, iret) &
         bind(c, name="parcel_hip_dgmres_dia")
      import :: c_int, c_ptr, c_double
      integer(c_int), value :: icomm, n, gn, istart, nnd, ipreflag, ILU_method, addL, iflagAS
      type(c_ptr), value :: vecx, vecb, diaA, offset, reshistory
      integer(c_int) :: itrmax, iret
      real(c_double), value :: rtolmax
      integer(c_int), value :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
      integer(c_int), value :: GMRES_m, GMRES_GSflag
    end function
    integer(c_int) function c_dcagmres_dia(icomm, vecx, vecb, n, gn, istart, diaA, offset, nnd, &
         ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, &
         precon_thblock, Neuman_itr, independ_nvec, nBlock, CAGMRES_sstep, CAGMRES_tstep, &
         CAGMRES_basis, CAGMRES_QRflag, iret) &
         bind(c, name="parcel_hip_dcagmres_dia")
      import :: c_int, c_ptr, c_double
      integer(c_int), value :: icomm, n, gn, istart, nnd, ipreflag, ILU_method, addL, iflagAS
      type(c_ptr), value :: vecx, vecb, diaA, offset, reshistory
      integer(c_int) :: itrmax, iret
      real(c_double), value :: rtolmax
      integer(c_int), value :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
      integer(c_int), value :: CAGMRES_sstep, CAGMRES_tstep, CAGMRES_basis, CAGMRES_QRflag
    end function
    integer(c_int) function c_dcbcg_dia(icomm, vecx, vecb, n, gn, istart, diaA, offset, nnd, &
         ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, &
         precon_thblock, Neuman_itr, independ_nvec, nBlock, CBCG_kstep, CBCG_Eigenflag, &
         power_method_itrmax, CAARNOLDI_sstep, CAARNOLDI_tstep, CAARNOLDI_basis, CAARNOLDI_QRflag, &
         iret) &
         bind(c, name="parcel_hip_dcbcg_dia")
      import :: c_int, c_ptr, c_double
      integer(c_int), value :: icomm, n, gn, istart, nnd, ipreflag, ILU_method, addL, iflagAS
      type(c_ptr), value :: vecx, vecb, diaA, offset, reshistory
      integer(c_int) :: itrmax, iret
      real(c_double), value :: rtolmax
      integer(c_int), value :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
      integer(c_int), value :: CBCG_kstep, CBCG_Eigenflag, power_method_itrmax, CAARNOLDI_sstep, CAARNOLDI_tstep, CAARNOLDI_basis, CAARNOLDI_QRflag
    end function
    integer(c_int) function c_dpower(icomm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, &
         ipreflag, ILU_method, addL, iflagAS, itrmax, iovlflag, precon_thblock, Neuman_itr, independ_nvec, &
         nBlock, pm_itrmax, ev_re, ev_im, ev_vec, ev_err, maxid, minid, iret) bind(c, name="parcel_hip_dpower")
      import :: c_int, c_ptr
      integer(c_int), value :: icomm, n, gn, nnz, istart, ipreflag, ILU_method, addL, iflagAS
      type(c_ptr), value :: vecx, vecb, crsA, crsRow_ptr, crsCol, ev_re, ev_im, ev_vec, ev_err
      integer(c_int) :: itrmax, maxid, minid, iret
      integer(c_int), value :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock, pm_itrmax
    end function
    integer(c_int) function c_dcaarnoldi(icomm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, &
         ipreflag, ILU_method, addL, iflagAS, iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock, &
         ca_s, ca_t, ca_basis, ca_qr, ev_re, ev_im, ev_vec, ev_err, maxid, minid, iret) &
         bind(c, name="parcel_hip_dcaarnoldi")
      import :: c_int, c_ptr
      integer(c_int), value :: icomm, n, gn, nnz, istart, ipreflag, ILU_method, addL, iflagAS
      type(c_ptr), value :: vecx, vecb, crsA, crsRow_ptr, crsCol, ev_re, ev_im, ev_vec, ev_err
      integer(c_int) :: maxid, minid, iret
      integer(c_int), value :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
      integer(c_int), value :: ca_s, ca_t, ca_basis, ca_qr
    end function
    integer(c_int) function c_ddcg(icomm, xh, xl, bh, bl, n, gn, nnz, istart, ah, al, crsRow_ptr, crsCol, &
         ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, precon_thblock, &
         independ_nvec, nBlock, pa, pb, px, pp, iret) bind(c, name="parcel_hip_ddcg")
      import :: c_int, c_ptr, c_double
      integer(c_int), value :: icomm, n, gn, nnz, istart, ipreflag, ILU_method, addL, iflagAS
      type(c_ptr), value :: xh, xl, bh, bl, ah, al, crsRow_ptr, crsCol, reshistory
      integer(c_int) :: itrmax, iret
      real(c_double), value :: rtolmax
      integer(c_int), value :: iovlflag, precon_thblock, independ_nvec, nBlock, pa, pb, px, pp
    end function
    integer(c_int) function c_dcg_ddm(icomm, x, b, nrow, nnd, ioff_dia, val_dia, num_neighbor_ranks, &
         neighbor_ranks, margin, ncomp_grids, ndiv_grids, num_grids, ioff_grids, div_direc_th, ipreflag, addL, &
         iflagAS, itrmax, rtolmax, reshistory, iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock, iret) &
         bind(c, name="parcel_hip_dcg_ddm")
      import :: c_int, c_ptr, c_double
      integer(c_int), value :: icomm, nrow, nnd, num_neighbor_ranks, margin, ncomp_grids, div_direc_th
      type(c_ptr), value :: x, b, ioff_dia, val_dia, neighbor_ranks, ndiv_grids, num_grids, ioff_grids, reshistory
      integer(c_int), value :: ipreflag, addL, iflagAS
      integer(c_int) :: itrmax, iret
      real(c_double), value :: rtolmax
      integer(c_int), value :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
    end function
  end interface

  procedure(c_krylov), bind(c, name="parcel_hip_dcg") :: c_dcg
  procedure(c_krylov), bind(c, name="parcel_hip_dbicgstab") :: c_dbicgstab

  interface parcel_dcg
    module procedure parcel_dcg_host, parcel_dcg_dev
  end interface
  interface parcel_dbicgstab
    module procedure parcel_dbicgstab_host, parcel_dbicgstab_dev
  end interface
  interface parcel_dgmres
    module procedure parcel_dgmres_host, parcel_dgmres_dev
  end interface
  interface parcel_dcagmres
    module procedure parcel_dcagmres_host, parcel_dcagmres_dev
  end interface
  interface parcel_dcbcg
    module procedure parcel_dcbcg_host, parcel_dcbcg_dev
  end interface
  interface parcel_ddcg
    module procedure parcel_ddcg_host, parcel_ddcg_dev
  end interface
  interface parcel_dcg_dia
    module procedure parcel_dcg_dia_host, parcel_dcg_dia_dev
  end interface
  interface parcel_dbicgstab_dia
    module procedure parcel_dbicgstab_dia_host, parcel_dbicgstab_dia_dev
  end interface
  interface parcel_dgmres_dia
    module procedure parcel_dgmres_dia_host, parcel_dgmres_dia_dev
  end interface
  interface parcel_dcagmres_dia
    module procedure parcel_dcagmres_dia_host, parcel_dcagmres_dia_dev
  end interface
  interface parcel_dcbcg_dia
    module procedure parcel_dcbcg_dia_host, parcel_dcbcg_dia_dev
  end interface
  interface parcel_dpower
    module procedure parcel_dpower_host, parcel_dpower_dev
  end interface
  interface parcel_dcaarnoldi
    module procedure parcel_dcaarnoldi_host, parcel_dcaarnoldi_dev
  end interface

contains

  subroutine check(rc)
    integer(c_int), intent(in) :: rc
    character(kind=c_char), pointer :: msg(:)
    integer :: k
    if (rc == 0) return
    call c_f_pointer(c_last_error(), msg, [1024])
    k = 0
    do while (k < 1024)
      if (msg(k + 1) == c_null_char) exit
      k = k + 1
    end do
    write (*, '(a,i0,a,1024a)') 'PARCEL error ', rc, ': ', msg(1:k)
    stop 1
  end subroutine check

  subroutine PARCEL_Init(myrank)
    integer, intent(in) :: myrank
    call check(c_init(int(myrank, c_int)))
  end subroutine PARCEL_Init

  subroutine PARCEL_Free()
    call check(c_finalize())
  end subroutine PARCEL_Free

  ! ---- CG ---------------------------------------------------------------
  subroutine parcel_dcg_host(icomm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, &
       ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, precon_thblock, &
       Neuman_itr, independ_nvec, nBlock, iret, machineflag)
    integer, intent(in) :: icomm, n, gn, nnz, istart, ipreflag, ILU_method, addL, iflagAS
    real(8), target, intent(inout) :: vecx(n)
    real(8), target, intent(in) :: vecb(n), crsA(nnz), rtolmax
    integer, target, intent(in) :: crsRow_ptr(n + 1), crsCol(nnz)
    integer, intent(inout) :: itrmax
    real(8), target, intent(inout) :: reshistory(*)
    integer, intent(in) :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
    integer, intent(out) :: iret
    integer, optional, intent(in) :: machineflag
    call check(c_dcg(icomm, c_loc(vecx), c_loc(vecb), n, gn, nnz, istart, c_loc(crsA), c_loc(crsRow_ptr), &
         c_loc(crsCol), ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, c_loc(reshistory), iovlflag, &
         precon_thblock, Neuman_itr, independ_nvec, nBlock, iret))
  end subroutine parcel_dcg_host

  subroutine parcel_dcg_dev(icomm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, &
       ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, precon_thblock, &
       Neuman_itr, independ_nvec, nBlock, iret, machineflag)
    integer, intent(in) :: icomm, n, gn, nnz, istart, ipreflag, ILU_method, addL, iflagAS
    type(c_ptr), intent(in) :: vecx, vecb, crsA, crsRow_ptr, crsCol
    integer, intent(inout) :: itrmax
    real(8), intent(in) :: rtolmax
    real(8), target, intent(inout) :: reshistory(*)
    integer, intent(in) :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
    integer, intent(out) :: iret
    integer, optional, intent(in) :: machineflag
    call check(c_dcg(icomm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, ipreflag, ILU_method, &
         addL, iflagAS, itrmax, rtolmax, c_loc(reshistory), iovlflag, precon_thblock, Neuman_itr, &
         independ_nvec, nBlock, iret))
  end subroutine parcel_dcg_dev

  ! ---- BiCGStab ---------------------------------------------------------
  subroutine parcel_dbicgstab_host(icomm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, &
       ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, precon_thblock, &
       Neuman_itr, independ_nvec, nBlock, iret, machineflag)
    integer, intent(in) :: icomm, n, gn, nnz, istart, ipreflag, ILU_method, addL, iflagAS
    real(8), target, intent(inout) :: vecx(n)
    real(8), target, intent(in) :: vecb(n), crsA(nnz), rtolmax
    integer, target, intent(in) :: crsRow_ptr(n + 1), crsCol(nnz)
    integer, intent(inout) :: itrmax
    real(8), target, intent(inout) :: reshistory(*)
    integer, intent(in) :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
    integer, intent(out) :: iret
    integer, optional, intent(in) :: machineflag
    call check(c_dbicgstab(icomm, c_loc(vecx), c_loc(vecb), n, gn, nnz, istart, c_loc(crsA), &
         c_loc(crsRow_ptr), c_loc(crsCol), ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, &
         c_loc(reshistory), iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock, iret))
  end subroutine parcel_dbicgstab_host

  subroutine parcel_dbicgstab_dev(icomm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, &
       ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, precon_thblock, &
       Neuman_itr, independ_nvec, nBlock, iret, machineflag)
    integer, intent(in) :: icomm, n, gn, nnz, istart, ipreflag, ILU_method, addL, iflagAS
    type(c_ptr), intent(in) :: vecx, vecb, crsA, crsRow_ptr, crsCol
    integer, intent(inout) :: itrmax
    real(8), intent(in) :: rtolmax
    real(8), target, intent(inout) :: reshistory(*)
    integer, intent(in) :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
    integer, intent(out) :: iret
    integer, optional, intent(in) :: machineflag
    call check(c_dbicgstab(icomm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, ipreflag, &
         ILU_method, addL, iflagAS, itrmax, rtolmax, c_loc(reshistory), iovlflag, precon_thblock, &
         Neuman_itr, independ_nvec, nBlock, iret))
  end subroutine parcel_dbicgstab_dev

  ! ---- GMRES(m) ---------------------------------------------------------
  subroutine parcel_dgmres_host(icomm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, &
       ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, precon_thblock, &
       Neuman_itr, independ_nvec, nBlock, GMRES_m, GMRES_GSflag, iret, machineflag)
    integer, intent(in) :: icomm, n, gn, nnz, istart, ipreflag, ILU_method, addL, iflagAS
    real(8), target, intent(inout) :: vecx(n)
    real(8), target, intent(in) :: vecb(n), crsA(nnz), rtolmax
    integer, target, intent(in) :: crsRow_ptr(n + 1), crsCol(nnz)
    integer, intent(inout) :: itrmax
    real(8), target, intent(inout) :: reshistory(*)
    integer, intent(in) :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
    integer, intent(in) :: GMRES_m, GMRES_GSflag
    integer, intent(out) :: iret
    integer, optional, intent(in) :: machineflag
    call check(c_dgmres(icomm, c_loc(vecx), c_loc(vecb), n, gn, nnz, istart, c_loc(crsA), c_loc(crsRow_ptr), &
         c_loc(crsCol), ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, c_loc(reshistory), iovlflag, &
         precon_thblock, Neuman_itr, independ_nvec, nBlock, GMRES_m, GMRES_GSflag, iret))
  end subroutine parcel_dgmres_host

  subroutine parcel_dgmres_dev(icomm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, &
       ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, precon_thblock, &
       Neuman_itr, independ_nvec, nBlock, GMRES_m, GMRES_GSflag, iret, machineflag)
    integer, intent(in) :: icomm, n, gn, nnz, istart, ipreflag, ILU_method, addL, iflagAS
    type(c_ptr), intent(in) :: vecx, vecb, crsA, crsRow_ptr, crsCol
    integer, intent(inout) :: itrmax
    real(8), intent(in) :: rtolmax
    real(8), target, intent(inout) :: reshistory(*)
    integer, intent(in) :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
    integer, intent(in) :: GMRES_m, GMRES_GSflag
    integer, intent(out) :: iret
    integer, optional, intent(in) :: machineflag
    call check(c_dgmres(icomm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, ipreflag, ILU_method, &
         addL, iflagAS, itrmax, rtolmax, c_loc(reshistory), iovlflag, precon_thblock, Neuman_itr, &
         independ_nvec, nBlock, GMRES_m, GMRES_GSflag, iret))
  end subroutine parcel_dgmres_dev

  ! ---- CA-GMRES(s,t) (src/wrapper.f90:1355-1454 GPU build, :2807-2906 CPU build)
  subroutine parcel_dcagmres_host(icomm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, &
       ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, precon_thblock, &
       Neuman_itr, independ_nvec, nBlock, CAGMRES_sstep, CAGMRES_tstep, CAGMRES_basis, CAGMRES_QRflag, iret)
    integer, intent(in) :: icomm, n, gn, nnz, istart, ipreflag, ILU_method, addL, iflagAS
    real(8), target, intent(inout) :: vecx(n)
    real(8), target, intent(in) :: vecb(n), crsA(nnz), rtolmax
    integer, target, intent(in) :: crsRow_ptr(n + 1), crsCol(nnz)
    integer, intent(inout) :: itrmax
    real(8), target, intent(inout) :: reshistory(*)
    integer, intent(in) :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
    integer, intent(in) :: CAGMRES_sstep, CAGMRES_tstep, CAGMRES_basis, CAGMRES_QRflag
    integer, intent(out) :: iret
    call check(c_dcagmres(icomm, c_loc(vecx), c_loc(vecb), n, gn, nnz, istart, c_loc(crsA), c_loc(crsRow_ptr), &
         c_loc(crsCol), ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, c_loc(reshistory), iovlflag, &
         precon_thblock, Neuman_itr, independ_nvec, nBlock, CAGMRES_sstep, CAGMRES_tstep, CAGMRES_basis, &
         CAGMRES_QRflag, iret))
  end subroutine parcel_dcagmres_host

  subroutine parcel_dcagmres_dev(icomm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, &
       ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, precon_thblock, &
       Neuman_itr, independ_nvec, nBlock, CAGMRES_sstep, CAGMRES_tstep, CAGMRES_basis, CAGMRES_QRflag, iret)
    integer, intent(in) :: icomm, n, gn, nnz, istart, ipreflag, ILU_method, addL, iflagAS
    type(c_ptr), intent(in) :: vecx, vecb, crsA, crsRow_ptr, crsCol
    integer, intent(inout) :: itrmax
    real(8), intent(in) :: rtolmax
    real(8), target, intent(inout) :: reshistory(*)
    integer, intent(in) :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
    integer, intent(in) :: CAGMRES_sstep, CAGMRES_tstep, CAGMRES_basis, CAGMRES_QRflag
    integer, intent(out) :: iret
    call check(c_dcagmres(icomm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, ipreflag, &
         ILU_method, addL, iflagAS, itrmax, rtolmax, c_loc(reshistory), iovlflag, precon_thblock, Neuman_itr, &
         independ_nvec, nBlock, CAGMRES_sstep, CAGMRES_tstep, CAGMRES_basis, CAGMRES_QRflag, iret))
  end subroutine parcel_dcagmres_dev

  ! ---- Chebyshev-basis CG (src/wrapper.f90:1139-1233) ---------------------
  subroutine parcel_dcbcg_host(icomm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, &
       ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, precon_thblock, &
       Neuman_itr, independ_nvec, nBlock, CBCG_kstep, CBCG_Eigenflag, power_method_itrmax, &
       CAARNOLDI_sstep, CAARNOLDI_tstep, CAARNOLDI_basis, CAARNOLDI_QRflag, iret)
    integer, intent(in) :: icomm, n, gn, nnz, istart, ipreflag, ILU_method, addL, iflagAS
    real(8), target, intent(inout) :: vecx(n)
    real(8), target, intent(in) :: vecb(n), crsA(nnz), rtolmax
    integer, target, intent(in) :: crsRow_ptr(n + 1), crsCol(nnz)
    integer, intent(inout) :: itrmax
    real(8), target, intent(inout) :: reshistory(*)
    integer, intent(in) :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
    integer, intent(in) :: CBCG_kstep, CBCG_Eigenflag, power_method_itrmax
    integer, intent(in) :: CAARNOLDI_sstep, CAARNOLDI_tstep, CAARNOLDI_basis, CAARNOLDI_QRflag
    integer, intent(out) :: iret
    call check(c_dcbcg(icomm, c_loc(vecx), c_loc(vecb), n, gn, nnz, istart, c_loc(crsA), c_loc(crsRow_ptr), &
         c_loc(crsCol), ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, c_loc(reshistory), iovlflag, &
         precon_thblock, Neuman_itr, independ_nvec, nBlock, CBCG_kstep, CBCG_Eigenflag, power_method_itrmax, &
         CAARNOLDI_sstep, CAARNOLDI_tstep, CAARNOLDI_basis, CAARNOLDI_QRflag, iret))
  end subroutine parcel_dcbcg_host

  subroutine parcel_dcbcg_dev(icomm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, &
       ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, precon_thblock, &
       Neuman_itr, independ_nvec, nBlock, CBCG_kstep, CBCG_Eigenflag, power_method_itrmax, &
       CAARNOLDI_sstep, CAARNOLDI_tstep, CAARNOLDI_basis, CAARNOLDI_QRflag, iret)
    integer, intent(in) :: icomm, n, gn, nnz, istart, ipreflag, ILU_method, addL, iflagAS
    type(c_ptr), intent(in) :: vecx, vecb, crsA, crsRow_ptr, crsCol
    integer, intent(inout) :: itrmax
    real(8), intent(in) :: rtolmax
    real(8), target, intent(inout) :: reshistory(*)
    integer, intent(in) :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
    integer, intent(in) :: CBCG_kstep, CBCG_Eigenflag, power_method_itrmax
    integer, intent(in) :: CAARNOLDI_sstep, CAARNOLDI_tstep, CAARNOLDI_basis, CAARNOLDI_QRflag
    integer, intent(out) :: iret
    call check(c_dcbcg(icomm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, ipreflag, ILU_method, &
         addL, iflagAS, itrmax, rtolmax, c_loc(reshistory), iovlflag, precon_thblock, Neuman_itr, &
         independ_nvec, nBlock, CBCG_kstep, CBCG_Eigenflag, power_method_itrmax, CAARNOLDI_sstep, &
         CAARNOLDI_tstep, CAARNOLDI_basis, CAARNOLDI_QRflag, iret))
  end subroutine parcel_dcbcg_dev

  ! ---- DIA forms (src/wrapper.f90:454, :635, :821, :1246, :1457): the CRS
  ! procedures' arguments with (n, gn, istart, diaA, offset, nnd)

  subroutine parcel_dcg_dia_host(icomm, vecx, vecb, n, gn, istart, diaA, offset, nnd, ipreflag, &
       ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, precon_thblock, &
       Neuman_itr, independ_nvec, nBlock, iret, machineflag)
    integer, intent(in) :: icomm, n, gn, istart, nnd, ipreflag, ILU_method, addL, iflagAS
    real(8), target, intent(inout) :: vecx(n)
    real(8), target, intent(in) :: vecb(n), diaA(nnd * n), rtolmax
    integer, target, intent(in) :: offset(nnd)
    integer, intent(inout) :: itrmax
    real(8), target, intent(inout) :: reshistory(*)
    integer, intent(in) :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
    integer, intent(out) :: iret
    integer, optional, intent(in) :: machineflag
    call check(c_dcg_dia(icomm, c_loc(vecx), c_loc(vecb), n, gn, istart, c_loc(diaA), &
         c_loc(offset), nnd, ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, &
         c_loc(reshistory), iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock, iret))
  end subroutine parcel_dcg_dia_host

  subroutine parcel_dcg_dia_dev(icomm, vecx, vecb, n, gn, istart, diaA, offset, nnd, ipreflag, &
       ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, precon_thblock, &
       Neuman_itr, independ_nvec, nBlock, iret, machineflag)
    integer, intent(in) :: icomm, n, gn, istart, nnd, ipreflag, ILU_method, addL, iflagAS
    type(c_ptr), intent(in) :: vecx, vecb, diaA, offset
    integer, intent(inout) :: itrmax
    real(8), intent(in) :: rtolmax
    real(8), target, intent(inout) :: reshistory(*)
    integer, intent(in) :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
    integer, intent(out) :: iret
    integer, optional, intent(in) :: machineflag
    call check(c_dcg_dia(icomm, vecx, vecb, n, gn, istart, diaA, offset, nnd, ipreflag, ILU_method, &
         addL, iflagAS, itrmax, rtolmax, c_loc(reshistory), iovlflag, precon_thblock, Neuman_itr, &
         independ_nvec, nBlock, iret))
  end subroutine parcel_dcg_dia_dev

  subroutine parcel_dbicgstab_dia_host(icomm, vecx, vecb, n, gn, istart, diaA, offset, nnd, &
       ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, precon_thblock, &
       Neuman_itr, independ_nvec, nBlock, iret, machineflag)
    integer, intent(in) :: icomm, n, gn, istart, nnd, ipreflag, ILU_method, addL, iflagAS
    real(8), target, intent(inout) :: vecx(n)
    real(8), target, intent(in) :: vecb(n), diaA(nnd * n), rtolmax
    integer, target, intent(in) :: offset(nnd)
    integer, intent(inout) :: itrmax
    real(8), target, intent(inout) :: reshistory(*)
    integer, intent(in) :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
    integer, intent(out) :: iret
    integer, optional, intent(in) :: machineflag
    call check(c_dbicgstab_dia(icomm, c_loc(vecx), c_loc(vecb), n, gn, istart, c_loc(diaA), &
         c_loc(offset), nnd, ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, &
         c_loc(reshistory), iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock, iret))
  end subroutine parcel_dbicgstab_dia_host

  subroutine parcel_dbicgstab_dia_dev(icomm, vecx, vecb, n, gn, istart, diaA, offset, nnd, &
       ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, precon_thblock, &
       Neuman_itr, independ_nvec, nBlock, iret, machineflag)
    integer, intent(in) :: icomm, n, gn, istart, nnd, ipreflag, ILU_method, addL, iflagAS
    type(c_ptr), intent(in) :: vecx, vecb, diaA, offset
    integer, intent(inout) :: itrmax
    real(8), intent(in) :: rtolmax
    real(8), target, intent(inout) :: reshistory(*)
    integer, intent(in) :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
    integer, intent(out) :: iret
    integer, optional, intent(in) :: machineflag
    call check(c_dbicgstab_dia(icomm, vecx, vecb, n, gn, istart, diaA, offset, nnd, ipreflag, &
         ILU_method, addL, iflagAS, itrmax, rtolmax, c_loc(reshistory), iovlflag, precon_thblock, &
         Neuman_itr, independ_nvec, nBlock, iret))
  end subroutine parcel_dbicgstab_dia_dev

  subroutine parcel_dgmres_dia_host(icomm, vecx, vecb, n, gn, istart, diaA, offset, nnd, ipreflag, &
       ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, precon_thblock, &
       Neuman_itr, independ_nvec, nBlock, GMRES_m, GMRES_GSflag, iret, machineflag)
    integer, intent(in) :: icomm, n, gn, istart, nnd, ipreflag, ILU_method, addL, iflagAS
    real(8), target, intent(inout) :: vecx(n)
    real(8), target, intent(in) :: vecb(n), diaA(nnd * n), rtolmax
    integer, target, intent(in) :: offset(nnd)
    integer, intent(inout) :: itrmax
    real(8), target, intent(inout) :: reshistory(*)
    integer, intent(in) :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
    integer, intent(in) :: GMRES_m, GMRES_GSflag
    integer, intent(out) :: iret
    integer, optional, intent(in) :: machineflag
    call check(c_dgmres_dia(icomm, c_loc(vecx), c_loc(vecb), n, gn, istart, c_loc(diaA), &
         c_loc(offset), nnd, ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, &
         c_loc(reshistory), iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock, GMRES_m, &
         GMRES_GSflag, iret))
  end subroutine parcel_dgmres_dia_host

  subroutine parcel_dgmres_dia_dev(icomm, vecx, vecb, n, gn, istart, diaA, offset, nnd, ipreflag, &
       ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, precon_thblock, &
       Neuman_itr, independ_nvec, nBlock, GMRES_m, GMRES_GSflag, iret, machineflag)
    integer, intent(in) :: icomm, n, gn, istart, nnd, ipreflag, ILU_method, addL, iflagAS
    type(c_ptr), intent(in) :: vecx, vecb, diaA, offset
    integer, intent(inout) :: itrmax
    real(8), intent(in) :: rtolmax
    real(8), target, intent(inout) :: reshistory(*)
    integer, intent(in) :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
    integer, intent(in) :: GMRES_m, GMRES_GSflag
    integer, intent(out) :: iret
    integer, optional, intent(in) :: machineflag
    call check(c_dgmres_dia(icomm, vecx, vecb, n, gn, istart, diaA, offset, nnd, ipreflag, &
         ILU_method, addL, iflagAS, itrmax, rtolmax, c_loc(reshistory), iovlflag, precon_thblock, &
         Neuman_itr, independ_nvec, nBlock, GMRES_m, GMRES_GSflag, iret))
  end subroutine parcel_dgmres_dia_dev

  subroutine parcel_dcagmres_dia_host(icomm, vecx, vecb, n, gn, istart, diaA, offset, nnd, &
       ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, precon_thblock, &
       Neuman_itr, independ_nvec, nBlock, CAGMRES_sstep, CAGMRES_tstep, CAGMRES_basis, &
       CAGMRES_QRflag, iret, machineflag)
    integer, intent(in) :: icomm, n, gn, istart, nnd, ipreflag, ILU_method, addL, iflagAS
    real(8), target, intent(inout) :: vecx(n)
    real(8), target, intent(in) :: vecb(n), diaA(nnd * n), rtolmax
    integer, target, intent(in) :: offset(nnd)
    integer, intent(inout) :: itrmax
    real(8), target, intent(inout) :: reshistory(*)
    integer, intent(in) :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
    integer, intent(in) :: CAGMRES_sstep, CAGMRES_tstep, CAGMRES_basis, CAGMRES_QRflag
    integer, intent(out) :: iret
    integer, optional, intent(in) :: machineflag
    call check(c_dcagmres_dia(icomm, c_loc(vecx), c_loc(vecb), n, gn, istart, c_loc(diaA), &
         c_loc(offset), nnd, ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, &
         c_loc(reshistory), iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock, &
         CAGMRES_sstep, CAGMRES_tstep, CAGMRES_basis, CAGMRES_QRflag, iret))
  end subroutine parcel_dcagmres_dia_host

  subroutine parcel_dcagmres_dia_dev(icomm, vecx, vecb, n, gn, istart, diaA, offset, nnd, ipreflag, &
       ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, precon_thblock, &
       Neuman_itr, independ_nvec, nBlock, CAGMRES_sstep, CAGMRES_tstep, CAGMRES_basis, &
       CAGMRES_QRflag, iret, machineflag)
    integer, intent(in) :: icomm, n, gn, istart, nnd, ipreflag, ILU_method, addL, iflagAS
    type(c_ptr), intent(in) :: vecx, vecb, diaA, offset
    integer, intent(inout) :: itrmax
    real(8), intent(in) :: rtolmax
    real(8), target, intent(inout) :: reshistory(*)
    integer, intent(in) :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
    integer, intent(in) :: CAGMRES_sstep, CAGMRES_tstep, CAGMRES_basis, CAGMRES_QRflag
    integer, intent(out) :: iret
    integer, optional, intent(in) :: machineflag
    call check(c_dcagmres_dia(icomm, vecx, vecb, n, gn, istart, diaA, offset, nnd, ipreflag, &
         ILU_method, addL, iflagAS, itrmax, rtolmax, c_loc(reshistory), iovlflag, precon_thblock, &
         Neuman_itr, independ_nvec, nBlock, CAGMRES_sstep, CAGMRES_tstep, CAGMRES_basis, &
         CAGMRES_QRflag, iret))
  end subroutine parcel_dcagmres_dia_dev

  subroutine parcel_dcbcg_dia_host(icomm, vecx, vecb, n, gn, istart, diaA, offset, nnd, ipreflag, &
       ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, precon_thblock, &
       Neuman_itr, independ_nvec, nBlock, CBCG_kstep, CBCG_Eigenflag, power_method_itrmax, &
       CAARNOLDI_sstep, CAARNOLDI_tstep, CAARNOLDI_basis, CAARNOLDI_QRflag, iret, machineflag)
    integer, intent(in) :: icomm, n, gn, istart, nnd, ipreflag, ILU_method, addL, iflagAS
    real(8), target, intent(inout) :: vecx(n)
    real(8), target, intent(in) :: vecb(n), diaA(nnd * n), rtolmax
    integer, target, intent(in) :: offset(nnd)
    integer, intent(inout) :: itrmax
    real(8), target, intent(inout) :: reshistory(*)
    integer, intent(in) :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
    integer, intent(in) :: CBCG_kstep, CBCG_Eigenflag, power_method_itrmax, CAARNOLDI_sstep, CAARNOLDI_tstep, CAARNOLDI_basis, CAARNOLDI_QRflag
    integer, intent(out) :: iret
    integer, optional, intent(in) :: machineflag
    call check(c_dcbcg_dia(icomm, c_loc(vecx), c_loc(vecb), n, gn, istart, c_loc(diaA), &
         c_loc(offset), nnd, ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, &
         c_loc(reshistory), iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock, &
         CBCG_kstep, CBCG_Eigenflag, power_method_itrmax, CAARNOLDI_sstep, CAARNOLDI_tstep, &
         CAARNOLDI_basis, CAARNOLDI_QRflag, iret))
  end subroutine parcel_dcbcg_dia_host

  subroutine parcel_dcbcg_dia_dev(icomm, vecx, vecb, n, gn, istart, diaA, offset, nnd, ipreflag, &
       ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, precon_thblock, &
       Neuman_itr, independ_nvec, nBlock, CBCG_kstep, CBCG_Eigenflag, power_method_itrmax, &
       CAARNOLDI_sstep, CAARNOLDI_tstep, CAARNOLDI_basis, CAARNOLDI_QRflag, iret, machineflag)
    integer, intent(in) :: icomm, n, gn, istart, nnd, ipreflag, ILU_method, addL, iflagAS
    type(c_ptr), intent(in) :: vecx, vecb, diaA, offset
    integer, intent(inout) :: itrmax
    real(8), intent(in) :: rtolmax
    real(8), target, intent(inout) :: reshistory(*)
    integer, intent(in) :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
    integer, intent(in) :: CBCG_kstep, CBCG_Eigenflag, power_method_itrmax, CAARNOLDI_sstep, CAARNOLDI_tstep, CAARNOLDI_basis, CAARNOLDI_QRflag
    integer, intent(out) :: iret
    integer, optional, intent(in) :: machineflag
    call check(c_dcbcg_dia(icomm, vecx, vecb, n, gn, istart, diaA, offset, nnd, ipreflag, &
         ILU_method, addL, iflagAS, itrmax, rtolmax, c_loc(reshistory), iovlflag, precon_thblock, &
         Neuman_itr, independ_nvec, nBlock, CBCG_kstep, CBCG_Eigenflag, power_method_itrmax, &
         CAARNOLDI_sstep, CAARNOLDI_tstep, CAARNOLDI_basis, CAARNOLDI_QRflag, iret))
  end subroutine parcel_dcbcg_dia_dev

  ! ---- power method / CA-Arnoldi (one eigenpair; the CRS drivers return
  ! the first component of the eigenvector) ------------------------------
  subroutine parcel_dpower_host(icomm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, &
       ipreflag, ILU_method, addL, iflagAS, itrmax, iovlflag, precon_thblock, Neuman_itr, independ_nvec, &
       nBlock, power_method_itrmax, EigenValueReal, EigenValueImag, EigenVector, EigenErrorNorm, &
       MaxEigenID, MinEigenID, iret, machineflag)
    integer, intent(in) :: icomm, n, gn, nnz, istart, ipreflag, ILU_method, addL, iflagAS
    real(8), target, intent(inout) :: vecx(n)
    real(8), target, intent(in) :: vecb(n), crsA(nnz)
    integer, target, intent(in) :: crsRow_ptr(n + 1), crsCol(nnz)
    integer, intent(inout) :: itrmax
    integer, intent(in) :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock, power_method_itrmax
    real(8), target, intent(inout) :: EigenValueReal(*), EigenValueImag(*), EigenVector(*), EigenErrorNorm(*)
    integer, intent(out) :: MaxEigenID, MinEigenID, iret
    integer, optional, intent(in) :: machineflag
    call check(c_dpower(icomm, c_loc(vecx), c_loc(vecb), n, gn, nnz, istart, c_loc(crsA), c_loc(crsRow_ptr), &
         c_loc(crsCol), ipreflag, ILU_method, addL, iflagAS, itrmax, iovlflag, precon_thblock, Neuman_itr, &
         independ_nvec, nBlock, power_method_itrmax, c_loc(EigenValueReal), c_loc(EigenValueImag), &
         c_loc(EigenVector), c_loc(EigenErrorNorm), MaxEigenID, MinEigenID, iret))
  end subroutine parcel_dpower_host

  subroutine parcel_dpower_dev(icomm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, &
       ipreflag, ILU_method, addL, iflagAS, itrmax, iovlflag, precon_thblock, Neuman_itr, independ_nvec, &
       nBlock, power_method_itrmax, EigenValueReal, EigenValueImag, EigenVector, EigenErrorNorm, &
       MaxEigenID, MinEigenID, iret, machineflag)
    integer, intent(in) :: icomm, n, gn, nnz, istart, ipreflag, ILU_method, addL, iflagAS
    type(c_ptr), intent(in) :: vecx, vecb, crsA, crsRow_ptr, crsCol
    integer, intent(inout) :: itrmax
    integer, intent(in) :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock, power_method_itrmax
    real(8), target, intent(inout) :: EigenValueReal(*), EigenValueImag(*), EigenVector(*), EigenErrorNorm(*)
    integer, intent(out) :: MaxEigenID, MinEigenID, iret
    integer, optional, intent(in) :: machineflag
    call check(c_dpower(icomm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, ipreflag, &
         ILU_method, addL, iflagAS, itrmax, iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock, &
         power_method_itrmax, c_loc(EigenValueReal), c_loc(EigenValueImag), c_loc(EigenVector), &
         c_loc(EigenErrorNorm), MaxEigenID, MinEigenID, iret))
  end subroutine parcel_dpower_dev

  subroutine parcel_dcaarnoldi_host(icomm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, &
       ipreflag, ILU_method, addL, iflagAS, iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock, &
       CAARNOLDI_sstep, CAARNOLDI_tstep, CAARNOLDI_basis, CAARNOLDI_QRflag, EigenValueReal, EigenValueImag, &
       EigenVector, EigenErrorNorm, MaxEigenID, MinEigenID, iret, machineflag)
    integer, intent(in) :: icomm, n, gn, nnz, istart, ipreflag, ILU_method, addL, iflagAS
    real(8), target, intent(inout) :: vecx(n)
    real(8), target, intent(in) :: vecb(n), crsA(nnz)
    integer, target, intent(in) :: crsRow_ptr(n + 1), crsCol(nnz)
    integer, intent(in) :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
    integer, intent(in) :: CAARNOLDI_sstep, CAARNOLDI_tstep, CAARNOLDI_basis, CAARNOLDI_QRflag
    real(8), target, intent(inout) :: EigenValueReal(*), EigenValueImag(*), EigenVector(*), EigenErrorNorm(*)
    integer, intent(out) :: MaxEigenID, MinEigenID, iret
    integer, optional, intent(in) :: machineflag
    call check(c_dcaarnoldi(icomm, c_loc(vecx), c_loc(vecb), n, gn, nnz, istart, c_loc(crsA), &
         c_loc(crsRow_ptr), c_loc(crsCol), ipreflag, ILU_method, addL, iflagAS, iovlflag, precon_thblock, &
         Neuman_itr, independ_nvec, nBlock, CAARNOLDI_sstep, CAARNOLDI_tstep, CAARNOLDI_basis, &
         CAARNOLDI_QRflag, c_loc(EigenValueReal), c_loc(EigenValueImag), c_loc(EigenVector), &
         c_loc(EigenErrorNorm), MaxEigenID, MinEigenID, iret))
  end subroutine parcel_dcaarnoldi_host

  subroutine parcel_dcaarnoldi_dev(icomm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, &
       ipreflag, ILU_method, addL, iflagAS, iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock, &
       CAARNOLDI_sstep, CAARNOLDI_tstep, CAARNOLDI_basis, CAARNOLDI_QRflag, EigenValueReal, EigenValueImag, &
       EigenVector, EigenErrorNorm, MaxEigenID, MinEigenID, iret, machineflag)
    integer, intent(in) :: icomm, n, gn, nnz, istart, ipreflag, ILU_method, addL, iflagAS
    type(c_ptr), intent(in) :: vecx, vecb, crsA, crsRow_ptr, crsCol
    integer, intent(in) :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
    integer, intent(in) :: CAARNOLDI_sstep, CAARNOLDI_tstep, CAARNOLDI_basis, CAARNOLDI_QRflag
    real(8), target, intent(inout) :: EigenValueReal(*), EigenValueImag(*), EigenVector(*), EigenErrorNorm(*)
    integer, intent(out) :: MaxEigenID, MinEigenID, iret
    integer, optional, intent(in) :: machineflag
    call check(c_dcaarnoldi(icomm, vecx, vecb, n, gn, nnz, istart, crsA, crsRow_ptr, crsCol, ipreflag, &
         ILU_method, addL, iflagAS, iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock, &
         CAARNOLDI_sstep, CAARNOLDI_tstep, CAARNOLDI_basis, CAARNOLDI_QRflag, c_loc(EigenValueReal), &
         c_loc(EigenValueImag), c_loc(EigenVector), c_loc(EigenErrorNorm), MaxEigenID, MinEigenID, iret))
  end subroutine parcel_dcaarnoldi_dev

  ! ---- double-double CG (src/ddwrapper.f90:184-282) ----------------------
  subroutine parcel_ddcg_host(icomm, vecx_hi, vecx_lo, vecb_hi, vecb_lo, n, gn, nnz, istart, crsA_hi, crsA_lo, &
       crsRow_ptr, crsCol, ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, &
       precon_thblock, independ_nvec, nBlock, PRECISION_A, PRECISION_b, PRECISION_x, PRECISION_PRECONDITION, iret)
    integer, intent(in) :: icomm, n, gn, nnz, istart, ipreflag, ILU_method, addL, iflagAS
    real(8), target, intent(inout) :: vecx_hi(n), vecx_lo(n)
    real(8), target, intent(in) :: vecb_hi(n), vecb_lo(n), crsA_hi(nnz), crsA_lo(nnz), rtolmax
    integer, target, intent(in) :: crsRow_ptr(n + 1), crsCol(nnz)
    integer, intent(inout) :: itrmax
    real(8), target, intent(inout) :: reshistory(*)
    integer, intent(in) :: iovlflag, precon_thblock, independ_nvec, nBlock
    integer, intent(in) :: PRECISION_A, PRECISION_b, PRECISION_x, PRECISION_PRECONDITION
    integer, intent(out) :: iret
    call check(c_ddcg(icomm, c_loc(vecx_hi), c_loc(vecx_lo), c_loc(vecb_hi), c_loc(vecb_lo), n, gn, nnz, istart, &
         c_loc(crsA_hi), c_loc(crsA_lo), c_loc(crsRow_ptr), c_loc(crsCol), ipreflag, ILU_method, addL, iflagAS, &
         itrmax, rtolmax, c_loc(reshistory), iovlflag, precon_thblock, independ_nvec, nBlock, PRECISION_A, &
         PRECISION_b, PRECISION_x, PRECISION_PRECONDITION, iret))
  end subroutine parcel_ddcg_host

  subroutine parcel_ddcg_dev(icomm, vecx_hi, vecx_lo, vecb_hi, vecb_lo, n, gn, nnz, istart, crsA_hi, crsA_lo, &
       crsRow_ptr, crsCol, ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, reshistory, iovlflag, &
       precon_thblock, independ_nvec, nBlock, PRECISION_A, PRECISION_b, PRECISION_x, PRECISION_PRECONDITION, iret)
    integer, intent(in) :: icomm, n, gn, nnz, istart, ipreflag, ILU_method, addL, iflagAS
    type(c_ptr), intent(in) :: vecx_hi, vecx_lo, vecb_hi, vecb_lo, crsA_hi, crsA_lo, crsRow_ptr, crsCol
    integer, intent(inout) :: itrmax
    real(8), intent(in) :: rtolmax
    real(8), target, intent(inout) :: reshistory(*)
    integer, intent(in) :: iovlflag, precon_thblock, independ_nvec, nBlock
    integer, intent(in) :: PRECISION_A, PRECISION_b, PRECISION_x, PRECISION_PRECONDITION
    integer, intent(out) :: iret
    call check(c_ddcg(icomm, vecx_hi, vecx_lo, vecb_hi, vecb_lo, n, gn, nnz, istart, crsA_hi, crsA_lo, &
         crsRow_ptr, crsCol, ipreflag, ILU_method, addL, iflagAS, itrmax, rtolmax, c_loc(reshistory), iovlflag, &
         precon_thblock, independ_nvec, nBlock, PRECISION_A, PRECISION_b, PRECISION_x, PRECISION_PRECONDITION, &
         iret))
  end subroutine parcel_ddcg_dev

  ! ---- stencil (DDM) CG (src/src_DDM/ddm_kernel.f90:2266-2360): x, b,
  ! ioff_dia and val_dia are type(c_ptr) in the reference in both builds
  subroutine parcel_dcg_ddm(icomm, x, b, nrow, nnd, ioff_dia, val_dia, num_neighbor_ranks, neighbor_ranks, &
       margin, ncomp_grids, ndiv_grids, num_grids, ioff_grids, div_direc_th, ipreflag, addL, iflagAS, itrmax, &
       rtolmax, reshistory, iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock, iret)
    integer, intent(in) :: icomm, nrow, nnd
    type(c_ptr) :: x, b, ioff_dia, val_dia
    integer, intent(in) :: num_neighbor_ranks, margin, ncomp_grids, div_direc_th
    integer, target, intent(in) :: neighbor_ranks(*)
    integer, target, intent(in) :: ndiv_grids(ncomp_grids), num_grids(ncomp_grids), ioff_grids(nnd, ncomp_grids)
    integer :: ipreflag
    integer, intent(inout) :: addL
    integer, intent(in) :: iflagAS
    integer, intent(inout) :: itrmax
    real(8), intent(inout) :: rtolmax
    real(8), target, intent(inout) :: reshistory(*)
    integer, intent(in) :: iovlflag, precon_thblock, Neuman_itr, independ_nvec, nBlock
    integer, intent(inout) :: iret
    call check(c_dcg_ddm(icomm, x, b, nrow, nnd, ioff_dia, val_dia, num_neighbor_ranks, c_loc(neighbor_ranks), &
         margin, ncomp_grids, c_loc(ndiv_grids), c_loc(num_grids), c_loc(ioff_grids), div_direc_th, ipreflag, &
         addL, iflagAS, itrmax, rtolmax, c_loc(reshistory), iovlflag, precon_thblock, Neuman_itr, &
         independ_nvec, nBlock, iret))
  end subroutine parcel_dcg_ddm

end module PARCEL_wrapper

! parcel_wrapper.f90 -- Fortran module PARCEL_wrapper over libparcel_hip.so.
!
! Keeps the procedure names and argument lists of the reference module
! (src/wrapper.f90): parcel_dcg (:544-632 GPU build, :1995-2083 CPU build),
! parcel_dbicgstab (:362, :1813), parcel_dgmres (:727, :2179), and
! PARCEL_Init / PARCEL_Free (src/c_krylov_wrapper.f90:6-22).  Each generic
! name takes either host arrays (the CPU build's real*8 / integer arrays) or
! device pointers (the GPU build's type(c_ptr)); both forward to the C ABI of
! include/parcel_hip.h.  A library error stops the program with its message,
! as the reference stops on set-up errors (src/krylov.f90:551-560).
module PARCEL_wrapper
  use iso_c_binding
  implicit none
  private
  public :: PARCEL_Init, PARCEL_Free, parcel_dcg, parcel_dbicgstab, parcel_dgmres

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

end module PARCEL_wrapper

! example_poisson.f90 -- calls the PARCEL_wrapper module the way the
! reference's example drivers do (src/example/main.f90): build the 3D Poisson
! system of src/example/problem.f90:448-618,772-875, solve it with
! parcel_dcg / parcel_dbicgstab / parcel_dgmres, print niter, iret and the
! residual history (one value per line, 17 significant digits).
!
! usage: example_poisson solver gnx ipreflag ILU_method nblocks itrmax rtol [m gsflag]
program example_poisson
  use iso_c_binding
  use PARCEL_wrapper
  implicit none
  interface
    integer(c_int) function gen_a(gnx, gny, gnz, nnz, rp, col, val) bind(c, name="parcel_gen_poisson3d_parcel")
      import :: c_int, c_ptr
      integer(c_int), value :: gnx, gny, gnz
      integer(c_int) :: nnz
      type(c_ptr), value :: rp, col, val
    end function
    integer(c_int) function gen_b(gnx, gny, gnz, b, xe) bind(c, name="parcel_gen_poisson3d_rhs_parcel")
      import :: c_int, c_ptr
      integer(c_int), value :: gnx, gny, gnz
      type(c_ptr), value :: b, xe
    end function
  end interface
  integer :: solver, g, ipreflag, ilu, nblk, itrmax, m, gs, n, nnz, iret, i, rc
  real(8) :: rtol
  character(len=32) :: arg
  integer, allocatable, target :: rp(:), col(:)
  real(8), allocatable, target :: val(:), b(:), xe(:), x(:), hist(:)

  call get_command_argument(1, arg); read (arg, *) solver
  call get_command_argument(2, arg); read (arg, *) g
  call get_command_argument(3, arg); read (arg, *) ipreflag
  call get_command_argument(4, arg); read (arg, *) ilu
  call get_command_argument(5, arg); read (arg, *) nblk
  call get_command_argument(6, arg); read (arg, *) itrmax
  call get_command_argument(7, arg); read (arg, *) rtol
  m = 4
  gs = 1
  if (command_argument_count() >= 9) then
    call get_command_argument(8, arg); read (arg, *) m
    call get_command_argument(9, arg); read (arg, *) gs
  end if
  n = g * g * g
  nnz = 0
  rc = gen_a(g, g, g, nnz, c_null_ptr, c_null_ptr, c_null_ptr)
  allocate (rp(n + 1), col(nnz), val(nnz), b(n), xe(n), x(n), hist(max(itrmax, m + 1)))
  rc = gen_a(g, g, g, nnz, c_loc(rp), c_loc(col), c_loc(val))
  rc = gen_b(g, g, g, c_loc(b), c_loc(xe))
  x = 0.0d0
  hist = 0.0d0
  call PARCEL_Init(0)
  select case (solver)
  case (1)
    call parcel_dcg(0, x, b, n, n, nnz, 1, val, rp, col, ipreflag, ilu, 0, 4, itrmax, rtol, hist, &
         0, nblk, 1, 1, 2000, iret)
  case (2)
    call parcel_dbicgstab(0, x, b, n, n, nnz, 1, val, rp, col, ipreflag, ilu, 0, 4, itrmax, rtol, hist, &
         0, nblk, 1, 1, 2000, iret)
  case (3)
    call parcel_dgmres(0, x, b, n, n, nnz, 1, val, rp, col, ipreflag, ilu, 0, 4, itrmax, rtol, hist, &
         0, nblk, 1, 1, 2000, m, gs, iret)
  end select
  call PARCEL_Free()
  write (*, '(a,i0,a,i0)') 'niter ', itrmax, ' iret ', iret
  do i = 1, itrmax
    write (*, '(es25.17e3)') hist(i)
  end do
  write (*, '(a,es25.17e3)') 'x1 ', x(1)
  write (*, '(a,es25.17e3)') 'xn ', x(n)
end program example_poisson

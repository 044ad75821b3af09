! example_poisson.f90 -- calls the PARCEL_wrapper module the way the
! reference's example drivers do (src/example/main.f90): build the 3D Poisson
! system of src/example/problem.f90:448-618,772-875, solve it with one of
! the module's solvers, print niter, iret and the residual history (one
! value per line, 17 significant digits) and the first and last entries of x.
!
! usage: example_poisson solver gnx ipreflag ILU_method nblocks itrmax rtol [p1 p2 p3 p4]
!   solver 1 parcel_dcg, 2 parcel_dbicgstab, 3 parcel_dgmres (p1 = m, p2 = GS flag),
!   4 parcel_dcagmres (p1 = s, p2 = t, p3 = basis, p4 = QR flag),
!   5 parcel_dcbcg (p1 = k, p2 = eigenvalue-bound flag),
!   6 parcel_ddcg (double-double, lo parts zero; x printed as hi),
!   7 parcel_dcg_ddm (the same grid as the ghost-padded stencil of
!     src/example/ddm_setup.f90, margin 1; x printed at the first and last
!     interior cells; ILU_method and nblocks unused, p1 = addL),
!   8 parcel_dcg_dia, 9 parcel_dgmres_dia (p1 = m, p2 = GS flag): the same
!     matrix as diagonals (make_poisson2D_matrix_DIA_from_CRS's layout)
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
    integer(c_int) function crs_to_dia(n, rp, col, val, nnd, offset, dia) bind(c, name="parcel_gen_crs_to_dia")
      import :: c_int, c_ptr
      integer(c_int), value :: n
      type(c_ptr), value :: rp, col, val, offset, dia
      integer(c_int) :: nnd
    end function
    integer(c_int) function gen_ddm(gnx, gny, gnz, margin, val_dia, ioff_dia, ioff_grids, b, xe) &
         bind(c, name="parcel_gen_poisson3d_ddm")
      import :: c_int, c_ptr
      integer(c_int), value :: gnx, gny, gnz, margin
      type(c_ptr), value :: val_dia, ioff_dia, ioff_grids, b, xe
    end function
  end interface
  integer :: solver, g, ipreflag, ilu, nblk, itrmax, n, nnz, iret, i, rc, k, m, addl, nnd
  integer :: p(4)
  real(8) :: rtol
  character(len=32) :: arg
  integer, allocatable, target :: rp(:), col(:), ioff(:), iog(:, :), doff(:)
  real(8), allocatable, target :: val(:), vlo(:), b(:), blo(:), xe(:), x(:), xlo(:), hist(:), vd(:), dval(:)
  integer, target :: nbr(1), ndiv(3), ngrid(3)

  call get_command_argument(1, arg); read (arg, *) solver
  call get_command_argument(2, arg); read (arg, *) g
  call get_command_argument(3, arg); read (arg, *) ipreflag
  call get_command_argument(4, arg); read (arg, *) ilu
  call get_command_argument(5, arg); read (arg, *) nblk
  call get_command_argument(6, arg); read (arg, *) itrmax
  call get_command_argument(7, arg); read (arg, *) rtol
  p = [4, 1, 0, 4]
  if (solver == 5) p = [4, 1, 0, 0]
  do k = 1, 4
    if (command_argument_count() >= 7 + k) then
      call get_command_argument(7 + k, arg); read (arg, *) p(k)
    end if
  end do
  n = g * g * g
  allocate (hist(itrmax + 64))
  hist = 0.0d0
  call PARCEL_Init(0)
  if (solver == 7) then
    m = (g + 2) ** 3
    addl = p(1)
    allocate (vd(7 * m), ioff(7), iog(7, 3), b(m), xe(m), x(m))
    rc = gen_ddm(g, g, g, 1, c_loc(vd), c_loc(ioff), c_loc(iog), c_loc(b), c_loc(xe))
    x = 0.0d0
    nbr = 0
    ndiv = 1
    ngrid = g
    call parcel_dcg_ddm(0, c_loc(x), c_loc(b), m, 7, c_loc(ioff), c_loc(vd), 0, nbr, 1, 3, ndiv, ngrid, iog, 1, &
         ipreflag, addl, 0, itrmax, rtol, hist, 0, -1, 1, 1, 2000, iret)
    call PARCEL_Free()
    write (*, '(a,i0,a,i0)') 'niter ', itrmax, ' iret ', iret
    do i = 1, itrmax
      write (*, '(es25.17e3)') hist(i)
    end do
    write (*, '(a,es25.17e3)') 'x1 ', x(1 + 1 + (g + 2) + (g + 2) ** 2)
    write (*, '(a,es25.17e3)') 'xn ', x(g + 1 + (g + 2) * g + (g + 2) ** 2 * g)
    stop
  end if
  nnz = 0
  rc = gen_a(g, g, g, nnz, c_null_ptr, c_null_ptr, c_null_ptr)
  allocate (rp(n + 1), col(nnz), val(nnz), vlo(nnz), b(n), blo(n), xe(n), x(n), xlo(n))
  rc = gen_a(g, g, g, nnz, c_loc(rp), c_loc(col), c_loc(val))
  rc = gen_b(g, g, g, c_loc(b), c_loc(xe))
  x = 0.0d0
  xlo = 0.0d0
  vlo = 0.0d0
  blo = 0.0d0
  if (solver == 8 .or. solver == 9) then
    nnd = 0
    rc = crs_to_dia(n, c_loc(rp), c_loc(col), c_loc(val), nnd, c_null_ptr, c_null_ptr)
    allocate (doff(nnd), dval(nnd * n))
    rc = crs_to_dia(n, c_loc(rp), c_loc(col), c_loc(val), nnd, c_loc(doff), c_loc(dval))
  end if
  select case (solver)
  case (1)
    call parcel_dcg(0, x, b, n, n, nnz, 1, val, rp, col, ipreflag, ilu, 0, 4, itrmax, rtol, hist, &
         0, nblk, 1, 1, 2000, iret)
  case (2)
    call parcel_dbicgstab(0, x, b, n, n, nnz, 1, val, rp, col, ipreflag, ilu, 0, 4, itrmax, rtol, hist, &
         0, nblk, 1, 1, 2000, iret)
  case (3)
    call parcel_dgmres(0, x, b, n, n, nnz, 1, val, rp, col, ipreflag, ilu, 0, 4, itrmax, rtol, hist, &
         0, nblk, 1, 1, 2000, p(1), p(2), iret)
  case (4)
    call parcel_dcagmres(0, x, b, n, n, nnz, 1, val, rp, col, ipreflag, ilu, 0, 4, itrmax, rtol, hist, &
         0, nblk, 1, 1, 2000, p(1), p(2), p(3), p(4), iret)
  case (5)
    call parcel_dcbcg(0, x, b, n, n, nnz, 1, val, rp, col, ipreflag, ilu, 0, 4, itrmax, rtol, hist, &
         0, nblk, 1, 1, 2000, p(1), p(2), 100, 4, 10, 1, 3, iret)
  case (6)
    call parcel_ddcg(0, x, xlo, b, blo, n, n, nnz, 1, val, vlo, rp, col, ipreflag, ilu, 0, 4, itrmax, rtol, &
         hist, 0, nblk, 1, 2000, 2, 2, 2, 2, iret)
  case (8)
    call parcel_dcg_dia(0, x, b, n, n, 1, dval, doff, nnd, ipreflag, ilu, 0, 4, itrmax, rtol, hist, &
         0, nblk, 1, 1, 2000, iret)
  case (9)
    call parcel_dgmres_dia(0, x, b, n, n, 1, dval, doff, nnd, ipreflag, ilu, 0, 4, itrmax, rtol, hist, &
         0, nblk, 1, 1, 2000, p(1), p(2), iret)
  end select
  call PARCEL_Free()
  write (*, '(a,i0,a,i0)') 'niter ', itrmax, ' iret ', iret
  do i = 1, itrmax
    write (*, '(es25.17e3)') hist(i)
  end do
  write (*, '(a,es25.17e3)') 'x1 ', x(1)
  write (*, '(a,es25.17e3)') 'xn ', x(n)
end program example_poisson

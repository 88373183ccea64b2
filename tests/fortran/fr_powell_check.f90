! fr_powell_check.f90 -- a Fortran host minimising with the library's VA04A
! (fr_powell_minimize, through frealign_hip_mod) and with the reference's own
! VA04A (va04.f, flang-compiled into oracle/_ref with a test CALCFX), on the
! same objective.  Prints the number of differing bits patterns; 0 means the
! drop-in reproduces the reference's iterates.
module fr_powell_objective
  use iso_c_binding
  implicit none
contains
  ! the Rosenbrock chain of the test CALCFX (IBUF = 2), same FP32 operations
  function rosen(n, x, user) bind(C) result(rf)
    integer(c_int), value :: n
    real(c_float) :: x(n)
    type(c_ptr), value :: user
    real(c_float) :: rf
    real(c_float) :: u, v
    integer :: i
    rf = 0.0
    if (n == 1) then
      rf = (1.0 - x(1)) * (1.0 - x(1))
    else
      do i = 1, n - 1
        u = x(i + 1) - x(i) * x(i)
        v = 1.0 - x(i)
        rf = rf + 100.0 * (u * u) + v * v
      end do
    end if
  end function
end module

program fr_powell_check
  use iso_c_binding
  use frealign_hip
  use fr_powell_objective
  implicit none
  interface
    subroutine ref_va04a(x, e, n, escale, icon, maxit, fn, f, nevals) bind(C, name="ref_va04a")
      import :: c_int, c_float
      integer(c_int), value :: n, icon, maxit, fn
      real(c_float), value :: escale
      real(c_float) :: x(*), e(*), f
      integer(c_int) :: nevals
    end subroutine
  end interface
  real(c_float) :: x(6), xr(6), e(6), f, fr
  integer(c_int) :: nev, nevr, n, nbad, icon
  nbad = 0
  e = [0.1, 0.1, 0.1, 0.01, 0.01, 0.001]
  do icon = 1, 2
    do n = 2, 6
      x = [-0.3, 0.2, 0.45, -0.1, 0.05, 0.3]
      xr = x
      call fr_check(fr_powell_minimize(n, x, e, 100.0, icon, 50, c_funloc(rosen), c_null_ptr, &
                                       f, nev), 'fr_powell_minimize')
      call ref_va04a(xr, e, n, 100.0, icon, 50, 2, fr, nevr)
      if (nev /= nevr) nbad = nbad + 1
      if (any(transfer(x(1:n), 1_c_int, n) /= transfer(xr(1:n), 1_c_int, n))) nbad = nbad + 1
      if (transfer(f, 1_c_int) /= transfer(fr, 1_c_int)) nbad = nbad + 1
    end do
  end do
  write(*, '(A,I0)') 'FR_POWELL_CHECK mismatches: ', nbad
end program

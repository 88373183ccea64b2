! ref_va04_obj.f90 -- test objectives for the reference's VA04A (va04.f)
! compiled in isolation (test infrastructure; built into oracle/_ref only).
!
! va04.f calls CALCFX with the reference's full argument list.  This CALCFX
! takes the same list and returns a closed-form function of X, selected by
! IBUF, so that the minimiser's iterates can be compared with the oracle's
! restatement (fr_oracle_powell.c, or_va04_testfn: same FP32 operation order).
module ref_va04_count
  implicit none
  integer :: ncalls = 0
end module

subroutine calcfx(nx, xparmap, rf, xpar, npar, nsam, maxr1, maxr2, mask, a3df, irad, pbuf, &
                  shx, shy, cs, wl, wgh1, wgh2, thetatr, ctff, amag, rih, halfw, ri2, ri3, &
                  ri4, phi, theta, psi, rmax1, rmax2, xstd, mbuf, ilst, datc, ibuf, b3dv, &
                  datd, sinclut, ivflag, rbfact, outd, outc, qbuc, ipad, rbuf, fsct, dfmid1, &
                  dfmid2, angast, iewald, tx, ty, xm, ym, sx, sy, sig2, dfstd, asym, thetam, &
                  stheta, psim, spsi, stif, pweights, pssnr, fftw_plans, sm)
  use iso_c_binding
  use ref_va04_count
  implicit none
  integer :: nx, npar, nsam, maxr1, maxr2, mask(5), irad, ilst(*), ibuf, ivflag, ipad, iewald
  real :: xparmap(*), rf, xpar(*), pbuf(*), shx(*), shy(*), cs, wl, wgh1, wgh2, thetatr
  real :: amag(*), rih, halfw, ri2, ri3, ri4, phi(*), theta(*), psi(*), rmax1, rmax2, xstd
  real :: mbuf(*), b3dv(*), datd(*), sinclut(*), rbfact, outd(*), rbuf(*), fsct(*)
  real :: dfmid1(*), dfmid2(*), angast(*), tx, ty, xm, ym, sx, sy, sig2, dfstd, thetam
  real :: stheta, psim, spsi, stif, pweights(*), pssnr(*), sm(9)
  complex :: a3df(*), ctff(*), datc(*), outc(*), qbuc(*)
  character(len=*) :: asym
  type(c_ptr) :: fftw_plans(*)
  integer :: i
  real :: t, u, v, q
  ncalls = ncalls + 1
  rf = 0.0
  if (ibuf == 1) then
    do i = 1, nx
      t = xparmap(i) - 0.1 * real(i)
      rf = rf + (real(i) * 0.7) * (t * t)
    end do
    if (nx >= 2) rf = rf + 0.4 * (xparmap(1) - 0.1) * (xparmap(2) - 0.2)
  else if (ibuf == 2) then
    if (nx == 1) then
      rf = (1.0 - xparmap(1)) * (1.0 - xparmap(1))
    else
      do i = 1, nx - 1
        u = xparmap(i + 1) - xparmap(i) * xparmap(i)
        v = 1.0 - xparmap(i)
        rf = rf + 100.0 * (u * u) + v * v
      end do
    end if
  else
    q = 0.0
    do i = 1, nx
      t = xparmap(i) - 0.05 * real(i)
      q = q + (1.0 + 0.5 * real(i)) * (t * t)
    end do
    rf = -1.0 / (1.0 + q)
  end if
end subroutine

! VA04A(X, E, N, F, ESCALE, IPRINT=0, ICON, MAXIT, ...) on test objective fn
! (1, 2 or 3); x in/out, returns F and the number of CALCFX calls.
subroutine ref_va04a(x, e, n, escale, icon, maxit, fn, f, nevals) bind(C, name="ref_va04a")
  use iso_c_binding
  use ref_va04_count
  implicit none
  integer(c_int), value :: n, icon, maxit, fn
  real(c_float), value :: escale
  real(c_float) :: x(*), e(*), f
  integer(c_int) :: nevals
  real :: xpar(6), shx(1), shy(1), pbuf(1), phi(1), theta(1), psi(1), mbuf(1), b3dv(1)
  real :: datd(1), sinclut(1), outd(1), rbuf(1), fsct(1), dfmid1(1), dfmid2(1), angast(1)
  real :: pweights(1), pssnr(1), sm(9), amag, rmax1, rmax2, xstd, rbfact, dumr
  complex :: c3df(1), ctff(1), datc(1), outc(1), qbuc(1)
  integer :: mask(5), ilst(1), npar, ibuf, ivflag
  character(len=3) :: asym
  type(c_ptr) :: plans(10)
  npar = n
  mask = 1
  ibuf = fn
  ivflag = 0
  asym = 'C  '
  xpar = 0.0
  sm = 0.0
  amag = 1.0
  rmax1 = 0.0
  rmax2 = 0.0
  xstd = 0.0
  rbfact = 0.0
  dumr = 0.0
  ncalls = 0
  call va04a(x, e, n, f, escale, 0, icon, maxit, xpar, npar, 64, 1, 2, mask, c3df, 1, pbuf, &
             shx, shy, 1, 0.0, 0.0, 0.0, 0.0, 0.0, ctff, amag, 0.0, 0.0, 0.0, 0.0, 0.0, phi, &
             theta, psi, rmax1, rmax2, xstd, mbuf, ilst, datc, ibuf, b3dv, datd, sinclut, ivflag, &
             rbfact, outd, outc, qbuc, rbuf, fsct, dfmid1, dfmid2, angast, 0, 0.0, 0.0, 0.0, 0.0, &
             1.0, 1.0, 0.0, dumr, asym, 0.0, 0.0, 0.0, 0.0, 0.0, pweights, pssnr, plans, sm)
  nevals = ncalls
end subroutine

! fr_percall_check.f90 -- a Fortran host that calls PEXTRACT, CC3M, PRESB and
! SIGMA exactly as LMAIN does (implicit interfaces, reference argument lists)
! and is linked against libfrealign_hip.so instead of the reference's
! routines.  Inputs come from a stream file written by the test; the results
! go to another stream file for comparison with the oracle.
program fr_percall_check
  implicit none
  integer :: nsam, jc, ic, nvol, maxr1, maxr2, irad, ipad, iewald, k
  real :: phi, theta, psi, shx, shy, amag, ri2, ri3, rih, halfw, rbfact, rclas, sigp, signset
  real :: dmask(4), sm(9), thetatr, sigma2n, sigma2, logpc, ccv, sinclut(1), pweights(1)
  real :: cs, wl, wgh1, wgh2, df1, df2, angast
  complex, allocatable :: a3df(:), spec(:), datc(:), ctff(:), outc(:), pbuc(:)
  real, allocatable :: fsct(:), rbin(:), rbuf(:), bufa(:), bufb(:), bufab(:)
  integer, allocatable :: ns(:)
  integer(8) :: plans(10)
  character(len=3) :: asym
  real :: cc3m
  external cc3m, pextract, presb, sigma
  character(len=512) :: fin, fout
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) nsam, maxr1, maxr2
  jc = nsam / 2 + 1
  ic = nsam * jc
  nvol = ic * nsam
  allocate(a3df(nvol), spec(ic), datc(ic), ctff(2 * ic), outc(ic), pbuc(ic), fsct(jc), &
           rbin(nsam), rbuf(4 * nsam * nsam + 16 * nsam), bufa(nsam + 2), bufb(nsam + 2), &
           bufab(nsam * nsam + 3 * nsam + 8), ns(nsam + 2))
  read(10) phi, theta, psi, shx, shy, amag, ri2, ri3, rih, halfw, rbfact, rclas, sigp, signset
  read(10) dmask, sm
  read(10) a3df, spec, datc, ctff, fsct, rbin
  close(10)
  irad = 1
  ipad = 1
  iewald = 0
  thetatr = 0.0
  asym = 'D  '
  cs = 0.0; wl = 0.0; wgh1 = 0.0; wgh2 = 0.0; df1 = 0.0; df2 = 0.0; angast = 0.0
  k = 1
  plans = 0
  call pextract(nsam, irad, outc, a3df, phi, theta, psi, sinclut, ipad, thetatr, iewald, ctff, sm)
  ccv = cc3m(nsam, irad, amag, spec, a3df, maxr1, maxr2, phi, theta, psi, shx, shy, 0, pbuc, &
             rbfact, sinclut, ipad, rbuf, fsct, iewald, thetatr, ctff, ri2, ri3, rih, halfw, &
             pweights, plans, sm)
  call presb(nsam, irad, amag, spec, a3df, maxr1, maxr2, phi, theta, psi, shx, shy, rbin, &
             sinclut, ipad, ns, bufa, bufb, bufab, cs, wl, wgh1, wgh2, df1, df2, angast, thetatr, &
             iewald, ctff, ri2, ri3, rih, halfw, asym, plans, sm)
  call sigma(nsam, irad, amag, a3df, datc, phi, theta, psi, shx, shy, sinclut, ipad, iewald, &
             thetatr, ctff, rbuf, ri2, maxr1, maxr2, signset, sigma2n, sigma2, logpc, dmask, &
             asym, k, sigp, pweights, rclas, plans, sm)
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) outc, ccv, rbin, sigma2n, sigma2, logpc
  close(11)
  write(*, '(A)') 'FR_PERCALL_CHECK done'
end program

! fr_percall_check.f90 -- a Fortran host that calls CTFAPPLY, PEXTRACT, CC3M,
! PRESB, SIGMA and PINSERT_TRILIN as LMAIN and PINSERT do (implicit
! interfaces, reference argument lists)
! and is linked against libfrealign_hip.so instead of the reference's
! routines.  Inputs come from a stream file written by the test; the results
! go to another stream file for comparison with the oracle.
program fr_percall_check
  implicit none
  integer :: nsam, jc, ic, nvol, maxr1, maxr2, irad, ipad, iewald, k
  real :: phi, theta, psi, shx, shy, amag, ri2, ri3, rih, halfw, rbfact, rclas, sigp, signset
  real :: dmask(4), sm(9), thetatr, sigma2n, sigma2, logpc, ccv, sinclut(1), pweights(1)
  real :: cs, wl, wgh1, wgh2, df1, df2, angast
  ! CTFAPPLY inputs
  real :: ccs, cwl, cwgh1, cwgh2, cdf1, cdf2, cang, cthetatr, camagp, crih, chalfw, cri2, cri3
  real :: ri4, xstd, tx, ty
  ! PINSERT_TRILIN inputs
  integer :: irrec2, ip, nsamh, irada, nnstat
  real :: occ, tmp, dx, dy, stw, rii, dmi(9)
  integer(8) :: nksum
  double precision :: ttd
  complex, allocatable :: a3df(:), spec(:), datc(:), ctff(:), outc(:), pbuc(:)
  complex, allocatable :: ctf2(:), outd(:), datc2(:), b3df(:), c3df(:)
  real, allocatable :: fsct(:), rbin(:), rbuf(:), bufa(:), bufb(:), bufab(:), pssnr(:)
  real, allocatable :: t3df(:), datd(:), b3dv(:), wbuf(:), asum(:)
  integer, allocatable :: ns(:), ksum(:)
  integer(8) :: plans(10)
  character(len=3) :: asym
  real :: cc3m
  external cc3m, pextract, presb, sigma, ctfapply, pinsert_trilin
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
  allocate(ctf2(2 * ic), outd(ic), datc2(ic), pssnr(jc + 1), b3df(nvol), t3df(nvol), &
           c3df(1), datd(1), b3dv(1), wbuf(1), asum(1), ksum(1))
  read(10) ccs, cwl, cwgh1, cwgh2, cdf1, cdf2, cang, cthetatr, camagp, crih, chalfw, cri2, cri3
  read(10) pssnr
  read(10) irrec2, ip
  read(10) occ, tmp, dx, dy, stw
  read(10) dmi, b3df, t3df, nksum, ttd
  close(10)
  irad = 1
  ipad = 1
  iewald = 0
  thetatr = 0.0
  asym = 'D  '
  cs = 0.0; wl = 0.0; wgh1 = 0.0; wgh2 = 0.0; df1 = 0.0; df2 = 0.0; angast = 0.0
  k = 1
  plans = 0
  ri4 = 0.0; xstd = 0.0; tx = 0.0; ty = 0.0
  call ctfapply(nsam, spec, shx, shy, ccs, cwl, cwgh1, cwgh2, cdf1, cdf2, cang, cthetatr, ctf2, &
                outd, outd, camagp, crih, chalfw, cri2, cri3, ri4, datd, datc2, b3dv, phi, theta, &
                psi, wbuf, xstd, 0, tx, ty, asym, pssnr, rbuf, pweights, 1, plans)
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
  nsamh = nsam / 2
  irada = 1
  nnstat = 0
  rii = 0.0
  call pinsert_trilin(nsam, rii, irrec2, spec, ctff, occ, b3df, t3df, ttd, shx, shy, asum, &
                      asum, asum, c3df, ksum, sinclut, irada, ipad, thetatr, iewald, nnstat, &
                      nksum, jc, dmi, tmp, nsamh, ic, dx, dy, stw, ip, ns)
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) outc, ccv, rbin, sigma2n, sigma2, logpc
  write(11) ctf2, outd, datc2, b3df, t3df, nksum, ttd
  close(11)
  write(*, '(A)') 'FR_PERCALL_CHECK done'
end program

! ref_bridge.f90 -- C-callable wrappers around the flang-compiled reference
! routines, for pinning the oracle (TEST INFRASTRUCTURE ONLY).
!
! This file is project code: it contains no reference source.  It opens the
! log unit SEARCHANG writes to (searchang.f:138-141) so that the routine can be
! called outside the reference's driver, and gives the COMPLEX FUNCTIONs
! CTF / AINTERPO3DS a subroutine form with plain by-reference outputs.

subroutine ref_searchang(rmm, ri, dangin, nsang, iquadmax, asymc, nasym, sang, &
                         asymstore, mask, store) bind(C, name="ref_searchang")
  use iso_c_binding
  implicit none
  real(c_float), intent(in) :: rmm, ri, dangin
  integer(c_int), intent(inout) :: nsang, iquadmax
  character(kind=c_char), intent(in) :: asymc
  integer(c_int), intent(in) :: nasym, asymstore, store
  real(c_float), intent(inout) :: sang(3, *)
  integer(c_int), intent(in) :: mask(5)
  character(len=3) :: asym
  logical :: lstore, lasym
  integer :: ndoc1, nset, m(5)
  logical, save :: opened = .false.
  asym = asymc // '  '
  ndoc1 = 70
  nset = 1
  m = mask
  lstore = (store /= 0)
  lasym = (asymstore /= 0)
  if (.not. opened) then
    open(unit=ndoc1 + nset, file='/dev/null', status='unknown')
    opened = .true.
  end if
  call searchang(rmm, ri, dangin, nsang, iquadmax, asym, nasym, sang, ndoc1, nset, &
                 lasym, m, lstore)
end subroutine

subroutine ref_ctf(cs, wl, wgh1, wgh2in, df1, df2, angast, thetatr, ix, iy, tx, ty, out) &
    bind(C, name="ref_ctf")
  use iso_c_binding
  implicit none
  real(c_float), intent(in) :: cs, wl, wgh1, wgh2in, df1, df2, angast, thetatr, tx, ty
  integer(c_int), intent(in) :: ix, iy
  real(c_float), intent(out) :: out(2)
  complex :: ctf, c
  external ctf
  integer :: jx, jy
  real :: a1, a2, a3, a4, a5, a6, a7, a8, a9, a10
  a1 = cs; a2 = wl; a3 = wgh1; a4 = wgh2in; a5 = df1; a6 = df2; a7 = angast
  a8 = thetatr; a9 = tx; a10 = ty
  jx = ix; jy = iy
  c = ctf(a1, a2, a3, a4, a5, a6, a7, a8, jx, jy, a9, a10)
  out(1) = real(c)
  out(2) = aimag(c)
end subroutine

subroutine ref_ainterpo3ds(nsam, a3df, xx, yy, zz, out) bind(C, name="ref_ainterpo3ds")
  use iso_c_binding
  implicit none
  integer(c_int), intent(in) :: nsam
  complex(c_float_complex), intent(in) :: a3df(*)
  real(c_float), intent(in) :: xx, yy, zz
  real(c_float), intent(out) :: out(2)
  complex :: ainterpo3ds, c
  external ainterpo3ds
  real :: sinclut(1), x, y, z
  integer :: n, irad, ipad
  n = nsam; irad = 1; ipad = 1
  x = xx; y = yy; z = zz
  sinclut(1) = 0.0
  c = ainterpo3ds(n, irad, a3df, x, y, z, sinclut, ipad)
  out(1) = real(c)
  out(2) = aimag(c)
end subroutine

subroutine ref_random(is, out) bind(C, name="ref_random")
  use iso_c_binding
  implicit none
  integer(c_int), intent(inout) :: is
  real(c_float), intent(out) :: out
  real :: random
  external random
  integer :: s
  s = is
  out = random(s)
  is = s
end subroutine

subroutine ref_maskcos(nsam, outd, ri2, ri3, rih, halfw, amagp, origin) bind(C, name="ref_maskcos")
  use iso_c_binding
  implicit none
  integer(c_int), intent(in) :: nsam, origin
  real(c_float), intent(inout) :: outd(*)
  real(c_float), intent(in) :: ri2, ri3, rih, halfw, amagp
  integer :: n
  real :: a, b, c, d, e
  n = nsam; a = ri2; b = ri3; c = rih; d = halfw; e = amagp
  if (origin /= 0) then
    call maskcos2(n, outd, a, b, c, d, e)
  else
    call maskcos(n, outd, a, b, c, d, e)
  end if
end subroutine

subroutine ref_pweight(nsam, outc, ssnr, ctff, ictf) bind(C, name="ref_pweight")
  use iso_c_binding
  implicit none
  integer(c_int), intent(in) :: nsam, ictf
  complex(c_float_complex), intent(inout) :: outc(*)
  real(c_float), intent(in) :: ssnr(*)
  complex(c_float_complex), intent(in) :: ctff(*)
  real, allocatable :: pweights(:), rbuf(:)
  integer :: n, ic
  n = nsam; ic = ictf
  allocate(pweights(nsam + 2), rbuf(2 * nsam + 8))
  rbuf = 0.0
  ! CTFAPPLY passes OUTC for both SPEC and OUTC (ctfapply.f:157)
  call pweight(n, outc, outc, ssnr, ctff, pweights, rbuf, ic)
  deallocate(pweights, rbuf)
end subroutine

subroutine ref_pextract(nsam, outc, a3df, phi, theta, psi, ctff, sm) bind(C, name="ref_pextract")
  use iso_c_binding
  implicit none
  integer(c_int), intent(in) :: nsam
  complex(c_float_complex), intent(out) :: outc(*)
  complex(c_float_complex), intent(in) :: a3df(*), ctff(*)
  real(c_float), intent(in) :: phi, theta, psi, sm(9)
  real :: sinclut(1), thet, p1, p2, p3, smc(9)
  integer :: n, irad, ipad, iewald
  n = nsam; irad = 1; ipad = 1; iewald = 0; thet = 0.0; sinclut(1) = 0.0
  p1 = phi; p2 = theta; p3 = psi; smc = sm
  call pextract(n, irad, outc, a3df, p1, p2, p3, sinclut, ipad, thet, iewald, ctff, smc)
end subroutine

! pinsert.f:8 with A3D3_S's argument mapping (a3d3_s.f:65-74): the same
! volume pair is passed as A3DF/S3DF and B3DF/T3DF, IMP=1, INTERP=1,
! IEWALD=0, NNSTAT=0.  nksum and ttd accumulate (NKSUM1(I), TTD1(IS)).
subroutine ref_pinsert(nsam, rrec, amag, spec, ctff, b3df, t3df, ttd, phi, theta, psi, &
                       shx, shy, presa, pbc, boff, nsym, jsym, symop, nksum, psize, occ, &
                       sm, nonsym) bind(C, name="ref_pinsert")
  use iso_c_binding
  implicit none
  integer(c_int), intent(in) :: nsam, nsym, nonsym
  real(c_float), intent(in) :: rrec, amag, phi, theta, psi, shx, shy, presa, pbc, boff, psize, occ
  complex(c_float_complex), intent(in) :: spec(*), ctff(*)
  complex(c_float_complex), intent(inout) :: b3df(*)
  real(c_float), intent(inout) :: t3df(*)
  real(c_double), intent(inout) :: ttd
  integer(c_int), intent(in) :: jsym(*)
  real(c_float), intent(in) :: symop(3, 3, *), sm(9)
  integer(c_int64_t), intent(inout) :: nksum
  integer :: n, irad, irada, ipad, iewald, isymax, nnstat, imp, interp, ns(8), ns1(8), ksum(1), js(8)
  real :: ri, asum(1), vsum(1), psum(1), sinclut(1), thet, smc(9), symc(3, 3, 8)
  real :: a_rrec, a_amag, a_phi, a_theta, a_psi, a_shx, a_shy, a_presa, a_pbc, a_boff, a_psize, a_occ
  complex :: c3df(1)
  integer(8) :: nk, nk1(1)
  double precision :: std, ttd1(1)
  logical :: lnonsym
  integer :: nsy, i
  n = nsam; irad = 1; irada = 1; ipad = 1; iewald = 0; nnstat = 0; imp = 1; interp = 1
  ri = 0.0; thet = 0.0; sinclut(1) = 0.0
  nsy = nsym; isymax = max(nsym, 1)
  do i = 1, nsym
    js(i) = jsym(i)
    symc(:, :, i) = symop(:, :, i)
  end do
  smc = sm
  a_rrec = rrec; a_amag = amag; a_phi = phi; a_theta = theta; a_psi = psi
  a_shx = shx; a_shy = shy; a_presa = presa; a_pbc = pbc; a_boff = boff
  a_psize = psize; a_occ = occ
  lnonsym = (nonsym /= 0)
  std = 0.0d0
  nk = 0
  nk1(1) = nksum
  ttd1(1) = ttd
  call pinsert(n, irad, ri, a_rrec, a_amag, spec, ctff, b3df, t3df, b3df, t3df, std, &
               a_phi, a_theta, a_psi, a_shx, a_shy, a_presa, a_pbc, a_boff, asum, vsum, psum, &
               c3df, ksum, sinclut, irada, ipad, thet, iewald, nsy, isymax, js, symc, nnstat, &
               nk, imp, a_psize, nk1, ttd1, ns, ns1, a_occ, interp, smc, lnonsym)
  nksum = nk1(1)
  ttd = ttd1(1)
end subroutine

! pinsert.f:8 as ref_pinsert, with NNSTAT=1: PINSERT dispatches to
! PINSERT_TRILIN_SE (pinsert.f:137-176), which adds the per-voxel statistics
! KSUM, VSUM, PSUM, ASUM against the reference transform C3DF (IPAD=1).
subroutine ref_pinsert_se(nsam, rrec, amag, spec, ctff, b3df, t3df, ttd, phi, theta, psi, &
                          shx, shy, presa, pbc, boff, nsym, jsym, symop, nksum, psize, occ, &
                          sm, nonsym, c3df, ksum, vsum, psum, asum) bind(C, name="ref_pinsert_se")
  use iso_c_binding
  implicit none
  integer(c_int), intent(in) :: nsam, nsym, nonsym
  real(c_float), intent(in) :: rrec, amag, phi, theta, psi, shx, shy, presa, pbc, boff, psize, occ
  complex(c_float_complex), intent(in) :: spec(*), ctff(*), c3df(*)
  complex(c_float_complex), intent(inout) :: b3df(*)
  real(c_float), intent(inout) :: t3df(*), vsum(*), psum(*), asum(*)
  integer(c_int), intent(inout) :: ksum(*)
  real(c_double), intent(inout) :: ttd
  integer(c_int), intent(in) :: jsym(*)
  real(c_float), intent(in) :: symop(3, 3, *), sm(9)
  integer(c_int64_t), intent(inout) :: nksum
  integer :: n, irad, irada, ipad, iewald, isymax, nnstat, imp, interp, ns(8), ns1(8), js(8)
  real :: ri, sinclut(1), thet, smc(9), symc(3, 3, 8)
  real :: a_rrec, a_amag, a_phi, a_theta, a_psi, a_shx, a_shy, a_presa, a_pbc, a_boff, a_psize, a_occ
  integer(8) :: nk, nk1(1)
  double precision :: std, ttd1(1)
  logical :: lnonsym
  integer :: nsy, i
  n = nsam; irad = 1; irada = 1; ipad = 1; iewald = 0; nnstat = 1; imp = 1; interp = 1
  ri = 0.0; thet = 0.0; sinclut(1) = 0.0
  nsy = nsym; isymax = max(nsym, 1)
  do i = 1, nsym
    js(i) = jsym(i)
    symc(:, :, i) = symop(:, :, i)
  end do
  smc = sm
  a_rrec = rrec; a_amag = amag; a_phi = phi; a_theta = theta; a_psi = psi
  a_shx = shx; a_shy = shy; a_presa = presa; a_pbc = pbc; a_boff = boff
  a_psize = psize; a_occ = occ
  lnonsym = (nonsym /= 0)
  std = 0.0d0
  nk = 0
  nk1(1) = nksum
  ttd1(1) = ttd
  call pinsert(n, irad, ri, a_rrec, a_amag, spec, ctff, b3df, t3df, b3df, t3df, std, &
               a_phi, a_theta, a_psi, a_shx, a_shy, a_presa, a_pbc, a_boff, asum, vsum, psum, &
               c3df, ksum, sinclut, irada, ipad, thet, iewald, nsy, isymax, js, symc, nnstat, &
               nk, imp, a_psize, nk1, ttd1, ns, ns1, a_occ, interp, smc, lnonsym)
  nksum = nk1(1)
  ttd = ttd1(1)
end subroutine

subroutine ref_getsymmat(asymc, nasym, symop, nsym, jsym, iredun) bind(C, name="ref_getsymmat")
  use iso_c_binding
  implicit none
  character(kind=c_char), intent(in) :: asymc
  integer(c_int), intent(in) :: nasym
  real(c_float), intent(out) :: symop(3, 3, *)
  integer(c_int), intent(out) :: nsym, jsym(*), iredun
  character(len=1) :: a
  integer :: na, isymax
  a = asymc; na = nasym; isymax = 8
  call getsymmat(a, na, symop, nsym, jsym, isymax, iredun)
end subroutine

subroutine ref_storeshift(nsam, std, vtd, iredun, a3df, s3df, d3df, v3df, b3df, c3df) &
    bind(C, name="ref_storeshift")
  use iso_c_binding
  implicit none
  integer(c_int), intent(in) :: nsam, iredun
  real(c_double), intent(in) :: std, vtd
  complex(c_float_complex), intent(in) :: a3df(*), d3df(*)
  real(c_float), intent(in) :: s3df(*), v3df(*)
  complex(c_float_complex), intent(inout) :: b3df(*), c3df(*)
  integer :: n, jc, nsamh, ir, nn1, ipad, nnstat, ifsc, ksum(1)
  real :: qcp(1), asum(1), psum(1), vsum(1), ssnr(1)
  complex :: ref3df(1)
  double precision :: s, v
  n = nsam; jc = nsam / 2 + 1; nsamh = nsam / 2; ir = iredun; nn1 = nsam
  ipad = 1; nnstat = 0; ifsc = 0; s = std; v = vtd
  call storeshift(s, v, jc, n, nsamh, ir, nn1, qcp, asum, ksum, psum, vsum, ssnr, &
                  a3df, b3df, c3df, d3df, s3df, v3df, ref3df, ipad, nnstat, ifsc)
end subroutine

subroutine ref_sheltest(nsam, istp, a3df, d3df, s3df, v3df, fmask, npix, npixt, dfsc, fsc, &
                        pssnr, nstp) bind(C, name="ref_sheltest")
  use iso_c_binding
  implicit none
  integer(c_int), intent(in) :: nsam, istp
  complex(c_float_complex), intent(in) :: a3df(*), d3df(*)
  real(c_float), intent(in) :: s3df(*), v3df(*), fmask
  integer(c_int), intent(out) :: npix(*), npixt(*), nstp
  real(c_double), intent(out) :: dfsc(*)
  real(c_float), intent(out) :: fsc(*), pssnr(*)
  integer :: n, is, ns(1024), ksum(1), ir, nnstat, nst
  real :: qcp(1), asum(1), psum(1), pr(1024), acc(1024), qf(1024), qc(1024), pf(1024), fm
  n = nsam; is = istp; ir = 1; nnstat = 0; fm = fmask
  call sheltest(n, is, a3df, d3df, asum, psum, ksum, qcp, nst, pr, fsc, acc, qf, qc, pf, &
                ns, npix, npixt, dfsc, ir, nnstat, pssnr, s3df, v3df, fm)
  nstp = nst
end subroutine

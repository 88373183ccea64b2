! Writes and reads parameter-file records with the v9 record format of the
! reference (lmain.f:922-923) so that the library's formatter and parser can
! be compared with a Fortran runtime's F/I editing.
!   fr_par_check write <in> <out>: <in> holds lines of 16 values (C PSI THETA
!     PHI SHX SHY MAG FILM DF1 DF2 ANGAST OCC LogP SIGMA SCORE CHANGE, the
!     integers as reals); each becomes one formatted record in <out>.
!   fr_par_check read <in> <out>: each record of <in> is read with the same
!     format and its 16 values written to <out>, one record per line.
program fr_par_check
  implicit none
  character(len=512) :: mode, fin, fout, line
  real :: v(16), psi, theta, phi, shx, shy, df1, df2, ang, occ, sig, score, change
  integer :: c, mag, film, logp, ios
  call get_command_argument(1, mode)
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)
  open(10, file=trim(fin), status='old')
  open(11, file=trim(fout), status='replace')
  do
    read(10, '(A)', iostat=ios) line
    if (ios /= 0) exit
    if (trim(mode) == 'write') then
      read(line, *) v
      write(11, 7011) nint(v(1)), v(2), v(3), v(4), v(5), v(6), nint(v(7)), nint(v(8)), &
           v(9), v(10), v(11), v(12), nint(v(13)), v(14), v(15), v(16)
    else
      read(line, 7011) c, psi, theta, phi, shx, shy, mag, film, df1, df2, ang, occ, logp, &
           sig, score, change
      write(11, '(16ES26.16E3)') real(c, 8), real(psi, 8), real(theta, 8), real(phi, 8), &
           real(shx, 8), real(shy, 8), real(mag, 8), real(film, 8), real(df1, 8), &
           real(df2, 8), real(ang, 8), real(occ, 8), real(logp, 8), real(sig, 8), &
           real(score, 8), real(change, 8)
    end if
  end do
  close(10)
  close(11)
7011 format(I7,3F8.2,2F10.2,I8,I6,2F9.1,2F8.2,I10,F11.4,2F8.2)
end program fr_par_check

! fr_grid_check.f90 -- a Fortran host calling the C-ABI through
! frealign_hip_mod, next to the reference's own SEARCHANG and GETSYMMAT
! (linked from the flang-compiled reference routines).  Prints the number of
! differing values; 0 means the drop-in reproduces the reference bit for bit.
program fr_grid_check
  use iso_c_binding
  use frealign_hip
  implicit none
  real(c_float), allocatable :: sang_ref(:, :), sang_gpu(:, :)
  real(c_float) :: symop_ref(3, 3, 8), symop_gpu(3, 3, 8)
  integer(c_int) :: jsym_ref(8), jsym_gpu(8), nsym_ref, nsym_gpu, ired_ref, ired_gpu
  integer(c_int) :: pmask(5), n, n2, iq, iq2, nbad, k
  character(len=3) :: asym
  character(len=1) :: groups(3) = ['C', 'D', 'O']
  integer :: nasyms(3) = [1, 7, 0], g
  logical :: lstore
  pmask = 1
  nbad = 0
  open(unit=71, file='/dev/null', status='unknown')
  do g = 1, 3
    asym = groups(g) // '  '
    ! the reference: count, then store (frealign_v9.f:513-523)
    lstore = .false.
    allocate(sang_ref(3, 1))
    call searchang(0.1, 50.0, 5.0, n, iq, asym, nasyms(g), sang_ref, 70, 1, .true., pmask, lstore)
    deallocate(sang_ref)
    allocate(sang_ref(3, n), sang_gpu(3, n))
    lstore = .true.
    call searchang(0.1, 50.0, 5.0, n, iq, asym, nasyms(g), sang_ref, 70, 1, .true., pmask, lstore)
    call fr_check(fr_search_grid(0.1, 50.0, 5.0, groups(g), nasyms(g), pmask, 1, sang_gpu, n, n2, &
                                 iq2), 'fr_search_grid')
    if (n2 /= n .or. iq2 /= iq) nbad = nbad + 1
    do k = 1, n
      if (any(transfer(sang_ref(:, k), 1_c_int, 3) /= transfer(sang_gpu(:, k), 1_c_int, 3))) &
        nbad = nbad + 1
    end do
    deallocate(sang_ref, sang_gpu)
    call getsymmat(groups(g), nasyms(g), symop_ref, nsym_ref, jsym_ref, 8, ired_ref)
    call fr_check(fr_point_group(groups(g), nasyms(g), symop_gpu, jsym_gpu, nsym_gpu, ired_gpu), &
                  'fr_point_group')
    if (nsym_ref /= nsym_gpu .or. ired_ref /= ired_gpu) nbad = nbad + 1
    if (any(jsym_ref(1:nsym_ref) /= jsym_gpu(1:nsym_gpu))) nbad = nbad + 1
    if (any(transfer(symop_ref(:, :, 1:nsym_ref), 1_c_int, 9 * nsym_ref) /= &
            transfer(symop_gpu(:, :, 1:nsym_gpu), 1_c_int, 9 * nsym_gpu))) nbad = nbad + 1
  end do
  write(*, '(A,I0)') 'FR_GRID_CHECK mismatches: ', nbad
end program

! frealign_hip_mod.f90 -- ISO_C_BINDING interface to libfrealign_hip.so for a
! Fortran host that keeps Frealign's card / par-file operator surface.
!
! Each interface is the C entry point of include/frealign_hip.h; the comment
! names the reference routine it replaces.  Device buffers are TYPE(C_PTR)
! (allocate them with fr_dev_alloc); host arrays are passed by reference as
! the reference passes them.  A non-zero status is fatal, as STOP is in the
! reference: use fr_check.
module frealign_hip
  use iso_c_binding
  implicit none

  interface
    function fr_version() bind(C, name="fr_version")
      import :: c_ptr
      type(c_ptr) :: fr_version
    end function

    function fr_last_error() bind(C, name="fr_last_error")
      import :: c_ptr
      type(c_ptr) :: fr_last_error
    end function

    function fr_create(device, ctx) bind(C, name="fr_create")
      import :: c_int, c_ptr
      integer(c_int), value :: device
      type(c_ptr) :: ctx
      integer(c_int) :: fr_create
    end function

    function fr_destroy(ctx) bind(C, name="fr_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: fr_destroy
    end function

    function fr_synchronize(ctx) bind(C, name="fr_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: fr_synchronize
    end function

    ! SEARCHANG + LIMITSYMM (searchang.f:8-150)
    function fr_search_grid(rmm, ri, dangin, asym, nasym, pmask, asymstore, sang, cap, nsang, &
                            iquadmax) bind(C, name="fr_search_grid")
      import :: c_int, c_float, c_char
      real(c_float), value :: rmm, ri, dangin
      character(kind=c_char), value :: asym
      integer(c_int), value :: nasym, asymstore, cap
      integer(c_int) :: pmask(5), nsang, iquadmax
      real(c_float) :: sang(3, *)
      integer(c_int) :: fr_search_grid
    end function

    ! GETSYMMAT (getsymmat.f:8-97)
    function fr_point_group(asym, nasym, symop, jsym, nsym, iredun) bind(C, name="fr_point_group")
      import :: c_int, c_float, c_char
      character(kind=c_char), value :: asym
      integer(c_int), value :: nasym
      real(c_float) :: symop(3, 3, 8)
      integer(c_int) :: jsym(8), nsym, iredun
      integer(c_int) :: fr_point_group
    end function

    ! rotation matrices of cc3m.f:48-65 / pextract.f:30-51
    function fr_rotation(n, euler, amag, sm, dm) bind(C, name="fr_rotation")
      import :: c_int, c_float
      integer(c_int), value :: n
      real(c_float) :: euler(3, *), sm(9), dm(9, *)
      real(c_float), value :: amag
      integer(c_int) :: fr_rotation
    end function

    ! LMAIN's r2c (lmain.f:459) + CTFAPPLY (ctfapply.f:8-161), batched
    function fr_prep_batch(ctx, nsam, nb, d_data, d_ppar, d_pssnr, gpar, ictf, d_ctff, d_outc, &
                           d_datc) bind(C, name="fr_prep_batch")
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: ctx, d_data, d_ppar, d_pssnr, d_ctff, d_outc, d_datc
      integer(c_int), value :: nsam, nb, ictf
      real(c_float) :: gpar(11)
      integer(c_int) :: fr_prep_batch
    end function

    ! PEXTRACT (pextract.f:8-101)
    function fr_pextract_batch(ctx, nsam, n, d_a3df, d_dm, d_pidx, d_ctff, d_outc) &
        bind(C, name="fr_pextract_batch")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, d_a3df, d_dm, d_pidx, d_ctff, d_outc
      integer(c_int), value :: nsam, n
      integer(c_int) :: fr_pextract_batch
    end function

    ! CC3M (cc3m.f:8-169)
    function fr_cc3m_batch(ctx, nsam, n, d_a3df, d_spec, d_ctff, d_pidx, d_dm, d_shift, d_amag, &
                           maxr1, maxr2, rbfact, d_fsct, ring, d_score) bind(C, name="fr_cc3m_batch")
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: ctx, d_a3df, d_spec, d_ctff, d_pidx, d_dm, d_shift, d_amag, d_fsct, &
                            d_score
      integer(c_int), value :: nsam, n, maxr1, maxr2
      real(c_float), value :: rbfact
      real(c_float) :: ring(4)
      integer(c_int) :: fr_cc3m_batch
    end function

    ! PRESB (presb.f:8-160)
    function fr_presb_batch(ctx, nsam, n, d_a3df, d_spec, d_ctff, d_pidx, d_dm, d_shift, d_amag, &
                            ring, d_rbin) bind(C, name="fr_presb_batch")
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: ctx, d_a3df, d_spec, d_ctff, d_pidx, d_dm, d_shift, d_amag, d_rbin
      integer(c_int), value :: nsam, n
      real(c_float) :: ring(4)
      integer(c_int) :: fr_presb_batch
    end function

    ! SIGMA (sigma.f:8-221)
    function fr_sigma_batch(ctx, nsam, n, d_a3df, d_datc, d_ctff, pidx, par, amag, maxr1, rclas, &
                            dmask, sigp, signset, sm, out) bind(C, name="fr_sigma_batch")
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: ctx, d_a3df, d_datc, d_ctff
      real(c_float) :: sm(9)
      integer(c_int), value :: nsam, n, maxr1
      real(c_float), value :: rclas, sigp
      integer(c_int) :: pidx(*)
      real(c_float) :: par(5, *), amag(*), dmask(4), signset(*), out(3, *)
      integer(c_int) :: fr_sigma_batch
    end function

    ! VA04A (va04.f:8-369) on a host objective: f = fn(n, x, user)
    function fr_powell_minimize(n, x, e, escale, icon, maxit, fn, user, f, nevals) &
        bind(C, name="fr_powell_minimize")
      import :: c_int, c_float, c_ptr, c_funptr
      integer(c_int), value :: n, icon, maxit
      real(c_float) :: x(*), e(*), f
      real(c_float), value :: escale
      type(c_funptr), value :: fn
      type(c_ptr), value :: user
      integer(c_int) :: nevals
      integer(c_int) :: fr_powell_minimize
    end function

    ! PREFINE (prefine.f:8-162) for n items, lock-stepped
    function fr_refine_batch(ctx, nsam, n, d_a3df, d_spec, d_ctff, pidx, amag, par, mask, maxr1, &
                             maxr2, rbfact, d_fsct, ring, prior, maxit, nevals) &
        bind(C, name="fr_refine_batch")
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: ctx, d_a3df, d_spec, d_ctff, d_fsct
      integer(c_int), value :: nsam, n, maxr1, maxr2, maxit
      real(c_float), value :: rbfact
      integer(c_int) :: pidx(*), mask(5), nevals(*)
      real(c_float) :: amag(*), par(6, *), ring(4), prior(5, *)
      integer(c_int) :: fr_refine_batch
    end function

    ! A3D3 flush -> PINSERT -> PINSERT_TRILIN (a3d3.f:107-143, pinsert.f, pinsert_trilin.f)
    function fr_insert_batch(ctx, nsam, nb, d_datc, d_ctff, recb, half, nsym, jsym, symop, pbc, &
                             boff, psize, d_num0, d_ctf2_0, d_num1, d_ctf2_1, d_nksum, d_ttd) &
        bind(C, name="fr_insert_batch")
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: ctx, d_datc, d_ctff, d_num0, d_ctf2_0, d_num1, d_ctf2_1, d_nksum, d_ttd
      integer(c_int), value :: nsam, nb, nsym
      real(c_float) :: recb(10, *), symop(3, 3, *)
      integer(c_int) :: half(*), jsym(*)
      real(c_float), value :: pbc, boff, psize
      integer(c_int) :: fr_insert_batch
    end function

    ! SHELTEST (sheltest.f:8-122)
    function fr_sheltest(ctx, nsam, istp, d_a, d_d, d_s, d_v, fmask, npix, npixt, dfsc, pssnr, &
                         nstp) bind(C, name="fr_sheltest")
      import :: c_int, c_float, c_double, c_ptr
      type(c_ptr), value :: ctx, d_a, d_d, d_s, d_v
      integer(c_int), value :: nsam, istp
      real(c_float), value :: fmask
      integer(c_int) :: npix(*), npixt(*), nstp
      real(c_double) :: dfsc(*)
      real(c_float) :: pssnr(*)
      integer(c_int) :: fr_sheltest
    end function

    ! STORESHIFT (storeshift.f:43-66)
    function fr_storeshift(ctx, nsam, std, vtd, iredun, d_a, d_s, d_d, d_v, d_b, d_c) &
        bind(C, name="fr_storeshift")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, d_a, d_s, d_d, d_v, d_b, d_c
      integer(c_int), value :: nsam, iredun
      real(c_double), value :: std, vtd
      integer(c_int) :: fr_storeshift
    end function

    ! COMBINE_ARRAYS (combine_arrays.f:8-69): copies nn3 elements apart
    function fr_combine_arrays(ctx, nsam, nn3, nvol1, nvol2, ifsc, d_a3df, d_s3df, d_d3df, &
                               d_v3df, imp, nksum1, ttd1, ttd2, nksum, std, vtd) &
        bind(C, name="fr_combine_arrays")
      import :: c_int, c_long_long, c_double, c_ptr
      type(c_ptr), value :: ctx, d_a3df, d_s3df, d_d3df, d_v3df
      integer(c_int), value :: nsam, nvol1, nvol2, ifsc, imp
      integer(c_long_long), value :: nn3
      integer(c_long_long), intent(in) :: nksum1(*)
      real(c_double), intent(in) :: ttd1(*), ttd2(*)
      integer(c_long_long) :: nksum
      real(c_double) :: std, vtd
      integer(c_int) :: fr_combine_arrays
    end function

    ! SHIFTVOL (shiftvol.f:8-77), in place; std is STD in/out
    function fr_shiftvol(ctx, nsam, std, vtd, rlim, iredun, d_a3df, d_d3df, d_s3df, d_v3df, &
                         ifsc, pssnr, fpart, fsc, fmask, nshell) bind(C, name="fr_shiftvol")
      import :: c_int, c_float, c_double, c_ptr
      type(c_ptr), value :: ctx, d_a3df, d_d3df, d_s3df, d_v3df
      integer(c_int), value :: nsam, iredun, ifsc, nshell
      real(c_double) :: std
      real(c_double), value :: vtd
      real(c_float), value :: rlim, fpart, fmask
      real(c_float), intent(in) :: pssnr(*), fsc(*)
      integer(c_int) :: fr_shiftvol
    end function

    ! CORRECT3D (correct3d.f:19-47) on the padded real volume
    function fr_correct3d(ctx, nsam, d_a3dv, interp, ipad, sinclut, nlut) &
        bind(C, name="fr_correct3d")
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: ctx, d_a3dv
      integer(c_int), value :: nsam, interp, ipad, nlut
      real(c_float), intent(in) :: sinclut(*)
      integer(c_int) :: fr_correct3d
    end function

    ! 3D FFTs of one padded volume (FFTW_PLANS(3), (4)), in place
    function fr_fft3d_r2c(ctx, nsam, d_data) bind(C, name="fr_fft3d_r2c")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, d_data
      integer(c_int), value :: nsam
      integer(c_int) :: fr_fft3d_r2c
    end function

    function fr_fft3d_c2r(ctx, nsam, d_data) bind(C, name="fr_fft3d_c2r")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, d_data
      integer(c_int), value :: nsam
      integer(c_int) :: fr_fft3d_c2r
    end function

    ! MASK3D (mask3d.f:8-142) on the padded real volume
    function fr_mask3d(ctx, nsam, d_a3dv, rio, ric, halfw, fmask, ave, std) &
        bind(C, name="fr_mask3d")
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: ctx, d_a3dv
      integer(c_int), value :: nsam
      real(c_float), value :: rio, ric, halfw
      real(c_float) :: fmask, ave, std
      integer(c_int) :: fr_mask3d
    end function

    ! FIND_FPART closed form for MW /= 0 (find_fpart.f:141-149)
    function fr_fpart_from_mw(nsam, mw, dalt, psize, ssnr, fpart) bind(C, name="fr_fpart_from_mw")
      import :: c_int, c_float
      integer(c_int), value :: nsam
      real(c_float), value :: mw, dalt, psize
      real(c_float) :: ssnr(*), fpart
      integer(c_int) :: fr_fpart_from_mw
    end function

    ! PSEARCH -> CCP (psearch.f:8-121, ccp.f:8-275)
    function fr_search_layout(ctx, nsam, maxr1, maxr2, kc, kpad) bind(C, name="fr_search_layout")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsam, maxr1, maxr2
      integer(c_int) :: kc, kpad
      integer(c_int) :: fr_search_layout
    end function

    function fr_search_bank(ctx, nsam, maxr1, maxr2, d_a3df, norient, d_dm, dtype, d_bank) &
        bind(C, name="fr_search_bank")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, d_a3df, d_dm, d_bank
      integer(c_int), value :: nsam, maxr1, maxr2, norient, dtype
      integer(c_int) :: fr_search_bank
    end function

    function fr_search_batch(ctx, nsam, maxr1, maxr2, rbfact, norient, d_bank, dtype, group, np, &
                             d_outc, d_ctff, topk, d_top_o, d_top_row, d_top_score) &
        bind(C, name="fr_search_batch")
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: ctx, d_bank, d_outc, d_ctff, d_top_o, d_top_row, d_top_score
      integer(c_int), value :: nsam, maxr1, maxr2, norient, dtype, group, np, topk
      real(c_float), value :: rbfact
      integer(c_int) :: fr_search_batch
    end function

    ! PSEARCH over per-particle orientation lists d_cand(d_loff(p)+1 : d_loff(p+1))
    function fr_search_ragged_batch(ctx, nsam, maxr1, maxr2, rbfact, norient, d_bank, dtype, &
                                    np, d_outc, d_ctff, d_loff, d_grp, d_cand, d_info, cap, topk, &
                                    d_top_o, d_top_row, d_top_score) &
        bind(C, name="fr_search_ragged_batch")
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: ctx, d_bank, d_outc, d_ctff, d_loff, d_grp, d_cand, d_info
      type(c_ptr), value :: d_top_o, d_top_row, d_top_score
      integer(c_int), value :: nsam, maxr1, maxr2, norient, dtype, np, cap, topk
      real(c_float), value :: rbfact
      integer(c_int) :: fr_search_ragged_batch
    end function

    ! learned proposal: network input, forward pass, top classes, candidate lists
    function fr_cnn_images(ctx, nsam, s, nb, d_datc, d_ctff, d_img) bind(C, name="fr_cnn_images")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, d_datc, d_ctff, d_img
      integer(c_int), value :: nsam, s, nb
      integer(c_int) :: fr_cnn_images
    end function

    function fr_cnn_forward(ctx, s, c, nconv, k, d_wq, nb, d_img, d_logits) &
        bind(C, name="fr_cnn_forward")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, d_wq, d_img, d_logits
      integer(c_int), value :: s, c, nconv, k, nb
      integer(c_int) :: fr_cnn_forward
    end function

    function fr_cnn_forward_bf16(ctx, s, c, nconv, k, d_wq, nb, d_img, d_logits) &
        bind(C, name="fr_cnn_forward_bf16")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, d_wq, d_img, d_logits
      integer(c_int), value :: s, c, nconv, k, nb
      integer(c_int) :: fr_cnn_forward_bf16
    end function

    function fr_cnn_propose(ctx, k, nb, d_logits, topk, d_top, d_prob) &
        bind(C, name="fr_cnn_propose")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, d_logits, d_top, d_prob
      integer(c_int), value :: k, nb, topk
      integer(c_int) :: fr_cnn_propose
    end function

    function fr_build_lists(ctx, np, ktop, d_top, d_prob, threshold, ncls, d_cell_off, d_cells, &
                            group, nbank, bank_group, maxlen, d_loff, d_grp, d_cand, d_info) &
        bind(C, name="fr_build_lists")
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: ctx, d_top, d_prob, d_cell_off, d_cells, d_loff, d_grp, d_cand, d_info
      integer(c_int), value :: np, ktop, ncls, group, nbank, bank_group, maxlen
      real(c_float), value :: threshold
      integer(c_int) :: fr_build_lists
    end function

    ! CC3M of every candidate (lmain.f:615-642) and each particle's winner
    function fr_rescore_batch(ctx, nsam, np, k, d_a3df, d_outc, d_ctff, d_top_o, d_top_row, &
                              nbank, d_dm_bank, nsamr, maxr1, maxr2, rbfact, d_fsct, ring, &
                              d_score, d_best) bind(C, name="fr_rescore_batch")
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: ctx, d_a3df, d_outc, d_ctff, d_top_o, d_top_row, d_dm_bank, d_fsct
      type(c_ptr), value :: d_score, d_best
      integer(c_int), value :: nsam, np, k, nbank, nsamr, maxr1, maxr2
      real(c_float), value :: rbfact
      real(c_float), intent(in) :: ring(4)
      integer(c_int) :: fr_rescore_batch
    end function
    ! PINSERT_TRILIN_SE with NNSTAT = 1 (pinsert_trilin_se.f:186-194): as
    ! fr_insert_batch plus the per-voxel KSUM, VSUM, PSUM, ASUM
    function fr_insert_batch_se(ctx, nsam, nb, d_datc, d_ctff, recb, half, nsym, jsym, symop, &
                                pbc, boff, psize, d_num0, d_ctf2_0, d_num1, d_ctf2_1, d_nksum, &
                                d_ttd, d_c3df, d_ksum, d_vsum, d_psum, d_asum) &
        bind(C, name="fr_insert_batch_se")
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: ctx, d_datc, d_ctff, d_num0, d_ctf2_0, d_num1, d_ctf2_1, d_nksum, d_ttd
      type(c_ptr), value :: d_c3df, d_ksum, d_vsum, d_psum, d_asum
      integer(c_int), value :: nsam, nb, nsym
      real(c_float) :: recb(10, *), symop(3, 3, *)
      integer(c_int) :: half(*), jsym(*)
      real(c_float), value :: pbc, boff, psize
      integer(c_int) :: fr_insert_batch_se
    end function

    ! SHELTEST with its accumulators in a fixed device array (no allocation)
    function fr_sheltest_noalloc(ctx, nsam, istp, d_a, d_d, d_s, d_v, fmask, npix, npixt, dfsc, &
                                 pssnr, nstp) bind(C, name="fr_sheltest_noalloc")
      import :: c_int, c_float, c_double, c_ptr
      type(c_ptr), value :: ctx, d_a, d_d, d_s, d_v
      integer(c_int), value :: nsam, istp
      real(c_float), value :: fmask
      integer(c_int) :: npix(*), npixt(*), nstp
      real(c_double) :: dfsc(*)
      real(c_float) :: pssnr(*)
      integer(c_int) :: fr_sheltest_noalloc
    end function

    ! FIND_FPART (find_fpart.f:8-150, fpart_cc.f, volmeasure.f); pssnr in/out
    function fr_find_fpart(ctx, nsam, std, vtd, d_a3df, d_d3df, d_s3df, d_v3df, iredun, ri, ric, &
                           psize, mw, dalt, fsc, pssnr, nshell, fpart) bind(C, name="fr_find_fpart")
      import :: c_int, c_float, c_double, c_ptr
      type(c_ptr), value :: ctx, d_a3df, d_d3df, d_s3df, d_v3df
      integer(c_int), value :: nsam, iredun, nshell
      real(c_double), value :: std, vtd
      real(c_float), value :: ri, ric, psize, mw, dalt
      real(c_float) :: fsc(*), pssnr(*), fpart
      integer(c_int) :: fr_find_fpart
    end function

    ! the end of a refinement cycle (frealign_v9.f:790-950) in one call
    function fr_finalize(ctx, nsam, d_a3df, d_s3df, d_d3df, d_v3df, nksum, ttd, iredun, psize, &
                         ri, ric, mw, dalt, ffilt, rlim, d_b, d_c, nshell, fsc, pssnr, npix, &
                         npixt, nstp, fmask, fpart) bind(C, name="fr_finalize")
      import :: c_int, c_float, c_double, c_long_long, c_ptr
      type(c_ptr), value :: ctx, d_a3df, d_s3df, d_d3df, d_v3df, d_b, d_c
      integer(c_int), value :: nsam, iredun, ffilt, nshell
      integer(c_long_long) :: nksum(2)
      real(c_double) :: ttd(2)
      real(c_float), value :: psize, ri, ric, mw, dalt, rlim
      real(c_float) :: fsc(*), pssnr(*), fmask, fpart
      integer(c_int) :: npix(*), npixt(*), nstp
      integer(c_int) :: fr_finalize
    end function

    ! CC3 (cc3.f:8-159), the score of CTFREFINE and MAGREFINE
    function fr_cc3_batch(ctx, nsam, n, d_a3df, d_spec, d_ctff, d_pidx, d_dm, d_shift, d_amag, &
                          maxr1, maxr2, rbfact, d_fsct, ring, d_score) bind(C, name="fr_cc3_batch")
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: ctx, d_a3df, d_spec, d_ctff, d_pidx, d_dm, d_shift, d_amag, d_fsct, &
                            d_score
      integer(c_int), value :: nsam, n, maxr1, maxr2
      real(c_float), value :: rbfact
      real(c_float) :: ring(4)
      integer(c_int) :: fr_cc3_batch
    end function

    ! CTFREFINE (ctfrefine.f:8-100) and MAGREFINE (magrefine.f:8-73) per group
    function fr_ctfrefine_batch(ctx, nsam, np, d_spec, ppar, euler, ng, gstart, fastig, dfstd, &
                                sig2, d_a3df, d_pssnr, gpar, rmax1, rmax2, rbfact, ring, out, &
                                nevals) bind(C, name="fr_ctfrefine_batch")
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: ctx, d_spec, d_a3df, d_pssnr
      integer(c_int), value :: nsam, np, ng, fastig
      real(c_float), value :: dfstd, rmax1, rmax2, rbfact
      real(c_float) :: ppar(8, *), euler(3, *), sig2(*), gpar(11), ring(4), out(4, *)
      integer(c_int) :: gstart(*), nevals(*)
      integer(c_int) :: fr_ctfrefine_batch
    end function

    function fr_magrefine_batch(ctx, nsam, np, d_spec, ppar, euler, ng, gstart, d_a3df, d_pssnr, &
                                gpar, rmax1, rmax2, rbfact, ring, out, nevals) &
        bind(C, name="fr_magrefine_batch")
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: ctx, d_spec, d_a3df, d_pssnr
      integer(c_int), value :: nsam, np, ng
      real(c_float), value :: rmax1, rmax2, rbfact
      real(c_float) :: ppar(8, *), euler(3, *), gpar(11), ring(4), out(4, *)
      integer(c_int) :: gstart(*), nevals(*)
      integer(c_int) :: fr_magrefine_batch
    end function

    ! mode-1 local grid search before PREFINE (lmain.f:743-828); par(6, np) in/out
    function fr_local_search_batch(ctx, nsam, np, d_a3df, d_spec, d_ctff, amag, par, nang, dang, &
                                   nsh, dsh, maxr1, maxr2, rbfact, d_fsct, ring, sm, d_scores, &
                                   d_codes) bind(C, name="fr_local_search_batch")
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: ctx, d_a3df, d_spec, d_ctff, d_fsct, sm, d_scores, d_codes
      integer(c_int), value :: nsam, np, nang, nsh, maxr1, maxr2
      real(c_float), value :: dang, dsh, rbfact
      real(c_float) :: amag(*), par(6, *), ring(4)
      integer(c_int) :: fr_local_search_batch
    end function

    ! reference-exact PSEARCH -> CCP with PAD and ROTATE2D (psearch.f:8-121,
    ! ccp.f:8-275); sm may be c_null_ptr (identity)
    function fr_psearch(ctx, nsam, np, amag, d_a3df, d_spec, d_ctff, dangin, maxr1, maxr2, par, &
                        sang, nsang, iquadmax, mask, rbfact, testpar, ipmax, iran, sm, asym) &
        bind(C, name="fr_psearch")
      import :: c_int, c_float, c_ptr, c_char
      type(c_ptr), value :: ctx, d_a3df, d_spec, d_ctff, sm
      integer(c_int), value :: nsam, np, maxr1, maxr2, nsang, iquadmax, ipmax
      real(c_float), value :: dangin, rbfact
      real(c_float) :: amag(*), par(5, *), sang(3, *), testpar(6, ipmax, *)
      integer(c_int) :: mask(5), iran
      character(kind=c_char), value :: asym
      integer(c_int) :: fr_psearch
    end function

    function fr_ccp_view_batch(ctx, nsam, np, nview, d_a3df, d_spec, d_ctff, d_dm, maxr1, maxr2, &
                               rbfact, dang, iquadmax, mask, d_score, d_code) &
        bind(C, name="fr_ccp_view_batch")
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: ctx, d_a3df, d_spec, d_ctff, d_dm, d_score, d_code
      integer(c_int), value :: nsam, np, nview, maxr1, maxr2, iquadmax
      real(c_float), value :: rbfact, dang
      integer(c_int) :: mask(5)
      integer(c_int) :: fr_ccp_view_batch
    end function

    ! PSEARCH over lists of one length (ncand per particle, in groups)
    function fr_search_list_batch(ctx, nsam, maxr1, maxr2, rbfact, norient, d_bank, dtype, np, &
                                  d_outc, d_ctff, ncand, d_cand, group, topk, d_top_o, d_top_row, &
                                  d_top_score) bind(C, name="fr_search_list_batch")
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: ctx, d_bank, d_outc, d_ctff, d_cand, d_top_o, d_top_row, d_top_score
      integer(c_int), value :: nsam, maxr1, maxr2, norient, dtype, np, ncand, group, topk
      real(c_float), value :: rbfact
      integer(c_int) :: fr_search_list_batch
    end function

    ! learned proposal: parameters, training images, one training step, Adam
    function fr_cnn_param_count(s, c, nconv, k, nparam) bind(C, name="fr_cnn_param_count")
      import :: c_int, c_long_long
      integer(c_int), value :: s, c, nconv, k
      integer(c_long_long) :: nparam
      integer(c_int) :: fr_cnn_param_count
    end function

    function fr_cnn_init(ctx, s, c, nconv, k, seed, d_master, d_wq) bind(C, name="fr_cnn_init")
      import :: c_int, c_long_long, c_ptr
      type(c_ptr), value :: ctx, d_master, d_wq
      integer(c_int), value :: s, c, nconv, k
      integer(c_long_long), value :: seed
      integer(c_int) :: fr_cnn_init
    end function

    function fr_cnn_synth(ctx, nsam, s, nb, d_a3df, d_dm, d_spar, gpar, disc, seed, d_img) &
        bind(C, name="fr_cnn_synth")
      import :: c_int, c_float, c_long_long, c_ptr
      type(c_ptr), value :: ctx, d_a3df, d_dm, d_spar, d_img
      integer(c_int), value :: nsam, s, nb
      real(c_float) :: gpar(11)
      real(c_float), value :: disc
      integer(c_long_long), value :: seed
      integer(c_int) :: fr_cnn_synth
    end function

    function fr_cnn_gradients(ctx, s, c, nconv, k, d_wq, nb, d_img, d_label, d_loss, d_grad) &
        bind(C, name="fr_cnn_gradients")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, d_wq, d_img, d_label, d_loss, d_grad
      integer(c_int), value :: s, c, nconv, k, nb
      integer(c_int) :: fr_cnn_gradients
    end function

    function fr_cnn_gradients_bf16(ctx, s, c, nconv, k, d_wq, nb, d_img, d_label, d_loss, d_grad) &
        bind(C, name="fr_cnn_gradients_bf16")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, d_wq, d_img, d_label, d_loss, d_grad
      integer(c_int), value :: s, c, nconv, k, nb
      integer(c_int) :: fr_cnn_gradients_bf16
    end function

    function fr_cnn_adam(ctx, s, c, nconv, k, d_master, d_wq, d_grad, d_m, d_v, step, lr) &
        bind(C, name="fr_cnn_adam")
      import :: c_int, c_float, c_ptr
      type(c_ptr), value :: ctx, d_master, d_wq, d_grad, d_m, d_v
      integer(c_int), value :: s, c, nconv, k, step
      real(c_float), value :: lr
      integer(c_int) :: fr_cnn_adam
    end function

    function fr_orient_nearest(ctx, n, d_dm, ncls, d_dmc, nsym, d_sym, d_label) &
        bind(C, name="fr_orient_nearest")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, d_dm, d_dmc, d_sym, d_label
      integer(c_int), value :: n, ncls, nsym
      integer(c_int) :: fr_orient_nearest
    end function

    ! CALC_OCC's occupancy update (calc_occ.f:179-217); arrays (np, nc)
    function fr_calc_occ(nc, np, f, avg, occ, algp, sig) bind(C, name="fr_calc_occ")
      import :: c_int, c_float
      integer(c_int), value :: nc, np
      real(c_float), value :: f
      real(c_float) :: avg(*), occ(*), algp(*), sig(*)
      integer(c_int) :: fr_calc_occ
    end function
  end interface

contains

  ! STOP on a non-zero status with the library's message, as the reference
  ! stops on its own fatal conditions.
  subroutine fr_check(status, what)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (status == 0) return
    call c_f_pointer(fr_last_error(), msg, [512])
    n = 0
    do while (n < 512)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    end do
    write(*, '(A,": ",512A1)') what, msg(1:n)
    stop 'frealign_hip: fatal error'
  end subroutine
end module frealign_hip

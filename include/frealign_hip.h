/*
 * frealign_hip.h -- C-ABI of the MI355X (gfx950) implementation of Frealign
 * v9.12's per-particle hot path.
 *
 * Conventions
 *   - Plain C types only.  Arguments named d_* are device pointers (hipMalloc
 *     or any allocator on the context's device); everything else is host memory.
 *   - Work is enqueued on the context's stream (fr_set_stream; default: the
 *     null stream) and is asynchronous unless a function returns host results.
 *   - Return value: 0 on success, negative on error; fr_last_error() gives the
 *     message (per thread).  The Fortran glue maps a non-zero status to STOP,
 *     as the reference does for its own fatal conditions.
 *   - Storage follows the reference (column-major, x fastest, 0-based here):
 *       half-plane  JC x NSAM complex64, JC = NSAM/2+1          (ctfapply.f:64-76)
 *       CTFF        2 half-planes: CTFR then CTFL               (ctfapply.f:78-79)
 *       image       NSAM x NSAM float32, row pitch NSAM+2 (in-place r2c layout)
 *       half-volume JC x NSAM x NSAM complex64 (+ float32 CTF^2) (shift.f:34)
 *     complex arrays are interleaved float pairs and passed as float*.
 *   - Angles in radians; shifts in phase units (2*pi*pixels/NSAM), as after
 *     the unit conversion of card10.f:275-291.
 */
#ifndef FREALIGN_HIP_H
#define FREALIGN_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fr_ctx fr_ctx;

const char *fr_version(void);
const char *fr_last_error(void);

/* Context: one per device and host thread.  Owns grow-only scratch memory. */
int fr_create(int device, fr_ctx **out);
int fr_destroy(fr_ctx *ctx);
int fr_set_stream(fr_ctx *ctx, void *hip_stream);
int fr_synchronize(fr_ctx *ctx);

/* Per-kernel timing with HIP events on the context's stream (off by default).
 * which: 0 search GEMM (exhaustive), 1 insertion, 2 CC3M, 3 particle
 * preparation, 4 search GEMM (candidate lists), 5 proposal-network forward,
 * 6 a graph-captured batch (fr_plan_run).
 * fr_profile_read synchronises, returns the summed duration and count of the
 * recorded launches of that family, and forgets them.  No reference
 * counterpart (the reference only prints wall-clock stamps, lmain.f:238-246). */
int fr_profile_enable(fr_ctx *ctx, int on);
int fr_profile_read(fr_ctx *ctx, int which, double *total_ms, long long *count);

/* ---------------------------------------------------------------- geometry
 * Host functions; their FP32 results are bit-identical to the reference's. */

/* Replaces SEARCHANG + LIMITSYMM (searchang.f:8-150, limitsymm.f:8-51).
 * pmask: PMASK(1..5).  Writes up to cap triples (phi, theta, psi) to sang and
 * the full count to *nsang (call with cap=0 to size). */
int fr_search_grid(float rmm, float ri, float dangin, char asym, int nasym, const int *pmask,
                   int asymstore, float *sang, int cap, int *nsang, int *iquadmax);

/* Replaces GETSYMMAT + CHECKSYM (getsymmat.f:8-97, checksym.f:8-44):
 * symop 8 x 9 floats (column-major 3x3), jsym 8 ints. */
int fr_point_group(char asym, int nasym, float *symop, int *jsym, int *nsym, int *iredun);

/* Rotation matrices DM of cc3m.f:48-65 (amag != 0: divided by |AMAG|) or
 * pextract.f:30-51 (amag == 0), then DM <- MATMUL_T(DM, SM) when sm != NULL
 * (matmul_t.f).  euler: n x (phi, theta, psi); dm: n x 9. */
int fr_rotation(int n, const float *euler, float amag, const float *sm, float *dm);

/* RANDOM (random.f:8-32): the L'Ecuyer generator PSEARCH jitters its grid with. */
float fr_lcg_random(int *iran);

/* ---------------------------------------------------------------- FFTs
 * Replaces fftwf_plan_dft_r2c_2d / _c2r_2d as used through FFTW_FWD / FFTW_BWD
 * (fftw_plans.f:12-13,47,62): nimg images back to back (stride NSAM*(NSAM+2)
 * floats), in place, unnormalised. */
int fr_fft2d_r2c(fr_ctx *ctx, int nsam, int nimg, float *d_data);
int fr_fft2d_c2r(fr_ctx *ctx, int nsam, int nimg, float *d_data);

/* Replaces the 3D plans FFTW_PLANS(3), (4) (frealign_v9.f:91,490; used
 * through FFTW_FWD / FFTW_BWD on the maps, e.g. frealign_v9.f:825,839): one
 * padded volume (NSAM+2 row pitch, NSAM^2 rows), in
 * place, unnormalised. */
int fr_fft3d_r2c(fr_ctx *ctx, int nsam, float *d_data);
int fr_fft3d_c2r(fr_ctx *ctx, int nsam, float *d_data);

/* ---------------------------------------------------- particle preparation
 * Replaces, per particle, the r2c FFT in LMAIN (lmain.f:459) and CTFAPPLY
 * (ctfapply.f:8-161; called at lmain.f:509-514) with MASKCOS and PWEIGHT.
 *   d_data  in: nb images (destroyed; holds the r2c spectra on return)
 *   d_ppar  nb x 8: SHX, SHY (phase units), DFMID1, DFMID2 (A), ANGAST (rad),
 *           AMAGP, 0, 0
 *   d_pssnr PSSNR(1..JC+1) (pweight.f reads one past JC; see card10.f:81-90)
 *   gpar    11 host floats: CS (A), WL (A), WGH1, WGH2, THETATR, RIH, HALFW,
 *           RI2, RI3, TX, TY  (lmain.f:275, frealign_v9.f:526-531)
 *   outputs d_ctff (nb x 2 half-planes), d_outc (nb, masked + SNR-weighted),
 *           d_datc (nb, shifted, unmasked, origin-centred) */
int fr_prep_batch(fr_ctx *ctx, int nsam, int nb, float *d_data, const float *d_ppar,
                  const float *d_pssnr, const float *gpar, int ictf, float *d_ctff,
                  float *d_outc, float *d_datc);

/* ---------------------------------------------------- extraction / scoring */

/* Replaces PEXTRACT (pextract.f:8-101; IEWALD=0, IPAD=1, IRAD>0): n central
 * sections, item k uses d_dm[9k..] (from fr_rotation with amag=0, sm=identity)
 * and the CTF of particle d_pidx[k] (d_pidx may be NULL: particle k). */
int fr_pextract_batch(fr_ctx *ctx, int nsam, int n, const float *d_a3df, const float *d_dm,
                      const int *d_pidx, const float *d_ctff, float *d_outc);

/* Replaces CC3M (cc3m.f:8-169; IEWALD=0, IBUF=0): n scores.  Item k compares
 * particle d_pidx[k] (d_spec: OUTC of fr_prep_batch, d_ctff) with the section
 * d_dm[9k..] (fr_rotation with amag=AMAG, sm=identity) at shift
 * d_shift[2k..] and masks with AMAG d_amag[k].  ring: RI2, RI3, RIH, HALFW.
 * Requires MAXR2 < NSAM/2. */
int fr_cc3m_batch(fr_ctx *ctx, int nsam, int n, const float *d_a3df, const float *d_spec,
                  const float *d_ctff, const int *d_pidx, const float *d_dm,
                  const float *d_shift, const float *d_amag, int maxr1, int maxr2, float rbfact,
                  const float *d_fsct, const float *ring, float *d_score);

/* Replaces CC3 (cc3.f:8-159; IEWALD=0, IBUF=0), the score of the defocus and
 * magnification refinements (calcfx.f:63-140): as fr_cc3m_batch with the
 * section multiplied by the sign of REAL(CTFR+CONJG(CTFL)) instead of the CTF
 * (cc3.f:86-89) and R_k signed in every shell (no FSCT branch; d_fsct is not
 * read). */
int fr_cc3_batch(fr_ctx *ctx, int nsam, int n, const float *d_a3df, const float *d_spec,
                 const float *d_ctff, const int *d_pidx, const float *d_dm, const float *d_shift,
                 const float *d_amag, int maxr1, int maxr2, float rbfact, const float *d_fsct,
                 const float *ring, float *d_score);

/* Replaces PRESB (presb.f:8-160; IEWALD=0, non-helical): per-shell
 * correlations of n (particle, section, shift) items over 1 <= r^2 < (NSAM/2)^2.
 * Arguments as fr_cc3m_batch.  d_rbin is n x NSAM floats, in/out: for every
 * shell i = 1..NSAM/2 that holds a point, RBIN(2i-1) = sum Re(SAMP conj P) /
 * sqrt(sum|SAMP|^2 sum|P|^2) and RBIN(2i) = 1; other entries are untouched. */
int fr_presb_batch(fr_ctx *ctx, int nsam, int n, const float *d_a3df, const float *d_spec,
                   const float *d_ctff, const int *d_pidx, const float *d_dm,
                   const float *d_shift, const float *d_amag, const float *ring, float *d_rbin);

/* -------------------------------------------------------------- insertion
 * Replaces the A3D3 flush (a3d3.f:107-143) -> A3D3_S -> PINSERT ->
 * PINSERT_TRILIN (pinsert.f:8-195, pinsert_trilin.f:8-244; INTERP=1,
 * IEWALD=0, NNSTAT=0) for nb buffered particles.
 *   recb  host nb x 10, the RECB record of a3d3.f:81-90:
 *         PHI, THETA, PSI, DSHX, DSHY, AMAGP, OCC, SIG2N, PRESA, RREC
 *   half  host nb: 0 -> first half-map (A3DF/S3DF), 1 -> second (D3DF/V3DF)
 *   symop/jsym/nsym from fr_point_group; PBC, BOFF, PSIZE as in pinsert.f.
 * Accumulates into the half-map volumes and into d_nksum[half] (NKSUM1) and
 * d_ttd[half] (TTD1, FP64). */
int fr_insert_batch(fr_ctx *ctx, int nsam, int nb, const float *d_datc, const float *d_ctff,
                    const float *recb, const int *half, int nsym, const int *jsym,
                    const float *symop, float pbc, float boff, float psize, float *d_num0,
                    float *d_ctf2_0, float *d_num1, float *d_ctf2_1, long long *d_nksum,
                    double *d_ttd);

/* Replaces the same flush with NNSTAT = 1 (FSTAT), where PINSERT dispatches to
 * PINSERT_TRILIN_SE (pinsert.f:137-176, pinsert_trilin_se.f:8-359; IEWALD=0,
 * IPAD=1): as fr_insert_batch, and every inserted term BSAMP also adds, at its
 * voxel, to the per-voxel statistics shared by both half-maps
 * (pinsert_trilin_se.f:186-194): d_ksum += 1 (int), d_vsum += |BSAMP|^2,
 * d_psum += PDIFF(BSAMP, C3DF) (pdiff.f), d_asum += |BSAMP|.  d_c3df is the
 * reference transform C3DF in the half-volume layout (complex, as d_num0);
 * the four statistics arrays have one entry per half-volume voxel. */
int fr_insert_batch_se(fr_ctx *ctx, int nsam, int nb, const float *d_datc, const float *d_ctff,
                       const float *recb, const int *half, int nsym, const int *jsym,
                       const float *symop, float pbc, float boff, float psize, float *d_num0,
                       float *d_ctf2_0, float *d_num1, float *d_ctf2_1, long long *d_nksum,
                       double *d_ttd, const float *d_c3df, int *d_ksum, float *d_vsum,
                       float *d_psum, float *d_asum);

/* ------------------------------------------------------------- statistics */

/* Replaces COMBINE_ARRAYS (combine_arrays.f:8-69, NNSTAT=0): copies 2..nvol1
 * of d_a3df/d_s3df (nn3 elements apart, the A3DF(NN3,NVOL1) layout) are added
 * into copy 1 in copy order, and for ifsc == 0 copies 2..nvol2 of
 * d_d3df/d_v3df.  Host counters: *nksum += nksum1[0..imp-1]; *std_ (and *vtd
 * for ifsc == 0) += ttd1 (ttd2) over nvol1 entries when nvol1 != 1, else imp. */
int fr_combine_arrays(fr_ctx *ctx, int nsam, long long nn3, int nvol1, int nvol2, int ifsc,
                      float *d_a3df, float *d_s3df, float *d_d3df, float *d_v3df, int imp,
                      const long long *nksum1, const double *ttd1, const double *ttd2,
                      long long *nksum, double *std_, double *vtd);

/* Replaces FIND_FPART (find_fpart.f:8-150) with FPART_CC (fpart_cc.f:8-94)
 * and VOLMEASURE (volmeasure.f:8-69): the particle fraction FPART of the
 * Wiener filter.  mw != 0: the closed form 1000 MW / DALT / (PSIZE NSAM)^3.
 * mw == 0: the search, on the device -- B = A/(S + STD IREDUN) and the
 * low-passed C = (A+D)/(S+V + (STD+VTD) IREDUN) to real space, MASK3D of C
 * (RI, RIC, HALFW 1), the voxels above the first threshold MAVE + (1.5 -
 * 0.1K) MSTD that holds more than 0.2 N^3 FMASK of them, VOLMEASURE's start
 * and FPART_CC's correlation of B with D/(V + FPART/PSSNR) over them,
 * stepped by 0.2, 0.067, ... while it rises.  The half-maps d_a3df, d_d3df
 * (complex), d_s3df, d_v3df are read only; fsc and pssnr are host arrays of
 * nshell >= NSAM/2+1 shells (SHELTEST's); pssnr is divided by FPART in place
 * (find_fpart.f:143-149) and *fpart returned.  STD or VTD zero is an error
 * (find_fpart.f:37-43).  VOLMEASURE's class sums are taken in double where
 * the reference adds in FP32. */
int fr_find_fpart(fr_ctx *ctx, int nsam, double std_, double vtd, const float *d_a3df,
                  const float *d_d3df, const float *d_s3df, const float *d_v3df, int iredun,
                  float ri, float ric, float psize, float mw, float dalt, const float *fsc,
                  float *pssnr, int nshell, float *fpart);

/* The end of a refinement cycle (frealign_v9.f:790-950, IFSC 0, INTERP 1,
 * IPAD 1) in one call, by the entry points of this header in the reference's
 * order: STD, VTD = TTD / (NKSUM(1) + NKSUM(2)) * 0.1; STORESHIFT into d_b,
 * d_c; each half: 3D c2r, CORRECT3D (CALCSINC's table), MASK3D at |ri| with
 * ric and HALFW 6 (FMASK from the first), 3D r2c; SHELTEST; FIND_FPART (mw,
 * dalt; PSSNR divided by FPART); SHIFTVOL into d_a3df/d_s3df (overwritten, as
 * in the reference) with the Wiener filter when ffilt, RLIM rlim; the final
 * map's 3D c2r and CORRECT3D, and MASK3D when ri > 0.  On return d_a3df holds
 * the final real-space map (padded layout), d_b/d_c the masked half-maps'
 * transforms.  Host: nksum[2], ttd[2] (A3D3's counters), and per shell
 * (nshell >= NSAM/2 + 3 entries) fsc, pssnr (divided by FPART), npix, npixt;
 * *nstp shells are set; *fmask, *fpart. */
int fr_finalize(fr_ctx *ctx, int nsam, float *d_a3df, float *d_s3df, const float *d_d3df,
                const float *d_v3df, const long long *nksum, const double *ttd, int iredun,
                float psize, float ri, float ric, float mw, float dalt, int ffilt, float rlim,
                float *d_b, float *d_c, int nshell, float *fsc, float *pssnr, int *npix,
                int *npixt, int *nstp, float *fmask, float *fpart);

/* Replaces SHIFTVOL (shiftvol.f:8-77, NNSTAT=0), in place: *std_ += vtd; for
 * ifsc == 0 A += D and S += V; then within INT(NSAM*RLIM) A = A(-1)^(l+m+n) /
 * (S + STD*IREDUN) (fpart == 0) or that / (1 + FPART/A1/FMASK) with
 * A1 = 2|FSC|/(1-|FSC|) (1000 at |FSC| >= 1), zero where PSSNR or S is zero;
 * zero outside.  pssnr/fsc: host, nshell >= INT(NSAM*RLIM)+1 entries (read
 * only when fpart != 0). */
int fr_shiftvol(fr_ctx *ctx, int nsam, double *std_, double vtd, float rlim, int iredun,
                float *d_a3df, float *d_d3df, float *d_s3df, const float *d_v3df, int ifsc,
                const float *pssnr, float fpart, const float *fsc, float fmask, int nshell);

/* Replaces CORRECT3D (correct3d.f:19-47) with BOXFT_LUT (boxft_lut.f:26-37):
 * the real volume d_a3dv (NSAM+2 row pitch) divided by the sinc product
 * (interp == 0) or its square (interp == 1).  sinclut: host SINCLUT of
 * CALCSINC (calcsinc.f:17-19), nlut >= 90 entries. */
int fr_correct3d(fr_ctx *ctx, int nsam, float *d_a3dv, int interp, int ipad, const float *sinclut,
                 int nlut);

/* Replaces MASK3D (mask3d.f:8-142), in place on the real volume d_a3dv (NSAM+2
 * row pitch): cosine edge of width HALFW at RIO (clamped to NSAM/2-1-HALFW/2)
 * about (1+NSAM/2)^3 with the edge shell's mean outside, and for ric > 0 the
 * inner mask at RIC likewise.  Host outputs: FMASK (the mask's mean square),
 * AVE and STD of the volume over RIC^2 <= r^2 <= RIO^2 (before masking). */
int fr_mask3d(fr_ctx *ctx, int nsam, float *d_a3dv, float rio, float ric, float halfw,
              float *fmask, float *ave, float *std_);

/* FIND_FPART's closed form for MW != 0 (find_fpart.f:141-149): FPART =
 * |1000 MW / DALT / (PSIZE NSAM)^3| and SSNR(1..JC) /= FPART (host).
 * fr_find_fpart covers both branches, the MW = 0 search on the device. */
int fr_fpart_from_mw(int nsam, float mw, float dalt, float psize, float *ssnr, float *fpart);

/* Replaces SHELTEST's FSC/PSSNR part (sheltest.f:8-122, NNSTAT=0) with FP64
 * shell sums on the GPU.  Host outputs per shell: npix, npixt, dfsc, pssnr. */
int fr_sheltest(fr_ctx *ctx, int nsam, int istp, const float *d_a, const float *d_d,
                const float *d_s, const float *d_v, float fmask, int *npix, int *npixt,
                double *dfsc, float *pssnr, int *nstp);

/* No counterpart in the reference.  fr_sheltest for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same arguments and results, with the FP64 shell
 * accumulators in a fixed device array of 520 shells x 6 doubles (24,960 bytes
 * per device, held until the process ends) instead of context scratch.  It
 * calls no hipMalloc/hipFree; the accumulators are cleared by a fill kernel.
 * Returns an error when NSAM/2+1 over ISTP exceeds 520 shells.  Every call on
 * one device shares that array: calls that overlap on one device race. */
int fr_sheltest_noalloc(fr_ctx *ctx, int nsam, int istp, const float *d_a, const float *d_d,
                        const float *d_s, const float *d_v, float fmask, int *npix, int *npixt,
                        double *dfsc, float *pssnr, int *nstp);

/* Replaces STORESHIFT's half-map normalisation (storeshift.f:43-66, IFSC=0,
 * NNSTAT=0): B = A*(-1)^(l+m+n)/(S+STD*IREDUN)/NSAM^3, C likewise from D, V. */
int fr_storeshift(fr_ctx *ctx, int nsam, double std_, double vtd, int iredun, const float *d_a,
                  const float *d_s, const float *d_d, const float *d_v, float *d_b, float *d_c);

/* ------------------------------------------------------------------ search
 * Replaces PSEARCH -> CCP (psearch.f:8-121, ccp.f:8-275; called at
 * lmain.f:570-575) for a batch of particles against a shared bank of
 * orientations.  The correlation annulus is ccp.f's: I >= 0,
 * max(MAXR1^2,1) <= I^2+J^2 < MAXR2^2, with MAXR2 <= NSAM/4 (the NSAMR grid).
 * Shifts are the 15x15 NSAMR-grid hypotheses -7..7; each (particle,
 * orientation) is scored as CCP scores a view: argmax over -6..6 (x outer,
 * first max wins), 3x3 box sum /9, normalised by sqrt(sum|cT|^2 sum|X|^2).
 * dtype 1 runs the correlation as a BF16 MFMA GEMM with FP32 accumulation,
 * dtype 0 as an FP32 MFMA GEMM.  dtype 2 (MAXR2 <= 32) evaluates the same
 * 15x15 map on the BF16 matrix cores in two separable stages (the J phase,
 * then the I phase) from a bank in the separable layout. */

/* Annulus size kc and the padded real row length kpad of a bank row. */
int fr_search_layout(fr_ctx *ctx, int nsam, int maxr1, int maxr2, int *kc, int *kpad);

/* Bank of CTF-free central sections T_o (AINTERPO3DS of the plane
 * d_dm[9o..]*(I,J), dm from fr_rotation with amag=AMAG): norient x kpad
 * elements of 4 (dtype 0) or 2 (dtype 1) bytes, the annulus in order; or
 * (dtype 2) norient x 4096 bf16 in the separable layout: grid point
 * (I, J), I in 0..31, J in -32..31, is complex element
 * ((q * 64 + lane) * 4 + t) of the row with q = (I / 16) * 4 + (J + 32) / 16,
 * lane = ((J + 32) % 16 / 4) * 16 + I % 16, t = (J + 32) % 4; zero off the
 * annulus.  The bf16 values equal dtype 1's. */
int fr_search_bank(fr_ctx *ctx, int nsam, int maxr1, int maxr2, const float *d_a3df, int norient,
                   const float *d_dm, int dtype, void *d_bank);

/* Scores np particles (d_outc: OUTC of fr_prep_batch, d_ctff) against the
 * bank.  Orientations come in views of `group` consecutive in-plane angles;
 * each view contributes its first maximum (ccp.f:144-264) and the topk
 * (<= 32) views per particle are returned by strict '>' in grid order
 * (psearch.f:89-102): orientation index, shift row ((sx+7)*15+(sy+7), NSAMR
 * pixels) and score.  rbfact: RBFACT of lmain.f:556. */
int fr_search_batch(fr_ctx *ctx, int nsam, int maxr1, int maxr2, float rbfact, int norient,
                    const void *d_bank, int dtype, int group, int np, const float *d_outc,
                    const float *d_ctff, int topk, int *d_top_o, int *d_top_row,
                    float *d_top_score);

/* Replaces PSEARCH -> CCP (psearch.f:8-121, ccp.f:8-275) with ROTATE2D and
 * PAD (rotate2d.f:8-77, pad.f:8-68) as the reference computes them, for np
 * particles (IRAD > 0, IPAD = 1, IEWALD = 0, non-helical ASYM): each view's
 * CTF-multiplied section is padded to 4 NSAM with its edge mean, rotated
 * through the NPSI in-plane steps by nearest-neighbour lookup, turned through
 * the IQUADMAX quadrants by index permutation and correlated with the
 * B-factor-weighted particle on the NSAMR grid (argmax over -6..6, first
 * maximum, 3x3 box / 9, normalised by sqrt(SUM) sqrt(SUM2)).  Call order and
 * arithmetic of the reference's bookkeeping are kept: particle k draws its
 * jitter as the k-th PSEARCH call (RANDOM, *iran advanced), XPAR, TESTPAR by
 * strict '>' with CC > 0 admission (columns of unadmitted slots keep their
 * contents, CC = 0), and the running best.
 *   amag     host np: AMAGP of each particle (the DM is divided by |AMAG|)
 *   par      host np x 5, in/out: PHI THETA PSI DSHX DSHY -> the best view
 *   sang     host nsang x 3 from fr_search_grid; dangin as given to SEARCHANG
 *   mask     host PMASK(1..5); sm: SM (NULL: identity)
 *   testpar  host np x ipmax x 6 (TESTPAR(6,IPMAX) per particle), in/out
 * MAXR2 <= min(32, NSAMR/2). */
int fr_psearch(fr_ctx *ctx, int nsam, int np, const float *amag, const float *d_a3df,
               const float *d_spec, const float *d_ctff, float dangin, int maxr1, int maxr2,
               float *par, const float *sang, int nsang, int iquadmax, const int *mask,
               float rbfact, float *testpar, int ipmax, int *iran, const float *sm, char asym);

/* The device part of fr_psearch: CCP (ccp.f:8-275 with rotate2d.f, pad.f) for
 * np x nview (particle, view) items.  d_dm holds each item's DM (np x nview x
 * 9, fr_rotation with amag = AMAG and SM at the view's XPAR, psearch.f:69-80);
 * dang is CCP's DANGIN in degrees (NPSI = NINT(90/DANG), ccp.f:62-64).  Per
 * item: d_score the CCMAX of the view and d_code its argument,
 * ((IPSI-1)*4 + IQUAD-1)*225 + (IXM+7)*15 + (IYM+7). */
int fr_ccp_view_batch(fr_ctx *ctx, int nsam, int np, int nview, const float *d_a3df,
                      const float *d_spec, const float *d_ctff, const float *d_dm, int maxr1,
                      int maxr2, float rbfact, float dang, int iquadmax, const int *mask,
                      float *d_score, int *d_code);

/* ------------------------------------------------------------------ SIGMA */

/* Replaces SIGMA (sigma.f:8-221; non-helical, IEWALD=0, IPAD=1) for n items:
 * item k compares particle pidx[k] (d_datc: DATC of fr_prep_batch, d_ctff)
 * with the section at par[5k..5k+4] = PHI, THETA, PSI, SHX, SHY (radians,
 * phase units) for magnification amag[k], in the band MAXR1^2 <= r^2 <
 * (INT(NSAM*RCLAS*|AMAG|))^2 and the real-space disc DMASK(4)/AMAG around
 * the projection of the 3D mask centre DMASK(1..3).  signset (may be NULL):
 * the previous cycle's sigma per item (0: none).  sm (may be NULL: identity)
 * is the symmetry-related matrix SM of the DMs.  out is host n x 3:
 * SIGMA2N, SIGMA2, LOGPC.  pidx, par, amag, dmask, signset are host arrays. */
int fr_sigma_batch(fr_ctx *ctx, int nsam, int n, const float *d_a3df, const float *d_datc,
                   const float *d_ctff, const int *pidx, const float *par, const float *amag,
                   int maxr1, float rclas, const float *dmask, float sigp, const float *signset,
                   const float *sm, float *out);

/* ---------------------------------------------------------------- PREFINE */

/* Replaces VA04A (va04.f:8-369), Powell's minimiser, for a host objective:
 * minimises fn over x[0..n-1] (n <= 6) with initial steps e, ESCALE, ICON and
 * MAXIT as VA04A takes them (PREFINE: ESCALE 100, ICON 1, MAXIT 50).  FP32 in
 * the reference's order: for the same objective values the iterates are the
 * reference's.  Returns F in *f and the number of objective calls. */
int fr_powell_minimize(int n, float *x, const float *e, float escale, int icon, int maxit,
                       float (*fn)(int n, const float *x, void *user), void *user, float *f,
                       int *nevals);

/* Replaces CTFREFINE (ctfrefine.f:8-100) for ng groups at once: VA04A over
 * the defocus change (x1, x2, x3) = (dDFMID1, dDFMID2, dANGAST) -- three
 * parameters when fastig (FASTIG), else one with x2 = x1, x3 = 0 -- of
 * CALCFX's CTF section (calcfx.f:63-110): the mean over the group's particles
 * of -CC3 (cc3.f) after CTFAPPLY at DF1+x1, DF2+x2, ANGAST+x3 (ICTF 0), minus
 * FD(SIG2, x1, x2, DFSTD) (fd.f).  Group g holds particles gstart[g] ..
 * gstart[g+1]-1 (a film, or one particle when defocus is refined per particle).
 *   d_spec  np spectra, NSAM x JC complex each (the r2c of fr_prep_batch)
 *   ppar    host np x 8 as fr_prep_batch's d_ppar; euler host np x 3 (rad)
 *   sig2    host ng: SIG2 of FD (0 for a film, lmain.f:400-411; SIG2N per
 *           particle, lmain.f:492-500)
 *   gpar    11 host floats as fr_prep_batch (CTFAPPLY); ring RI2 RI3 RIH
 *           HALFW of CC3's MASKCOS2; MAXR = INT(NSAM RMAX |AMAG|) per particle
 *   out     host ng x 4: dDFMID1, dDFMID2, dANGAST and the final objective
 * EPAR (1, 1, 1), ESCALE 1000, MAXIT 50 (ctfrefine.f:32-35); the caller adds
 * the changes to the group's particles (ctfrefine.f:64-84). */
int fr_ctfrefine_batch(fr_ctx *ctx, int nsam, int np, const float *d_spec, const float *ppar,
                       const float *euler, int ng, const int *gstart, int fastig, float dfstd,
                       const float *sig2, const float *d_a3df, const float *d_pssnr,
                       const float *gpar, float rmax1, float rmax2, float rbfact,
                       const float *ring, float *out, int *nevals);

/* Replaces MAGREFINE (magrefine.f:8-73) for ng groups (films): VA04A over the
 * relative magnification AMAG from the group's first particle's AMAGP (EPAR
 * 0.01, ESCALE 100, MAXIT 50) of CALCFX's magnification section
 * (calcfx.f:112-140): the mean of -CC3 after CTFAPPLY at AMAG, MAXR =
 * INT(NSAM RMAX |AMAG|).  Arguments as fr_ctfrefine_batch; out host ng x 4:
 * AMAG, 0, 0, the final objective. */
int fr_magrefine_batch(fr_ctx *ctx, int nsam, int np, const float *d_spec, const float *ppar,
                       const float *euler, int ng, const int *gstart, const float *d_a3df,
                       const float *d_pssnr, const float *gpar, float rmax1, float rmax2,
                       float rbfact, const float *ring, float *out, int *nevals);

/* The local grid search of a mode-1 refinement, before PREFINE (no routine of
 * the reference scans a grid here: LMAIN calls PREFINE from the particle's
 * parameters, lmain.f:743-828).  For each of np particles (pidx 0..np-1 of
 * d_spec/d_ctff) the (2 nang + 1)^3 orientations PHI + a DANG, THETA + b DANG,
 * PSI + c DANG (a, b, c = -nang..nang; angles summed in FP32, radians) are
 * scored by CC3M (cc3m.f:8-169, the score PREFINE maximises) at every shift
 * of the (2 nsh + 1)^2 window SHX + ix DSH, SHY + iy DSH (phase units,
 * ix, iy = -nsh..nsh), and the first maximum in grid order (PHI outer, then
 * THETA, PSI, then the shift with SHX outer) is kept.  par: host np x 6, in:
 * PHI THETA PSI SHX SHY (the centre), out: the best grid point and its CC3M.
 * amag: host np (DM divided by |AMAG|, cc3m.f:60-64); sm: host 9 or NULL
 * (DM <- DM^T SM, matmul_t.f).  NSAM a multiple of 32 in 64..256, MAXR2 <=
 * 32, nang <= 12, nsh <= 5.  d_scores / d_codes (device, np x (2 nang + 1)^3,
 * may be NULL) receive every orientation's best score and shift index. */
int fr_local_search_batch(fr_ctx *ctx, int nsam, int np, const float *d_a3df,
                          const float *d_spec, const float *d_ctff, const float *amag, float *par,
                          int nang, float dang, int nsh, float dsh, int maxr1, int maxr2,
                          float rbfact, const float *d_fsct, const float *ring, const float *sm,
                          float *d_scores, int *d_codes);

/* Replaces PREFINE (prefine.f:8-162; IVFLAG=0, non-helical, AMAG > 0, IEWALD=0)
 * for n items refined in lock-step: each round evaluates every active item's
 * CALCFX (calcfx.f:152-191: -CC3M - FSH + range penalties) in one batched CC3M
 * launch.  Item k uses particle pidx[k] of d_spec/d_ctff (as fr_cc3m_batch)
 * and magnification amag[k].  par is host n x 6: PHI THETA PSI DSHX DSHY in
 * (radians, phase units), refined values and PRESA (the final score) out.
 * mask: the 5 refinement flags of PMASK.  prior is host n x 5 per item: XM-SHX,
 * YM-SHY, SX, SY, SIG2 of FSH (fsh.f).  nevals (may be NULL) receives each
 * item's number of CC3M evaluations. */
int fr_refine_batch(fr_ctx *ctx, int nsam, int n, const float *d_a3df, const float *d_spec,
                    const float *d_ctff, const int *pidx, const float *amag, float *par,
                    const int *mask, int maxr1, int maxr2, float rbfact, const float *d_fsct,
                    const float *ring, const float *prior, int maxit, int *nevals);

/* Candidate-list form of fr_search_batch (no reference counterpart: the
 * learned proposal prunes PSEARCH's view loop, psearch.f:69-87, to the
 * neighbourhoods of the proposed classes).  Particle p is scored against the
 * bank rows d_cand[p*ncand .. p*ncand+ncand-1] (-1: empty slot), which come
 * in groups of `group` in-plane candidates; each group contributes its first
 * maximum and the topk best groups are returned as for fr_search_batch, with
 * d_top_o holding bank orientation indices. */
int fr_search_list_batch(fr_ctx *ctx, int nsam, int maxr1, int maxr2, float rbfact, int norient,
                         const void *d_bank, int dtype, int np, const float *d_outc,
                         const float *d_ctff, int ncand, const int *d_cand, int group, int topk,
                         int *d_top_o, int *d_top_row, float *d_top_score);

/* Ragged-list form of fr_search_batch (PSEARCH/CCP, psearch.f:8-121 and
 * ccp.f:8-275, each particle restricted to its own list): particle p is
 * scored against bank rows d_cand[d_loff[p] .. d_loff[p+1]-1] (-1: empty
 * slot), in view groups of d_grp[p] entries; all device arrays.  A list that
 * is the whole bank in order with the bank's view group gives exactly
 * fr_search_batch's answer for that particle.  cap is the length of d_cand;
 * list positions at or beyond it are ignored.  One persistent GEMM launch
 * covers every list.  d_info (np x 2 from fr_build_lists, or NULL): with the
 * separable BF16 bank (dtype 2), the particles flagged as falling back
 * (d_info[2p] == 0; their lists are the whole bank in order) are searched
 * four to a workgroup against shared templates instead of one list tile at
 * a time, which reads the bank once per four such particles; the scores are
 * the same. */
int fr_search_ragged_batch(fr_ctx *ctx, int nsam, int maxr1, int maxr2, float rbfact,
                           int norient, const void *d_bank, int dtype, int np,
                           const float *d_outc, const float *d_ctff, const int *d_loff,
                           const int *d_grp, const int *d_cand, const int *d_info, int cap,
                           int topk, int *d_top_o, int *d_top_row, float *d_top_score);

/* Candidate lists of the learned search (no reference counterpart; the
 * fallback list is PSEARCH's full view loop, psearch.f:73-102).  d_top and
 * d_prob are np x ktop classes and probabilities from fr_cnn_propose.  A
 * particle with d_prob[p*ktop] >= threshold gets the concatenated cells of
 * its classes (class c's cell: d_cells[d_cell_off[c] .. d_cell_off[c+1]-1],
 * -1 padded to groups of `group`); any other particle gets the whole bank
 * 0..nbank-1 in groups of bank_group.  Lists are clipped to maxlen entries
 * (whole groups) and written to d_cand (np*maxlen ints) at d_loff (np+1);
 * d_grp (np) receives the group sizes and d_info (np x 2) the confidence flag
 * and the number of non-padding entries. */
int fr_build_lists(fr_ctx *ctx, int np, int ktop, const int *d_top, const float *d_prob,
                   float threshold, int ncls, const int *d_cell_off, const int *d_cells, int group,
                   int nbank, int bank_group, int maxlen, int *d_loff, int *d_grp, int *d_cand,
                   int *d_info);

/* FP32 rescoring of each particle's k search candidates (the score LMAIN
 * refines them with, lmain.f:615-642): CC3M (as fr_cc3m_batch, AMAG = 1) at
 * bank orientation d_top_o[p*k+q] (rotation d_dm_bank, nbank x 9) and the
 * CCP shift of map row d_top_row[p*k+q], SX = IXM/NSAMR*2pi (ccp.f:254-256).
 * d_score (np x k) receives the scores, -inf for empty candidates (-1), and
 * d_best (np) the first maximum's position (strict '>', lmain.f:634). */
int fr_rescore_batch(fr_ctx *ctx, int nsam, int np, int k, const float *d_a3df,
                     const float *d_outc, const float *d_ctff, const int *d_top_o,
                     const int *d_top_row, int nbank, const float *d_dm_bank, int nsamr,
                     int maxr1, int maxr2, float rbfact, const float *d_fsct, const float *ring,
                     float *d_score, int *d_best);

/* ------------------------------------------- learned orientation proposal
 * No counterpart in the reference.  A network of nconv 3x3 / stride 2 / pad 1
 * convolutions with ReLU (1 -> C -> ... -> C channels on an S x S input) and
 * one fully connected layer to K class logits.  Parameters are one flat
 * array (per convolution W[C_out][C_in][3][3], b[C_out]; then W_fc[K][F],
 * b_fc[K], F = C (S >> nconv)^2): an FP32 master copy and d_wq, the values
 * the kernels use (weights rounded to BF16, biases exact).  Images are
 * nb x S x S FP32, logits nb x K. */
int fr_cnn_param_count(int S, int C, int nconv, int K, long long *nparam);

/* He-normal weights from a counter-based stream of `seed`, zero biases. */
int fr_cnn_init(fr_ctx *ctx, int S, int C, int nconv, int K, unsigned long long seed,
                float *d_master, float *d_wq);

int fr_cnn_forward(fr_ctx *ctx, int S, int C, int nconv, int K, const float *d_wq, int nb,
                   const float *d_img, float *d_logits);

/* The same logits with the activations stored as BF16 and the convolutions
 * after the first on the BF16 matrix cores (FP32 accumulation); C must be
 * 64.  The proposal's inference; no counterpart in the reference. */
int fr_cnn_forward_bf16(fr_ctx *ctx, int S, int C, int nconv, int K, const float *d_wq, int nb,
                        const float *d_img, float *d_logits);

/* Softmax cross-entropy against d_label (class per image): per-image losses
 * in d_loss[nb] and the gradient of their mean in d_grad (param layout). */
int fr_cnn_gradients(fr_ctx *ctx, int S, int C, int nconv, int K, const float *d_wq, int nb,
                     const float *d_img, const int *d_label, float *d_loss, float *d_grad);

/* The same step in mixed precision, as training uses it: every GEMM of the
 * forward and backward passes takes its operands rounded to BF16 (nearest
 * even) and accumulates in FP32; elementwise work, the loss and the outputs
 * are FP32. */
int fr_cnn_gradients_bf16(fr_ctx *ctx, int S, int C, int nconv, int K, const float *d_wq,
                          int nb, const float *d_img, const int *d_label, float *d_loss,
                          float *d_grad);

/* One Adam step (beta 0.9/0.999, eps 1e-8, bias-corrected; step >= 1) on the
 * master copy; d_wq is refreshed from it. */
int fr_cnn_adam(fr_ctx *ctx, int S, int C, int nconv, int K, float *d_master, float *d_wq,
                const float *d_grad, float *d_m, float *d_v, int step, float lr);

/* Softmax of each logit row; the topk classes (descending probability, lower
 * class on ties) and their probabilities, nb x topk each. */
int fr_cnn_propose(fr_ctx *ctx, int K, int nb, const float *d_logits, int topk, int *d_top,
                   float *d_prob);

/* Network input of nb prepared particles (DATC and CTFF of fr_prep_batch):
 * the S x S Fourier crop, phase-flipped by sign(Re(CTFR + conj CTFL)),
 * brought to real space with the particle centre at S/2 and normalised to
 * zero mean and unit variance. */
int fr_cnn_images(fr_ctx *ctx, int nsam, int S, int nb, const float *d_datc, const float *d_ctff,
                  float *d_img);

/* Training images of the same form from the current reference: central
 * section d_dm[9b..] (fr_rotation with amag=0) x CTF (ctf.f, gpar as in
 * fr_prep_batch) x the ramp exp(-i(SHX I + SHY J)) + white noise at SNR.
 * d_spar is nb x 8: SHX, SHY, DF1, DF2, ANGAST, SNR, 0, 0.  The SNR is the
 * signal variance over `disc` pixels of the NSAM^2 image over the noise
 * variance per pixel.  Noise comes from a counter-based stream of `seed`. */
int fr_cnn_synth(fr_ctx *ctx, int nsam, int S, int nb, const float *d_a3df, const float *d_dm,
                 const float *d_spar, const float *gpar, float disc, unsigned long long seed,
                 float *d_img);

/* Orientation classes under a point group.  Orientation matrices are the
 * DM arrays of fr_rotation (amag=0); d_sym holds nsym row-major 3x3
 * operators.  fr_orient_nearest: class of each of n orientations (largest
 * symmetry-maximised trace).  fr_orient_within: bit i of d_mask[b*W + w]
 * (W = ceil(na/32)) is set when orientation a = 32w+i of set A lies within
 * max_deg of orientation b of set B. */
int fr_orient_nearest(fr_ctx *ctx, int n, const float *d_dm, int ncls, const float *d_dmc,
                      int nsym, const float *d_sym, int *d_label);
int fr_orient_within(fr_ctx *ctx, int na, const float *d_dma, int nb, const float *d_dmb,
                     int nsym, const float *d_sym, float max_deg, unsigned *d_mask);

/* ------------------------------------------------------------ file formats
 * The reference's MRC/CCP4 images and maps (iof.f:177-780) and parameter-file
 * records (card10.f:92-300, lmain.f:913-923), host side. */

/* IOPEN's header read (SET_MHDR, iof.f:994-1060): NX NY NZ, MODE, PSIZE =
 * CELLA(1)/MX, (DMIN, DMAX, DMEAN, RMS), NSYMBT; any pointer but nxyz may be
 * NULL.  Modes 0, 1, 2; byte-swapped files by the machine stamp. */
int fr_mrc_info(const char *path, int *nxyz, int *mode, float *psize, float *stats, int *nsymbt);
/* IREAD (iof.f:505-576) of sections first .. first+count-1 into float32 (mode
 * 0 bytes unsigned, mode 1 INTEGER*2, mode 2 REAL*4). */
int fr_mrc_read(const char *path, long long first, long long count, float *out);
/* IOPEN/IWRITE/ICLOSE of a new mode-2 file: SET_MHDR's fields (CELLA = PSIZE
 * N, CELLB 90, MAPCRS 1 2 3, ISPG 1, 'MAP ', little-endian stamp) and ICLOSE's
 * statistics (FP64 sums of the values and of their FP32 squares); label is
 * the first 80-character title (may be NULL). */
int fr_mrc_write(const char *path, int nx, int ny, int nz, float psize, const float *data,
                 const char *label);
/* The v9 record of lmain.f:913-923, FORMAT(I7,3F8.2,2F10.2,I8,I6,2F9.1,2F8.2,
 * I10,F11.4,2F8.2), from the printed values (MAG and LogP as the REALs whose
 * NINT is printed); out receives 136 characters and a terminating 0. */
int fr_par_format(int c, float psi, float theta, float phi, float shx, float shy, float mag,
                  int film, float df1, float df2, float angast, float occ, float logp,
                  float sigma, float score, float change, char *out, int cap);
/* card10.f:136-172: one record into rec[15] = C PSI THETA PHI SHX SHY MAG FILM
 * DF1 DF2 ANGAST OCC LogP SIGMA SCORE as written; *format 0 (pre-v9, up to 95
 * characters: OCC 100, SIGMA 0, SCORE the old phase residual), 1 (v9 fixed
 * columns, up to 128: SIGMA 0) or 2 (list-directed).  Returns 1 for a comment
 * line ('C' in column 1). */
int fr_par_parse(const char *line, double *rec, int *format);

/* calc_occ.f:179-217 (the CALC_OCC program's update): np particles in nc
 * classes, arrays [class][particle] in FP32: OCC (in/out, %), LogP (the
 * integers of the par files), SIGMA (in/out); avg[i] = A(I), class i's
 * average occupancy; f the OCC change multiplier. */
int fr_calc_occ(int nc, int np, float f, const float *avg, float *occ, const float *algp,
                float *sig);

/* --------------------------------- per-call entry points, Fortran symbols
 * The reference's own routine names and argument lists (Fortran 77 implicit
 * interfaces: every argument by reference, the CHARACTER length of ASYM
 * appended), so a Fortran LMAIN links against this library unchanged.  Each
 * call runs one item through the batched kernels on a library-wide context
 * on device 0; not thread-safe, as the reference's own routines are called
 * serially.  Per-particle arrays are copied per call.  The reference
 * transform (C3DF/A3DF) stays resident on the device and is uploaded again
 * only when a call names another host array or the array's fingerprint
 * (1,024 sampled words) changed, or after fr_percall_invalidate.  A3D3's
 * half-maps accumulate on the device and are copied back into the host
 * arrays at the flush of the last particle (K = ILAST) or by
 * fr_percall_flush.  Supported: IPAD = 1, IRAD > 0, IEWALD = 0, non-helical
 * ASYM, IBUF = 0.  Other cases and device errors stop the program (message
 * on stdout, exit(1)), as the reference's STOPs. */

/* CTFAPPLY (ctfapply.f:8-161; XSTD >= 0, non-helical): CTFF, DATC and the
 * masked, SNR-weighted transform in OUTD (OUTD and OUTC are one array,
 * frealign_v9.f:679).  CTFF is recomputed on every call. */
void ctfapply_(const int *nsam, const float *spec, const float *shx, const float *shy,
               const float *cs, const float *wl, const float *wgh1, const float *wgh2,
               const float *dfmid1, const float *dfmid2, const float *angast,
               const float *thetatr, float *ctff, float *outd, float *outc,
               const float *amagp, const float *rih, const float *halfw, const float *ri2,
               const float *ri3, const float *ri4, float *datd, float *datc, const float *b3dv,
               const float *phi, const float *theta, const float *psi, float *w,
               const float *xstd, const int *ibuf, const float *tx, const float *ty,
               const char *asym, const float *pssnr, float *rbuf, float *pweights,
               const int *ictf, const void *fftw_plans, size_t asym_len);

/* CC3M (cc3m.f:8-169): REAL FUNCTION */
float cc3m_(const int *nsam, const int *irad, const float *amag, const float *spec,
            const float *a3df, const int *maxr1, const int *maxr2, const float *phi,
            const float *theta, const float *psi, const float *shx, const float *shy,
            const int *ibuf, float *pbuc, const float *rbfact, const float *sinclut,
            const int *ipad, float *rbuf, const float *fsct, const int *iewald,
            const float *thetatr, const float *ctff, const float *ri2, const float *ri3,
            const float *rih, const float *halfw, const float *pweights,
            const void *fftw_plans, const float *sm);

/* PEXTRACT (pextract.f:8-101) */
void pextract_(const int *nsam, const int *irad, float *outc, const float *a3df, const float *phi,
               const float *theta, const float *psi, const float *sinclut, const int *ipad,
               const float *thet, const int *iewald, const float *ctff, const float *sm);

/* PRESB (presb.f:8-160) */
void presb_(const int *nsam, const int *irad, const float *amag, const float *spec,
            const float *a3df, const int *maxr1, const int *maxr2, const float *phi,
            const float *theta, const float *psi, const float *shx, const float *shy, float *rbin,
            const float *sinclut, const int *ipad, int *ns, float *bufa, float *bufb,
            float *bufab, const float *cs, const float *wl, const float *wgh1, const float *wgh2,
            const float *df1, const float *df2, const float *angast, const float *thet,
            const int *iewald, const float *ctff, const float *ri2, const float *ri3,
            const float *rih, const float *halfw, const char *asym, const void *fftw_plans,
            const float *sm, size_t asym_len);

/* SIGMA (sigma.f:8-221) */
void sigma_(const int *nsam, const int *irad, const float *amag, const float *a3df,
            const float *datc, const float *phi, const float *theta, const float *psi,
            const float *shx, const float *shy, const float *sinclut, const int *ipad,
            const int *iewald, const float *thetatr, const float *ctff, float *rbuf,
            const float *ri2, const int *maxr1, const int *maxr2, const float *signset,
            float *sigma2n, float *sigma2, float *logpc, const float *dmask, const char *asym,
            const int *k, const float *sigp, const float *pweights, const float *rclas,
            const void *fftw_plans, const float *sm, size_t asym_len);

/* PREFINE (prefine.f:8-162; IVFLAG = 0, AMAG > 0, non-helical): refined
 * PHI, THETA, PSI, DSHX, DSHY (as MASK selects), PRESA and NPAR out. */
void prefine_(const int *nsam, const int *irad, float *amag, const int *maxr1, const int *maxr2,
              float *phi, float *theta, float *psi, float *dshx, float *dshy, float *presa,
              const int *mask, const float *rbfact, int *npar, const float *c3df, float *pbuf,
              const float *shx, const float *shy, const int *ipad, const float *cs,
              const float *wl, const float *wgh1, const float *wgh2, const float *thetatr,
              const float *ctff, const float *amagp, const float *rih, const float *halfw,
              const float *ri2, const float *ri3, const float *ri4, const float *rmax1,
              const float *rmax2, const float *xstd, float *mbuf, const int *ilst,
              const float *datc, const int *ibuf, const float *b3dv, float *datd,
              const float *sinclut, const int *ivflag, float *outd, const float *outc,
              float *qbuc, float *rbuf, const float *fsct, const int *iewald, const float *xm,
              const float *ym, const float *sx, const float *sy, const float *sig2,
              const char *asym, const float *thetam, const float *stheta, const float *psim,
              const float *spsi, const float *stif, const float *pweights,
              const void *fftw_plans, const float *sm, size_t asym_len);

/* PINSERT_TRILIN (pinsert_trilin.f:8-244; IEWALD = 0, IPAD = 1): the
 * stripe IP of width ST of the plane DM(1..6), phase shift SHX+DX, SHY+DY,
 * added into B3DF, T3DF, NKSUM and TTD (volumes in the batched layout); with
 * NNSTAT != 0 also into KSUM, VSUM, PSUM, ASUM against C3DF, as the _SE
 * variant does. */
void pinsert_trilin_(const int *nsam, const float *rii, const int *irrec2, const float *spec,
                     const float *ctff, const float *occ, float *b3df, float *t3df, double *ttd,
                     const float *shx, const float *shy, float *asum, float *vsum, float *psum,
                     float *c3df, int *ksum, const float *sinclut, const int *irada,
                     const int *ipad, const float *thet, const int *iewald, const int *nnstat,
                     long long *nksum, const int *jc, const float *dm, const float *tmp,
                     const int *nsamh, const int *ic, const float *dx, const float *dy,
                     const float *st, const int *ip, int *ns1);

/* PINSERT_TRILIN_SE (pinsert_trilin_se.f:8-359; IEWALD = 0, IPAD = 1): the
 * same stripe, with NNSTAT != 0 adding every term BSAMP to the per-voxel
 * statistics (pinsert_trilin_se.f:186-194): KSUM += 1, VSUM += |BSAMP|^2,
 * PSUM += PDIFF(BSAMP, C3DF), ASUM += |BSAMP|.  Same arguments as
 * PINSERT_TRILIN. */
void pinsert_trilin_se_(const int *nsam, const float *rii, const int *irrec2, const float *spec,
                        const float *ctff, const float *occ, float *b3df, float *t3df,
                        double *ttd, const float *shx, const float *shy, float *asum,
                        float *vsum, float *psum, float *c3df, int *ksum, const float *sinclut,
                        const int *irada, const int *ipad, const float *thet, const int *iewald,
                        const int *nnstat, long long *nksum, const int *jc, const float *dm,
                        const float *tmp, const int *nsamh, const int *ic, const float *dx,
                        const float *dy, const float *st, const int *ip, int *ns1);

/* A3D3 (a3d3.f:8-146; INTERP = 1, non-helical, NONSYM false): records the
 * particle in RECBF/CTFBF/RECB/ISW, alternates ISWITCH, and when NVOL
 * particles are buffered or K = ILAST flushes slot I into private volume IS
 * (column IS of A3DF/S3DF(NN3, NVOL1) or D3DF/V3DF(NN3, NVOL2); IS = I, or
 * (I-1)/2+1 when IFSC = 0, a3d3_s.f:39-40), NKSUM1(I) and TTD1/TTD2(IS);
 * with NNSTAT != 0 also into column I of KSUM/VSUM/PSUM/ASUM(NN3+1, NVOL)
 * against C3DF (the _SE insertion).  PBC is REAL storage (pinsert.f:31 reads
 * it so). */
void a3d3_(const int *nset, int *iswitch, const int *nsam, const int *irad, const float *ri,
           const float *rrec, const int *k, const int *ilast, const float *amagp,
           const float *datc, const float *ctff, float *asum, float *vsum, float *psum,
           float *c3df, int *ksum, const float *phi, const float *theta, const float *psi,
           const float *dshx, const float *dshy, const float *presa, const float *pbc,
           const float *boff, float *a3df, float *s3df, double *std_, float *d3df, float *v3df,
           double *vtd, const int *nn1, const int *maxset, const float *sinclut,
           const int *irada, const int *ipad, const float *thetatr, const int *iewald,
           const int *nsym, const int *isymax, const int *jsym, const float *symop,
           const int *nnstat, long long *nksum, const int *ifsc, const int *imp,
           const int *nvol, const int *nvol1, const int *nvol2, const float *occ,
           const float *sig2n, const float *psize, const char *asym, const float *alpha,
           const float *rise, const int *nu, const int *hstart, long long *nksum1,
           double *ttd1, double *ttd2, int *ns, int *ns1, const int *interp, float *recbf,
           float *recb, int *isw, int *ibuf, float *ctfbf, const int *nn2, const int *nn3,
           const float *sm, const int *nonsym, const int *ism, size_t asym_len);

/* PSEARCH (psearch.f:8-121; non-helical, IPAD = 1, IEWALD = 0): the
 * reference-exact global search of fr_psearch for one particle.  Returns the
 * best view in PHI, THETA, PSI, DSHX, DSHY, the IPMAX best views in TESTPAR
 * (strict '>', CC > 0 only) and IRAN after its three RANDOM draws. */
void psearch_(const int *nsam, const int *irad, const float *amag, const float *spec,
              const float *a3df, const float *dangin, const int *maxr1, const int *maxr2,
              float *phi, float *theta, float *psi, float *dshx, float *dshy, float *presa,
              const float *sang, const int *nsang, float *ccd, float *ccc, float *cbuf,
              const int *iquadmax, const int *mask, const float *rbfact, const float *sinclut,
              const int *ipad, float *rbuf, const int *iewald, const float *thetatr,
              const float *ctff, const float *ri2, const float *ri3, const float *rih,
              const float *halfw, float *testpar, const int *ipmax, const char *asym, int *iran,
              const void *fftw_plans, const float *sm, size_t asym_len);

/* No counterpart in the reference: the residency controls of the per-call
 * symbols, for a host that changes the reference transform in place
 * (invalidate: the next call uploads it again) or reads A3D3's half-maps
 * before the last particle (flush: copy them back now).  The trailing-
 * underscore forms are CALL FR_PERCALL_INVALIDATE / CALL FR_PERCALL_FLUSH
 * from Fortran. */
void fr_percall_invalidate(void);
void fr_percall_invalidate_(void);
void fr_percall_flush(void);
void fr_percall_flush_(void);

/* ------------------------------------------- batches of uneven box sizes
 * No counterpart in the reference (its particles go one at a time through
 * LMAIN, lmain.f:509-885).  A plan serves one box size and batch size: the
 * score-and-insert pipeline of a batch -- fr_prep_batch, fr_cc3m_batch at
 * each particle's current parameters, the A3D3 flush of fr_insert_batch into
 * the bound half-maps -- captured once as a HIP graph and replayed per batch.
 * Inputs per batch: the images, written into fr_plan_images() (nb x NSAM x
 * (NSAM+2) floats, device), and the host records ppar (nb x 8, as
 * fr_prep_batch) and recb/half (as fr_insert_batch; the CC3M uses PHI THETA
 * PSI DSHX DSHY AMAGP of recb).  score (host, may be NULL) receives the CC3M
 * of each particle.  Results equal the eager sequence of the three entry
 * points (NKSUM exactly; scores and half-maps up to the order of float
 * atomics in the shell and insertion sums).
 * The CC3M launch variant is the context's choice for NSAM at creation
 * (fr_autotune_cc3m). */
typedef struct fr_plan fr_plan;
int fr_plan_create(fr_ctx *ctx, int nsam, int nb, const float *d_a3df, const float *d_pssnr,
                   const float *gpar, int ictf, int maxr1, int maxr2, float rbfact,
                   const float *d_fsct, const float *ring, int nsym, const int *jsym,
                   const float *symop, float pbc, float boff, float psize, float *d_num0,
                   float *d_ctf2_0, float *d_num1, float *d_ctf2_1, long long *d_nksum,
                   double *d_ttd, fr_plan **plan);
float *fr_plan_images(fr_plan *plan);
/* Copies nb images (device, nb x NSAM x (NSAM+2) floats) into the plan. */
int fr_plan_load(fr_plan *plan, const float *d_images);
int fr_plan_run(fr_plan *plan, const float *ppar, const float *recb, const int *half,
                float *score);
int fr_plan_destroy(fr_plan *plan);

/* Launch autotuner (no reference counterpart): times the CC3M variants that
 * apply at NSAM -- the band-pruned fused kernel with one workgroup per item
 * or a persistent grid of 1/2/4/8 workgroups per CU, and the FFT path -- on
 * the caller's items (arguments as fr_cc3m_batch), keeps the fastest for
 * NSAM in the context and writes the timings and the choice as JSON into
 * log (logcap bytes, may be NULL). */
/* Sets the CC3M variant for NSAM by hand: fused (1) or FFT path (0), and the
 * fused kernel's persistent workgroup count (0: one per item).  With no
 * choice made for NSAM the fused kernel runs one full wave of workgroups (as
 * many as the device holds at once), each looping over the items. */
int fr_set_cc3m_variant(fr_ctx *ctx, int nsam, int fused, int workgroups);
int fr_autotune_cc3m(fr_ctx *ctx, int nsam, int n, const float *d_a3df, const float *d_spec,
                     const float *d_ctff, const int *d_pidx, const float *d_dm,
                     const float *d_shift, const float *d_amag, int maxr1, int maxr2,
                     float rbfact, const float *d_fsct, const float *ring, float *d_score,
                     char *log, int logcap);

#ifdef __cplusplus
}
#endif
#endif

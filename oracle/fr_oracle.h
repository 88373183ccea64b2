/*
 * fr_oracle.h -- CPU restatement of the Frealign v9.12 per-particle hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in frealign_amd/csrc.  Only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py may load it; the product path never
 * calls it.
 *
 * Layout conventions (identical to the reference's Fortran storage, 0-based):
 *   half-plane  : JC x NSAM complex64, index l + JC*m, l = x-frequency 0..NSAM/2,
 *                 m = wrapped y-frequency (m >= JC means m-NSAM); ctfapply.f:64-76
 *   CTFF        : 2 half-planes, CTFR then CTFL (offset IC = NSAM*JC); ctfapply.f:78-79
 *   real image  : NSAM x NSAM float, row pitch NSAM+2 (in-place r2c layout)
 *   half-volume : JC x NSAM x NSAM complex64, index l + JC*(m + NSAM*n); shift.f:34
 * Complex arrays are interleaved (re, im).
 *
 * Every value-producing routine exists in two precisions:
 *   or32_*  FP32 arithmetic in the reference's operation order (the restatement)
 *   or64_*  the same algorithm with FP64 values (the error yardstick)
 * Coordinates, rotation matrices and Miller indices are always FP32, exactly as
 * in the reference, so both variants visit the same voxels.
 *
 * Parity pinning (tests/test_oracle_*.py):
 *   pinned against the flang-compiled reference routines (oracle/_ref):
 *     ctf, maskcos, maskcos2, pweight, ainterpo3ds, pextract, pinsert
 *     (pinsert_trilin + symmetry loop), searchang/limitsymm, random,
 *     getsymmat, sheltest, storeshift
 *   pinned by closed forms and invariants (the reference's copies need FFTW,
 *   which this image lacks): the 2D FFTs (vs numpy.fft), ctfapply (composition
 *   of pinned parts + FFT), cc3m (self-correlation closed form, scale
 *   invariance, FP64 agreement), presb (FSC-style closed form), ccp/psearch
 *   (brute-force direct DFT on small boxes, quadrant permutation identities).
 */
#ifndef FR_ORACLE_H
#define FR_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- precision-independent pieces (fr_oracle_common.c) ---- */

/* random.f:8-32 (L'Ecuyer LCG). *is is updated in place. */
float or_random(int32_t *is);

/* Rotation matrix of pextract.f:30-51 / cc3m.f:48-65: ZYZ Euler, divided by
 * |amag| unless amag == 0 (PEXTRACT does not divide), then DM <- MATMUL_T(DM,SM)
 * when sm != NULL (matmul_t.f:17-24).  dm is column-major 3x3 (Fortran DM(1..9)). */
void or_dm(float phi, float theta, float psi, float amag, const float *sm, float dm[9]);

/* matmul_t.f:8-26: C <- (A^T B)^T computed in the reference order. */
void or_matmul_t(const float a[9], const float b[9], float c[9]);
/* calcsinc.f:17-19: SINCLUT(1..n) */
void or_calcsinc(float *sinclut, int n);
/* combine_arrays.f:53-67: the particle-count and CTF^2-sum counters */
void or_combine_counters(int ifsc, int imp, int nvol1, const int64_t *nksum1, const double *ttd1,
                         const double *ttd2, int64_t *nksum, double *std, double *vtd);

/* getsymmat.f:8-97 for C<n> and D<n> (and T/O/I tables).  Returns nsym;
 * fills symop (3x3 column-major per operator), jsym (rotational order per
 * operator, checksym.f) and *iredun.  asym: 'C','D','T','O','I'. */
int or_getsymmat(char asym, int nasym, float *symop, int *jsym, int *iredun);

/* searchang.f:8-150 + limitsymm.f:8-51.  Fills sang (3 x nsang radians, may be
 * NULL to count only) and returns nsang; *iquadmax and *dang_out as computed.
 * mask: PMASK(1..5). */
int or_searchang(float rmm, float ri, float dangin, char asym, int nasym,
                 const int mask[5], int asymstore, float *sang, int *iquadmax,
                 float *dang_out);

/* Unnormalised 2D real FFTs in the reference's FFTW layout (fftw_plans.f):
 * forward e^{-i}, backward e^{+i}; real data NSAM x NSAM, row pitch NSAM+2,
 * in place.  Computed in FP64; the float entry points round on output.
 * The backward transform follows the row-column c2r semantic: the imaginary
 * parts of the x=0 and x=NSAM/2 columns are ignored after the y pass. */
void or_rfft2_f(int nsam, float *data);   /* r2c in place */
void or_irfft2_f(int nsam, float *data);  /* c2r in place */
void or_rfft2_d(int nsam, double *data);
void or_irfft2_d(int nsam, double *data);

/* ---- precision-dependent pieces (fr_oracle.c compiled as or32_ / or64_) ---- */
#define OR_DECLARE(P, R)                                                        \
  /* ctf.f:8-64; out[2] = (re, im) */                                            \
  void P##ctf(float cs, float wl, float wgh1, float wgh2in, float df1,          \
              float df2, float angast, float thetatr, int ix, int iy,           \
              float tx, float ty, R *out);                                      \
  /* maskcos.f:8-55 (centre NSAM/2+1); outd pitch NSAM+2 */                     \
  void P##maskcos(int nsam, R *outd, float ri2, float ri3, float rih,           \
                  float halfw, float amagp);                                     \
  /* maskcos2.f:8-64 (centre at the FFT origin) */                              \
  void P##maskcos2(int nsam, R *outd, float ri2, float ri3, float rih,          \
                   float halfw, float amagp);                                    \
  /* pweight.f:8-98; spec and outc may alias */                                 \
  void P##pweight(int nsam, const R *spec, R *outc, const float *ssnr,          \
                  const R *ctff, int ictf);                                     \
  /* ctfapply.f:8-161 (non-helical, XSTD >= 0).  spec: half-plane of the r2c   \
   * FFT of the particle.  Outputs ctff (2 half-planes), outc (masked,          \
   * SNR-weighted), datc (shifted, unmasked, origin-centred). */                \
  void P##ctfapply(int nsam, const float *spec, float shx, float shy,           \
                   float cs, float wl, float wgh1, float wgh2, float df1,       \
                   float df2, float angast, float thetatr, float amagp,         \
                   float rih, float halfw, float ri2, float ri3,                \
                   const float *pssnr, int ictf, R *ctff, R *outc, R *datc);    \
  /* ainterpo3ds.f:8-99 (trilinear, Hermitian half-volume); nsam = IPAD*NSAM */ \
  void P##ainterpo3ds(int nsam, const float *a3df, float xx, float yy,          \
                      float zz, R *out);                                        \
  /* pextract.f:8-101 (IEWALD=0, IRAD!=0, IPAD=1) */                            \
  void P##pextract(int nsam, R *outc, const float *a3df, float phi,             \
                   float theta, float psi, const R *ctff, const float *sm);     \
  /* cc3m.f:8-169 (IEWALD=0, IBUF=0, IPAD=1, IRAD!=0) */                        \
  double P##cc3m(int nsam, float amag, const R *spec, const float *a3df,        \
                 int maxr1, int maxr2, float phi, float theta, float psi,       \
                 float shx, float shy, float rbfact, const float *fsct,         \
                 const R *ctff, float ri2, float ri3, float rih, float halfw,   \
                 const float *sm);                                              \
  /* cc3.f:8-159 (IEWALD=0, IBUF=0, IPAD=1, IRAD!=0): CC3M with the section    \
   * phase-flipped by the CTF sign and signed R_k in every shell */             \
  double P##cc3(int nsam, float amag, const R *spec, const float *a3df,         \
                int maxr1, int maxr2, float phi, float theta, float psi,        \
                float shx, float shy, float rbfact, const R *ctff, float ri2,   \
                float ri3, float rih, float halfw, const float *sm);            \
  /* presb.f:8-160 (non-helical).  rbin[2*NSAM] updated like the reference. */  \
  void P##presb(int nsam, float amag, const R *spec, const float *a3df,         \
                float phi, float theta, float psi, float shx, float shy,        \
                R *rbin, const R *ctff, float ri2, float ri3, float rih,        \
                float halfw, const float *sm);                                  \
  /* pinsert.f:8-195 with pinsert_trilin.f:8-244 (INTERP=1, IEWALD=0,          \
   * NNSTAT=0, IMP=1).  Accumulates into b3df (complex), t3df, *nksum, *ttd.    \
   * occ is the ON = OCC/SIG2N factor of a3d3_s.f:44. */                        \
  void P##pinsert(int nsam, float rrec, float amag, const R *spec,              \
                  const R *ctff, R *b3df, R *t3df, double *ttd, float phi,      \
                  float theta, float psi, float shx, float shy, float presa,    \
                  float pbc, float boff, int nsym, const int *jsym,             \
                  const float *symop, int64_t *nksum, float psize, float occ,   \
                  const float *sm, int nonsym);                                 \
  /* pinsert_trilin.f:8-244 as one call (IRAD=2, IEWALD=0, NNSTAT=0): the      \
   * stripe IP of width ST, phase shift SHX+DX, SHY+DY, rotation DM(1..6) */    \
  void P##pinsert_trilin(int nsam, int irrec2, const R *spec, const R *ctff,    \
                         float occ, R *b3df, R *t3df, double *ttd, float shx,   \
                         float shy, int64_t *nksum, const float *dm, float tmp, \
                         float dx, float dy, float st, int ip);                 \
  /* pinsert.f:8-195 -> pinsert_trilin_se.f:8-359 (NNSTAT=1, IEWALD=0,       \
   * IPAD=1): as pinsert, plus the per-voxel statistics of every inserted      \
   * term BSAMP (pinsert_trilin_se.f:186-194): ksum += 1, vsum += |BSAMP|^2,   \
   * psum += PDIFF(BSAMP, c3df), asum += |BSAMP|; c3df is the reference        \
   * transform in the half-volume layout. */                                   \
  void P##pinsert_se(int nsam, float rrec, float amag, const R *spec,           \
                     const R *ctff, R *b3df, R *t3df, double *ttd, float phi,   \
                     float theta, float psi, float shx, float shy, float presa, \
                     float pbc, float boff, int nsym, const int *jsym,          \
                     const float *symop, int64_t *nksum, float psize,           \
                     float occ, const float *sm, int nonsym, const R *c3df,     \
                     int *ksum, R *vsum, R *psum, R *asum);                     \
  /* pinsert_trilin_se.f:8-359 as one call (IEWALD=0, NNSTAT=1, IPAD=1) */      \
  void P##pinsert_trilin_se(int nsam, int irrec2, const R *spec, const R *ctff, \
                            float occ, R *b3df, R *t3df, double *ttd,           \
                            float shx, float shy, int64_t *nksum,               \
                            const float *dm, float tmp, float dx, float dy,     \
                            float st, int ip, const R *c3df, int *ksum,         \
                            R *vsum, R *psum, R *asum);                         \
  /* storeshift.f:43-102 (IFSC=0, NNSTAT=0): half-maps A,D -> B,C */            \
  void P##storeshift(int nsam, double std, double vtd, int iredun,              \
                     const R *a3df, const R *s3df, const R *d3df,               \
                     const R *v3df, R *b3df, R *c3df);                          \
  /* combine_arrays.f:33-52 (NNSTAT=0): the private copies (stride nn3)       \
   * added into copy 1 in order */                                             \
  void P##combine_arrays(int nsam, long long nn3, int nvol1, int nvol2,         \
                         int ifsc, R *a3df, R *s3df, R *d3df, R *v3df);         \
  /* shiftvol.f:8-77 (NNSTAT=0), in place; *std += vtd first */                 \
  void P##shiftvol(int nsam, double *std, double vtd, float rlim, int iredun,   \
                   R *a3df, R *d3df, R *s3df, const R *v3df, int ifsc,          \
                   const float *pssnr, float fpart, const float *fsc,           \
                   float fmask);                                                \
  /* correct3d.f:19-47 with boxft_lut.f:26-37; a3dv has NSAM+2 row pitch */     \
  void P##correct3d(int nsam, R *a3dv, int interp, int ipad,                    \
                    const float *sinclut);                                      \
  /* mask3d.f:8-142, in place on the padded real volume; FMASK, AVE, STD */    \
  void P##mask3d(int nsam, R *a3dv, float rio, float ric, float halfw,          \
                 float *fmask, float *ave, float *std);                         \
  /* sheltest.f:8-122 (FSC part, NNSTAT=0); returns nstp */                     \
  int P##sheltest(int nsam, int istp, const R *a3df, const R *d3df,             \
                  const R *s3df, const R *v3df, float fmask, int *npix,         \
                  int *npixt, double *dfsc, R *pssnr);                          \
  /* ccp.f:8-275 with rotate2d.f and pad.f (IRAD!=0, IPAD=1, IEWALD=0).        \
   * spec: B-factor weighted particle half-plane.  On return *psi, *shx, *shy  \
   * hold the best in-plane angle and shift and the score is returned. */       \
  double P##ccp(int nsam, float amag, const R *spec, const float *a3df,         \
                float dangin, int maxr1, int maxr2, float phi, float theta,     \
                float *psi, float *shx, float *shy, const int mask[5],          \
                const R *ctff, int iquadmax, const float *sm);                  \
  /* sigma.f:8-221 (non-helical, IEWALD=0): out = SIGMA2N, SIGMA2, LOGPC.     \
   * datc: the particle's DATC; dmask: DMASK(1..4); rclas: RCLAS. */            \
  void P##sigma(int nsam, float amag, const float *a3df, const R *datc,        \
                float phi, float theta, float psi, float shx, float shy,        \
                const R *ctff, int maxr1, float rclas, float signset,           \
                const float *dmask, float sigp, const float *sm, R *out);       \
  /* psearch.f:8-121.  testpar: 6 x ipmax; returns CCMAX. */                   \
  double P##psearch(int nsam, float amag, const R *spec, const float *a3df,     \
                    float dangin, int maxr1, int maxr2, float *phi,             \
                    float *theta, float *psi, float *dshx, float *dshy,         \
                    const float *sang, int nsang, int iquadmax,                 \
                    const int mask[5], float rbfact, const R *ctff,             \
                    float *testpar, int ipmax, int32_t *iran, const float *sm);

/* va04.f:8-369 (VA04A) with the objective as a callback: minimises fx over
 * x[0..n-1] (n <= 6) with steps e, ESCALE, IPRINT (printing only), ICON and
 * MAXIT as the reference takes them; returns F and the number of objective
 * evaluations in *nevals.  FP32 throughout, as the Fortran. */
typedef float (*or_objective)(int n, const float *x, void *user);
float or_va04a(float *x, const float *e, int n, float escale, int iprint, int icon, int maxit,
               or_objective fx, void *user, int *nevals);

/* prefine.f:51-145 (non-helical, AMAG > 0, IVFLAG = 0): VA04A over the
 * MASK-selected PHI THETA PSI DSHX DSHY of par[5] (in/out) minimising CALCFX
 * (calcfx.f:152-191: -or32_cc3m - FSH + the range penalties); prior =
 * (XM-SHX, YM-SHY, SX, SY, SIG2) of FSH (fsh.f).  Returns PRESA. */
float or32_prefine(int nsam, float amag, const float *spec, const float *a3df, int maxr1,
                   int maxr2, float *par, const int mask[5], float rbfact, const float *fsct,
                   const float *ctff, float ri2, float ri3, float rih, float halfw,
                   const float prior[5], const float *sm, int *nevals);

/* ctfrefine.f:8-100 with calcfx.f:63-110 (IVFLAG 1) for one group of n
 * particles: spec n half-plane spectra (r2c, interleaved complex), ppar n x 8
 * (SHX SHY DF1 DF2 ANGAST AMAGP . .), euler n x 3, gpar the 11 CTFAPPLY
 * globals (CS WL WGH1 WGH2 THETATR RIH HALFW RI2 RI3 TX TY), ring CC3's RI2
 * RI3 RIH HALFW.  xout: dDF1 dDF2 dANGAST; returns the final objective. */
float or32_ctfrefine(int nsam, int n, const float *spec, const float *ppar, const float *euler,
                     int fastig, float dfstd, float sig2, const float *a3df, const float *pssnr,
                     const float *gpar, float rmax1, float rmax2, float rbfact, const float *ring,
                     float *xout, int *nevals);
/* magrefine.f:8-73 with calcfx.f:112-140 (IVFLAG 2); xout[0] = AMAG. */
float or32_magrefine(int nsam, int n, const float *spec, const float *ppar, const float *euler,
                     const float *a3df, const float *pssnr, const float *gpar, float rmax1,
                     float rmax2, float rbfact, const float *ring, float *xout, int *nevals);

/* Closed-form test objectives shared with oracle/ref_va04_obj.f90 (same FP32
 * operation order): 1 coupled quadratic, 2 Rosenbrock chain, 3 rational bump. */
float or_va04_testfn(int n, const float *x, void *which);

OR_DECLARE(or32_, float)
OR_DECLARE(or64_, double)

#ifdef __cplusplus
}
#endif
#endif

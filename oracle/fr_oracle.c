/*
 * fr_oracle.c -- CPU restatement of the Frealign v9.12 per-particle hot path.
 * TEST INFRASTRUCTURE ONLY (see fr_oracle.h): the checker for the HIP path.
 *
 * Compiled twice: -DOR_FP64=0 gives or32_* (FP32 values, reference operation
 * order), -DOR_FP64=1 gives or64_* (FP64 values).  Geometry (radii tests,
 * rotation matrices, coordinates, Miller indices) stays FP32 in both, exactly
 * as in the reference, so both variants touch the same voxels and bins.
 * Built with -ffp-contract=off: no FMA contraction, as in the reference build.
 */
#include "fr_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#if OR_FP64
typedef double R;
#define P(n) or64_##n
#define RSQRT(x) sqrt(x)
#define RCOS(x) cos(x)
#define RSIN(x) sin(x)
#define REXP(x) exp(x)
#define RATAN2(y, x) atan2(y, x)
#define RABS(x) fabs(x)
/* CABS(Z)**2 */
#define CABS2(a, b) ((a) * (a) + (b) * (b))
#define RFFT2(n, d) or_rfft2_d(n, d)
#define IRFFT2(n, d) or_irfft2_d(n, d)
#else
typedef float R;
#define P(n) or32_##n
#define RSQRT(x) sqrtf(x)
#define RCOS(x) cosf(x)
#define RSIN(x) sinf(x)
#define REXP(x) expf(x)
#define RATAN2(y, x) atan2f(y, x)
#define RABS(x) fabsf(x)
/* flang evaluates CABS(Z)**2 as cabsf(z)**2 */
static inline float cabs2f(float a, float b) {
  float h = hypotf(a, b);
  return h * h;
}
#define CABS2(a, b) cabs2f(a, b)
#define RFFT2(n, d) or_rfft2_f(n, d)
#define IRFFT2(n, d) or_irfft2_f(n, d)
#endif

#if OR_FP64
#define RCABS(a, b) hypot(a, b)
#else
#define RCABS(a, b) hypotf(a, b)
#endif

#define PI_F 3.1415926535897f
#define TWOPI_F 6.2831853071796f

typedef struct { R re, im; } C;
static inline C cmul(C a, C b) {
  C r;
  r.re = a.re * b.re - a.im * b.im;
  r.im = a.re * b.im + a.im * b.re;
  return r;
}
static inline C cconj(C a) { a.im = -a.im; return a; }
static inline C cadd(C a, C b) { a.re = a.re + b.re; a.im = a.im + b.im; return a; }
static inline C cscale(C a, R s) { a.re = a.re * s; a.im = a.im * s; return a; }
static inline C cld(const R *p, size_t i) { C c = {p[2 * i], p[2 * i + 1]}; return c; }
static inline C cldf(const float *p, size_t i) { C c = {(R)p[2 * i], (R)p[2 * i + 1]}; return c; }
static inline void cst(R *p, size_t i, C c) { p[2 * i] = c.re; p[2 * i + 1] = c.im; }

/* Fortran MOD(J+NSAM) style wrap used everywhere: JJ=J+1; IF (JJ.LT.1) JJ=JJ+NSAM; 0-based */
static inline int wrapn(int j, int n) { return j < 0 ? j + n : j; }

/* ---------------------------------------------------------------- ctf.f */
static C ctf_c(float cs, float wl, float wgh1, float wgh2in, float df1, float df2, float angast,
               float thetatr, int ix, int iy, float tx, float ty) {
  C out;
  R wgh2 = RABS((R)wgh2in);
  if (wgh2in != -1.0f) {
    float radf = (float)(ix * ix + iy * iy);
    if (radf != 0.0f) {
#if OR_FP64
      double rad = sqrt((double)radf);
      double angle = rad * thetatr;
      double angspt = atan2((double)iy, (double)ix);
      double c1 = 6.2831853071796 * angle * angle / (2.0 * wl);
      double c2 = c1 * cs * angle * angle / 2.0;
      double angdif = angspt - angast;
      double ccos = cos(2.0 * angdif);
      double df = -0.5 * (df1 + df2 + ccos * (df1 - df2));
      double chi = c1 * df + c2;
      chi = chi + 6.2831853071796 * cs * angle * angle * ((double)ix * tx + (double)iy * ty) *
                      thetatr / wl / 1000.0;
      double schi = sin(chi), cchi = cos(chi);
      out.re = wgh1 * schi - wgh2 * cchi;
      out.im = -wgh1 * cchi - wgh2 * schi;
#else
      float rad = sqrtf(radf);
      float angle = rad * thetatr;
      float angspt = atan2f((float)iy, (float)ix);
      float c1 = TWOPI_F * angle * angle / (2.0f * wl);
      float c2 = c1 * cs * angle * angle / 2.0f;
      float angdif = angspt - angast;
      float ccos = cosf(2.0f * angdif);
      float df = -0.5f * (df1 + df2 + ccos * (df1 - df2));
      float chi = c1 * df + c2;
      chi = chi + TWOPI_F * cs * angle * angle * ((float)ix * tx + (float)iy * ty) * thetatr / wl /
                      1000.0f;
      float schi = sinf(chi), cchi = cosf(chi);
      out.re = wgh1 * schi - wgh2 * cchi;
      out.im = -wgh1 * cchi - wgh2 * schi;
#endif
    } else {
      out.re = -wgh2;
      out.im = 0;
    }
    if (wgh2in < 0.0f) { out.re = -out.re; out.im = -out.im; }
  } else {
    out.re = 1;
    out.im = 0;
  }
  return out;
}

void P(ctf)(float cs, float wl, float wgh1, float wgh2in, float df1, float df2, float angast,
            float thetatr, int ix, int iy, float tx, float ty, R *out) {
  C c = ctf_c(cs, wl, wgh1, wgh2in, df1, df2, angast, thetatr, ix, iy, tx, ty);
  out[0] = c.re;
  out[1] = c.im;
}

/* ------------------------------------------------------ maskcos(2).f */
static void maskcos_impl(int nsam, R *outd, float ri2, float ri3, float rih, float halfw,
                         float amagp, int origin) {
  float rihp = rih / amagp;
  float ri2p = ri2 / amagp / amagp;
  float ri3p = ri3 / amagp / amagp;
  int nmid = origin ? 1 : nsam / 2 + 1;
  int nsamh = nsam / 2;
  R sum = 0;
  int isum = 0;
  for (int pass = 0; pass < 2; pass++) {
    for (int i = 1; i <= nsam; i++) {
      int ii = i;
      if (origin && ii > nsamh) ii -= nsam;
      float d1 = (float)(ii - nmid);
      float fi1 = d1 * d1;
      for (int j = 1; j <= nsam; j++) {
        int jj = j;
        if (origin && jj > nsamh) jj -= nsam;
        size_t id = (size_t)(j - 1) + (size_t)(nsam + 2) * (i - 1);
        float d2 = (float)(jj - nmid);
        float crad2 = d2 * d2 + fi1;
        if (pass == 0) {
          if (crad2 > ri2p && crad2 <= ri3p) {
            sum = sum + outd[id];
            isum++;
          }
        } else {
          if (crad2 > ri2p && crad2 <= ri3p) {
            R srad2 = RSQRT((R)crad2);
            R edge = ((R)1.0 + RCOS((R)PI_F * (srad2 - (R)rihp) / (R)halfw)) / (R)2.0;
            outd[id] = outd[id] * edge + ((R)1.0 - edge) * sum;
          } else if (crad2 > ri3p) {
            outd[id] = sum;
          }
        }
      }
    }
    if (pass == 0 && isum != 0) sum = sum / (R)isum;
  }
}

void P(maskcos)(int nsam, R *outd, float ri2, float ri3, float rih, float halfw, float amagp) {
  maskcos_impl(nsam, outd, ri2, ri3, rih, halfw, amagp, 0);
}
void P(maskcos2)(int nsam, R *outd, float ri2, float ri3, float rih, float halfw, float amagp) {
  maskcos_impl(nsam, outd, ri2, ri3, rih, halfw, amagp, 1);
}

/* ------------------------------------------------------------ pweight.f
 * CTFAPPLY calls PWEIGHT(NSAM,OUTC,OUTC,...): input and output alias, and the
 * J=+-NSAM/2 passes both address the Nyquist row, so that row is weighted
 * twice.  The shell whose index rounds to JC+1 uses RBUF(JC+1), i.e. bin 1's
 * count slot, and SSNR(JC+1) (ssnr must hold JC+1 values).  All reproduced. */
void P(pweight)(int nsam, const R *spec, R *outc, const float *ssnr, const R *ctff, int ictf) {
  int jc = nsam / 2 + 1;
  size_t ic = (size_t)nsam * jc;
  if (spec != outc) memcpy(outc, spec, sizeof(R) * 2 * ic);
  int maxr12 = 0, maxr22 = jc * jc;
  R *rbuf = (R *)calloc((size_t)(2 * jc + 2), sizeof(R)); /* 1-based RBUF(1..2JC+1) */
  for (int i = 0; i <= jc - 1; i++)
    for (int j = -jc + 1; j <= jc - 1; j++) {
      int itemp = i * i + j * j;
      if (itemp >= maxr12 && itemp < maxr22) {
        int ibin = (int)(sqrtf((float)itemp) + 0.5f) + 1;
        size_t id = (size_t)i + (size_t)jc * wrapn(j, nsam);
        C z = cld(outc, id);
        rbuf[ibin] = rbuf[ibin] + CABS2(z.re, z.im);
        rbuf[ibin + jc] = rbuf[ibin + jc] + (R)1.0;
      }
    }
  for (int i = 1; i <= jc; i++) {
    if (rbuf[i + jc] != 0) rbuf[i] = RSQRT(rbuf[i] / rbuf[i + jc]);
    if (rbuf[i] == 0) rbuf[i] = 1;
  }
  for (int i = 0; i <= jc - 1; i++)
    for (int j = -jc + 1; j <= jc - 1; j++) {
      int itemp = i * i + j * j;
      size_t id = (size_t)i + (size_t)jc * wrapn(j, nsam);
      if (itemp >= maxr12 && itemp < maxr22) {
        int ibin = (int)(sqrtf((float)itemp) + 0.5f) + 1;
        R snr;
        if (ictf == 1) {
          C cr = cld(ctff, id), cl = cld(ctff, id + ic);
          C s = cadd(cr, cconj(cl));
          snr = (R)ssnr[ibin - 1] * CABS2(s.re, s.im);
        } else {
          snr = (R)ssnr[ibin - 1];
        }
        R wgt = snr / (snr + (R)1.0) / rbuf[ibin];
        C z = cscale(cld(outc, id), wgt);
        cst(outc, id, z);
      } else if (itemp >= maxr22) {
        C zero = {0, 0};
        cst(outc, id, zero);
      }
    }
  free(rbuf);
}

/* ----------------------------------------------------------- ctfapply.f */
void P(ctfapply)(int nsam, const float *spec, float shx, float shy, float cs, float wl,
                 float wgh1, float wgh2, float df1, float df2, float angast, float thetatr,
                 float amagp, float rih, float halfw, float ri2, float ri3, const float *pssnr,
                 int ictf, R *ctff, R *outc, R *datc) {
  float thetatrp = thetatr / amagp;
  R scal = (R)1.0 / (R)nsam / (R)nsam;
  int jc = nsam / 2 + 1;
  size_t ic = (size_t)nsam * jc;
  /* ctfapply.f:64-80: phase shift and CTF table */
  for (int l = 1; l <= jc; l++) {
    int ll = l - 1;
    float px = shx * (float)ll;
    for (int m = 1; m <= nsam; m++) {
      int mm = m - 1;
      if (mm >= jc) mm -= nsam;
      size_t id = (size_t)(l - 1) + (size_t)jc * (m - 1);
#if OR_FP64
      double phase = (double)shx * ll + (double)shy * mm;
#else
      float phase = px + shy * (float)mm;
#endif
      C pshft = {RCOS(phase), RSIN(phase)};
      C cr = ctf_c(cs, wl, wgh1, wgh2, df1, df2, angast, thetatrp, ll, mm, 0, 0);
      C cl = ctf_c(cs, wl, wgh1, wgh2, df1, df2, angast, thetatrp, -ll, -mm, 0, 0);
      C s = cldf(spec, id);
      C v = cmul(s, pshft);
      cst(outc, id, v);
      cst(datc, id, v);
      cst(ctff, id, cr);
      cst(ctff, id + ic, cl);
      (void)px;
    }
  }
  /* ctfapply.f:107-113: c2r, scale */
  IRFFT2(nsam, outc);
  for (int i = 0; i < nsam; i++)
    for (int ii = 0; ii < nsam; ii++) {
      size_t jj = (size_t)ii + (size_t)(nsam + 2) * i;
      outc[jj] = outc[jj] * scal;
    }
  /* ctfapply.f:116-122 */
  P(maskcos)(nsam, outc, ri2, ri3, rih, halfw, amagp);
  /* ctfapply.f:140 */
  RFFT2(nsam, outc);
  /* ctfapply.f:144-155: (-1)^(l+m) */
  for (int l = 1; l <= jc; l++) {
    int ll = l - 1;
    for (int m = 1; m <= nsam; m++) {
      int mm = m - 1;
      if (mm >= jc) mm -= nsam;
      R sgn = ((ll + mm) % 2 != 0) ? (R)-1.0 : (R)1.0;
      size_t id = (size_t)(l - 1) + (size_t)jc * (m - 1);
      cst(datc, id, cscale(cld(datc, id), sgn));
      cst(outc, id, cscale(cld(outc, id), sgn));
    }
  }
  /* ctfapply.f:157-158 */
  P(pweight)(nsam, outc, outc, pssnr, ctff, ictf);
}

/* -------------------------------------------------------- ainterpo3ds.f */
static C ainterpo3ds_c(int nsam, const float *a3df, float xx, float yy, float zz) {
  int jc = nsam / 2 + 1;
  float x, y, z;
  int sign;
  if (xx < 0.0f) {
    x = -xx; y = -yy; z = -zz;
    sign = -1;
  } else {
    x = xx; y = yy; z = zz;
    sign = 1;
  }
  int ls = (int)x, lt = ls + 1;
  if (lt >= jc) lt = jc - 1;
  int ms = (int)y, mt = ms + 1;
  if (y < 0.0f) { ms--; mt--; }
  if (ms <= -jc) ms = -jc + 1;
  if (mt >= jc) mt = jc - 1;
  int ns = (int)z, nt = ns + 1;
  if (z < 0.0f) { ns--; nt--; }
  if (ns <= -jc) ns = -jc + 1;
  if (nt >= jc) nt = jc - 1;
  /* ainterpo3ds.f:57-64: corner k=1..8 -> x:(k odd LS, even LT), y:(1,2,5,6 MS; 3,4,7,8 MT),
   * z:(1..4 NS; 5..8 NT) */
  float ax[2] = {1.0f - fabsf(x - (float)ls), 1.0f - fabsf(x - (float)lt)};
  float ay[2] = {1.0f - fabsf(y - (float)ms), 1.0f - fabsf(y - (float)mt)};
  float az[2] = {1.0f - fabsf(z - (float)ns), 1.0f - fabsf(z - (float)nt)};
  int lx[2] = {ls, lt};
  int my[2] = {wrapn(ms, nsam), wrapn(mt, nsam)};
  int nz[2] = {wrapn(ns, nsam), wrapn(nt, nsam)};
  C samp = {0, 0};
  for (int k = 0; k < 8; k++) {
    int bx = k & 1, by = (k >> 1) & 1, bz = (k >> 2) & 1;
#if OR_FP64
    R w = (R)ax[bx] * (R)ay[by] * (R)az[bz];
#else
    float w = ax[bx] * ay[by] * az[bz];
#endif
    size_t id = (size_t)lx[bx] + (size_t)jc * ((size_t)my[by] + (size_t)nsam * nz[bz]);
    C v = cldf(a3df, id);
    samp = cadd(samp, cscale(v, w));
  }
  if (sign < 0) samp = cconj(samp);
  return samp;
}

void P(ainterpo3ds)(int nsam, const float *a3df, float xx, float yy, float zz, R *out) {
  C c = ainterpo3ds_c(nsam, a3df, xx, yy, zz);
  out[0] = c.re;
  out[1] = c.im;
}

/* ----------------------------------------------------------- pextract.f */
void P(pextract)(int nsam, R *outc, const float *a3df, float phi, float theta, float psi,
                 const R *ctff, const float *sm) {
  int jc = nsam / 2 + 1;
  size_t ic = (size_t)nsam * jc;
  int nsamh = nsam / 2, ir2lim = nsamh * nsamh;
  float dm[9];
  or_dm(phi, theta, psi, 0.0f, sm, dm);
  for (int i = 0; i <= jc - 1; i++)
    for (int j = -jc + 1; j <= jc - 1; j++) {
      int ir2 = i * i + j * j;
      size_t id = (size_t)i + (size_t)jc * wrapn(j, nsam);
      if (ir2 <= ir2lim) {
        C cr = cld(ctff, id), cl = cld(ctff, id + ic);
        float x3 = dm[0] * (float)i + dm[3] * (float)j;
        float y3 = dm[1] * (float)i + dm[4] * (float)j;
        float z3 = dm[2] * (float)i + dm[5] * (float)j;
        C a = ainterpo3ds_c(nsam, a3df, x3, y3, z3);
        a = cmul(a, cadd(cr, cconj(cl)));
        cst(outc, id, a);
      } else {
        C zero = {0, 0};
        cst(outc, id, zero);
      }
    }
}

/* ------------------------------------------------------ cc3m.f and cc3.f */
/* flip = 0: CC3M (cc3m.f:8-169).  flip = 1: CC3 (cc3.f:8-159), whose section
 * is negated where REAL(CTFR+CONJG(CTFL)) < 0 instead of multiplied by the
 * CTF (cc3.f:86-89) and whose R_k keeps its sign in every shell. */
static double P(cc3_core)(int nsam, float amag, const R *spec, const float *a3df, int maxr1,
                          int maxr2, float phi, float theta, float psi, float shx, float shy,
                          float rbfact, const float *fsct, const R *ctff, float ri2, float ri3,
                          float rih, float halfw, const float *sm, int flip) {
  int jc = nsam / 2 + 1;
  size_t ic = (size_t)nsam * jc;
  int maxr12 = maxr1 * maxr1, maxr22 = maxr2 * maxr2;
  float dm[9];
  or_dm(phi, theta, psi, amag, sm, dm);
  /* cc3m.f:67-72: RBUF(1..IB1) shell bins, RBUF(IB1+1..) FFT buffer */
  int ib1 = 4 * nsam / 2;
  size_t nrbuf = (size_t)4 * nsam / 2 + (size_t)nsam * nsam + 4 * nsam;
  R *rbuf = (R *)calloc(nrbuf + 8, sizeof(R));
  R *buf = rbuf + ib1; /* RBUF(IB1+1) 0-based */
  R scal = (R)1.0 / (R)nsam / (R)nsam;
  for (int i = 0; i <= jc - 1; i++)
    for (int j = -jc + 1; j <= jc - 1; j++) {
      int itemp = i * i + j * j;
      if (itemp >= maxr12 && itemp < maxr22) {
        size_t id = (size_t)i + (size_t)jc * wrapn(j, nsam);
        C cr = cld(ctff, id), cl = cld(ctff, id + ic);
        float x3 = dm[0] * (float)i + dm[3] * (float)j;
        float y3 = dm[1] * (float)i + dm[4] * (float)j;
        float z3 = dm[2] * (float)i + dm[5] * (float)j;
        C a = ainterpo3ds_c(nsam, a3df, x3, y3, z3);
        if (flip) {
          if (cadd(cr, cconj(cl)).re < 0) {
            a.re = -a.re;
            a.im = -a.im;
          }
        } else {
          a = cmul(a, cadd(cr, cconj(cl)));
        }
        buf[2 * id] = a.re * scal;
        buf[2 * id + 1] = a.im * scal;
      }
    }
  IRFFT2(nsam, buf);
  P(maskcos2)(nsam, buf, ri2, ri3, rih, halfw, amag);
  RFFT2(nsam, buf);
  for (int i = 0; i <= jc - 1; i++)
    for (int j = -jc + 1; j <= jc - 1; j++) {
      int itemp = i * i + j * j;
      if (itemp >= maxr12 && itemp < maxr22) {
#if OR_FP64
        double phase = (double)shx * i + (double)shy * j;
#else
        float phase = shx * (float)i + shy * (float)j;
#endif
        C pshft = {RCOS(phase), RSIN(phase)};
        size_t id = (size_t)i + (size_t)jc * wrapn(j, nsam);
        C samp = cmul(cld(spec, id), pshft);
        C a = {buf[2 * id], buf[2 * id + 1]};
        int ibin = 4 * ((int)(sqrtf((float)itemp) + 0.5f) + 1);
        C q = cmul(samp, cconj(a));
        rbuf[ibin - 4] = rbuf[ibin - 4] + q.re;
        rbuf[ibin - 3] = rbuf[ibin - 3] + CABS2(samp.re, samp.im);
        rbuf[ibin - 2] = rbuf[ibin - 2] + CABS2(a.re, a.im);
        rbuf[ibin - 1] = rbuf[ibin - 1] + (R)1.0;
      }
    }
  double sum1 = 0, sum2 = 0, sum3 = 0;
  for (int i = 2; i <= nsam / 2; i++) {
    int ibin = 4 * i;
    if (rbuf[ibin - 1] != 0) {
      R r = rbuf[ibin - 4] / RSQRT(rbuf[ibin - 2]);
      if (!flip && fsct[i - 1] < 0.95f) r = RABS(r);
      float e = -rbfact * (float)((i - 1) * (i - 1));
      R wgt = REXP((R)e);
      sum3 = sum3 + (double)(r * wgt);
      sum2 = sum2 + (double)(wgt * wgt * rbuf[ibin - 3]);
      sum1 = sum1 + 1.0;
    }
  }
  if (sum1 * sum2 != 0.0) sum3 = sum3 / sqrt(sum1 * sum2);
  free(rbuf);
#if OR_FP64
  return sum3;
#else
  return (double)(float)sum3;
#endif
}

double P(cc3m)(int nsam, float amag, const R *spec, const float *a3df, int maxr1, int maxr2,
               float phi, float theta, float psi, float shx, float shy, float rbfact,
               const float *fsct, const R *ctff, float ri2, float ri3, float rih, float halfw,
               const float *sm) {
  return P(cc3_core)(nsam, amag, spec, a3df, maxr1, maxr2, phi, theta, psi, shx, shy, rbfact,
                     fsct, ctff, ri2, ri3, rih, halfw, sm, 0);
}

double P(cc3)(int nsam, float amag, const R *spec, const float *a3df, int maxr1, int maxr2,
              float phi, float theta, float psi, float shx, float shy, float rbfact,
              const R *ctff, float ri2, float ri3, float rih, float halfw, const float *sm) {
  return P(cc3_core)(nsam, amag, spec, a3df, maxr1, maxr2, phi, theta, psi, shx, shy, rbfact,
                     NULL, ctff, ri2, ri3, rih, halfw, sm, 1);
}

/* --------------------------------------------------------------- presb.f */
void P(presb)(int nsam, float amag, const R *spec, const float *a3df, float phi, float theta,
              float psi, float shx, float shy, R *rbin, const R *ctff, float ri2, float ri3,
              float rih, float halfw, const float *sm) {
  int jc = nsam / 2 + 1;
  size_t ic = (size_t)nsam * jc;
  int nsamh = nsam / 2;
  int maxr12 = 1, maxr22 = nsamh * nsamh;
  float dm[9];
  or_dm(phi, theta, psi, amag, sm, dm);
  int *ns = (int *)calloc((size_t)nsam + 2, sizeof(int));
  R *bufa = (R *)calloc((size_t)nsam + 2, sizeof(R));
  R *bufb = (R *)calloc((size_t)nsam + 2, sizeof(R));
  R *bufab = (R *)calloc((size_t)nsam * nsam + 3 * nsam + 8, sizeof(R));
  R *buf = bufab + nsam; /* BUFAB(IB1+1), IB1 = NSAM */
  R scal = (R)1.0 / (R)nsam / (R)nsam;
  for (int i = 0; i <= jc - 1; i++)
    for (int j = -jc + 1; j <= jc - 1; j++) {
      int itemp = i * i + j * j;
      if (itemp >= maxr12 && itemp < maxr22) {
        size_t id = (size_t)i + (size_t)jc * wrapn(j, nsam);
        C cr = cld(ctff, id), cl = cld(ctff, id + ic);
        float x3 = dm[0] * (float)i + dm[3] * (float)j;
        float y3 = dm[1] * (float)i + dm[4] * (float)j;
        float z3 = dm[2] * (float)i + dm[5] * (float)j;
        C a = ainterpo3ds_c(nsam, a3df, x3, y3, z3);
        a = cmul(a, cadd(cr, cconj(cl)));
        buf[2 * id] = a.re * scal;
        buf[2 * id + 1] = a.im * scal;
      }
    }
  IRFFT2(nsam, buf);
  P(maskcos2)(nsam, buf, ri2, ri3, rih, halfw, amag);
  RFFT2(nsam, buf);
  for (int i = 0; i <= jc - 1; i++)
    for (int j = -jc + 1; j <= jc - 1; j++) {
      int itemp = i * i + j * j;
      if (itemp >= maxr12 && itemp < maxr22) {
#if OR_FP64
        double phase = (double)shx * i + (double)shy * j;
#else
        float phase = shx * (float)i + shy * (float)j;
#endif
        C pshft = {RCOS(phase), RSIN(phase)};
        size_t id = (size_t)i + (size_t)jc * wrapn(j, nsam);
        C samp = cmul(cld(spec, id), pshft);
        C a = {buf[2 * id], buf[2 * id + 1]};
        int ibin = (int)(sqrtf((float)itemp) + 0.5f) + 1;
        C q = cmul(samp, cconj(a));
        bufab[ibin - 1] = bufab[ibin - 1] + q.re;
        bufa[ibin - 1] = bufa[ibin - 1] + CABS2(samp.re, samp.im);
        bufb[ibin - 1] = bufb[ibin - 1] + CABS2(a.re, a.im);
        ns[ibin - 1]++;
      }
    }
  for (int i = 1; i <= nsam / 2; i++)
    if (ns[i - 1] != 0) {
      rbin[2 * i - 2] = bufab[i - 1] / RSQRT(bufa[i - 1] * bufb[i - 1]);
      rbin[2 * i - 1] = 1;
    }
  free(ns);
  free(bufa);
  free(bufb);
  free(bufab);
}

/* ------------------------------------------------- pinsert_trilin.f */
/* pinsert_trilin.f:44-244 for the stripe I = i1..i3 (loops 30 and 31 cover it
 * in order); the phase uses SHX+DX, SHY+DY (pinsert_trilin.f:55,63) */
/* pdiff.f:8-24: |phase of C1 conj(C2)|, 0 where the product is 0 */
static R pdiff(C c1, C c2) {
  C c = cmul(c1, cconj(c2));
  if (c.re == (R)0 && c.im == (R)0) return (R)0;
  return RABS(RATAN2(c.im, c.re));
}

/* the per-voxel statistics of the _SE variants (NNSTAT = 1,
 * pinsert_trilin_se.f:186-194): KSUM += 1, VSUM += |BSAMP|^2,
 * PSUM += PDIFF(BSAMP, C3DF), ASUM += |BSAMP|, with IPAD = 1 (C3DF has the
 * half-volume's layout, ID4 = ID) */
typedef struct {
  const R *c3df;
  int *ksum;
  R *vsum, *psum, *asum;
} Stats;

static void pinsert_trilin(int nsam, int irrec2, const R *spec, const R *ctff, R occ, R *b3df,
                           R *t3df, double *ttd, float shx, float shy, int64_t *nksum, int jc,
                           const float dm[9], float tmp, int nsamh, size_t ic, int i1, int i3,
                           float dx, float dy, const Stats *st) {
  const float ri12 = 0.5f, ri2 = 1.0f; /* IRAD=2: (IRAD-1)/2, IRAD/2 */
  const float sx = shx + dx, sy = shy + dy;
  for (int i = i1; i <= i3; i++) {
    float dm1 = dm[0] * (float)i, dm2 = dm[1] * (float)i, dm3 = dm[2] * (float)i;
    int iq = i * i;
    float px = sx * (float)i;
    for (int j = -jc + 1; j <= jc - 1; j++) {
      float g2 = (float)(iq + j * j);
      if (!(g2 <= (float)irrec2)) continue;
      float x3 = dm1 + dm[3] * (float)j;
      float y3 = dm2 + dm[4] * (float)j;
      float z3 = dm3 + dm[5] * (float)j;
#if OR_FP64
      double wgt = exp((double)tmp * (double)g2) * occ;
      double phase = (double)sx * i + (double)sy * j;
#else
      float wgt = expf(tmp * g2) * occ;
      float phase = px + sy * (float)j;
#endif
      C pshft = {RCOS(phase), RSIN(phase)};
      C samp, cr, cl;
      if (i >= 0) {
        size_t id = (size_t)i + (size_t)jc * wrapn(j, nsam);
        cr = cld(ctff, id);
        cl = cld(ctff, id + ic);
        samp = cmul(cld(spec, id), pshft);
      } else {
        size_t id = (size_t)(-i) + (size_t)jc * wrapn(-j, nsam);
        cr = cld(ctff, id);
        cl = cld(ctff, id + ic);
        samp = cmul(cconj(cld(spec, id)), pshft);
      }
      C cs = cadd(cr, cconj(cl));
      samp = cmul(samp, cs);
      R ctfv2 = (R)0.25 * CABS2(cs.re, cs.im) * wgt;
      samp = cscale(samp, wgt);
      /* pinsert_trilin.f:88-111 bounds, with the reference's INT()/negative rule */
      float a;
      int ls, lt, ms, mt, ns, nt;
      a = x3 + 0.5f - ri12; ls = (int)a; if (a < 0.0f) ls--; if (ls < 0) ls = 0;
      a = x3 + ri2;         lt = (int)a; if (a < 0.0f) lt--; if (lt > nsamh) lt = nsamh;
      a = y3 + 0.5f - ri12; ms = (int)a; if (a < 0.0f) ms--; if (ms < -nsamh) ms = -nsamh;
      a = y3 + ri2;         mt = (int)a; if (a < 0.0f) mt--; if (mt > nsamh) mt = nsamh;
      a = z3 + 0.5f - ri12; ns = (int)a; if (a < 0.0f) ns--; if (ns < -nsamh) ns = -nsamh;
      a = z3 + ri2;         nt = (int)a; if (a < 0.0f) nt--; if (nt > nsamh) nt = nsamh;
      for (int n = ns; n <= nt; n++) {
        float arg3 = fabsf(z3 - (float)n);
        size_t n1 = (size_t)nsam * wrapn(n, nsam);
        for (int m = ms; m <= mt; m++) {
          float arg2 = fabsf(y3 - (float)m);
          size_t m1 = (size_t)jc * (wrapn(m, nsam) + n1);
          for (int l = ls; l <= lt; l++) {
            float arg1 = fabsf(x3 - (float)l);
#if OR_FP64
            double box = (1.0 - arg1) * (1.0 - arg2) * (1.0 - arg3);
#else
            float box = (1.0f - arg1) * (1.0f - arg2) * (1.0f - arg3);
#endif
            R c2 = ctfv2 * box;
            C bs = cscale(samp, box);
            size_t id = (size_t)l + m1;
            b3df[2 * id] = b3df[2 * id] + bs.re;
            b3df[2 * id + 1] = b3df[2 * id + 1] + bs.im;
            t3df[id] = t3df[id] + c2;
            (*nksum)++;
            if (st) {
              R ab = RCABS(bs.re, bs.im);
              st->ksum[id] = st->ksum[id] + 1;
              st->vsum[id] = st->vsum[id] + ab * ab;
              st->psum[id] = st->psum[id] + pdiff(bs, cld(st->c3df, id));
              st->asum[id] = st->asum[id] + ab;
            }
            *ttd = *ttd + (double)c2;
          }
        }
      }
    }
  }
}

/* ---------------------------------------------------------- pinsert.f */
static void pinsert_common(int nsam, float rrec, float amag, const R *spec, const R *ctff,
                           R *b3df, R *t3df, double *ttd, float phi, float theta, float psi,
                           float shx, float shy, float presa, float pbc, float boff, int nsym,
                           const int *jsym, const float *symop, int64_t *nksum, float psize,
                           float occ, const float *sm, int nonsym, const Stats *st) {
  if (occ == 0.0f) return;
  float pref = pbc / 100.0f;
  int jc = nsam / 2 + 1;
  size_t ic = (size_t)nsam * jc;
  int nsamh = nsam / 2;
  int irrec = (int)((float)nsam * rrec);
  int irrec2 = irrec * irrec;
  float tmp = (presa / pref - boff / pbc) / (psize * psize);
  if (tmp > 2.0f) tmp = 2.0f;
  tmp = 2.0f * tmp / ((float)nsamh * (float)nsamh);
  float tm[9];
  or_dm(phi, theta, psi, amag, NULL, tm);
  int isym[8];
  for (int i = 0; i < nsym; i++) isym[i] = 1;
  for (;;) {
    float dm[9];
    memcpy(dm, tm, sizeof dm);
    if (nonsym) {
      or_matmul_t(dm, sm, dm);
    } else {
      for (int i = 0; i < nsym; i++)
        if (isym[i] != jsym[i])
          for (int j = 0; j < isym[i]; j++) or_matmul_t(dm, symop + 9 * i, dm);
    }
    /* IP=1, IMP=1 (pinsert.f): the whole plane, I = -JC+1 .. JC-1, DX = DY = 0 */
    pinsert_trilin(nsam, irrec2, spec, ctff, (R)occ, b3df, t3df, ttd, shx, shy, nksum, jc, dm,
                   tmp, nsamh, ic, -jc + 1, jc - 1, 0.0f, 0.0f, st);
    if (nonsym) break;
    int i = 0, done = 0;
    for (;;) {
      isym[i]++;
      if (isym[i] > jsym[i]) {
        isym[i] = 1;
        i++;
        if (i < nsym) continue;
        done = 1;
      }
      break;
    }
    if (done) break;
  }
}

void P(pinsert)(int nsam, float rrec, float amag, const R *spec, const R *ctff, R *b3df, R *t3df,
                double *ttd, float phi, float theta, float psi, float shx, float shy, float presa,
                float pbc, float boff, int nsym, const int *jsym, const float *symop,
                int64_t *nksum, float psize, float occ, const float *sm, int nonsym) {
  pinsert_common(nsam, rrec, amag, spec, ctff, b3df, t3df, ttd, phi, theta, psi, shx, shy, presa,
                 pbc, boff, nsym, jsym, symop, nksum, psize, occ, sm, nonsym, NULL);
}

void P(pinsert_se)(int nsam, float rrec, float amag, const R *spec, const R *ctff, R *b3df,
                   R *t3df, double *ttd, float phi, float theta, float psi, float shx, float shy,
                   float presa, float pbc, float boff, int nsym, const int *jsym,
                   const float *symop, int64_t *nksum, float psize, float occ, const float *sm,
                   int nonsym, const R *c3df, int *ksum, R *vsum, R *psum, R *asum) {
  Stats st = {c3df, ksum, vsum, psum, asum};
  pinsert_common(nsam, rrec, amag, spec, ctff, b3df, t3df, ttd, phi, theta, psi, shx, shy, presa,
                 pbc, boff, nsym, jsym, symop, nksum, psize, occ, sm, nonsym, &st);
}

/* pinsert_trilin.f:8-244 as one call: stripe IP of width ST,
 * I1 = -JC+1+NINT((IP-1)*ST), I3 = -JC+NINT(IP*ST) (pinsert_trilin.f:44-45) */
void P(pinsert_trilin)(int nsam, int irrec2, const R *spec, const R *ctff, float occ, R *b3df,
                       R *t3df, double *ttd, float shx, float shy, int64_t *nksum,
                       const float *dm, float tmp, float dx, float dy, float st, int ip) {
  int jc = nsam / 2 + 1;
  int i1 = -jc + 1 + (int)lroundf((float)(ip - 1) * st);
  int i3 = -jc + (int)lroundf((float)ip * st);
  pinsert_trilin(nsam, irrec2, spec, ctff, (R)occ, b3df, t3df, ttd, shx, shy, nksum, jc, dm, tmp,
                 nsam / 2, (size_t)nsam * jc, i1, i3, dx, dy, NULL);
}

void P(pinsert_trilin_se)(int nsam, int irrec2, const R *spec, const R *ctff, float occ,
                          R *b3df, R *t3df, double *ttd, float shx, float shy, int64_t *nksum,
                          const float *dm, float tmp, float dx, float dy, float st, int ip,
                          const R *c3df, int *ksum, R *vsum, R *psum, R *asum) {
  int jc = nsam / 2 + 1;
  int i1 = -jc + 1 + (int)lroundf((float)(ip - 1) * st);
  int i3 = -jc + (int)lroundf((float)ip * st);
  Stats s = {c3df, ksum, vsum, psum, asum};
  pinsert_trilin(nsam, irrec2, spec, ctff, (R)occ, b3df, t3df, ttd, shx, shy, nksum, jc, dm, tmp,
                 nsam / 2, (size_t)nsam * jc, i1, i3, dx, dy, &s);
}

/* ------------------------------------------------------- storeshift.f */
void P(storeshift)(int nsam, double std, double vtd, int iredun, const R *a3df, const R *s3df,
                   const R *d3df, const R *v3df, R *b3df, R *c3df) {
  int jc = nsam / 2 + 1;
  double nsam3 = (double)nsam * nsam * nsam;
  for (int n = 1; n <= nsam; n++) {
    int nn = n - 1;
    if (nn >= jc) nn -= nsam;
    for (int m = 1; m <= nsam; m++) {
      int mm = m - 1;
      if (mm >= jc) mm -= nsam;
      for (int l = 1; l <= jc; l++) {
        int ll = l - 1;
        double sg = ((ll + mm + nn) % 2 != 0) ? -1.0 : 1.0;
        size_t id = (size_t)(l - 1) + (size_t)jc * ((size_t)(m - 1) + (size_t)nsam * (n - 1));
        double db = (double)s3df[id] + std * iredun;
        double dv = (double)v3df[id] + vtd * iredun;
        b3df[2 * id] = (R)((double)(R)(a3df[2 * id] * (R)sg) / db / nsam3);
        b3df[2 * id + 1] = (R)((double)(R)(a3df[2 * id + 1] * (R)sg) / db / nsam3);
        c3df[2 * id] = (R)((double)(R)(d3df[2 * id] * (R)sg) / dv / nsam3);
        c3df[2 * id + 1] = (R)((double)(R)(d3df[2 * id + 1] * (R)sg) / dv / nsam3);
      }
    }
  }
}

/* --------------------------------------------------- combine_arrays.f */
/* combine_arrays.f:33-52 (NNSTAT=0): copies 2..NVOL1 of A3DF/S3DF (stride
 * nn3 elements) added into copy 1 in order; for IFSC=0 likewise D3DF/V3DF
 * over NVOL2.  The counters (combine_arrays.f:53-67) are host arithmetic. */
void P(combine_arrays)(int nsam, long long nn3, int nvol1, int nvol2, int ifsc, R *a3df, R *s3df,
                       R *d3df, R *v3df) {
  size_t nv = (size_t)nsam * nsam * (nsam / 2 + 1);
  for (size_t i = 0; i < nv; i++) {
    for (int is = 2; is <= nvol1; is++) {
      size_t o = (size_t)(is - 1) * (size_t)nn3;
      a3df[2 * i] = a3df[2 * i] + a3df[2 * (o + i)];
      a3df[2 * i + 1] = a3df[2 * i + 1] + a3df[2 * (o + i) + 1];
      s3df[i] = s3df[i] + s3df[o + i];
    }
    if (ifsc == 0)
      for (int is = 2; is <= nvol2; is++) {
        size_t o = (size_t)(is - 1) * (size_t)nn3;
        d3df[2 * i] = d3df[2 * i] + d3df[2 * (o + i)];
        d3df[2 * i + 1] = d3df[2 * i + 1] + d3df[2 * (o + i) + 1];
        v3df[i] = v3df[i] + v3df[o + i];
      }
  }
}

/* --------------------------------------------------------- shiftvol.f */
/* shiftvol.f:8-77 (NNSTAT=0).  *std is STD in/out (STD+VTD on return).  The
 * divisions by S3DF+STD*IREDUN (a DOUBLE) run in double per component, as
 * COMPLEX / DOUBLE does (storeshift above follows the same rule). */
void P(shiftvol)(int nsam, double *std, double vtd, float rlim, int iredun, R *a3df, R *d3df,
                 R *s3df, const R *v3df, int ifsc, const float *pssnr, float fpart,
                 const float *fsc, float fmask) {
  int jc = nsam / 2 + 1;
  *std = *std + vtd;
  int ir = (int)((float)nsam * rlim);
  int irrec2 = ir * ir;
  for (int l = 1; l <= jc; l++) {
    int ll = l - 1;
    for (int m = 1; m <= nsam; m++) {
      int mm = m - 1;
      if (mm >= jc) mm -= nsam;
      for (int n = 1; n <= nsam; n++) {
        int nn = n - 1;
        if (nn >= jc) nn -= nsam;
        int jrad2 = ll * ll + mm * mm + nn * nn;
        R sg = ((ll + mm + nn) % 2 != 0) ? (R)-1.0 : (R)1.0;
        int i = (int)(sqrtf((float)jrad2) + 0.5f) + 1;
        size_t id = (size_t)(l - 1) + (size_t)jc * ((size_t)(m - 1) + (size_t)nsam * (n - 1));
        if (ifsc == 0) {
          a3df[2 * id] = a3df[2 * id] + d3df[2 * id];
          a3df[2 * id + 1] = a3df[2 * id + 1] + d3df[2 * id + 1];
          s3df[id] = s3df[id] + v3df[id];
        }
        R re = 0, im = 0;
        if (jrad2 <= irrec2) {
          double den = (double)s3df[id] + *std * iredun;
          if (fpart == 0.0f) {
            re = (R)((double)(R)(a3df[2 * id] * sg) / den);
            im = (R)((double)(R)(a3df[2 * id + 1] * sg) / den);
          } else if (pssnr[i - 1] != 0.0f && s3df[id] != (R)0) {
            float af = fabsf(fsc[i - 1]);
            float a1 = af < 1.0f ? 2.0f * af / (1.0f - af) : 1000.0f;
            float w = 1.0f + fpart / a1 / fmask;
            re = (R)((double)(R)(a3df[2 * id] * sg) / den / (double)w);
            im = (R)((double)(R)(a3df[2 * id + 1] * sg) / den / (double)w);
          }
        }
        a3df[2 * id] = re;
        a3df[2 * id + 1] = im;
      }
    }
  }
}

/* ------------------------------------------- correct3d.f, boxft_lut.f */
/* correct3d.f:19-47: A3DV (NSAM+2 row pitch) divided by BOXFT_LUT (INTERP=0)
 * or its square (INTERP=1); BOXFT_LUT (boxft_lut.f:26-37) is the product of
 * SINCLUT(NINT(180|a|)) over the three axes, 1 where |a| < 0.02. */
void P(correct3d)(int nsam, R *a3dv, int interp, int ipad, const float *sinclut) {
  int nmid = nsam / 2 + 1;
  float scal = 1.0f / (float)nsam;
  if (ipad != 0) scal = scal / (float)ipad;
  for (int i = 1; i <= nsam; i++)
    for (int j = 1; j <= nsam; j++)
      for (int k = 1; k <= nsam; k++) {
        float a[3] = {(float)(i - nmid) * scal, (float)(j - nmid) * scal,
                      (float)(k - nmid) * scal};
        float prod = 1.0f;
        for (int q = 0; q < 3; q++) {
          float x = fabsf(a[q]);
          float sn = x < 2.0e-2f ? 1.0f : sinclut[lroundf(180.0f * x) - 1];
          prod = prod * sn;
        }
        size_t id = (size_t)(i - 1) + (size_t)(nsam + 2) * ((size_t)(j - 1) + (size_t)nsam * (k - 1));
        if (interp == 0) a3dv[id] = a3dv[id] / (R)prod;
        else a3dv[id] = a3dv[id] / (R)(prod * prod);
      }
}

/* ----------------------------------------------------------- mask3d.f */
/* mask3d.f:8-142: cosine-edged spherical mask of the real volume (NSAM+2 row
 * pitch) about (1+NSAM/2)^3, the outside set to the mean BG of the edge
 * shell; with RIC > 0 an inner mask of radius RIC likewise.  BG and the
 * SUM1/SUM2 statistics of RIC^2 <= r^2 <= RIO^2 are DOUBLE PRECISION sums
 * in the reference's loop order (L outer, N inner). */
static void mask3d_bg(int nsam, const R *a, float xc, float ri2, float ri3, double *bg) {
  int nbg = 0;
  *bg = 0.0;
  for (int l = 1; l <= nsam; l++) {
    float xl = (float)l - xc;
    for (int m = 1; m <= nsam; m++) {
      float ym = (float)m - xc;
      for (int n = 1; n <= nsam; n++) {
        float zn = (float)n - xc;
        size_t id = (size_t)(l - 1) + (size_t)(nsam + 2) * ((size_t)(m - 1) + (size_t)nsam * (n - 1));
        float rad2 = xl * xl + ym * ym + zn * zn;
        if (rad2 > ri2 && rad2 < ri3) {
          *bg = *bg + (double)a[id];
          nbg++;
        }
      }
    }
  }
  if (nbg != 0) *bg = *bg / nbg;
}

void P(mask3d)(int nsam, R *a, float rio, float ric, float halfw, float *fmask, float *ave,
               float *std) {
  const float pi = 3.1415926535897f;
  float ri = rio, ri4 = rio * rio, ri5 = ric * ric;
  float rm = fminf(ri, (float)(nsam / 2) - 1.0f - halfw / 2.0f);
  float rii = rm + halfw / 2.0f, rih = rm - halfw / 2.0f;
  if (rih < 0.0f) rih = 0.0f;
  float ri2 = rih * rih, ri3 = rii * rii;
  float xc = 1.0f + 0.5f * (float)nsam;
  double bg = 0.0, sum1 = 0.0, sum2 = 0.0, sum3 = 0.0;
  int nbg = 0, na = 0;
  for (int l = 1; l <= nsam; l++) {
    float xl = (float)l - xc;
    for (int m = 1; m <= nsam; m++) {
      float ym = (float)m - xc;
      for (int n = 1; n <= nsam; n++) {
        float zn = (float)n - xc;
        size_t id = (size_t)(l - 1) + (size_t)(nsam + 2) * ((size_t)(m - 1) + (size_t)nsam * (n - 1));
        float rad2 = xl * xl + ym * ym + zn * zn;
        if (rad2 > ri2 && rad2 < ri3) {
          bg = bg + (double)a[id];
          nbg++;
        }
        if (rad2 <= ri4 && rad2 >= ri5) {
          sum1 = sum1 + (double)a[id];
          sum2 = sum2 + (double)(a[id] * a[id]);
          na++;
        }
      }
    }
  }
  if (nbg != 0) bg = bg / nbg;
  if (na != 0) {
    sum1 = sum1 / na;
    sum2 = sum2 / na;
  }
  *ave = (float)sum1;
  *std = (float)sqrt(fabs(sum2 - sum1 * sum1));
  for (int l = 1; l <= nsam; l++) {
    float xl = (float)l - xc;
    for (int m = 1; m <= nsam; m++) {
      float ym = (float)m - xc;
      for (int n = 1; n <= nsam; n++) {
        float zn = (float)n - xc;
        size_t id = (size_t)(l - 1) + (size_t)(nsam + 2) * ((size_t)(m - 1) + (size_t)nsam * (n - 1));
        float rad2 = xl * xl + ym * ym + zn * zn;
        float edge = 1.0f;
        if (rad2 > ri2) {
          float rad = sqrtf(rad2);
          if (rad > rih && rad < rii) edge = (1.0f + cosf(pi * (rad - rih) / halfw)) / 2.0f;
          else edge = 0.0f;
          a[id] = (R)((double)(a[id] * (R)edge) + bg * (double)(1.0f - edge));
        }
        sum3 = sum3 + (double)(edge * edge);
      }
    }
  }
  if (ric > 0.0f) {
    ri = ric;
    rm = fmaxf(ri, halfw / 2.0f);
    rii = rm + halfw / 2.0f;
    rih = rm - halfw / 2.0f;
    ri2 = rih * rih;
    ri3 = rii * rii;
    mask3d_bg(nsam, a, xc, ri2, ri3, &bg);
    for (int l = 1; l <= nsam; l++) {
      float xl = (float)l - xc;
      for (int m = 1; m <= nsam; m++) {
        float ym = (float)m - xc;
        for (int n = 1; n <= nsam; n++) {
          float zn = (float)n - xc;
          size_t id = (size_t)(l - 1) + (size_t)(nsam + 2) * ((size_t)(m - 1) + (size_t)nsam * (n - 1));
          float rad2 = xl * xl + ym * ym + zn * zn;
          float edge = 1.0f;
          if (rad2 < ri3) {
            float rad = sqrtf(rad2);
            if (rad > rih && rad < rii) edge = (1.0f + cosf(pi * (rii - rad) / halfw)) / 2.0f;
            else edge = 0.0f;
            a[id] = (R)((double)(a[id] * (R)edge) + bg * (double)(1.0f - edge));
          }
          if (edge != 1.0f) sum3 = sum3 - 1.0 + (double)(edge * edge);
        }
      }
    }
  }
  *fmask = (float)(sum3 / nsam / nsam / nsam);
}

/* -------------------------------------------------------- sheltest.f */
int P(sheltest)(int nsam, int istp, const R *a3df, const R *d3df, const R *s3df, const R *v3df,
                float fmask, int *npix, int *npixt, double *dfsc, R *pssnr) {
  int jc = nsam / 2 + 1;
  int nstp = jc / istp;
  if (nstp * istp < jc) nstp++;
  for (int is_ = 1; is_ <= nstp; is_++) {
    npix[is_ - 1] = 0;
    npixt[is_ - 1] = 0;
    dfsc[is_ - 1] = 0.0;
    double mean12 = 0, mean22 = 0;
    R ps = 0;
    int lo = (is_ - 1) * istp, hi = is_ * istp;
    if (is_ == nstp) hi = jc;
    int lo2 = lo * lo, hi2 = hi * hi;
    for (int l = 1; l <= jc; l++) {
      int ll = l - 1;
      for (int m = 1; m <= nsam; m++) {
        int mm = m - 1;
        if (mm >= jc) mm -= nsam;
        for (int n = 1; n <= nsam; n++) {
          int nn = n - 1;
          if (nn >= jc) nn -= nsam;
          float frad2 = (float)(ll * ll + mm * mm + nn * nn);
          if (frad2 >= (float)lo2 && frad2 < (float)hi2) {
            size_t id = (size_t)(l - 1) + (size_t)jc * ((size_t)(m - 1) + (size_t)nsam * (n - 1));
            C a = cld(a3df, id), d = cld(d3df, id);
            R a1 = RSQRT(CABS2(a.re, a.im)), a2 = RSQRT(CABS2(d.re, d.im));
#if !OR_FP64
            a1 = hypotf(a.re, a.im);
            a2 = hypotf(d.re, d.im);
#endif
            R prod = a1 * a2;
            if (prod != 0) {
              C q = cmul(a, cconj(d));
              dfsc[is_ - 1] += (double)q.re;
              mean12 += (double)a1 * (double)a1;
              mean22 += (double)a2 * (double)a2;
              npix[is_ - 1]++;
              ps = ps + s3df[id] + v3df[id];
            }
            npixt[is_ - 1]++;
          }
        }
      }
    }
    if (npix[is_ - 1] != 0) {
      dfsc[is_ - 1] = dfsc[is_ - 1] / sqrt(mean12 * mean22);
      R f = (R)dfsc[is_ - 1];
      /* sheltest.f:100-104: 2.0*ABS(DFSC) is evaluated in DOUBLE, then stored in REAL A1 */
      R a1 = (RABS(f) < 1) ? (R)(2.0 * fabs(dfsc[is_ - 1]) / (1.0 - fabs(dfsc[is_ - 1])))
                           : (R)1000.0;
      if (ps != 0) ps = a1 * (R)npix[is_ - 1] / ps * (R)fmask;
    }
    pssnr[is_ - 1] = ps;
  }
  return nstp;
}

/* ------------------------------------------------- pad.f, rotate2d.f, ccp.f */
/* pad.f:23-65.  data: NSAM half-plane (complex); outd: (4NSAM)x(4NSAM+2) reals.
 * The c2r output's row-padding slots are taken as zero (their content is
 * FFT-library scratch in the reference). */
static void pad_impl(int nsam, int ipad, const R *data, R *outd) {
  int nsam4 = ipad * nsam, nsamh = nsam / 2;
  R scal = (R)1.0 / (R)nsam / (R)nsam;
  size_t tot = (size_t)nsam4 * (nsam4 + 2);
  memset(outd, 0, sizeof(R) * tot);
  size_t n1 = (size_t)nsam * (nsam + 2);
  for (size_t i = 0; i < n1; i++) outd[i] = data[i] * scal;
  IRFFT2(nsam, outd);
  R circ = 0;
  for (int i = 1; i <= nsam; i++)
    for (int j = 1; j <= nsam; j++)
      if (i == nsam / 2 + 1 || j == nsam / 2 + 1 || i == nsam / 2 || j == nsam / 2)
        circ = circ + outd[(size_t)(i - 1) + (size_t)(nsam + 2) * (j - 1)];
  circ = circ / (R)(4 * nsam - 4);
  for (int j = 1; j <= nsam4; j++)
    for (int i = 1; i <= nsam4; i++) {
      size_t id4 = (size_t)(i - 1) + (size_t)(nsam4 + 2) * (j - 1);
      outd[id4] = outd[id4] + circ;
    }
  for (int j = nsam; j >= 1; j--) {
    int jj = j;
    if (jj > nsamh) jj += nsam4 - nsam;
    for (int i = nsam; i >= 1; i--) {
      int ii = i;
      if (ii > nsamh) ii += nsam4 - nsam;
      size_t id = (size_t)(i - 1) + (size_t)(nsam + 2) * (j - 1);
      size_t id4 = (size_t)(ii - 1) + (size_t)(nsam4 + 2) * (jj - 1);
      R a = outd[id] - circ;
      outd[id] = circ;
      outd[id4] = a;
    }
  }
  RFFT2(nsam4, outd);
}

/* rotate2d.f:20-75; NINT = round half away from zero */
static void rotate2d_impl(int nsam, int ipada, const R *spec, R *outc, float psi, R *bufc) {
  int ipad = abs(ipada), nsam4 = ipad * nsam, nsamh = nsam / 2, jc = nsam / 2 + 1;
  int irad2 = (nsamh - 1) * (nsamh - 1), jc4 = nsam4 / 2 + 1;
  memset(outc, 0, sizeof(R) * 2 * (size_t)nsam * jc);
  float cpsi = cosf(psi), spsi = sinf(psi);
  float d1 = cpsi, d2 = spsi, d4 = -spsi, d5 = cpsi;
  if (ipada < 0) pad_impl(nsam, ipad, spec, bufc);
  for (int i = 0; i <= nsamh; i++)
    for (int j = -nsamh; j <= nsamh; j++) {
      int itemp = i * i + j * j;
      if (itemp < irad2) {
        size_t id = (size_t)i + (size_t)jc * wrapn(j, nsam);
        float x = d1 * (float)i + d4 * (float)j;
        float y = d2 * (float)i + d5 * (float)j;
        int l = (int)lroundf((float)ipad * x);
        int m = (int)lroundf((float)ipad * y);
        if (l >= 0) {
          size_t id4 = (size_t)l + (size_t)jc4 * wrapn(m, nsam4);
          cst(outc, id, cld(bufc, id4));
        } else {
          size_t id4 = (size_t)(-l) + (size_t)jc4 * wrapn(-m, nsam4);
          cst(outc, id, cconj(cld(bufc, id4)));
        }
      }
    }
}

double P(ccp)(int nsam, float amag, const R *spec, const float *a3df, float dangin, int maxr1,
              int maxr2, float phi, float theta, float *psi, float *shx, float *shy,
              const int mask[5], const R *ctff, int iquadmax, const float *sm) {
  int jc = nsam / 2 + 1;
  size_t ic = (size_t)nsam * (nsam + 2) / 2;
  int nsamh = nsam / 2, nsamrh = nsamh / 2;
  if (2 * nsamrh != nsamh) nsamrh = (nsamh + 1) / 2;
  int jcr = nsamrh + 1, nsamr = 2 * nsamrh;
  int maxr12 = maxr1 * maxr1, maxr22 = maxr2 * maxr2;
  float dang = 0.25f;
  if (dangin != 0.0f) dang = dangin;
  int npsi = (int)lroundf(90.0f / dang);
  float dm[9];
  or_dm(phi, theta, *psi, amag, sm, dm);
  size_t nccc = (size_t)nsam * jc;
  R *ccc = (R *)calloc(2 * nccc + 8, sizeof(R));
  R *cbuf = (R *)calloc(2 * nccc + 8, sizeof(R));
  int nsam4 = 4 * nsam;
  R *rbuf = (R *)calloc((size_t)nsam4 * (nsam4 + 2) + 8, sizeof(R));
  /* ccp.f:110-131 */
  for (int i = 0; i <= jc - 2; i++)
    for (int j = -jc + 1; j <= jc - 1; j++) {
      int itemp = i * i + j * j;
      size_t id = (size_t)i + (size_t)jc * wrapn(j, nsam);
      if (itemp >= maxr12 && itemp < maxr22 && abs(j) != jc - 1) {
        C cr = cld(ctff, id), cl = cld(ctff, id + ic);
        float x3 = dm[0] * (float)i + dm[3] * (float)j;
        float y3 = dm[1] * (float)i + dm[4] * (float)j;
        float z3 = dm[2] * (float)i + dm[5] * (float)j;
        C a = ainterpo3ds_c(nsam, a3df, x3, y3, z3);
        cst(ccc, id, cmul(a, cadd(cr, cconj(cl))));
      }
    }
  double ccmax = -1e30;
  int np = npsi, iq = iquadmax;
  if (mask[2] == 0) { np = 1; iq = 1; }
  float sx = 0, sy = 0, psim = *psi;
  R *ccd = ccc; /* real alias used for the NSAMR correlation map */
  for (int ipsi = 1; ipsi <= np; ipsi++) {
    float psis = (float)(ipsi - 1) * PI_F / 2.0f / (float)npsi;
    rotate2d_impl(nsam, ipsi == 1 ? -4 : 4, ccc, cbuf, psis, rbuf);
    for (int iquad = 1; iquad <= iq; iquad++) {
      R sum = 0, sum2 = 0;
      for (int i = 0; i <= nsamrh; i++)
        for (int j = -nsamrh; j <= nsamrh - 1; j++) {
          int itemp = i * i + j * j;
          size_t id = (size_t)i + (size_t)jc * wrapn(j, nsam);
          size_t id2 = (size_t)i + (size_t)jcr * wrapn(j, nsamr);
          if (itemp >= 1 && itemp < maxr22 && abs(j) != jc - 1) {
            C ainpo;
            if (iquad == 1) {
              ainpo = cld(cbuf, id);
            } else {
              int iii, jjj, icon;
              if (iquad == 2) {
                if (j <= 0) { iii = -j; jjj = i; icon = 1; }
                else { iii = j; jjj = wrapn(-i, nsam); icon = -1; }
              } else if (iquad == 3) {
                iii = i; jjj = wrapn(j, nsam); icon = -1;
              } else {
                if (j > 0) { iii = j; jjj = wrapn(-i, nsam); icon = 1; }
                else { iii = -j; jjj = i; icon = -1; }
              }
              ainpo = cld(cbuf, (size_t)iii + (size_t)jc * jjj);
              if (icon == -1) ainpo = cconj(ainpo);
            }
            C samp = cld(spec, id);
            cst(ccc, id2, cmul(samp, cconj(ainpo)));
            sum = sum + CABS2(ainpo.re, ainpo.im);
            sum2 = sum2 + CABS2(samp.re, samp.im);
          } else {
            C zero = {0, 0};
            cst(ccc, id2, zero);
          }
        }
      IRFFT2(nsamr, ccc);
      R ccm = (R)-1e30;
      int ixm = 0, iym = 0;
      for (int ix = -6 * mask[3]; ix <= 6 * mask[3]; ix++) {
        int jx = wrapn(ix, nsamr);
        for (int iy = -6 * mask[4]; iy <= 6 * mask[4]; iy++) {
          int iy_ = wrapn(iy, nsamr);
          R v = ccd[(size_t)jx + (size_t)(nsamr + 2) * iy_];
          if (v > ccm) { ccm = v; ixm = ix; iym = iy; }
        }
      }
      ccm = 0;
      for (int ix = ixm - 1; ix <= ixm + 1; ix++) {
        int jx = wrapn(ix, nsamr);
        for (int iy = iym - 1; iy <= iym + 1; iy++) {
          int iy_ = wrapn(iy, nsamr);
          ccm = ccm + ccd[(size_t)jx + (size_t)(nsamr + 2) * iy_];
        }
      }
      R cc = ccm / RSQRT(sum) / RSQRT(sum2) / (R)9;
      if ((double)cc > ccmax) {
        sx = (float)ixm / (float)nsamr * PI_F * 2.0f;
        sy = (float)iym / (float)nsamr * PI_F * 2.0f;
        psim = *psi + (float)(iquad - 1) * PI_F / 2.0f + psis;
        ccmax = (double)cc;
      }
    }
  }
  *shx = sx;
  *shy = sy;
  *psi = psim;
  free(ccc);
  free(cbuf);
  free(rbuf);
  return ccmax;
}

/* ------------------------------------------------------------ psearch.f */
double P(psearch)(int nsam, float amag, const R *spec, const float *a3df, float dangin,
                  int maxr1, int maxr2, float *phi, float *theta, float *psi, float *dshx,
                  float *dshy, const float *sang, int nsang, int iquadmax, const int mask[5],
                  float rbfact, const R *ctff, float *testpar, int ipmax, int32_t *iran,
                  const float *sm) {
  int jc = nsam / 2 + 1, nsamh = nsam / 2;
  size_t nc = (size_t)nsam * jc;
  double ccmax = -1e30;
  for (int k = 0; k < ipmax; k++) testpar[6 * k + 5] = 0.0f;
  R *cbuf = (R *)calloc(2 * nc + 8, sizeof(R));
  for (int i = 0; i <= nsamh; i++)
    for (int j = -nsamh; j <= nsamh; j++) {
      int itemp = i * i + j * j;
      size_t id = (size_t)i + (size_t)jc * wrapn(j, nsam);
      R w = REXP((R)(-rbfact * (float)itemp));
      cst(cbuf, id, cscale(cld(spec, id), w));
    }
  float dang = dangin / 180.0f * PI_F;
  float phir = (or_random(iran) - 0.5f) * dang * (float)mask[0];
  float thetar = (or_random(iran) - 0.5f) * dang * (float)mask[1];
  float psir = (or_random(iran) - 0.5f) * dang * (float)mask[2];
  dang = dangin / 5.0f;
  if (dang < 1.0f) dang = 1.0f;
  float phi1 = *phi, theta1 = *theta, psi1 = *psi;
  for (int i = 0; i < nsang; i++) {
    float xp[5];
    xp[0] = (float)mask[0] * (sang[3 * i + 0] + phir) + *phi;
    xp[1] = (float)mask[1] * (sang[3 * i + 1] + thetar) + *theta;
    xp[2] = (float)mask[2] * (sang[3 * i + 2] + psir) + *psi;
    for (int m = 0; m < 3; m++)
      if (xp[m] > PI_F) xp[m] = xp[m] - 2.0f * PI_F;
    xp[3] = 0.0f;
    xp[4] = 0.0f;
    double cc = P(ccp)(nsam, amag, cbuf, a3df, dang, maxr1, maxr2, xp[0], xp[1], &xp[2], &xp[3],
                       &xp[4], mask, ctff, iquadmax, sm);
    float ccf = (float)cc;
    if (xp[2] > PI_F) xp[2] = xp[2] - 2.0f * PI_F;
    for (int k = 0; k < ipmax; k++) {
      if (ccf > testpar[6 * k + 5]) {
        for (int kk = ipmax - 1; kk >= k + 1; kk--)
          for (int j = 0; j < 6; j++) testpar[6 * kk + j] = testpar[6 * (kk - 1) + j];
        for (int j = 0; j < 5; j++) testpar[6 * k + j] = xp[j];
        testpar[6 * k + 5] = ccf;
        break;
      }
    }
    if ((double)ccf > ccmax) {
      ccmax = ccf;
      phi1 = xp[0];
      theta1 = xp[1];
      psi1 = xp[2];
      *dshx = xp[3];
      *dshy = xp[4];
    }
  }
  *phi = phi1;
  *theta = theta1;
  *psi = psi1;
  free(cbuf);
  return ccmax;
}

/* ---------------------------------------------------------- phaseflip.f */
/* SPEC(ID) = SPEC(ID) * CTF / |CTF| over I = 0..JC-1, J = -JC+1..JC-1 (the
 * Nyquist row is reached from J = +-(JC-1) and so multiplied twice). */
static void phaseflip_c(int nsam, R *spec, const R *ctff) {
  int jc = nsam / 2 + 1;
  size_t ic = (size_t)nsam * jc;
  for (int i = 0; i <= jc - 1; i++)
    for (int j = -jc + 1; j <= jc - 1; j++) {
      size_t id = (size_t)i + (size_t)jc * wrapn(j, nsam);
      C c = cadd(cld(ctff, id), cconj(cld(ctff, id + ic)));
#if OR_FP64
      R amp = sqrt(c.re * c.re + c.im * c.im);
#else
      R amp = hypotf(c.re, c.im);
#endif
      if (amp != 0) {
        C z = cmul(cld(spec, id), c);
        z.re = z.re / amp;
        z.im = z.im / amp;
        cst(spec, id, z);
      }
    }
}

/* --------------------------------------------------------------- sigma.f */
void P(sigma)(int nsam, float amag, const float *a3df, const R *datc, float phi, float theta,
              float psi, float shx, float shy, const R *ctff, int maxr1, float rclas,
              float signset, const float *dmask, float sigp, const float *sm, R *out) {
  int jc = nsam / 2 + 1, nsamh = nsam / 2;
  size_t ic = (size_t)nsam * jc;
  int nsam2 = nsam * nsam;
  size_t ib1 = (size_t)nsam * (nsam + 2);
  R *rbuf = (R *)calloc(2 * ib1 + 8, sizeof(R));
  R *part = rbuf + ib1;
  float t = dmask[3] / amag;
  float ri2f = t * t;
  int mrclas = (int)((float)nsam * rclas * fabsf(amag));
  int maxr12 = maxr1 * maxr1, maxr22 = mrclas * mrclas;
  P(pextract)(nsam, rbuf, a3df, phi, theta, psi, ctff, sm);
  for (int l = 0; l <= jc - 1; l++)
    for (int m = 0; m <= nsam - 1; m++) {
      int mm = m >= jc ? m - nsam : m;
      int itemp = l * l + mm * mm;
      size_t id = (size_t)l + (size_t)jc * m;
      if (itemp >= maxr12 && itemp < maxr22) {
#if OR_FP64
        double phase = -(double)shx * l - (double)shy * mm;
#else
        float phase = -shx * (float)l - shy * (float)mm;
#endif
        C pshft = {RCOS(phase), RSIN(phase)};
        cst(part, id, cmul(pshft, cld(datc, id)));
      } else {
        C zero = {0, 0};
        cst(part, id, zero);
        cst(rbuf, id, zero);
      }
    }
  phaseflip_c(nsam, part, ctff);
  phaseflip_c(nsam, rbuf, ctff);
  IRFFT2(nsam, part);
  IRFFT2(nsam, rbuf);
  float dm[9];
  or_dm(-psi, -theta, -phi, amag, sm, dm);
  float xm = dm[0] * (dmask[0] - (float)jc) + dm[3] * (dmask[1] - (float)jc) +
             dm[6] * (dmask[2] - (float)jc) + 1.0f;
  float ym = dm[1] * (dmask[0] - (float)jc) + dm[4] * (dmask[1] - (float)jc) +
             dm[7] * (dmask[2] - (float)jc) + 1.0f;
  double sum1 = 0, sum2 = 0, sum5 = 0, cc = 0, ave1 = 0, ave2 = 0, ave5 = 0;
  int n1 = 0;
  for (int L = 1; L <= nsam; L++) {
    float dx = (float)L - xm;
    if (L >= nsamh) dx = (float)L - xm - (float)nsam;
    float x2 = dx * dx;
    for (int M = 1; M <= nsam; M++) {
      float dy = (float)M - ym;
      if (M >= nsamh) dy = (float)M - ym - (float)nsam;
      float y2 = dy * dy;
      size_t id = (size_t)(L - 1) + (size_t)(nsam + 2) * (M - 1);
      R r1 = rbuf[id], r2 = part[id];
      if (x2 + y2 <= ri2f) {
        cc = cc + (double)(r1 * r2);
        sum1 = sum1 + (double)(r1 * r1);
        sum2 = sum2 + (double)(r2 * r2);
        ave1 = ave1 + (double)r1;
        ave2 = ave2 + (double)r2;
        n1 = n1 + 1;
      }
      sum5 = sum5 + (double)(r2 * r2);
      ave5 = ave5 + (double)r2;
    }
  }
  cc = cc - ave1 * ave2 / n1;
  sum1 = sum1 - ave1 * ave1 / n1;
  sum2 = sum2 - ave2 * ave2 / n1;
  sum5 = sum5 - ave5 * ave5 / nsam2;
  double a = fabs(cc / sum1);
  double s2 = fabs(sum2 - a * a * sum1) / n1;
#if OR_FP64
  double sigma2 = s2 / nsam2 / nsam2;
  double s2old = sum5 / nsam2;
  double sigma2old = s2old / nsam2 / nsam2;
  if (signset != 0.0f) {
    double rr = sqrt(sigma2old) / signset;
    if (rr <= 3.0 && rr >= 0.33) {
      s2old = (double)signset * signset * nsam2 * nsam2;
      sigma2old = (double)signset * signset;
    }
  }
  double sigma2n = s2old / a / sqrt(sum1 * sum2);
  int n1w = (int)(n1 * sigma2old / ((double)sigp * sigp));
  double logpc = -n1w * sigma2 / sigma2old / 2.0 - n1w * log(2.0 * 3.1415926535897 * sigma2old) / 2.0;
#else
  float sigma2 = (float)(s2 / nsam2 / nsam2);
  float s2old = (float)(sum5 / nsam2);
  float sigma2old = s2old / (float)nsam2 / (float)nsam2;
  if (signset != 0.0f) {
    if (sqrtf(sigma2old) / signset <= 3.0f && sqrtf(sigma2old) / signset >= 0.33f) {
      s2old = signset * signset * (float)nsam2 * (float)nsam2;
      sigma2old = signset * signset;
    }
  }
  float sigma2n = (float)((double)s2old / a / sqrt(sum1 * sum2));
  int n1w = (int)((float)n1 * sigma2old / (sigp * sigp));
  float logpc = -((float)n1w * sigma2 / sigma2old / 2.0f) -
                (float)n1w * logf(2.0f * PI_F * sigma2old) / 2.0f;
#endif
  out[0] = sigma2n;
  out[1] = sigma2;
  out[2] = logpc;
  free(rbuf);
}

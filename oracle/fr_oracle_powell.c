/* fr_oracle_powell.c -- CPU restatement of VA04A (va04.f:8-369), Powell's
 * conjugate-direction minimiser, with the objective as a callback.
 * Test infrastructure only: the library never links it.
 *
 * Every quantity is REAL (FP32) as in the Fortran; the control flow follows
 * the reference's labels one for one (arithmetic IFs become three-way
 * branches on the FP32 difference, computed GOTOs fall through when out of
 * range), so that the iterates are bit-identical to the flang-compiled
 * reference on the same objective.  Pinned in tests/test_oracle_ref.py
 * against oracle/_ref (va04.f + the test objective of ref_va04_obj.f90). */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>

#include "fr_oracle.h"

/* Fortran SIGN(A, B): |A| with the sign of B */
static float fsign(float a, float b) { return copysignf(fabsf(a), b); }

float or_va04a(float *x0, const float *e0, int n, float escale, int iprint, int icon, int maxit,
               or_objective fx, void *user, int *nevals) {
  float *X = x0 - 1;        /* 1-based views */
  const float *E = e0 - 1;
  float W[61], XS[11];
  float F = 0, DDMAG, SCER, FKEEP, FP = 0, SUM = 0, DMAX = 0, DACC = 0, DMAG, DDMAX = 0, DL = 0,
        D = 0, FPREV = 0, FA = 0, DA = 0, DD = 0, FB = 0, DB = 0, FC = 0, DC = 0, A = 0, B = 0,
        DI = 0, FI = 0, FHOLD = 0, AAA = 0, TMP;
  int JJ, JJJ, K, NFCC, IND, INN, ITERC, ISGRAD, ITONE = 1, IXP, IDIRN = 0, ILINE = 0, IS = 5,
      JIL = 1, I, J, ICNT, MAXX;
  int calls = 0;
#define CALC()                        \
  do {                                \
    F = fx(n, X + 1, user);           \
    calls++;                          \
  } while (0)

  DDMAG = 0.1f * escale;
  SCER = 0.05f / escale;
  ICNT = 0;
  MAXX = 100 * maxit;
  for (I = 1; I <= n; I++) XS[I] = X[I];
  JJ = n * n + n;
  JJJ = JJ + n;
  K = n + 1;
  NFCC = 1;
  IND = 1;
  INN = 1;
  for (I = 1; I <= n; I++)
    for (J = 1; J <= n; J++) {
      W[K] = 0.f;
      if (I == J) {
        W[K] = fabsf(E[I]);
        W[I] = escale;
      }
      K = K + 1;
    }
  ITERC = 1;
  ISGRAD = 2;
  ICNT = ICNT + 1;
  if (ICNT > MAXX) goto L998;
  CALC();
  FKEEP = fabsf(F) + fabsf(F);
L5:
  ITONE = 1;
  FP = F;
  SUM = 0.f;
  IXP = JJ;
  for (I = 1; I <= n; I++) {
    IXP = IXP + 1;
    W[IXP] = X[I];
  }
  IDIRN = n + 1;
  ILINE = 1;
L7:
  DMAX = W[ILINE];
  DACC = DMAX * SCER;
  DMAG = fminf(DDMAG, 0.1f * DMAX);
  DMAG = fmaxf(DMAG, 20.f * DACC);
  DDMAX = 10.f * DMAG;
  if (ITONE == 3) goto L71;
  /* L70 */
  DL = 0.f;
  D = DMAG;
  FPREV = F;
  IS = 5;
  FA = F;
  DA = DL;
L8:
  DD = D - DL;
  DL = D;
L58:
  K = IDIRN;
  for (I = 1; I <= n; I++) {
    X[I] = X[I] + DD * W[K];
    K = K + 1;
  }
  ICNT = ICNT + 1;
  if (ICNT > MAXX) goto L998;
  CALC();
  NFCC = NFCC + 1;
  switch (IS) {
    case 1: goto L10;
    case 2: goto L11;
    case 3: goto L12;
    case 4: goto L13;
    case 6: goto L96;
    default: break; /* 5, or out of range: the next statement */
  }
  /* L14 */
  if (F - FA < 0.f) goto L15;
  if (F - FA == 0.f) goto L16;
  goto L24;
L16:
  if (fabsf(D) - DMAX <= 0.f) {
    D = D + D; /* L17 */
    goto L8;
  }
  goto L20; /* L18: "maximum change does not alter function" */
L15:
  FB = F;
  DB = D;
  goto L21;
L24:
  FB = FA;
  DB = DA;
  FA = F;
  DA = D;
L21:
  if (ISGRAD == 1) goto L83;
L23:
  D = DB + DB - DA;
  IS = 1;
  goto L8;
L83:
  D = 0.5f * (DA + DB - (FA - FB) / (DA - DB));
  IS = 4;
  if ((DA - D) * (D - DB) < 0.f) goto L25;
  goto L8;
L25:
  IS = 1;
  if (fabsf(D - DB) - DDMAX <= 0.f) goto L8;
L26:
  D = DB + fsign(DDMAX, DB - DA);
  IS = 1;
  DDMAX = DDMAX + DDMAX;
  DDMAG = DDMAG + DDMAG;
  if (DDMAX - DMAX <= 0.f) goto L8;
  DDMAX = DMAX; /* L27 */
  goto L8;
L13:
  if (F - FA < 0.f) goto L28;
  goto L23;
L28:
  FC = FB;
  DC = DB;
L29:
  FB = F;
  DB = D;
  goto L30;
L12:
  if (F - FB <= 0.f) goto L28;
  FA = F; /* L31 */
  DA = D;
  goto L30;
L11:
  if (F - FB < 0.f) {
    FA = FB; /* L32 */
    DA = DB;
    goto L29;
  }
  goto L10;
L71:
  DL = 1.f;
  DDMAX = 5.f;
  FA = FP;
  DA = -1.f;
  FB = FHOLD;
  DB = 0.f;
  D = 1.f;
L10:
  FC = F;
  DC = D;
L30:
  A = (DB - DC) * (FA - FC);
  B = (DC - DA) * (FB - FC);
  if ((A + B) * (DA - DC) <= 0.f) {
    FA = FB; /* L33 */
    DA = DB;
    FB = FC;
    DB = DC;
    goto L26;
  }
  /* L34 */
  D = 0.5f * (A * (DB + DC) + B * (DA + DC)) / (A + B);
  DI = DB;
  FI = FB;
  if (FB - FC > 0.f) {
    DI = DC; /* L43 */
    FI = FC;
  }
  /* L44 */
  if (ITONE == 3) {
    ITONE = 2; /* L85 */
    goto L45;
  }
  /* L86 */
  if (fabsf(D - DI) - DACC <= 0.f) goto L41;
  if (fabsf(D - DI) - 0.03f * fabsf(D) <= 0.f) goto L41; /* L93 */
L45:
  if ((DA - DC) * (DC - D) >= 0.f) {
    FA = FB; /* L46 */
    DA = DB;
    FB = FC;
    DB = DC;
    goto L25;
  }
  IS = 2; /* L47 */
  if ((DB - D) * (D - DC) < 0.f) IS = 3; /* L48 */
  goto L8;
L41:
  F = FI;
  D = DI - DL;
  DD = (DC - DB) * (DC - DA) * (DA - DB) / (A + B);
  if (DD < 0.f) DD = 1e-10f;
  DD = sqrtf(DD);
  for (I = 1; I <= n; I++) {
    X[I] = X[I] + D * W[IDIRN];
    W[IDIRN] = DD * W[IDIRN];
    IDIRN = IDIRN + 1;
  }
  if (DD == 0.f) DD = 1e-10f;
  W[ILINE] = W[ILINE] / DD;
  ILINE = ILINE + 1;
  if (iprint == 1) {
    /* L50 -> GO TO (51,53),IPRINT: IPRINT 1 -> L51 */
  }
  /* L51 */
  if (ITONE == 2) goto L38;
  /* L55 (ITONE 1, or 3 falling through) */
  if (FPREV - F - SUM >= 0.f) {
    SUM = FPREV - F; /* L95 */
    JIL = ILINE;
  }
  /* L94 */
  if (IDIRN - JJ <= 0) goto L7;
  /* L84 */
  if (IND == 2) goto L72;
L92:
  FHOLD = F;
  IS = 6;
  IXP = JJ;
  for (I = 1; I <= n; I++) {
    IXP = IXP + 1;
    W[IXP] = X[I] - W[IXP];
  }
  DD = 1.f;
  goto L58;
L96:
  if (IND == 2) goto L87;
  /* L112 */
  if (FP - F <= 0.f) goto L37;
  /* L91 */
  D = 2.f * (FP + F - 2.f * FHOLD) / ((FP - F) * (FP - F));
  if (D * ((FP - FHOLD - SUM) * (FP - FHOLD - SUM)) - SUM >= 0.f) goto L37;
L87:
  J = JIL * n + 1;
  if (J - JJ <= 0) {
    for (I = J; I <= JJ; I++) { /* L60 */
      K = I - n;
      W[K] = W[I];
    }
    for (I = JIL; I <= n; I++) W[I - 1] = W[I];
  }
  /* L61 */
  IDIRN = IDIRN - n;
  ITONE = 3;
  K = IDIRN;
  IXP = JJ;
  AAA = 0.f;
  for (I = 1; I <= n; I++) {
    IXP = IXP + 1;
    W[K] = W[IXP];
    if (AAA - fabsf(W[K] / E[I]) < 0.f) AAA = fabsf(W[K] / E[I]);
    K = K + 1;
  }
  DDMAG = 1.f;
  if (AAA == 0.f) AAA = 1e-10f;
  W[n] = escale / AAA;
  ILINE = n;
  goto L7;
L37:
  IXP = JJ;
  AAA = 0.f;
  F = FHOLD;
  for (I = 1; I <= n; I++) {
    IXP = IXP + 1;
    X[I] = X[I] - W[IXP];
    if (AAA * fabsf(E[I]) - fabsf(W[IXP]) < 0.f) AAA = fabsf(W[IXP] / E[I]);
  }
  goto L72;
L38:
  AAA = AAA * (1.f + DI);
  if (IND == 2) goto L106;
L72:
  /* IF (IPRINT-2) 53,50,50: printing only */
  /* L53 */
  if (IND == 2) goto L88;
  /* L109 */
  if (AAA - 0.1f > 0.f) goto L76;
  /* L89: GO TO (20,116),ICON */
  if (icon == 1) goto L20;
  /* L116 */
  IND = 2;
  if (INN == 2) goto L101;
  /* L100 */
  INN = 2;
  K = JJJ;
  for (I = 1; I <= n; I++) {
    K = K + 1;
    W[K] = X[I];
    X[I] = X[I] + 10.f * E[I];
  }
  FKEEP = F;
  ICNT = ICNT + 1;
  if (ICNT > MAXX) goto L998;
  CALC();
  NFCC = NFCC + 1;
  DDMAG = 0.f;
  goto L108;
L76:
  if (F - FP < 0.f) goto L35;
  goto L20; /* L78: "accuracy limited by errors in F" */
L88:
  IND = 1;
L35:
  TMP = FP - F;
  if (TMP > 0.f)
    DDMAG = 0.4f * sqrtf(TMP);
  else
    DDMAG = 0.f;
  ISGRAD = 1;
L108:
  ITERC = ITERC + 1;
  if (ITERC - maxit <= 0) goto L5;
  /* L81 */
  if (F - FKEEP <= 0.f) goto L20;
  F = FKEEP; /* L110 */
  for (I = 1; I <= n; I++) {
    JJJ = JJJ + 1;
    X[I] = W[JJJ];
  }
  goto L20;
L101:
  JIL = 1;
  FP = FKEEP;
  if (F - FKEEP == 0.f) goto L20; /* -> L78 */
  if (F - FKEEP > 0.f) {
    JIL = 2; /* L104 */
    FP = F;
    F = FKEEP;
  }
  /* L105 */
  IXP = JJ;
  for (I = 1; I <= n; I++) {
    IXP = IXP + 1;
    K = IXP + n;
    if (JIL == 1) {
      W[IXP] = W[K];
    } else {
      W[IXP] = X[I];
      X[I] = W[K];
    }
  }
  JIL = 2;
  goto L92;
L106:
  if (AAA - 0.1f <= 0.f) goto L20;
  INN = 1; /* L107 */
  goto L35;
L20:
  if (nevals) *nevals = calls;
  return F;
L998:
  for (I = 1; I <= n; I++) X[I] = XS[I];
  CALC();
  if (nevals) *nevals = calls;
  return F;
#undef CALC
}

/* The test objectives of oracle/ref_va04_obj.f90, term for term. */
float or_va04_testfn(int n, const float *x, void *which) {
  const int fn = *(const int *)which;
  float rf = 0.f;
  if (fn == 1) {
    for (int i = 1; i <= n; i++) {
      float t = x[i - 1] - 0.1f * (float)i;
      rf = rf + ((float)i * 0.7f) * (t * t);
    }
    if (n >= 2) rf = rf + 0.4f * (x[0] - 0.1f) * (x[1] - 0.2f);
  } else if (fn == 2) {
    if (n == 1) {
      rf = (1.f - x[0]) * (1.f - x[0]);
    } else {
      for (int i = 1; i <= n - 1; i++) {
        float u = x[i] - x[i - 1] * x[i - 1];
        float v = 1.f - x[i - 1];
        rf = rf + 100.f * (u * u) + v * v;
      }
    }
  } else {
    float q = 0.f;
    for (int i = 1; i <= n; i++) {
      float t = x[i - 1] - 0.05f * (float)i;
      q = q + (1.f + 0.5f * (float)i) * (t * t);
    }
    rf = -1.f / (1.f + q);
  }
  return rf;
}

/* ------------------------------------------------------------ prefine.f */
/* PREFINE (prefine.f:51-145; non-helical, AMAG > 0, IVFLAG = 0) over the
 * MASK-selected PHI THETA PSI DSHX DSHY: VA04A with steps EPAR (prefine.f:48),
 * ESCALE 100, ICON 1, MAXIT NCYCLS = 50, minimising CALCFX
 * (calcfx.f:152-191): -CC3M(|AMAG|) - FSH(SIG2, DSHX, DSHY, XM-SHX, YM-SHY,
 * SX, SY) (fsh.f) plus the out-of-range penalties.  par[5] in/out; prior =
 * (XM-SHX, YM-SHY, SX, SY, SIG2).  Returns PRESA = -F; *nevals the CC3M
 * evaluations. */
struct prefine_ctx {
  int nsam, maxr1, maxr2;
  float amag, rbfact, ri2, ri3, rih, halfw;
  const float *spec, *a3df, *fsct, *ctff, *sm, *prior;
  const int *mask;
  float xpar[5];
};

static float fsh(float sig2, float x, float y, float xm, float ym, float sx, float sy) {
  const float dx = x - xm, dy = y - ym;
  return sig2 * (-(dx * dx) / 2.0f / (sx * sx) - (dy * dy) / 2.0f / (sy * sy));
}

static float calcfx(int n, const float *xmap, void *user) {
  struct prefine_ctx *c = (struct prefine_ctx *)user;
  const float pi = 3.1415926535897f;
  float x[5];
  int ip = 0;
  for (int j = 0; j < 5; j++) x[j] = c->xpar[j];
  for (int j = 0; j < 5; j++)
    if (c->mask[j] == 1) x[j] = xmap[ip++];
  (void)n;
  float cc = (float)or32_cc3m(c->nsam, c->amag, c->spec, c->a3df, c->maxr1, c->maxr2, x[0], x[1],
                              x[2], x[3], x[4], c->rbfact, c->fsct, c->ctff, c->ri2, c->ri3,
                              c->rih, c->halfw, c->sm);
  float rft = -cc - fsh(c->prior[4], x[3], x[4], c->prior[0], c->prior[1], c->prior[2],
                        c->prior[3]);
  for (int j = 0; j < 3; j++)
    if (fabsf(x[j]) > pi) rft = rft + (pi - fabsf(x[j])) * (pi - fabsf(x[j]));
  for (int j = 3; j < 5; j++)
    if (fabsf(x[j]) > 1.0f) rft = rft + (1.0f - fabsf(x[j])) * (1.0f - fabsf(x[j]));
  return rft;
}

float or32_prefine(int nsam, float amag, const float *spec, const float *a3df, int maxr1,
                   int maxr2, float *par, const int mask[5], float rbfact, const float *fsct,
                   const float *ctff, float ri2, float ri3, float rih, float halfw,
                   const float prior[5], const float *sm, int *nevals) {
  static const float epar[5] = {0.1f, 0.1f, 0.1f, 0.01f, 0.01f};
  struct prefine_ctx c = {nsam, maxr1, maxr2, fabsf(amag), rbfact, ri2, ri3, rih, halfw,
                          spec, a3df, fsct, ctff, sm, prior, mask, {0}};
  float xmap[5], emap[5];
  int npar = 0;
  for (int j = 0; j < 5; j++) c.xpar[j] = par[j];
  for (int j = 0; j < 5; j++)
    if (mask[j] == 1) {
      xmap[npar] = par[j];
      emap[npar] = epar[j];
      npar++;
    }
  float f;
  int ne = 0;
  if (npar != 0) {
    f = or_va04a(xmap, emap, npar, 100.0f, 0, 1, 50, calcfx, &c, &ne);
  } else {
    f = calcfx(0, xmap, &c);
    ne = 1;
  }
  int ip = 0;
  for (int j = 0; j < 5; j++)
    if (mask[j] == 1) par[j] = xmap[ip++];
  if (nevals) *nevals = ne;
  return -f;
}

/* ---------------------------------------------- ctfrefine.f, magrefine.f */
/* CTFREFINE (ctfrefine.f:8-100, one group of n particles) and MAGREFINE
 * (magrefine.f:8-73): VA04A over CALCFX's CTF section (IVFLAG 1,
 * calcfx.f:63-110) or magnification section (IVFLAG 2, calcfx.f:112-140), each
 * evaluation re-running CTFAPPLY (ICTF 0) on every particle of the group and
 * scoring it with CC3 at zero shift, MAXR = INT(NSAM RMAX |AMAG|). */
struct ctf_ctx {
  int mode, nsam, n, npar;
  const float *spec, *ppar, *euler, *a3df, *pssnr, *gpar, *ring;
  float dfstd, sig2, rmax1, rmax2, rbfact;
  float *ctff, *outc, *datc;
};

static float fd(float sig2, float d1, float d2, float dfstd) {
  return sig2 * (-(d1 * d1) / 2.0f / (dfstd * dfstd) - (d2 * d2) / 2.0f / (dfstd * dfstd));
}

static float calcfx_ctf(int nx, const float *x, void *user) {
  struct ctf_ctx *c = (struct ctf_ctx *)user;
  const int nsam = c->nsam, jc = nsam / 2 + 1;
  const size_t ic = (size_t)nsam * jc;
  const float *g = c->gpar;
  float rft = 0.0f;
  (void)nx;
  for (int k = 0; k < c->n; k++) {
    const float *p = c->ppar + 8 * (size_t)k;
    const float *e = c->euler + 3 * (size_t)k;
    float df1 = p[2], df2 = p[3], ast = p[4], amag = p[5];
    float x1 = 0, x2 = 0;
    if (c->mode == 1) {
      x1 = x[0];
      x2 = c->npar == 1 ? x[0] : x[1];
      df1 = p[2] + x1;
      df2 = p[3] + x2;
      ast = p[4] + (c->npar == 1 ? 0.0f : x[2]);
    } else {
      amag = x[0];
    }
    const int maxr1 = (int)((float)nsam * c->rmax1 * fabsf(amag));
    const int maxr2 = (int)((float)nsam * c->rmax2 * fabsf(amag));
    or32_ctfapply(nsam, c->spec + 2 * ic * (size_t)k, p[0], p[1], g[0], g[1], g[2], g[3], df1, df2,
                  ast, g[4], amag, g[5], g[6], g[7], g[8], c->pssnr, 0, c->ctff, c->outc,
                  c->datc);
    const float cc = (float)or32_cc3(nsam, amag, c->outc, c->a3df, maxr1, maxr2, e[0], e[1], e[2],
                                     0.0f, 0.0f, c->rbfact, c->ctff, c->ring[0], c->ring[1],
                                     c->ring[2], c->ring[3], NULL);
    if (c->mode == 1)
      rft = rft - cc - fd(c->sig2, x1, x2, c->dfstd);
    else
      rft = rft + (-cc);
  }
  return rft / (float)c->n;
}

static float refine_ctf(int mode, int nsam, int n, const float *spec, const float *ppar,
                        const float *euler, int fastig, float dfstd, float sig2,
                        const float *a3df, const float *pssnr, const float *gpar, float rmax1,
                        float rmax2, float rbfact, const float *ring, float *xout, int *nevals) {
  const int jc = nsam / 2 + 1;
  const size_t ic = (size_t)nsam * jc;
  struct ctf_ctx c = {mode, nsam, n, mode == 1 && fastig ? 3 : 1, spec, ppar, euler, a3df, pssnr,
                      gpar, ring, dfstd, sig2, rmax1, rmax2, rbfact, NULL, NULL, NULL};
  c.ctff = (float *)calloc(4 * ic, sizeof(float));
  c.outc = (float *)calloc(2 * ic, sizeof(float));
  c.datc = (float *)calloc(2 * ic, sizeof(float));
  float x[3] = {0.0f, 0.0f, 0.0f};
  float e[3] = {1.0f, 1.0f, 1.0f};
  float escale = 1000.0f;
  if (mode == 2) {
    x[0] = ppar[5];
    e[0] = 0.01f;
    escale = 100.0f;
  }
  int ne = 0;
  float f = or_va04a(x, e, c.npar, escale, 0, 1, 50, calcfx_ctf, &c, &ne);
  free(c.ctff);
  free(c.outc);
  free(c.datc);
  if (mode == 1) {
    xout[0] = x[0];
    xout[1] = c.npar == 1 ? x[0] : x[1];
    xout[2] = c.npar == 1 ? 0.0f : x[2];
  } else {
    xout[0] = x[0];
    xout[1] = xout[2] = 0.0f;
  }
  if (nevals) *nevals = ne;
  return f;
}

float or32_ctfrefine(int nsam, int n, const float *spec, const float *ppar, const float *euler,
                     int fastig, float dfstd, float sig2, const float *a3df, const float *pssnr,
                     const float *gpar, float rmax1, float rmax2, float rbfact, const float *ring,
                     float *xout, int *nevals) {
  return refine_ctf(1, nsam, n, spec, ppar, euler, fastig, dfstd, sig2, a3df, pssnr, gpar, rmax1,
                    rmax2, rbfact, ring, xout, nevals);
}

float or32_magrefine(int nsam, int n, const float *spec, const float *ppar, const float *euler,
                     const float *a3df, const float *pssnr, const float *gpar, float rmax1,
                     float rmax2, float rbfact, const float *ring, float *xout, int *nevals) {
  return refine_ctf(2, nsam, n, spec, ppar, euler, 0, 1.0f, 0.0f, a3df, pssnr, gpar, rmax1, rmax2,
                    rbfact, ring, xout, nevals);
}

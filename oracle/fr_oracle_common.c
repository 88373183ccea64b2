/*
 * fr_oracle_common.c -- precision-independent parts of the CPU restatement.
 * TEST INFRASTRUCTURE ONLY (see fr_oracle.h).
 * Built with -O2 -ffp-contract=off so that FP32 expressions round exactly
 * as written (the reference is compiled without FMA contraction).
 */
#include "fr_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI_F 3.1415926535897f /* PARAMETER (PI=3.1415926535897) as REAL */

/* random.f:18-29 */
float or_random(int32_t *is) {
  int32_t s = *is;
  if (s < 0) s = -s * 127773;
  int32_t i1 = s / 127773;
  int32_t i2 = s % 127773;
  int32_t i3 = 16807 * i2 - 2836 * i1;
  s = (i3 <= 0) ? i3 + 2147483647 : i3;
  *is = s;
  return (float)s / (float)2147483647;
}

/* matmul_t.f:17-24: T(J,I) = A(1,I)*B(1,J) + A(2,I)*B(2,J) + A(3,I)*B(3,J) */
void or_matmul_t(const float a[9], const float b[9], float c[9]) {
  float t[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      float s = a[0 + 3 * i] * b[0 + 3 * j];
      s = s + a[1 + 3 * i] * b[1 + 3 * j];
      s = s + a[2 + 3 * i] * b[2 + 3 * j];
      t[j + 3 * i] = s;
    }
  memcpy(c, t, sizeof t);
}

/* matmul.f: C = A*B (used by checksym.f) */
static void matmul3(const float a[9], const float b[9], float c[9]) {
  float t[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      float s = a[i + 0] * b[0 + 3 * j];
      s = s + a[i + 3] * b[1 + 3 * j];
      s = s + a[i + 6] * b[2 + 3 * j];
      t[i + 3 * j] = s;
    }
  memcpy(c, t, sizeof t);
}

/* cc3m.f:48-65 (amag != 0) and pextract.f:30-51 (amag == 0: no division). */
void or_dm(float phi, float theta, float psi, float amag, const float *sm, float dm[9]) {
  float cphi = cosf(phi), sphi = sinf(phi);
  float cthe = cosf(theta), sthe = sinf(theta);
  float cpsi = cosf(psi), spsi = sinf(psi);
  dm[0] = cphi * cthe * cpsi - sphi * spsi;
  dm[1] = sphi * cthe * cpsi + cphi * spsi;
  dm[2] = -sthe * cpsi;
  dm[3] = -cphi * cthe * spsi - sphi * cpsi;
  dm[4] = -sphi * cthe * spsi + cphi * cpsi;
  dm[5] = sthe * spsi;
  dm[6] = sthe * cphi;
  dm[7] = sthe * sphi;
  dm[8] = cthe;
  if (amag != 0.0f) {
    float a = fabsf(amag);
    for (int i = 0; i < 9; i++) dm[i] = dm[i] / a;
  }
  if (sm) or_matmul_t(dm, sm, dm);
}

/* checksym.f:8-44 */
static int checksym(const float symop[9]) {
  float tm[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  int count = 0;
  for (;;) {
    matmul3(tm, symop, tm);
    count++;
    float diff = 0.0f;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        float v = tm[i + 3 * j];
        diff = diff + ((i == j) ? fabsf(v - 1.0f) : fabsf(v));
      }
    if (diff < 0.01f) return count;
    if (count > 50) return -1;
  }
}

/* getsymmat.f:30-40: the eight stored operators (column-major 3x3). */
static const float SYMSTORE[8][9] = {
    {0.309017f, 0.809017f, -0.5f, -0.809017f, 0.5f, 0.309017f, 0.5f, 0.309017f, 0.809017f},
    {0.809017f, 0.309017f, 0.5f, 0.309017f, 0.5f, -0.809017f, -0.5f, 0.809017f, 0.309017f},
    {0.f, 1.f, 0.f, -1.f, 0.f, 0.f, 0.f, 0.f, 1.f},
    {0.f, 0.f, 1.f, 1.f, 0.f, 0.f, 0.f, 1.f, 0.f},
    {1.f, 0.f, 0.f, 0.f, -1.f, 0.f, 0.f, 0.f, -1.f},
    {-1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, -1.f},
    {0.f, 1.f, 0.f, 1.f, 0.f, 0.f, 0.f, 0.f, -1.f},
    {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 1.f}};
static const int NSYM_CDTOI[6] = {1, 2, 3, 3, 4, 4};
static const int NSYMTYPE[6][4] = {{8, 0, 0, 0}, {8, 5, 0, 0}, {4, 5, 6, 0},
                                   {3, 7, 4, 0}, {1, 4, 5, 6}, {2, 4, 5, 6}};

/* getsymmat.f:41-96 */
int or_getsymmat(char asym, int nasym, float *symop, int *jsym, int *iredun) {
  float store[8][9];
  memcpy(store, SYMSTORE, sizeof store);
  int n = 0;
  switch (asym) {
    case 'C': n = 1; break;
    case 'D': n = 2; break;
    case 'T': n = 3; break;
    case 'O': n = 4; break;
    case 'I': n = (nasym == 2) ? 6 : 5; break;
    default: return -1;
  }
  if (n <= 2) {
    if (nasym == 0) return -1;
    float angle = 2.0f * PI_F / (float)nasym;
    store[7][0] = cosf(angle);            /* SYMSTORE(1,1,8) */
    store[7][3] = sinf(angle);            /* SYMSTORE(1,2,8) */
    store[7][1] = -sinf(angle);           /* SYMSTORE(2,1,8) */
    store[7][4] = cosf(angle);            /* SYMSTORE(2,2,8) */
  }
  int nsym = NSYM_CDTOI[n - 1];
  *iredun = 1;
  for (int j = 0; j < nsym; j++) {
    int nget = NSYMTYPE[n - 1][j];
    memcpy(symop + 9 * j, store[nget - 1], 9 * sizeof(float));
    jsym[j] = checksym(symop + 9 * j);
    *iredun *= jsym[j];
  }
  return nsym;
}

/* limitsymm.f:31-49 */
static int limitsymm(char asym, int nasym, float *thetamax, float *phimax, int *jmax,
                     int *iquadmax) {
  static const float TH[5] = {90.0f, 90.0f, 54.7f, 54.7f, 31.7f};
  static const float PH[5] = {360.0f, 360.0f, 180.0f, 90.0f, 180.0f};
  static const int JS[5] = {2, 1, 1, 1, 1};
  int n;
  switch (asym) {
    case 'C': n = 1; break;
    case 'D': n = 2; break;
    case 'T': n = 3; break;
    case 'O': n = 4; break;
    case 'I': n = 5; break;
    default: return -1;
  }
  *thetamax = TH[n - 1];
  *jmax = JS[n - 1];
  *iquadmax = 4;
  *phimax = (n <= 2) ? PH[n - 1] / (float)nasym : PH[n - 1];
  return 0;
}

/* Fortran REAL DO-loop trip count: MAX(INT((m2-m1+m3)/m3),0), evaluated in REAL. */
static int real_trip(float m1, float m2, float m3) {
  float t = (m2 - m1 + m3) / m3;
  int n = (int)t;
  return n > 0 ? n : 0;
}

/* searchang.f:29-135 (non-helical). */
int or_searchang(float rmm, float ri, float dangin, char asym, int nasym, const int mask[5],
                 int asymstore, float *sang, int *iquadmax, float *dang_out) {
  float dang = 1.0f / rmm / ri / PI_F * 180.0f;
  if (dangin != 0.0f) dang = dangin;
  float tmp = (float)(int)(90.0f / dang + 0.5f);
  tmp = fmaxf(0.25f, tmp);
  dang = 90.0f / tmp;
  int nsang = 0;
  float thetamax = 90.0f * (float)mask[1];
  float thetamin = 0.0f;
  float thetamax1 = 90.0f;
  float phimax = 359.999f * (float)mask[2];
  int jmax = 2;
  *iquadmax = 4;
  if (asymstore) limitsymm(asym, nasym, &thetamax, &phimax, &jmax, iquadmax);
  int nth = real_trip(thetamin, thetamax, dang);
  float theta1 = thetamin;
  for (int it = 0; it < nth; it++, theta1 = theta1 + dang) {
    float dphi;
    if (theta1 == 0.0f || theta1 == 180.0f) {
      dphi = 360.0f;
    } else {
      dphi = dang / sinf(theta1 * PI_F / 180.0f);
      int k = (int)(359.99f / dphi) - 1;
      if (k < 1) k = 1;
      dphi = 360.0f / (float)k;
    }
    int nph = real_trip(0.0f, phimax, dphi);
    float phi1 = 0.0f;
    for (int ip = 0; ip < nph; ip++, phi1 = phi1 + dphi) {
      if (!((fabsf(theta1 - 90.0f) > 1.0f) || (phi1 < 179.0f))) continue;
      if (asym == 'T' || asym == 'O') {
        float pttt = fabsf(fmodf(phi1, 90.0f));
        pttt = fminf(pttt, 90.0f - pttt);
        thetamax1 = atan2f(1.0f, cosf(pttt * PI_F / 180.0f)) * 180.0f / PI_F;
      } else if (asym == 'I') {
        float a25 = atan2f(2.0f, 1.0f + sqrtf(5.0f));
        float a23 = asinf(tanf(30.0f * PI_F / 180.0f) * tanf(a25));
        float a253 = atan2f(tanf(30.0f * PI_F / 180.0f) * cosf(a23), 1.0f);
        float pc = (nasym == 2) ? fabsf(fmodf(phi1, 180.0f)) : fabsf(fmodf(phi1 + 90.0f, 180.0f));
        pc = fminf(pc, 180.0f - pc) * PI_F / 180.0f;
        float d = sinf(a25) / (cosf(pc) + sinf(pc) / tanf(a253));
        float c = cosf(a25) - d * sinf(pc) * tanf(a23);
        thetamax1 = atan2f(d, c) * 180.0f / PI_F;
      }
      if (theta1 > thetamax1) continue;
      float psi1 = 0.0f;
      for (int j = 1; j <= jmax; j++) {
        if (sang) {
          float *s = sang + 3 * nsang;
          if (j == 1) {
            s[0] = phi1; s[1] = theta1; s[2] = psi1;
          } else {
            s[0] = phi1 + 180.0f; s[1] = 180.0f - theta1; s[2] = psi1 + 180.0f;
          }
          for (int m = 0; m < 3; m++)
            if (s[m] > 180.0f) s[m] = s[m] - 360.0f;
          for (int m = 0; m < 3; m++) s[m] = s[m] / 180.0f * PI_F;
        }
        nsang++;
      }
    }
  }
  if (dang_out) *dang_out = dang;
  return nsang;
}

/* ------------------------------------------------------------------------ */
/* FFT: recursive mixed-radix complex DFT in FP64 (any length; radices 2..n). */
typedef struct { double re, im; } zc;

/* Per-thread cache of e^{-2 pi i k / n} tables (the recursion reuses the
 * top-level table with a stride, so no trigonometry inside the butterflies). */
#define TW_SLOTS 8
static __thread int tw_n[TW_SLOTS];
static __thread zc *tw_tab[TW_SLOTS];
static const zc *twiddles(int n) {
  for (int i = 0; i < TW_SLOTS; i++)
    if (tw_n[i] == n) return tw_tab[i];
  int slot = 0;
  for (int i = 0; i < TW_SLOTS; i++)
    if (tw_n[i] == 0) { slot = i; break; }
  free(tw_tab[slot]);
  zc *t = (zc *)malloc(sizeof(zc) * (size_t)n);
  for (int k = 0; k < n; k++) {
    double ang = -2.0 * M_PI * (double)k / (double)n;
    t[k].re = cos(ang);
    t[k].im = sin(ang);
  }
  tw_n[slot] = n;
  tw_tab[slot] = t;
  return t;
}

/* out[q*m+k] = sum_r W_n^{sign*r(q*m+k)} Y_r[k], W_n = W_N^{N/n} */
static void dft_rec(const zc *in, zc *out, int n, int stride, int sign, zc *scratch,
                    const zc *tw, int N) {
  if (n == 1) { out[0] = in[0]; return; }
  int p = 2;
  while (p * p <= n && n % p) p++;
  if (n % p) p = n;
  int m = n / p;
  for (int r = 0; r < p; r++)
    dft_rec(in + r * stride, out + r * m, m, stride * p, sign, scratch, tw, N);
  int step = N / n;
  for (int k = 0; k < m; k++) {
    for (int q = 0; q < p; q++) {
      double sr = 0.0, si = 0.0;
      for (int r = 0; r < p; r++) {
        long e = ((long)r * (long)(q * m + k)) % n;
        zc w = tw[(size_t)e * step];
        double wi = sign < 0 ? w.im : -w.im;
        zc y = out[r * m + k];
        sr += y.re * w.re - y.im * wi;
        si += y.re * wi + y.im * w.re;
      }
      scratch[q].re = sr;
      scratch[q].im = si;
    }
    for (int q = 0; q < p; q++) out[q * m + k] = scratch[q];
  }
}

static void zfft(zc *x, int n, int sign) {
  zc *tmp = (zc *)malloc(sizeof(zc) * (size_t)n);
  zc *scr = (zc *)malloc(sizeof(zc) * (size_t)n);
  dft_rec(x, tmp, n, 1, sign, scr, twiddles(n), n);
  memcpy(x, tmp, sizeof(zc) * (size_t)n);
  free(tmp);
  free(scr);
}

/* r2c: rows along x (fast), then columns along y.  fftw_plans.f:12 */
static void rfft2_core(int n, double *data) {
  int jc = n / 2 + 1, pitch = n + 2;
  zc *row = (zc *)malloc(sizeof(zc) * (size_t)n);
  for (int y = 0; y < n; y++) {
    double *r = data + (size_t)pitch * y;
    for (int x = 0; x < n; x++) { row[x].re = r[x]; row[x].im = 0.0; }
    zfft(row, n, -1);
    for (int x = 0; x < jc; x++) { r[2 * x] = row[x].re; r[2 * x + 1] = row[x].im; }
  }
  zc *col = (zc *)malloc(sizeof(zc) * (size_t)n);
  for (int x = 0; x < jc; x++) {
    for (int y = 0; y < n; y++) {
      col[y].re = data[(size_t)pitch * y + 2 * x];
      col[y].im = data[(size_t)pitch * y + 2 * x + 1];
    }
    zfft(col, n, -1);
    for (int y = 0; y < n; y++) {
      data[(size_t)pitch * y + 2 * x] = col[y].re;
      data[(size_t)pitch * y + 2 * x + 1] = col[y].im;
    }
  }
  free(row);
  free(col);
}

/* c2r: columns along y (complex, e^{+i}), then each row c2r along x using only
 * the real parts of the x=0 and x=n/2 terms.  fftw_plans.f:13 */
static void irfft2_core(int n, double *data) {
  int jc = n / 2 + 1, pitch = n + 2;
  zc *col = (zc *)malloc(sizeof(zc) * (size_t)n);
  for (int x = 0; x < jc; x++) {
    for (int y = 0; y < n; y++) {
      col[y].re = data[(size_t)pitch * y + 2 * x];
      col[y].im = data[(size_t)pitch * y + 2 * x + 1];
    }
    zfft(col, n, +1);
    for (int y = 0; y < n; y++) {
      data[(size_t)pitch * y + 2 * x] = col[y].re;
      data[(size_t)pitch * y + 2 * x + 1] = col[y].im;
    }
  }
  zc *row = (zc *)malloc(sizeof(zc) * (size_t)n);
  for (int y = 0; y < n; y++) {
    double *r = data + (size_t)pitch * y;
    for (int x = 0; x < jc; x++) { row[x].re = r[2 * x]; row[x].im = r[2 * x + 1]; }
    row[0].im = 0.0;
    if (n % 2 == 0) row[n / 2].im = 0.0;
    for (int x = jc; x < n; x++) { row[x].re = row[n - x].re; row[x].im = -row[n - x].im; }
    zfft(row, n, +1);
    for (int x = 0; x < n; x++) r[x] = row[x].re;
    r[n] = 0.0;
    r[n + 1] = 0.0;
  }
  free(col);
  free(row);
}

void or_rfft2_d(int n, double *data) { rfft2_core(n, data); }
void or_irfft2_d(int n, double *data) { irfft2_core(n, data); }

void or_rfft2_f(int n, float *data) {
  size_t sz = (size_t)n * (n + 2);
  double *d = (double *)malloc(sizeof(double) * sz);
  for (size_t i = 0; i < sz; i++) d[i] = data[i];
  rfft2_core(n, d);
  for (size_t i = 0; i < sz; i++) data[i] = (float)d[i];
  free(d);
}

void or_irfft2_f(int n, float *data) {
  size_t sz = (size_t)n * (n + 2);
  double *d = (double *)malloc(sizeof(double) * sz);
  for (size_t i = 0; i < sz; i++) d[i] = data[i];
  irfft2_core(n, d);
  for (size_t i = 0; i < sz; i++) data[i] = (float)d[i];
  free(d);
}

/* calcsinc.f:17-19: SINCLUT(J) = SIN(J*PI/180)/(J*PI/180), J = 1..n (FP32) */
void or_calcsinc(float *sinclut, int n) {
  const float pi = 3.1415926535897f;
  for (int j = 1; j <= n; j++) {
    float x = (float)j * pi / 180.0f;
    sinclut[j - 1] = sinf(x) / x;
  }
}

/* combine_arrays.f:53-67: NKSUM += NKSUM1(1..IMP); STD (and VTD for IFSC=0)
 * += TTD1 (TTD2) over NVOL1 entries when NVOL1 != 1, else over IMP */
void or_combine_counters(int ifsc, int imp, int nvol1, const int64_t *nksum1, const double *ttd1,
                         const double *ttd2, int64_t *nksum, double *std, double *vtd) {
  for (int i = 0; i < imp; i++) *nksum += nksum1[i];
  int n = nvol1 != 1 ? nvol1 : imp;
  for (int i = 0; i < n; i++) {
    *std = *std + ttd1[i];
    if (ifsc == 0) *vtd = *vtd + ttd2[i];
  }
}

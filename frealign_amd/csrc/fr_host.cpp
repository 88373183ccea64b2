// fr_host.cpp -- host-side geometry for the hot path: the orientation grid,
// the point-group operators, rotation matrices and the jitter RNG.
//
// These run on the host so that the FP32 rotation matrices the kernels use
// are identical to the ones the reference computes (glibc cosf/sinf, the
// reference's operation order, no FMA contraction): integer Miller indices
// and grid indices then come out bit-identical on the GPU.
#include <math.h>
#include <string.h>

#include "fr_host.h"

#pragma clang fp contract(off)

namespace fr {

static const float kPi = 3.1415926535897f;  // PARAMETER (PI=3.1415926535897), REAL

float lcg_random(int* is) {  // random.f:18-29
  int s = *is;
  if (s < 0) s = -s * 127773;
  int i1 = s / 127773, i2 = s % 127773;
  int i3 = 16807 * i2 - 2836 * i1;
  s = i3 <= 0 ? i3 + 2147483647 : i3;
  *is = s;
  return (float)s / (float)2147483647;
}

void matmul_t(const float a[9], const float b[9], float c[9]) {  // matmul_t.f:17-24
  float t[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      float s = a[3 * i] * b[3 * j];
      s = s + a[1 + 3 * i] * b[1 + 3 * j];
      s = s + a[2 + 3 * i] * b[2 + 3 * j];
      t[j + 3 * i] = s;
    }
  memcpy(c, t, sizeof t);
}

static void matmul(const float a[9], const float b[9], float c[9]) {  // matmul.f
  float t[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      float s = a[i] * b[3 * j];
      s = s + a[i + 3] * b[1 + 3 * j];
      s = s + a[i + 6] * b[2 + 3 * j];
      t[i + 3 * j] = s;
    }
  memcpy(c, t, sizeof t);
}

// cc3m.f:48-65 (amag != 0 divides by |AMAG|), pextract.f:30-51 (amag == 0)
void rotation(float phi, float theta, float psi, float amag, const float* sm, float dm[9]) {
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
  if (sm) matmul_t(dm, sm, dm);
}

static int checksym(const float op[9]) {  // checksym.f:19-43
  float tm[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  for (int count = 1;; count++) {
    matmul(tm, op, tm);
    float diff = 0.0f;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        float v = tm[i + 3 * j];
        diff = diff + (i == j ? fabsf(v - 1.0f) : fabsf(v));
      }
    if (diff < 0.01f) return count;
    if (count > 50) return -1;
  }
}

int point_group(char asym, int nasym, float* symop, int* jsym, int* nsym, int* iredun) {
  // getsymmat.f:30-96; operators stored column-major, as SYMSTORE(3,3,8)
  float store[8][9] = {
      {0.309017f, 0.809017f, -0.5f, -0.809017f, 0.5f, 0.309017f, 0.5f, 0.309017f, 0.809017f},
      {0.809017f, 0.309017f, 0.5f, 0.309017f, 0.5f, -0.809017f, -0.5f, 0.809017f, 0.309017f},
      {0.f, 1.f, 0.f, -1.f, 0.f, 0.f, 0.f, 0.f, 1.f},
      {0.f, 0.f, 1.f, 1.f, 0.f, 0.f, 0.f, 1.f, 0.f},
      {1.f, 0.f, 0.f, 0.f, -1.f, 0.f, 0.f, 0.f, -1.f},
      {-1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, -1.f},
      {0.f, 1.f, 0.f, 1.f, 0.f, 0.f, 0.f, 0.f, -1.f},
      {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 1.f}};
  static const int kNsym[6] = {1, 2, 3, 3, 4, 4};
  static const int kType[6][4] = {{8, 0, 0, 0}, {8, 5, 0, 0}, {4, 5, 6, 0},
                                  {3, 7, 4, 0}, {1, 4, 5, 6}, {2, 4, 5, 6}};
  int g;
  switch (asym) {
    case 'C': g = 1; break;
    case 'D': g = 2; break;
    case 'T': g = 3; break;
    case 'O': g = 4; break;
    case 'I': g = nasym == 2 ? 6 : 5; break;
    default: return -1;
  }
  if (g <= 2) {
    if (nasym <= 0) return -1;
    float ang = 2.0f * kPi / (float)nasym;
    store[7][0] = cosf(ang);
    store[7][3] = sinf(ang);
    store[7][1] = -sinf(ang);
    store[7][4] = cosf(ang);
  }
  *nsym = kNsym[g - 1];
  *iredun = 1;
  for (int j = 0; j < *nsym; j++) {
    memcpy(symop + 9 * j, store[kType[g - 1][j] - 1], 9 * sizeof(float));
    jsym[j] = checksym(symop + 9 * j);
    if (jsym[j] <= 0) return -1;
    *iredun *= jsym[j];
  }
  return 0;
}

// The IREDUN symmetry copies of pinsert.f:74-93,178-190, in the reference's
// enumeration order (ISYM(1) fastest), each as the 9 numbers of DM.
int symmetry_copies(const float tm[9], int nsym, const int* jsym, const float* symop, float* out) {
  int isym[8];
  for (int i = 0; i < nsym; i++) isym[i] = 1;
  int ncopy = 0;
  for (;;) {
    float dm[9];
    memcpy(dm, tm, sizeof dm);
    for (int i = 0; i < nsym; i++)
      if (isym[i] != jsym[i])
        for (int j = 0; j < isym[i]; j++) matmul_t(dm, symop + 9 * i, dm);
    memcpy(out + 9 * ncopy, dm, sizeof dm);
    ncopy++;
    int i = 0;
    bool done = false;
    for (;;) {
      isym[i]++;
      if (isym[i] > jsym[i]) {
        isym[i] = 1;
        if (++i < nsym) continue;
        done = true;
      }
      break;
    }
    if (done || nsym == 0) break;
  }
  return ncopy;
}

static int real_trip(float m1, float m2, float m3) {  // Fortran REAL DO trip count
  float t = (m2 - m1 + m3) / m3;
  int n = (int)t;
  return n > 0 ? n : 0;
}

// searchang.f:29-135 with limitsymm.f:31-49 (non-helical)
int search_grid(float rmm, float ri, float dangin, char asym, int nasym, const int mask[5],
                int asymstore, float* sang, int cap, int* iquadmax) {
  float dang = 1.0f / rmm / ri / kPi * 180.0f;
  if (dangin != 0.0f) dang = dangin;
  float tmp = (float)(int)(90.0f / dang + 0.5f);
  tmp = fmaxf(0.25f, tmp);
  dang = 90.0f / tmp;
  float thetamax = 90.0f * (float)mask[1], thetamin = 0.0f, thetamax1 = 90.0f;
  float phimax = 359.999f * (float)mask[2];
  int jmax = 2;
  *iquadmax = 4;
  if (asymstore) {
    static const float TH[5] = {90.0f, 90.0f, 54.7f, 54.7f, 31.7f};
    static const float PH[5] = {360.0f, 360.0f, 180.0f, 90.0f, 180.0f};
    static const int JS[5] = {2, 1, 1, 1, 1};
    int g = asym == 'C' ? 1 : asym == 'D' ? 2 : asym == 'T' ? 3 : asym == 'O' ? 4 : asym == 'I' ? 5 : 0;
    if (g == 0) return -1;
    thetamax = TH[g - 1];
    jmax = JS[g - 1];
    phimax = g <= 2 ? PH[g - 1] / (float)nasym : PH[g - 1];
  }
  int nsang = 0;
  int nth = real_trip(thetamin, thetamax, dang);
  float theta1 = thetamin;
  for (int it = 0; it < nth; it++, theta1 = theta1 + dang) {
    float dphi;
    if (theta1 == 0.0f || theta1 == 180.0f) {
      dphi = 360.0f;
    } else {
      dphi = dang / sinf(theta1 * kPi / 180.0f);
      int k = (int)(359.99f / dphi) - 1;
      dphi = 360.0f / (float)(k < 1 ? 1 : k);
    }
    int nph = real_trip(0.0f, phimax, dphi);
    float phi1 = 0.0f;
    for (int ip = 0; ip < nph; ip++, phi1 = phi1 + dphi) {
      if (!((fabsf(theta1 - 90.0f) > 1.0f) || (phi1 < 179.0f))) continue;
      if (asym == 'T' || asym == 'O') {
        float pt = fabsf(fmodf(phi1, 90.0f));
        pt = fminf(pt, 90.0f - pt);
        thetamax1 = atan2f(1.0f, cosf(pt * kPi / 180.0f)) * 180.0f / kPi;
      } else if (asym == 'I') {
        float a25 = atan2f(2.0f, 1.0f + sqrtf(5.0f));
        float a23 = asinf(tanf(30.0f * kPi / 180.0f) * tanf(a25));
        float a253 = atan2f(tanf(30.0f * kPi / 180.0f) * cosf(a23), 1.0f);
        float pc = nasym == 2 ? fabsf(fmodf(phi1, 180.0f)) : fabsf(fmodf(phi1 + 90.0f, 180.0f));
        pc = fminf(pc, 180.0f - pc) * kPi / 180.0f;
        float d = sinf(a25) / (cosf(pc) + sinf(pc) / tanf(a253));
        float c = cosf(a25) - d * sinf(pc) * tanf(a23);
        thetamax1 = atan2f(d, c) * 180.0f / kPi;
      }
      if (theta1 > thetamax1) continue;
      for (int j = 1; j <= jmax; j++) {
        if (sang && nsang < cap) {
          float* s = sang + 3 * nsang;
          if (j == 1) {
            s[0] = phi1; s[1] = theta1; s[2] = 0.0f;
          } else {
            s[0] = phi1 + 180.0f; s[1] = 180.0f - theta1; s[2] = 0.0f + 180.0f;
          }
          for (int m = 0; m < 3; m++)
            if (s[m] > 180.0f) s[m] = s[m] - 360.0f;
          for (int m = 0; m < 3; m++) s[m] = s[m] / 180.0f * kPi;
        }
        nsang++;
      }
    }
  }
  return nsang;
}

}  // namespace fr

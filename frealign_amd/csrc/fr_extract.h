// fr_extract.h -- device-side trilinear gather from the Hermitian half-volume.
#pragma once
#include "fr_common.h"

namespace fr {

// AINTERPO3DS (ainterpo3ds.f:22-98, IPAD=1) evaluated at X3 = DM*(I,J): the
// plane coordinates are formed in the reference's FP32 operation order
// (X3 = DM(1)*I + DM(4)*J, no FMA), and the eight corners use the same
// integer rules, so the voxels read are bit-identical to the reference.
__device__ inline cf ainterpo_dev(int nsam, const float2* __restrict__ a3df, const float* d, int i,
                                  int j) {
  const int jc = nsam / 2 + 1;
  float x, y, z;
  int ls, lt, ms, mt, ns, nt;
  bool neg;
  {
#pragma clang fp contract(off)
    float fi = (float)i, fj = (float)j;
    float xx = d[0] * fi + d[3] * fj;
    float yy = d[1] * fi + d[4] * fj;
    float zz = d[2] * fi + d[5] * fj;
    neg = xx < 0.0f;
    x = neg ? -xx : xx;
    y = neg ? -yy : yy;
    z = neg ? -zz : zz;
  }
  ls = (int)x;
  lt = ls + 1;
  if (lt >= jc) lt = jc - 1;
  ms = (int)y;
  mt = ms + 1;
  if (y < 0.0f) { ms--; mt--; }
  if (ms <= -jc) ms = -jc + 1;
  if (mt >= jc) mt = jc - 1;
  ns = (int)z;
  nt = ns + 1;
  if (z < 0.0f) { ns--; nt--; }
  if (ns <= -jc) ns = -jc + 1;
  if (nt >= jc) nt = jc - 1;
  // Guards beyond the reference's clamps: they only act where the reference
  // itself would index outside the half-volume (|coordinate| >= JC, possible
  // for |AMAG| < 1 or non-finite input), so in-range results are unchanged.
  ls = min(max(ls, 0), jc - 1);
  lt = min(max(lt, 0), jc - 1);
  ms = min(max(ms, -jc + 1), jc - 1);
  mt = min(max(mt, -jc + 1), jc - 1);
  ns = min(max(ns, -jc + 1), jc - 1);
  nt = min(max(nt, -jc + 1), jc - 1);
  const float ax0 = 1.0f - fabsf(x - (float)ls), ax1 = 1.0f - fabsf(x - (float)lt);
  const float ay0 = 1.0f - fabsf(y - (float)ms), ay1 = 1.0f - fabsf(y - (float)mt);
  const float az0 = 1.0f - fabsf(z - (float)ns), az1 = 1.0f - fabsf(z - (float)nt);
  const size_t my0 = (size_t)jc * wrapn(ms, nsam), my1 = (size_t)jc * wrapn(mt, nsam);
  const size_t nz0 = (size_t)jc * nsam * wrapn(ns, nsam), nz1 = (size_t)jc * nsam * wrapn(nt, nsam);
  float2 c000 = a3df[ls + my0 + nz0], c100 = a3df[lt + my0 + nz0];
  float2 c010 = a3df[ls + my1 + nz0], c110 = a3df[lt + my1 + nz0];
  float2 c001 = a3df[ls + my0 + nz1], c101 = a3df[lt + my0 + nz1];
  float2 c011 = a3df[ls + my1 + nz1], c111 = a3df[lt + my1 + nz1];
  float w000 = ax0 * ay0 * az0, w100 = ax1 * ay0 * az0, w010 = ax0 * ay1 * az0, w110 = ax1 * ay1 * az0;
  float w001 = ax0 * ay0 * az1, w101 = ax1 * ay0 * az1, w011 = ax0 * ay1 * az1, w111 = ax1 * ay1 * az1;
  float re = c000.x * w000 + c100.x * w100 + c010.x * w010 + c110.x * w110 + c001.x * w001 +
             c101.x * w101 + c011.x * w011 + c111.x * w111;
  float im = c000.y * w000 + c100.y * w100 + c010.y * w010 + c110.y * w110 + c001.y * w001 +
             c101.y * w101 + c011.y * w011 + c111.y * w111;
  return neg ? cf{re, -im} : cf{re, im};
}

}  // namespace fr

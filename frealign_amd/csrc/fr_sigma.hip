// fr_sigma.hip -- SIGMA (sigma.f:8-221) for a batch of (particle, parameter)
// items: the noise sigma of the particle image and its likelihood LOGPC.
//
//   k_sigma_fill   PEXTRACT section and shifted particle in the band
//                  MAXR1^2 <= r^2 < MRCLAS^2 (zero elsewhere), both
//                  phase-flipped (phaseflip.f: x CTF/|CTF|, the Nyquist row
//                  twice, as the reference's J loop reaches it twice)
//   c2r            both images (the library's FFT)
//   k_sigma_stats  FP64 real-space sums in the mask disc and the closing
//                  formulas, one workgroup per item
#include <hip/hip_runtime.h>

#include "fr_common.h"
#include "fr_extract.h"
#include "fr_kernels.h"

namespace fr {

namespace {

constexpr int kThreads = 256;

__global__ void __launch_bounds__(kThreads)
    k_sigma_fill(int nsam, int n, const float2* __restrict__ a3df,
                 const SigmaItem* __restrict__ items, const float2* __restrict__ datc,
                 const float2* __restrict__ ctff, float* buf, size_t img) {
  const int jc = nsam / 2 + 1, nsamh = nsam / 2;
  const size_t ic = (size_t)nsam * jc;
  const size_t gid = blockIdx.x * (size_t)kThreads + threadIdx.x;
  if (gid >= ic * n) return;
  const int it = (int)(gid / ic);
  const int id = (int)(gid - (size_t)it * ic);
  const int l = id % jc, m = id / jc;
  const int mm = m >= jc ? m - nsam : m;
  const SigmaItem& s = items[it];
  const float2* cp = ctff + (size_t)s.pidx * 2 * ic;
  const float2 cr = cp[id], cl = cp[id + ic];
  const cf c = cf{cr.x + cl.x, cr.y - cl.y};
  cf sec = cf{0.f, 0.f}, par = cf{0.f, 0.f};
  const int itemp = l * l + mm * mm;
  if (itemp >= s.maxr12 && itemp < s.maxr22) {
    // pextract.f:30-51 for the storage row m (row NSAM/2 holds J = +NSAM/2)
    const int j = m >= jc ? m - nsam : m;
    if (l * l + j * j <= nsamh * nsamh) sec = cmul(ainterpo_dev(nsam, a3df, s.dm, l, j), c);
    float phase;
    {
#pragma clang fp contract(off)
      phase = -s.shx * (float)l - s.shy * (float)mm;
    }
    float sn, cs;
    sincosf(phase, &sn, &cs);
    const float2 d = datc[(size_t)s.pidx * ic + id];
    par = cmul(cf{cs, sn}, cf{d.x, d.y});
    const float amp = hypotf(c.re, c.im);
    if (amp != 0.0f) {
      const int times = (2 * m == nsam) ? 2 : 1;
      for (int q = 0; q < times; q++) {
        sec = cmul(sec, c);
        sec = cf{sec.re / amp, sec.im / amp};
        par = cmul(par, c);
        par = cf{par.re / amp, par.im / amp};
      }
    }
  }
  float2* o1 = reinterpret_cast<float2*>(buf + (size_t)it * img);
  float2* o2 = reinterpret_cast<float2*>(buf + (size_t)(n + it) * img);
  o1[id] = make_float2(sec.re, sec.im);
  o2[id] = make_float2(par.re, par.im);
}

__device__ inline double block_sum_d(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

__global__ void __launch_bounds__(kThreads)
    k_sigma_stats(int nsam, int n, const SigmaItem* __restrict__ items, const float* buf,
                  size_t img, float sigp, float* out) {
  __shared__ double red[kThreads / 64];
  const int it = blockIdx.x, nsamh = nsam / 2, pitch = nsam + 2;
  const SigmaItem& s = items[it];
  const float* r1p = buf + (size_t)it * img;
  const float* r2p = buf + (size_t)(n + it) * img;
  double cc = 0, s1 = 0, s2 = 0, a1 = 0, a2 = 0, s5 = 0, a5 = 0, cnt = 0;
  for (int e = threadIdx.x; e < nsam * nsam; e += kThreads) {
    const int L = e / nsam + 1, M = e % nsam + 1;  // sigma.f:223 loop indices
    float dx, dy, x2, y2, rr;
    {
#pragma clang fp contract(off)
      dx = (float)L - s.xm;
      if (L >= nsamh) dx = (float)L - s.xm - (float)nsam;
      dy = (float)M - s.ym;
      if (M >= nsamh) dy = (float)M - s.ym - (float)nsam;
      x2 = dx * dx;
      y2 = dy * dy;
      rr = x2 + y2;
    }
    const size_t id = (size_t)(L - 1) + (size_t)pitch * (M - 1);
    const float r1 = r1p[id], r2 = r2p[id];
    if (rr <= s.ri2f) {
      cc += (double)(r1 * r2);
      s1 += (double)(r1 * r1);
      s2 += (double)(r2 * r2);
      a1 += (double)r1;
      a2 += (double)r2;
      cnt += 1.0;
    }
    s5 += (double)(r2 * r2);
    a5 += (double)r2;
  }
  cc = block_sum_d(cc, red);
  s1 = block_sum_d(s1, red);
  s2 = block_sum_d(s2, red);
  a1 = block_sum_d(a1, red);
  a2 = block_sum_d(a2, red);
  s5 = block_sum_d(s5, red);
  a5 = block_sum_d(a5, red);
  cnt = block_sum_d(cnt, red);
  if (threadIdx.x != 0) return;
  // sigma.f:248-296
  const int n1 = (int)cnt;
  const int nsam2 = nsam * nsam;
  cc = cc - a1 * a2 / n1;
  s1 = s1 - a1 * a1 / n1;
  s2 = s2 - a2 * a2 / n1;
  s5 = s5 - a5 * a5 / nsam2;
  const double A = fabs(cc / s1);
  const double S2 = fabs(s2 - A * A * s1) / n1;
  const float sigma2 = (float)(S2 / nsam2 / nsam2);
  float s2old = (float)(s5 / nsam2);
  float sigma2old = s2old / (float)nsam2 / (float)nsam2;
  if (s.signset != 0.0f) {
    const float r = sqrtf(sigma2old) / s.signset;
    if (r <= 3.0f && r >= 0.33f) {
      s2old = s.signset * s.signset * (float)nsam2 * (float)nsam2;
      sigma2old = s.signset * s.signset;
    }
  }
  const float sigma2n = (float)((double)s2old / A / sqrt(s1 * s2));
  const int n1w = (int)((float)n1 * sigma2old / (sigp * sigp));
  const float logpc = -((float)n1w * sigma2 / sigma2old / 2.0f) -
                      (float)n1w * logf(2.0f * 3.1415926535897f * sigma2old) / 2.0f;
  out[3 * it + 0] = sigma2n;
  out[3 * it + 1] = sigma2;
  out[3 * it + 2] = logpc;
}

}  // namespace

int launch_sigma(Ctx& c, int nsam, int n, const float* a3df, const SigmaItem* items,
                 const float* datc, const float* ctff, float sigp, float* buf, float* out) {
  if (n <= 0) return 0;
  const size_t ic = (size_t)nsam * (nsam / 2 + 1);
  const size_t img = (size_t)nsam * (nsam + 2);
  const size_t tot = ic * n;
  hipLaunchKernelGGL(k_sigma_fill, dim3((unsigned)((tot + kThreads - 1) / kThreads)),
                     dim3(kThreads), 0, c.stream, nsam, n, reinterpret_cast<const float2*>(a3df),
                     items, reinterpret_cast<const float2*>(datc),
                     reinterpret_cast<const float2*>(ctff), buf, img);
  FR_HIP(hipGetLastError());
  if (int r = fft_c2r_2d(c, nsam, 2 * n, buf, img)) return r;
  hipLaunchKernelGGL(k_sigma_stats, dim3(n), dim3(kThreads), 0, c.stream, nsam, n, items, buf, img,
                     sigp, out);
  FR_HIP(hipGetLastError());
  return 0;
}

}  // namespace fr

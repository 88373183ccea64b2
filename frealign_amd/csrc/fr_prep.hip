// fr_prep.hip -- batched particle preparation.
//
// Replaces, for a batch of particles, the r2c FFT of the particle image in
// LMAIN (lmain.f:459) followed by CTFAPPLY (ctfapply.f:8-161): CTF table
// (ctf.f), shift phase ramp, c2r, scale, MASKCOS (maskcos.f), r2c,
// (-1)^(l+m) origin shift and PWEIGHT (pweight.f), including PWEIGHT's
// in-place double weighting of the Nyquist row and its JC+1 shell aliasing
// into bin 1's count (the reference calls PWEIGHT(NSAM,OUTC,OUTC,...)).
#include <hip/hip_runtime.h>

#include "fr_common.h"
#include "fr_kernels.h"
#include "fr_ctf.h"

namespace fr {

namespace {

constexpr int kThreads = 256;
constexpr float kPi = 3.1415926535897f;

// ctfapply.f:64-80: CTF table and shifted spectrum; one thread per half-plane point
__global__ void __launch_bounds__(kThreads)
    k_ctf_shift(int nsam, int nb, const float2* __restrict__ spec, const float* __restrict__ ppar,
                PrepGlobals g, float2* ctff, float2* outc, float2* datc) {
  const int jc = nsam / 2 + 1;
  const size_t ic = (size_t)nsam * jc;
  const size_t gid = blockIdx.x * (size_t)kThreads + threadIdx.x;
  if (gid >= ic * nb) return;
  const int b = (int)(gid / ic);
  const int id = (int)(gid - (size_t)b * ic);
  const int l = id % jc, m = id / jc;
  const int mm = m >= jc ? m - nsam : m;
  const float* pp = ppar + 8 * b;
  const float shx = pp[0], shy = pp[1], df1 = pp[2], df2 = pp[3], ang = pp[4], amagp = pp[5];
  const float thetatrp = g.thetatr / amagp;
  float phase = shx * (float)l + shy * (float)mm;
  float sn, cs;
  sincosf(phase, &sn, &cs);
  cf cr = ctf_dev(g, df1, df2, ang, thetatrp, l, mm);
  cf cl = ctf_dev(g, df1, df2, ang, thetatrp, -l, -mm);
  float2 s = spec[(size_t)b * ic + id];
  cf v = cmul(cf{s.x, s.y}, cf{cs, sn});
  outc[(size_t)b * ic + id] = make_float2(v.re, v.im);
  // DATC takes the (-1)^(l+m) origin shift of ctfapply.f:144-158 here (it
  // is not transformed again); OUTC takes it in PWEIGHT's weighting pass
  const float sg = ((l + mm) & 1) ? -1.0f : 1.0f;
  datc[(size_t)b * ic + id] = make_float2(v.re * sg, v.im * sg);
  ctff[(size_t)b * 2 * ic + id] = make_float2(cr.re, cr.im);
  ctff[(size_t)b * 2 * ic + ic + id] = make_float2(cl.re, cl.im);
}

__device__ inline float block_sum(float v, float* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  float t = 0.0f;
  for (int k = 0; k < (int)(blockDim.x >> 6); k++) t += red[k];
  return t;
}

}  // namespace

// ctfapply.f:107-122 and maskcos.f / maskcos2.f: scale, then cosine-edged
// real-space mask.  origin=0 centres the disc at NSAM/2+1 (MASKCOS), origin=1
// at the FFT origin (MASKCOS2).  One workgroup per image: a read-only pass for
// the mean of the scaled ring, then one pass that scales, masks and writes.
// Threads take fixed columns and step over rows (no per-pixel division).
__global__ void __launch_bounds__(kMaskThreads)
    k_scale_mask(int nsam, float* data, size_t img_stride, float scal, const float* amag_arr,
                 int amag_stride, float ri2, float ri3, float rih, float halfw, int origin) {
  __shared__ float red[kMaskThreads / 64];
  __shared__ int redi[kMaskThreads / 64];
  float* img = data + blockIdx.x * img_stride;
  const float amagp = amag_arr[(size_t)blockIdx.x * amag_stride];
  const float rihp = rih / amagp, ri2p = ri2 / amagp / amagp, ri3p = ri3 / amagp / amagp;
  const int nmid = origin ? 1 : nsam / 2 + 1, nsamh = nsam / 2, pitch = nsam + 2;
  const int nt = blockDim.x;
  const int cols = nsam < nt ? nsam : nt, rstep = nt / cols;
  const int c0 = threadIdx.x % cols, r0 = threadIdx.x / cols;
  const bool act = r0 < rstep;
  auto rad2 = [&](int i, int jx) {  // 1-based pixel (i, jx)
    int ii = i, jj = jx;
    if (origin) {
      if (ii > nsamh) ii -= nsam;
      if (jj > nsamh) jj -= nsam;
    }
    float d1 = (float)(ii - nmid), d2 = (float)(jj - nmid);
    float crad2;
    {
#pragma clang fp contract(off)
      crad2 = d2 * d2 + d1 * d1;
    }
    return crad2;
  };
  float sum = 0.0f;
  int isum = 0;
  if (act)
    for (int i0 = r0; i0 < nsam; i0 += rstep)
      for (int j0 = c0; j0 < nsam; j0 += cols) {
        float crad2 = rad2(i0 + 1, j0 + 1);
        if (crad2 > ri2p && crad2 <= ri3p) {
          sum += img[(size_t)i0 * pitch + j0] * scal;
          isum++;
        }
      }
  float tot = block_sum(sum, red);
  int ci = isum;
  for (int off = 32; off > 0; off >>= 1) ci += __shfl_down(ci, off, 64);
  if ((threadIdx.x & 63) == 0) redi[threadIdx.x >> 6] = ci;
  __syncthreads();
  int itot = 0;
  for (int k = 0; k < (int)(blockDim.x >> 6); k++) itot += redi[k];
  const float mean = itot ? tot / (float)itot : tot;
  if (act)
    for (int i0 = r0; i0 < nsam; i0 += rstep)
      for (int j0 = c0; j0 < nsam; j0 += cols) {
        float crad2 = rad2(i0 + 1, j0 + 1);
        float* px = img + (size_t)i0 * pitch + j0;
        float v = *px * scal;
        if (crad2 > ri2p && crad2 <= ri3p) {
          float edge = (1.0f + cosf(kPi * (sqrtf(crad2) - rihp) / halfw)) / 2.0f;
          v = v * edge + (1.0f - edge) * mean;
        } else if (crad2 > ri3p) {
          v = mean;
        }
        *px = v;
      }
}

namespace {

// ctfapply.f:144-158: PWEIGHT on OUTC, with OUTC's origin shift applied in
// the weighting pass (DATC's is applied in k_ctf_shift).
// One workgroup per particle.  Shell bins are 1-based as in pweight.f.  The
// shell sums run in a fixed order: wave w takes the 64-point runs w, w + 16,
// ... of the half-plane (row-major), each run reduces its points per shell
// with a segmented scan (a row's shells never decrease along L, so a shell's
// points in a run are contiguous), the run's last point of each shell adds
// the run's sum into the wave's own bins, and the waves' bins are added in
// wave order.
constexpr int kPwWaves = kMaskThreads / 64;

// LDS: kPwWaves wave bin sets and the combined bins, 2 JC + 2 floats each
size_t pweight_lds(int nsam) { return sizeof(float) * (kPwWaves + 1) * (size_t)(nsam + 4); }

__global__ void __launch_bounds__(kMaskThreads)
    k_origin_pweight(int nsam, const float* __restrict__ pssnr, int ictf,
                     const float2* __restrict__ ctff, float2* outc) {
  extern __shared__ float pw_lds[];
  const int jc = nsam / 2 + 1;
  const int ic = nsam * jc;
  const int nbin = 2 * jc + 2;
  float* rbuf = pw_lds + kPwWaves * nbin;
  float2* oc = outc + blockIdx.x * (size_t)ic;
  const float2* cf_ = ctff + blockIdx.x * 2 * (size_t)ic;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  float* wb = pw_lds + wave * nbin;
  for (int k = lane; k < nbin; k += 64) wb[k] = 0.0f;
  // segment keys: row * (jc + 3) + shell, shell jc + 2 marking points outside
  // the disc (they close each row), so keys never decrease within a run
  const int kstride = jc + 3;
  // the shell sums of pweight.f:31-42 over J=-JC+1..JC-1 (the Nyquist row
  // is addressed by both J=-(JC-1) and J=+(JC-1)); |X| is the same before
  // and after the origin shift, which the weighting pass applies
  for (int r0 = wave * 64; r0 < ic; r0 += kMaskThreads) {
    const int e = r0 + lane;
    int key = 0x7fffffff;  // past the end
    float v = 0.0f, cnt = 0.0f;
    if (e < ic) {
      const int m = e / jc, l = e - m * jc;
      int mm = m >= jc ? m - nsam : m;
      const float2 o = oc[e];
      int itemp = l * l + mm * mm;
      int ibin = jc + 2;
      if (itemp < jc * jc) {
        ibin = (int)(sqrtf((float)itemp) + 0.5f) + 1;
        float h = hypotf(o.x, o.y);
        float visits = (2 * m == nsam) ? 2.0f : 1.0f;
        v = visits * (h * h);
        cnt = visits;
      }
      key = m * kstride + ibin;
    }
    for (int off = 1; off < 64; off <<= 1) {
      float vo = __shfl_up(v, off, 64), co = __shfl_up(cnt, off, 64);
      int ko = __shfl_up(key, off, 64);
      if (lane >= off && ko == key) {
        v += vo;
        cnt += co;
      }
    }
    const int kn = __shfl_down(key, 1, 64);
    const bool tail = e < ic && (lane == 63 || kn != key);
    const int row = key / kstride, ibin = key - row * kstride;
    // one row of the run at a time: the shells closing in one row are
    // distinct, so no two lanes touch one address in a statement (bin JC+1's
    // sum aliases bin 1's count, but bin 1 lies only in row 0 and bin JC+1
    // never does); the empty asm keeps the compiler from moving LDS accesses
    // across statements, since lanes may alias one another
    const int mlo = r0 / jc, mhi = min(ic - 1, r0 + 63) / jc;
    for (int mr = mlo; mr <= mhi; mr++) {
      const bool w = tail && row == mr && ibin <= jc + 1;
      if (w) wb[ibin] += v;
      asm volatile("" ::: "memory");
      if (w) wb[ibin + jc] += cnt;
      asm volatile("" ::: "memory");
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nbin; k += kMaskThreads) {
    float t = 0.0f;
    for (int w = 0; w < kPwWaves; w++) t += pw_lds[w * nbin + k];
    rbuf[k] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i <= jc; i++) {
      if (rbuf[i + jc] != 0.0f) rbuf[i] = sqrtf(rbuf[i] / rbuf[i + jc]);
      if (rbuf[i] == 0.0f) rbuf[i] = 1.0f;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < ic; e += kMaskThreads) {
    const int m = e / jc, l = e - m * jc;
    int mm = m >= jc ? m - nsam : m;
    int itemp = l * l + mm * mm;
    float2 o = oc[e];
    const float sg = ((l + mm) & 1) ? -1.0f : 1.0f;  // origin shift, ctfapply.f:144-158
    o.x *= sg;
    o.y *= sg;
    if (itemp < jc * jc) {
      int ibin = (int)(sqrtf((float)itemp) + 0.5f) + 1;
      float snr;
      if (ictf == 1) {
        float2 cr = cf_[e], cl = cf_[e + ic];
        float re = cr.x + cl.x, im = cr.y - cl.y;
        float h = hypotf(re, im);
        snr = pssnr[ibin - 1] * (h * h);
      } else {
        snr = pssnr[ibin - 1];
      }
      float wgt = snr / (snr + 1.0f) / rbuf[ibin];
      o.x *= wgt;
      o.y *= wgt;
      if (2 * m == nsam) { o.x *= wgt; o.y *= wgt; }
    } else {
      o = make_float2(0.0f, 0.0f);
    }
    oc[e] = o;
  }
}

}  // namespace

int launch_prep(Ctx& c, int nsam, int nb, float* data, const float* ppar, const float* pssnr,
                const PrepGlobals& g, float* ctff, float* outc, float* datc) {
  if (nb <= 0) return 0;
  // lmain.f:459: r2c of the particle image, in place
  if (int r = fft_r2c_2d(c, nsam, nb, data, (size_t)nsam * (nsam + 2))) return r;
  return launch_ctfapply(c, nsam, nb, data, ppar, pssnr, g, ctff, outc, datc);
}

int launch_ctfapply(Ctx& c, int nsam, int nb, const float* spec, const float* ppar,
                    const float* pssnr, const PrepGlobals& g, float* ctff, float* outc,
                    float* datc) {
  if (nb <= 0) return 0;
  FR_CHECK(nsam / 2 + 1 <= 520, "box too large for PWEIGHT bins");
  const int jc = nsam / 2 + 1;
  const size_t ic = (size_t)nsam * jc;
  const size_t img = (size_t)nsam * (nsam + 2);
  size_t tot = ic * nb;
  hipLaunchKernelGGL(k_ctf_shift, dim3((unsigned)((tot + kThreads - 1) / kThreads)), dim3(kThreads),
                     0, c.stream, nsam, nb, reinterpret_cast<const float2*>(spec), ppar, g,
                     reinterpret_cast<float2*>(ctff), reinterpret_cast<float2*>(outc),
                     reinterpret_cast<float2*>(datc));
  FR_HIP(hipGetLastError());
  if (int r = fft_c2r_2d(c, nsam, nb, outc, img)) return r;
  // ctfapply.f:107-122; per-particle AMAGP is column 5 of ppar
  hipLaunchKernelGGL(k_scale_mask, dim3(nb), dim3(kMaskThreads), 0, c.stream, nsam, outc, img,
                     1.0f / (float)nsam / (float)nsam, ppar + 5, 8, g.ri2, g.ri3, g.rih, g.halfw,
                     0);
  FR_HIP(hipGetLastError());
  if (int r = fft_r2c_2d(c, nsam, nb, outc, img)) return r;
  const size_t lds = pweight_lds(nsam);
  FR_HIP(hipFuncSetAttribute((const void*)k_origin_pweight,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_origin_pweight, dim3(nb), dim3(kMaskThreads), lds, c.stream, nsam, pssnr, g.ictf,
                     reinterpret_cast<const float2*>(ctff), reinterpret_cast<float2*>(outc));
  FR_HIP(hipGetLastError());
  return 0;
}

}  // namespace fr

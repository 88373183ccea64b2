// fr_extract.hip -- central-section extraction and the CC3M score.
//
// ainterpo_dev: AINTERPO3DS (ainterpo3ds.f:8-99), trilinear gather from the
//   Hermitian half-volume, with the reference's integer rules (INT() plus the
//   negative-coordinate correction, the JC-1 clamps that double-weight the
//   x = NSAM/2 plane) reproduced exactly.
// k_pextract: PEXTRACT (pextract.f:8-101) for a batch of (particle, DM).
// launch_cc3m: CC3M (cc3m.f:8-169) for a batch of (particle, DM, shift)
//   items: annulus extraction x CTF x 1/NSAM^2 -> c2r -> MASKCOS2 -> r2c ->
//   per-shell sums -> FP64 shell combination.
#include <hip/hip_runtime.h>

#include "fr_common.h"
#include "fr_extract.h"
#include "fr_kernels.h"

namespace fr {

namespace {
constexpr int kThreads = 256;
}

// pextract.f:30-51 / cc3m.f:74-112: one point of a central section
__global__ void __launch_bounds__(kThreads)
    k_pextract(int nsam, int n, const float2* __restrict__ a3df, const float* __restrict__ dm,
               const int* __restrict__ pidx, const float2* __restrict__ ctff, float2* outc) {
  const int jc = nsam / 2 + 1, nsamh = nsam / 2;
  const size_t ic = (size_t)nsam * jc;
  const size_t gid = blockIdx.x * (size_t)kThreads + threadIdx.x;
  if (gid >= ic * n) return;
  const int it = (int)(gid / ic);
  const int id = (int)(gid - (size_t)it * ic);
  const int i = id % jc, m = id / jc;
  // storage row m = NSAM/2 is written last by J=+NSAM/2 in the reference loop
  const int j = m >= jc ? m - nsam : m;
  float2 out = make_float2(0.f, 0.f);
  if (i * i + j * j <= nsamh * nsamh) {
    const float* d = dm + 9 * (size_t)it;
    const int p = pidx ? pidx[it] : it;
    const float2* cp = ctff + (size_t)p * 2 * ic;
    cf a = ainterpo_dev(nsam, a3df, d, i, j);
    float2 cr = cp[id], cl = cp[id + ic];
    a = cmul(a, cf{cr.x + cl.x, cr.y - cl.y});
    out = make_float2(a.re, a.im);
  }
  outc[gid] = out;
}

int launch_pextract(Ctx& c, int nsam, int n, const float* a3df, const float* dm, const int* pidx,
                    const float* ctff, float* outc) {
  if (n <= 0) return 0;
  size_t tot = (size_t)nsam * (nsam / 2 + 1) * n;
  hipLaunchKernelGGL(k_pextract, dim3((unsigned)((tot + kThreads - 1) / kThreads)), dim3(kThreads),
                     0, c.stream, nsam, n, reinterpret_cast<const float2*>(a3df), dm, pidx,
                     reinterpret_cast<const float2*>(ctff), reinterpret_cast<float2*>(outc));
  FR_HIP(hipGetLastError());
  return 0;
}

namespace {

// cc3m.f:74-112: the CTF-multiplied annulus, scaled by 1/NSAM^2, into the
// real-FFT buffer of each item (zero elsewhere).
__global__ void __launch_bounds__(kThreads)
    k_cc3m_fill(int nsam, int n, int maxr12, int maxr22, const float2* __restrict__ a3df,
                const float* __restrict__ dm, const int* __restrict__ pidx,
                const float2* __restrict__ ctff, float* work) {
  const int jc = nsam / 2 + 1;
  const size_t ic = (size_t)nsam * jc;
  const size_t gid = blockIdx.x * (size_t)kThreads + threadIdx.x;
  if (gid >= ic * n) return;
  const int it = (int)(gid / ic);
  const int id = (int)(gid - (size_t)it * ic);
  const int i = id % jc, m = id / jc;
  const int j = m >= jc ? m - nsam : m;
  const int itemp = i * i + j * j;
  float2 out = make_float2(0.f, 0.f);
  if (itemp >= maxr12 && itemp < maxr22) {
    const float* d = dm + 9 * (size_t)it;
    const float2* cp = ctff + (size_t)pidx[it] * 2 * ic;
    cf a = ainterpo_dev(nsam, a3df, d, i, j);
    float2 cr = cp[id], cl = cp[id + ic];
    a = cmul(a, cf{cr.x + cl.x, cr.y - cl.y});
    const float scal = 1.0f / (float)nsam / (float)nsam;
    out = make_float2(a.re * scal, a.im * scal);
  }
  reinterpret_cast<float2*>(work)[gid] = out;
}

// cc3m.f:117-166: shifted particle x conj(masked projection), per-shell sums
// in FP32, then the FP64 combination.  One workgroup per item.
__global__ void __launch_bounds__(kThreads)
    k_cc3m_score(int nsam, int maxr12, int maxr22, float rbfact, const float2* __restrict__ spec,
                 const int* __restrict__ pidx, const float* __restrict__ shift,
                 const float* __restrict__ fsct, const float* __restrict__ work, float* score) {
  __shared__ float bins[4 * 260];
  const int jc = nsam / 2 + 1;
  const size_t ic = (size_t)nsam * jc;
  const int it = blockIdx.x;
  const int nb = nsam / 2 + 2;
  for (int k = threadIdx.x; k < 4 * nb; k += blockDim.x) bins[k] = 0.0f;
  __syncthreads();
  const float2* sp = spec + (size_t)pidx[it] * ic;
  const float2* pw = reinterpret_cast<const float2*>(work) + (size_t)it * ic;
  const float shx = shift[2 * it], shy = shift[2 * it + 1];
  // rows of the half-plane that can intersect the annulus
  const int r2 = maxr22 > 0 ? maxr22 : 0;
  for (int e = threadIdx.x; e < (int)ic; e += blockDim.x) {
    int i = e % jc, m = e / jc;
    int j = m >= jc ? m - nsam : m;
    int itemp = i * i + j * j;
    if (itemp < maxr12 || itemp >= r2) continue;
    float sn, cs;
    sincosf(shx * (float)i + shy * (float)j, &sn, &cs);
    float2 s = sp[e];
    cf samp = cmul(cf{s.x, s.y}, cf{cs, sn});
    float2 a = pw[e];
    float q = samp.re * a.x + samp.im * a.y;  // REAL(SAMP*CONJG(AINPO))
    int ibin = (int)(sqrtf((float)itemp) + 0.5f);  // 0-based shell
    atomicAdd(&bins[4 * ibin + 0], q);
    atomicAdd(&bins[4 * ibin + 1], samp.re * samp.re + samp.im * samp.im);
    atomicAdd(&bins[4 * ibin + 2], a.x * a.x + a.y * a.y);
    atomicAdd(&bins[4 * ibin + 3], 1.0f);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (int i = 2; i <= nsam / 2; i++) {
      const float* b = bins + 4 * (i - 1);
      if (b[3] != 0.0f) {
        float r = b[0] / sqrtf(b[2]);
        if (fsct[i - 1] < 0.95f) r = fabsf(r);
        float wgt = expf(-rbfact * (float)((i - 1) * (i - 1)));
        s3 += (double)(r * wgt);
        s2 += (double)(wgt * wgt * b[1]);
        s1 += 1.0;
      }
    }
    if (s1 * s2 != 0.0) s3 = s3 / sqrt(s1 * s2);
    score[it] = (float)s3;
  }
}

}  // namespace

int launch_cc3m(Ctx& c, const Cc3mArgs& a, const float* a3df, const float* spec, const float* ctff,
                const int* pidx, const float* dm, const float* shift, const float* amag,
                const float* fsct, float* work, float* score) {
  if (a.n <= 0) return 0;
  const int nsam = a.nsam;
  FR_CHECK(nsam / 2 + 2 <= 260, "box too large for CC3M bins");
  FR_CHECK(a.maxr2 < nsam / 2, "CC3M needs MAXR2 < NSAM/2 (got %d)", a.maxr2);
  const size_t ic = (size_t)nsam * (nsam / 2 + 1);
  const size_t img = (size_t)nsam * (nsam + 2);
  size_t tot = ic * a.n;
  hipLaunchKernelGGL(k_cc3m_fill, dim3((unsigned)((tot + kThreads - 1) / kThreads)), dim3(kThreads),
                     0, c.stream, nsam, a.n, a.maxr1 * a.maxr1, a.maxr2 * a.maxr2,
                     reinterpret_cast<const float2*>(a3df), dm, pidx,
                     reinterpret_cast<const float2*>(ctff), work);
  FR_HIP(hipGetLastError());
  if (int r = fft_c2r_2d(c, nsam, a.n, work, img)) return r;
  // cc3m.f:114: MASKCOS2 with AMAG of each item (scal 1: already scaled)
  hipLaunchKernelGGL(k_scale_mask, dim3(a.n), dim3(kThreads), 0, c.stream, nsam, work, img, 1.0f,
                     amag, 1, a.ri2, a.ri3, a.rih, a.halfw, 1);
  FR_HIP(hipGetLastError());
  if (int r = fft_r2c_2d(c, nsam, a.n, work, img)) return r;
  hipLaunchKernelGGL(k_cc3m_score, dim3(a.n), dim3(kThreads), 0, c.stream, nsam,
                     a.maxr1 * a.maxr1, a.maxr2 * a.maxr2, a.rbfact,
                     reinterpret_cast<const float2*>(spec), pidx, shift, fsct, work, score);
  FR_HIP(hipGetLastError());
  return 0;
}

}  // namespace fr

// fr_final.hip -- the volume finalisation after the particle loop: the sum of
// the private insertion copies (COMBINE_ARRAYS, combine_arrays.f:33-52), the
// final map filter (SHIFTVOL, shiftvol.f:8-77) and the real-space
// interpolation correction (CORRECT3D, correct3d.f:19-47 with BOXFT_LUT,
// boxft_lut.f:26-37).  All three are one pass over a half-volume (or the
// padded real volume) with no reuse: HBM-bound elementwise kernels, one
// thread per voxel, grid-strided.
#include <hip/hip_runtime.h>

#include <math.h>

#include "fr_common.h"
#include "fr_kernels.h"

namespace fr {

namespace {

constexpr int kThreads = 256;
constexpr int kMaxBlocks = 8192;

int blocks_for(size_t n) {
  size_t b = (n + kThreads - 1) / kThreads;
  return (int)(b > kMaxBlocks ? kMaxBlocks : b);
}

// combine_arrays.f:33-46: copy 1 += copies 2..ncopy, in copy order
__global__ void __launch_bounds__(kThreads)
    k_combine(size_t nv, size_t stride, int ncopy, float2* a, float* s) {
  for (size_t i = blockIdx.x * (size_t)kThreads + threadIdx.x; i < nv;
       i += (size_t)gridDim.x * kThreads) {
    float2 av = a[i];
    float sv = s[i];
    for (int is = 1; is < ncopy; is++) {
      const float2 o = a[is * stride + i];
      av.x = av.x + o.x;
      av.y = av.y + o.y;
      sv = sv + s[is * stride + i];
    }
    a[i] = av;
    s[i] = sv;
  }
}

// shiftvol.f:24-74 (NNSTAT = 0).  The divisions by S3DF + STD*IREDUN (a
// DOUBLE PRECISION value) run per component in double, as COMPLEX / DOUBLE
// does in the reference.
__global__ void __launch_bounds__(kThreads)
    k_shiftvol(int nsam, double stdi, int irrec2, int ifsc, float fpart, float fmask,
               float2* __restrict__ a, const float2* __restrict__ d, float* __restrict__ s,
               const float* __restrict__ v, const float* __restrict__ pssnr,
               const float* __restrict__ fsc) {
  const int jc = nsam / 2 + 1;
  const size_t nv = (size_t)jc * nsam * nsam;
  for (size_t id = blockIdx.x * (size_t)kThreads + threadIdx.x; id < nv;
       id += (size_t)gridDim.x * kThreads) {
    const int ll = (int)(id % jc);
    const size_t r = id / jc;
    int mm = (int)(r % nsam), nn = (int)(r / nsam);
    if (mm >= jc) mm -= nsam;
    if (nn >= jc) nn -= nsam;
    const int jrad2 = ll * ll + mm * mm + nn * nn;
    const float sg = ((ll + mm + nn) & 1) ? -1.0f : 1.0f;
    float2 av = a[id];
    float sv = s[id];
    if (ifsc == 0) {
      const float2 dv = d[id];
      av.x = av.x + dv.x;
      av.y = av.y + dv.y;
      sv = sv + v[id];
      s[id] = sv;
    }
    float2 out = make_float2(0.f, 0.f);
    if (jrad2 <= irrec2) {
      const double den = (double)sv + stdi;
      if (fpart == 0.0f) {
        out.x = (float)((double)(av.x * sg) / den);
        out.y = (float)((double)(av.y * sg) / den);
      } else {
        const int i = (int)(sqrtf((float)jrad2) + 0.5f);  // shell I - 1
        if (pssnr[i] != 0.0f && sv != 0.0f) {
          const float af = fabsf(fsc[i]);
          const float a1 = af < 1.0f ? 2.0f * af / (1.0f - af) : 1000.0f;
          const float w = 1.0f + fpart / a1 / fmask;
          out.x = (float)((double)(av.x * sg) / den / (double)w);
          out.y = (float)((double)(av.y * sg) / den / (double)w);
        }
      }
    }
    a[id] = out;
  }
}

struct SincLut {
  float v[kSincLut];
};

// correct3d.f:26-46: divide by BOXFT_LUT (INTERP = 0) or its square
__global__ void __launch_bounds__(kThreads)
    k_correct3d(int nsam, float scal, int interp, SincLut lut, float* a3dv) {
  const int nmid = nsam / 2 + 1;
  const size_t nv = (size_t)nsam * nsam * nsam;
  for (size_t e = blockIdx.x * (size_t)kThreads + threadIdx.x; e < nv;
       e += (size_t)gridDim.x * kThreads) {
    const int i = (int)(e % nsam);
    const size_t r = e / nsam;  // J - 1 + NSAM (K - 1)
    const int j = (int)(r % nsam), k = (int)(r / nsam);
    const float ax[3] = {(float)(i + 1 - nmid) * scal, (float)(j + 1 - nmid) * scal,
                         (float)(k + 1 - nmid) * scal};
    float prod = 1.0f;
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const float x = fabsf(ax[q]);
      const float sn = x < 2.0e-2f ? 1.0f : lut.v[(int)lroundf(180.0f * x) - 1];
      prod = prod * sn;
    }
    const size_t id = (size_t)i + (size_t)(nsam + 2) * r;
    const float div = interp == 0 ? prod : prod * prod;
    a3dv[id] = a3dv[id] / div;
  }
}

}  // namespace

int launch_combine(Ctx& c, int nsam, size_t stride, int ncopy, float* a, float* s) {
  if (ncopy <= 1) return 0;
  const size_t nv = (size_t)(nsam / 2 + 1) * nsam * nsam;
  hipLaunchKernelGGL(k_combine, dim3(blocks_for(nv)), dim3(kThreads), 0, c.stream, nv, stride,
                     ncopy, reinterpret_cast<float2*>(a), s);
  FR_HIP(hipGetLastError());
  return 0;
}

int launch_shiftvol(Ctx& c, int nsam, double stdi, int irrec2, int ifsc, float fpart, float fmask,
                    float* a, const float* d, float* s, const float* v, const float* pssnr,
                    const float* fsc) {
  const size_t nv = (size_t)(nsam / 2 + 1) * nsam * nsam;
  hipLaunchKernelGGL(k_shiftvol, dim3(blocks_for(nv)), dim3(kThreads), 0, c.stream, nsam, stdi,
                     irrec2, ifsc, fpart, fmask, reinterpret_cast<float2*>(a),
                     reinterpret_cast<const float2*>(d), s, v, pssnr, fsc);
  FR_HIP(hipGetLastError());
  return 0;
}

int launch_correct3d(Ctx& c, int nsam, float scal, int interp, const float* sinclut,
                     float* a3dv) {
  SincLut lut;
  for (int k = 0; k < kSincLut; k++) lut.v[k] = sinclut[k];
  const size_t nv = (size_t)nsam * nsam * nsam;
  hipLaunchKernelGGL(k_correct3d, dim3(blocks_for(nv)), dim3(kThreads), 0, c.stream, nsam, scal,
                     interp, lut, a3dv);
  FR_HIP(hipGetLastError());
  return 0;
}

}  // namespace fr

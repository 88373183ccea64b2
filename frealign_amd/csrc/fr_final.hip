// fr_final.hip -- the volume finalisation after the particle loop: the sum of
// the private insertion copies (COMBINE_ARRAYS, combine_arrays.f:33-52), the
// final map filter (SHIFTVOL, shiftvol.f:8-77), the real-space interpolation
// correction (CORRECT3D, correct3d.f:19-47 with BOXFT_LUT, boxft_lut.f:26-37)
// and the spherical cosine-edge mask (MASK3D, mask3d.f:8-142).  Each is one
// pass (MASK3D: two or four) over a half-volume or the padded real volume
// with no reuse: HBM-bound kernels, one thread per voxel, grid-strided.  The
// FP64 sums of MASK3D are block partials reduced in a fixed order.
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

// mask3d.f: the voxel's (X, Y, Z) - XC squared radius, in the reference's
// FP32 order (no contraction)
__device__ inline float mask3d_rad2(int nsam, size_t e, float xc) {
#pragma clang fp contract(off)
  const int l = (int)(e % nsam);
  const size_t r = e / nsam;
  const int m = (int)(r % nsam), n = (int)(r / nsam);
  const float xl = (float)(l + 1) - xc, ym = (float)(m + 1) - xc, zn = (float)(n + 1) - xc;
  return xl * xl + ym * ym + zn * zn;
}

// block sums in a fixed order (lane butterflies, then waves in order)
template <int K>
__device__ inline void block_sums(double (&v)[K], double* out) {
  __shared__ double red[kThreads / 64][K];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; k++)
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < K; k++) red[wave][k] = v[k];
  __syncthreads();
  if (threadIdx.x < K) {
    double t = 0.0;
    for (int w = 0; w < kThreads / 64; w++) t += red[w][threadIdx.x];
    out[threadIdx.x] = t;
  }
}

// mask3d.f:32-48 (and 83-95 with stats = false): the edge shell's sum and
// count, and the RIC^2 <= r^2 <= RIO^2 sums of A and A^2 (A^2 in FP32)
__global__ void __launch_bounds__(kThreads)
    k_mask3d_sums(int nsam, const float* __restrict__ a, float xc, float ri2, float ri3, float ri4,
                  float ri5, int stats, double* part) {
  const size_t nv = (size_t)nsam * nsam * nsam;
  double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (size_t e = blockIdx.x * (size_t)kThreads + threadIdx.x; e < nv;
       e += (size_t)gridDim.x * kThreads) {
    const float rad2 = mask3d_rad2(nsam, e, xc);
    const float x = a[e % nsam + (size_t)(nsam + 2) * (e / nsam)];
    if (rad2 > ri2 && rad2 < ri3) {
      v[0] += (double)x;
      v[1] += 1.0;
    }
    if (stats && rad2 <= ri4 && rad2 >= ri5) {
      v[2] += (double)x;
      v[3] += (double)(x * x);
      v[4] += 1.0;
    }
  }
  block_sums<5>(v, part + 5 * (size_t)blockIdx.x);
}

// mask3d.f:50-63 (outer, inner = 0) and 96-110 (inner = 1): the cosine edge,
// A = A*EDGE + BG*(1-EDGE) with BG in double; FMASK's SUM3 terms per block
__global__ void __launch_bounds__(kThreads)
    k_mask3d_apply(int nsam, float* a, float xc, float ri2, float ri3, float rih, float rii,
                   float halfw, double bg, int inner, double* part) {
  const float pi = 3.1415926535897f;
  const size_t nv = (size_t)nsam * nsam * nsam;
  double v[1] = {0.0};
  for (size_t e = blockIdx.x * (size_t)kThreads + threadIdx.x; e < nv;
       e += (size_t)gridDim.x * kThreads) {
#pragma clang fp contract(off)
    const float rad2 = mask3d_rad2(nsam, e, xc);
    const size_t id = e % nsam + (size_t)(nsam + 2) * (e / nsam);
    float edge = 1.0f;
    const bool in = inner ? rad2 < ri3 : rad2 > ri2;
    if (in) {
      const float rad = sqrtf(rad2);
      if (rad > rih && rad < rii)
        edge = (1.0f + cosf(pi * (inner ? rii - rad : rad - rih) / halfw)) / 2.0f;
      else
        edge = 0.0f;
      const double t = bg * (double)(1.0f - edge);
      a[id] = (float)((double)(a[id] * edge) + t);
    }
    if (!inner) v[0] += (double)(edge * edge);
    else if (edge != 1.0f) v[0] += -1.0 + (double)(edge * edge);
  }
  block_sums<1>(v, part + blockIdx.x);
}

// the partial sums of nblk blocks (k values each), added in block order
__global__ void __launch_bounds__(kThreads)
    k_sum_parts(int nblk, int k, const double* __restrict__ part, double* out) {
  double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nblk; b += kThreads)
    for (int q = 0; q < k; q++) v[q] += part[(size_t)b * k + q];
  block_sums<5>(v, out);
}

}  // namespace

int launch_mask3d_sums(Ctx& c, int nsam, const float* a, float xc, float ri2, float ri3, float ri4,
                       float ri5, int stats, double* part, double* out) {
  hipLaunchKernelGGL(k_mask3d_sums, dim3(kMask3dBlocks), dim3(kThreads), 0, c.stream, nsam, a, xc,
                     ri2, ri3, ri4, ri5, stats, part);
  FR_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kThreads), 0, c.stream, kMask3dBlocks, 5, part,
                     out);
  FR_HIP(hipGetLastError());
  return 0;
}

int launch_mask3d_apply(Ctx& c, int nsam, float* a, float xc, float ri2, float ri3, float rih,
                        float rii, float halfw, double bg, int inner, double* part, double* out) {
  hipLaunchKernelGGL(k_mask3d_apply, dim3(kMask3dBlocks), dim3(kThreads), 0, c.stream, nsam, a, xc,
                     ri2, ri3, rih, rii, halfw, bg, inner, part);
  FR_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kThreads), 0, c.stream, kMask3dBlocks, 1, part,
                     out);
  FR_HIP(hipGetLastError());
  return 0;
}

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

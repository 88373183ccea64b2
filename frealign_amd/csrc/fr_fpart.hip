// fr_fpart.hip -- FIND_FPART (find_fpart.f:8-150) with its MW = 0 search:
// FPART_CC (fpart_cc.f:8-94) and VOLMEASURE (volmeasure.f:8-69).
//
// The particle fraction FPART scales the PSSNR of the Wiener filter.  With a
// molecular mass it is a closed form; without one (MW = 0) it is searched:
// the reconstruction B = A/(S + STD IREDUN) and a low-passed sum C of both
// half-maps go to real space, C is masked (MASK3D) and thresholded to pick
// the particle's voxels, and FPART moves in steps of 0.2, 0.067, ... until
// the correlation over those voxels of B with D/(V + FPART/PSSNR) (FPART_CC)
// stops rising.  Every volume pass runs here on the device (the 3D FFTs and
// MASK3D are the library's); the step logic, scalars and reductions'
// combinations run on the host, in the reference's order.  Reductions are
// per-block partial sums (double) combined in block order: deterministic.
#include <hip/hip_runtime.h>

#include <math.h>

#include <vector>

#include "../../include/frealign_hip.h"
#include "fr_common.h"
#include "fr_ctx.h"
#include "fr_kernels.h"

namespace fr {
namespace {

constexpr int kT = 256;
constexpr int kBlocks = 1024;   // partial-sum blocks of the reductions
constexpr int kThresh = 16;     // find_fpart.f:71: K = 0 .. INT(T/0.1) = 15

// The half-volume voxel id = l + JC (m + NSAM n) and its Miller indices.
__device__ inline void hv_index(long long id, int nsam, int& ll, int& mm, int& nn) {
  const int jc = nsam / 2 + 1;
  const int l = (int)(id % jc);
  const long long r = id / jc;
  const int m = (int)(r % nsam), n = (int)(r / nsam);
  ll = l;
  mm = m >= jc ? m - nsam : m;
  nn = n >= jc ? n - nsam : n;
}

// find_fpart.f:46-61: B = A s / (S + STD IREDUN) / N^3 and
// C = (A + D) s / ((S + V) + (STD + VTD) IREDUN) / N^3 x EXP(-r^2 / FRAD2),
// s = (-1)^(l+m+n); the complex / DOUBLE divisions per component in double
__global__ void __launch_bounds__(kT)
    k_fpart_vols(int nsam, const float2* __restrict__ a, const float2* __restrict__ d,
                 const float* __restrict__ s3, const float* __restrict__ v3, double stdi,
                 double svi, float frad2, float2* b, float2* c) {
  const long long nv = (long long)nsam * nsam * (nsam / 2 + 1);
  const long long id = blockIdx.x * (long long)kT + threadIdx.x;
  if (id >= nv) return;
  int ll, mm, nn;
  hv_index(id, nsam, ll, mm, nn);
  const float sg = ((ll + mm + nn) % 2 != 0) ? -1.0f : 1.0f;
  const double n3 = (double)nsam * nsam * nsam;
  const float2 av = a[id], dv = d[id];
  const double den1 = (double)s3[id] + stdi;
  b[id] = make_float2((float)((double)(av.x * sg) / den1 / n3),
                      (float)((double)(av.y * sg) / den1 / n3));
  const double den2 = (double)(s3[id] + v3[id]) + svi;
  float2 cv = make_float2((float)((double)((av.x + dv.x) * sg) / den2 / n3),
                          (float)((double)((av.y + dv.y) * sg) / den2 / n3));
  const float g = expf(-(float)(ll * ll + mm * mm + nn * nn) / frad2);
  c[id] = make_float2(cv.x * g, cv.y * g);
}

// fpart_cc.f:31-47: C = D s / (V + FPART / SSNR(I)) / N^3 where SSNR(I) != 0
// and FSC(I) > 0.5 (I the shell INT(SQRT(r^2) + 0.5) + 1), else 0; FP32
__global__ void __launch_bounds__(kT)
    k_fpart_cc_vol(int nsam, const float2* __restrict__ d, const float* __restrict__ v3,
                   float fpart, const float* __restrict__ ssnr, const float* __restrict__ fsc,
                   int nshell, float2* c) {
  const long long nv = (long long)nsam * nsam * (nsam / 2 + 1);
  const long long id = blockIdx.x * (long long)kT + threadIdx.x;
  if (id >= nv) return;
  int ll, mm, nn;
  hv_index(id, nsam, ll, mm, nn);
  const float sg = ((ll + mm + nn) % 2 != 0) ? -1.0f : 1.0f;
  const int i = (int)(sqrtf((float)(ll * ll + mm * mm + nn * nn)) + 0.5f);  // I - 1
  const float n3 = (float)nsam * (float)nsam * (float)nsam;
  float2 out = make_float2(0.f, 0.f);
  if (i < nshell && ssnr[i] != 0.0f && fsc[i] > 0.5f) {
    const float den = v3[id] + fpart / ssnr[i];
    const float2 dv = d[id];
    out = make_float2(dv.x * sg / den / n3, dv.y * sg / den / n3);
  }
  c[id] = out;
}

// block-order combination of per-block partials: out[k] = sum over blocks
__global__ void __launch_bounds__(kT)
    k_combine(int nval, int nblk, const double* __restrict__ part, double* out) {
  const int k = threadIdx.x;
  if (k >= nval) return;
  double s = 0.0;
  for (int b = 0; b < nblk; b++) s += part[(size_t)b * nval + k];
  out[k] = s;
}

// wave then workgroup sums of nv values per thread into part[block][*]
template <int NV>
__device__ inline void block_sums(double (&v)[NV], double* part) {
  __shared__ double red[kT / 64][NV];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++)
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; k++) red[wave][k] = v[k];
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0.0;
    for (int w = 0; w < kT / 64; w++) s += red[w][threadIdx.x];
    part[(size_t)blockIdx.x * NV + threadIdx.x] = s;
  }
}

// the voxels of the padded real volume [N][N][N+2], first N of each row
__device__ inline bool real_voxel(long long e, int nsam, long long& at) {
  const long long row = e / nsam;
  at = row * (nsam + 2) + (e - row * nsam);
  return e < (long long)nsam * nsam * nsam;
}

// find_fpart.f:71-82: how many voxels of C reach each of the thresholds tt[k]
__global__ void __launch_bounds__(kT)
    k_count_thresh(int nsam, const float* __restrict__ cv, const float* __restrict__ tt,
                   double* part) {
  double cnt[kThresh];
#pragma unroll
  for (int k = 0; k < kThresh; k++) cnt[k] = 0.0;
  const long long nvox = (long long)nsam * nsam * nsam;
  for (long long e = blockIdx.x * (long long)kT + threadIdx.x; e < nvox;
       e += (long long)gridDim.x * kT) {
    long long at;
    real_voxel(e, nsam, at);
    const float x = cv[at];
#pragma unroll
    for (int k = 0; k < kThresh; k++) cnt[k] += x >= tt[k] ? 1.0 : 0.0;
  }
  block_sums<kThresh>(cnt, part);
}

// find_fpart.f:85-99: B = 0 where C < TT; sums of B and B^2 and the count of
// the others
__global__ void __launch_bounds__(kT)
    k_select(int nsam, float* bv, const float* __restrict__ cv, float tt, double* part) {
  double v[3] = {0.0, 0.0, 0.0};
  const long long nvox = (long long)nsam * nsam * nsam;
  for (long long e = blockIdx.x * (long long)kT + threadIdx.x; e < nvox;
       e += (long long)gridDim.x * kT) {
    long long at;
    real_voxel(e, nsam, at);
    if (cv[at] < tt) {
      bv[at] = 0.0f;
    } else {
      const float x = bv[at];
      v[0] += (double)x;
      v[1] += (double)(x * x);
      v[2] += 1.0;
    }
  }
  block_sums<3>(v, part);
}

// volmeasure.f:17-40: the sums and counts of C below and at/above thresh
__global__ void __launch_bounds__(kT)
    k_two_class(int nsam, const float* __restrict__ cv, float thresh, double* part) {
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  const long long nvox = (long long)nsam * nsam * nsam;
  for (long long e = blockIdx.x * (long long)kT + threadIdx.x; e < nvox;
       e += (long long)gridDim.x * kT) {
    long long at;
    real_voxel(e, nsam, at);
    const float x = cv[at];
    if (x >= thresh) {
      v[2] += (double)x;
      v[3] += 1.0;
    } else {
      v[0] += (double)x;
      v[1] += 1.0;
    }
  }
  block_sums<4>(v, part);
}

// fpart_cc.f:50-64 (mode 0: sums of C and C^2 and the count where B != 0)
// and 66-77 (mode 1: the centred cross and square sums)
__global__ void __launch_bounds__(kT)
    k_cc_sums(int nsam, const float* __restrict__ bv, const float* __restrict__ cv, int mode,
              float mb, float mc, double* part) {
  double v[3] = {0.0, 0.0, 0.0};
  const long long nvox = (long long)nsam * nsam * nsam;
  for (long long e = blockIdx.x * (long long)kT + threadIdx.x; e < nvox;
       e += (long long)gridDim.x * kT) {
    long long at;
    real_voxel(e, nsam, at);
    const float b = bv[at];
    if (b == 0.0f) continue;
    const float c = cv[at];
    if (mode == 0) {
      v[0] += (double)c;
      v[1] += (double)(c * c);
      v[2] += 1.0;
    } else {
      const float db = b - mb, dc = c - mc;
      v[0] += (double)(db * dc);
      v[1] += (double)(db * db);
      v[2] += (double)(dc * dc);
    }
  }
  block_sums<3>(v, part);
}

inline unsigned nblk(long long n) { return (unsigned)((n + kT - 1) / kT); }

}  // namespace
}  // namespace fr

using namespace fr;

extern "C" {

int fr_find_fpart(fr_ctx* c, int nsam, double std_, double vtd, const float* d_a3df,
                  const float* d_d3df, const float* d_s3df, const float* d_v3df, int iredun,
                  float ri, float ric, float psize, float mw, float dalt, const float* fsc,
                  float* pssnr, int nshell, float* fpart) {
  FR_CHECK(c != nullptr, "null context");
  FR_CHECK(nsam >= 8 && nsam % 2 == 0 && nsam <= 1024, "NSAM must be even, 8 .. 1024");
  FR_CHECK(fsc && pssnr && fpart, "null argument");
  const int jc = nsam / 2 + 1;
  FR_CHECK(nshell >= jc, "FSC and PSSNR need NSAM/2+1 shells (got %d)", nshell);
  // find_fpart.f:37-43
  FR_CHECK(std_ != 0.0 && vtd != 0.0,
           "STD or VTD is zero (at least two accepted particles are needed)");
  float f;
  if (mw == 0.0f) {
    FR_CHECK(d_a3df && d_d3df && d_s3df && d_v3df, "null argument");
    const long long nv = (long long)nsam * nsam * jc;
    const long long nvox = (long long)nsam * nsam * nsam;
    int err = 0;
    float* bvol = (float*)c->slot(29, sizeof(float) * 2 * (size_t)nv, &err);
    if (!bvol) return err;
    float* cvol = (float*)c->slot(30, sizeof(float) * 2 * (size_t)nv, &err);
    if (!cvol) return err;
    double* part = (double*)c->slot(31, sizeof(double) * ((size_t)kBlocks * kThresh + kThresh),
                                    &err);
    if (!part) return err;
    double* out = part + (size_t)kBlocks * kThresh;
    float* dtab = (float*)c->slot(32, sizeof(float) * (2 * (size_t)nshell + kThresh), &err);
    if (!dtab) return err;
    auto reduce = [&](int nval, double* h) -> int {
      hipLaunchKernelGGL(k_combine, dim3(1), dim3(kT), 0, c->stream, nval, kBlocks, part, out);
      FR_HIP(hipGetLastError());
      FR_HIP(hipMemcpyAsync(h, out, sizeof(double) * nval, hipMemcpyDeviceToHost, c->stream));
      FR_HIP(hipStreamSynchronize(c->stream));
      return 0;
    };
    // find_fpart.f:45-61, then FFTW_BWD of both
    const float frad2 = 2.0f * psize / 30.0f * (float)nsam;
    hipLaunchKernelGGL(k_fpart_vols, dim3(nblk(nv)), dim3(kT), 0, c->stream, nsam,
                       reinterpret_cast<const float2*>(d_a3df),
                       reinterpret_cast<const float2*>(d_d3df), d_s3df, d_v3df,
                       std_ * (double)iredun, (std_ + vtd) * (double)iredun, frad2,
                       reinterpret_cast<float2*>(bvol), reinterpret_cast<float2*>(cvol));
    FR_HIP(hipGetLastError());
    if (int r = fft_c2r_3d(*c, nsam, bvol)) return r;
    if (int r = fft_c2r_3d(*c, nsam, cvol)) return r;
    // find_fpart.f:62-69: MASK3D of C with HALFW 1
    float fmask = 0.f, mave = 0.f, mstd = 0.f;
    if (int r = fr_mask3d(c, nsam, cvol, ri, ric, 1.0f, &fmask, &mave, &mstd)) return r;
    // find_fpart.f:71-82: the first threshold TT = MAVE + (T - 0.1 K) MSTD
    // holding more than 0.2 N^3 FMASK voxels (else the last, K = 15)
    float tt[kThresh];
    for (int k = 0; k < kThresh; k++) tt[k] = mave + (1.5f - (float)k * 0.1f) * mstd;
    FR_HIP(hipMemcpyAsync(dtab + 2 * nshell, tt, sizeof tt, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_count_thresh, dim3(kBlocks), dim3(kT), 0, c->stream, nsam, cvol,
                       dtab + 2 * nshell, part);
    FR_HIP(hipGetLastError());
    double cnt[kThresh];
    if (int r = reduce(kThresh, cnt)) return r;
    float ttsel = tt[kThresh - 1];
    const float lim = 0.2f * (float)nvox * fmask;
    for (int k = 0; k < kThresh; k++)
      if ((float)cnt[k] > lim) {
        ttsel = tt[k];
        break;
      }
    // find_fpart.f:84-103: B outside the selection zeroed; its mean and SD
    hipLaunchKernelGGL(k_select, dim3(kBlocks), dim3(kT), 0, c->stream, nsam, bvol, cvol, ttsel,
                       part);
    FR_HIP(hipGetLastError());
    double sel[3];
    if (int r = reduce(3, sel)) return r;
    double sum1 = sel[0], sum2 = sel[1];
    if (sel[2] != 0.0) {
      sum1 /= sel[2];
      sum2 /= sel[2];
    }
    const float mb = (float)sum1;
    const float sb = (float)sqrt(fabs(sum2 - sum1 * sum1));
    (void)sb;
    // volmeasure.f:17-60: ten two-class iterations from threshold 0
    float thresh = 0.0f, d1 = 0.0f, d2 = 0.0f;
    for (int j = 0; j < 10; j++) {
      hipLaunchKernelGGL(k_two_class, dim3(kBlocks), dim3(kT), 0, c->stream, nsam, cvol, thresh,
                         part);
      FR_HIP(hipGetLastError());
      double h[4];
      if (int r = reduce(4, h)) return r;
      d1 = h[1] != 0.0 ? (float)(h[0] / h[1]) : (float)h[0];
      d2 = h[3] != 0.0 ? (float)(h[2] / h[3]) : (float)h[2];
      thresh = (d2 - d1) / 2.0f + d1;
    }
    thresh = (d2 - d1) / 2.0f + d1;
    {
      hipLaunchKernelGGL(k_two_class, dim3(kBlocks), dim3(kT), 0, c->stream, nsam, cvol, thresh,
                         part);
      FR_HIP(hipGetLastError());
      double h[4];
      if (int r = reduce(4, h)) return r;
      f = (float)h[3] / (float)nvox;
    }
    // fpart_cc.f for one FPART
    FR_HIP(hipMemcpyAsync(dtab, pssnr, sizeof(float) * nshell, hipMemcpyHostToDevice, c->stream));
    FR_HIP(hipMemcpyAsync(dtab + nshell, fsc, sizeof(float) * nshell, hipMemcpyHostToDevice,
                          c->stream));
    auto fpart_cc = [&](float fp, float* cc) -> int {
      hipLaunchKernelGGL(k_fpart_cc_vol, dim3(nblk(nv)), dim3(kT), 0, c->stream, nsam,
                         reinterpret_cast<const float2*>(d_d3df), d_v3df, fp, dtab, dtab + nshell,
                         nshell, reinterpret_cast<float2*>(cvol));
      FR_HIP(hipGetLastError());
      if (int r = fft_c2r_3d(*c, nsam, cvol)) return r;
      hipLaunchKernelGGL(k_cc_sums, dim3(kBlocks), dim3(kT), 0, c->stream, nsam, bvol, cvol, 0,
                         0.f, 0.f, part);
      FR_HIP(hipGetLastError());
      double h[3];
      if (int r = reduce(3, h)) return r;
      double s1 = h[0], s2 = h[1];
      if (h[2] != 0.0) {
        s1 /= h[2];
        s2 /= h[2];
      }
      const float mc = (float)s1;
      hipLaunchKernelGGL(k_cc_sums, dim3(kBlocks), dim3(kT), 0, c->stream, nsam, bvol, cvol, 1,
                         mb, mc, part);
      FR_HIP(hipGetLastError());
      if (int r = reduce(3, h)) return r;
      *cc = (float)(h[0] / sqrt(h[1] * h[2]));
      return 0;
    };
    // find_fpart.f:111-135: steps of DF = 0.2, reversed and divided by three
    // when the correlation stops rising, until |step| <= EP = 0.005
    float df1 = -0.2f, ccmax = 0.f, cc = 0.f;
    if (int r = fpart_cc(f, &ccmax)) return r;
    for (int guard = 0; guard < 10000; guard++) {
      f = f + df1;
      if (int r = fpart_cc(f, &cc)) return r;
      if (cc > ccmax) {
        ccmax = cc;
        continue;
      }
      f = f - df1;
      if (fabsf(df1) > 0.005f) {
        df1 = -df1 / 3.0f;
        continue;
      }
      break;
    }
  } else {
    FR_CHECK(dalt != 0.0f && psize != 0.0f, "DALT and PSIZE must be non-zero");
    const float box = psize * (float)nsam;  // find_fpart.f:141
    f = 1000.0f * mw / dalt / (box * box * box);
  }
  // find_fpart.f:143-149
  f = fabsf(f);
  if (f != 0.0f)
    for (int i = 0; i < jc; i++) pssnr[i] = pssnr[i] / f;
  *fpart = f;
  return 0;
}

}  // extern "C"

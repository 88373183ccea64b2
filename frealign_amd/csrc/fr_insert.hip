// fr_insert.hip -- CTF-weighted trilinear Fourier insertion.
//
// Replaces PINSERT_TRILIN (pinsert_trilin.f:8-244) as driven by PINSERT
// (pinsert.f:8-195) for a buffer of particles, the A3D3/A3D3_S flush
// (a3d3.f:107-143).  One task = (particle, symmetry copy); one thread = one
// point (I,J) of the full central plane with I^2+J^2 <= IRREC^2.  The plane
// coordinates X3 = DM*(I,J) and the voxel range follow the reference's FP32
// operation order exactly (no FMA contraction), so the set of voxels touched
// and NKSUM are bit-identical to the reference; the numerator and CTF^2 sums
// use hardware FP32 atomics (order-dependent rounding only).
#include <hip/hip_runtime.h>

#include "fr_common.h"
#include "fr_kernels.h"

namespace fr {

namespace {

constexpr int kThreads = 256;

__device__ inline void wave_reduce(double& d, unsigned long long& n) {
  for (int off = 32; off > 0; off >>= 1) {
    d += __shfl_down(d, off, 64);
    n += __shfl_down(n, off, 64);
  }
}

__global__ void __launch_bounds__(kThreads)
    k_insert(const float2* __restrict__ datc, const float2* __restrict__ ctff,
             const InsertTask* __restrict__ tasks, int nsam, float2* num0, float* ctf2_0,
             float2* num1, float* ctf2_1, unsigned long long* nksum, double* ttd) {
  const InsertTask t = tasks[blockIdx.y];
  const int jc = nsam / 2 + 1, nsamh = nsam / 2, w = 2 * jc - 1;
  const size_t ic = (size_t)nsam * jc;
  const int p = blockIdx.x * kThreads + threadIdx.x;
  double my_ttd = 0.0;
  unsigned long long my_n = 0;
  if (p < w * w) {
    const int i = p % w - (jc - 1);
    const int j = p / w - (jc - 1);
    float x3, y3, z3, g2;
    int ls, lt, ms, mt, ns, nt;
    bool inside;
    {
#pragma clang fp contract(off)
      g2 = (float)(i * i + j * j);
      inside = g2 <= (float)t.irrec2;
      float fi = (float)i, fj = (float)j;
      float dm1 = t.dm[0] * fi, dm2 = t.dm[1] * fi, dm3 = t.dm[2] * fi;
      x3 = dm1 + t.dm[3] * fj;
      y3 = dm2 + t.dm[4] * fj;
      z3 = dm3 + t.dm[5] * fj;
      // pinsert_trilin.f:88-111: A=X3+0.5-RI12 (RI12=0.5), A=X3+RI2 (RI2=1)
      float a;
      a = (x3 + 0.5f) - 0.5f; ls = (int)a; if (a < 0.0f) ls--; if (ls < 0) ls = 0;
      a = x3 + 1.0f;          lt = (int)a; if (a < 0.0f) lt--; if (lt > nsamh) lt = nsamh;
      a = (y3 + 0.5f) - 0.5f; ms = (int)a; if (a < 0.0f) ms--; if (ms < -nsamh) ms = -nsamh;
      a = y3 + 1.0f;          mt = (int)a; if (a < 0.0f) mt--; if (mt > nsamh) mt = nsamh;
      a = (z3 + 0.5f) - 0.5f; ns = (int)a; if (a < 0.0f) ns--; if (ns < -nsamh) ns = -nsamh;
      a = z3 + 1.0f;          nt = (int)a; if (a < 0.0f) nt--; if (nt > nsamh) nt = nsamh;
    }
    if (inside) {
      const float2* sp = datc + (size_t)t.part * ic;
      const float2* cp = ctff + (size_t)t.part * 2 * ic;
      size_t id;
      bool mate = i < 0;
      if (!mate) id = (size_t)i + (size_t)jc * wrapn(j, nsam);
      else id = (size_t)(-i) + (size_t)jc * wrapn(-j, nsam);
      float2 s = sp[id];
      if (mate) s.y = -s.y;
      float2 cr = cp[id], cl = cp[id + ic];
      cf ctfs = cf{cr.x + cl.x, cr.y - cl.y};  // CTFR + CONJG(CTFL)
      float phase = t.shx * (float)i + t.shy * (float)j;
      float sn, cs;
      sincosf(phase, &sn, &cs);
      float wgt = expf(t.tmp * g2) * t.occ;
      cf samp = cmul(cmul(cf{s.x, s.y}, cf{cs, sn}), ctfs);
      float ctfv2 = 0.25f * (ctfs.re * ctfs.re + ctfs.im * ctfs.im) * wgt;
      samp = cscale(samp, wgt);
      float2* num = t.half ? num1 : num0;
      float* c2 = t.half ? ctf2_1 : ctf2_0;
      for (int n = ns; n <= nt; n++) {
        float b3 = 1.0f - fabsf(z3 - (float)n);
        size_t n1 = (size_t)nsam * wrapn(n, nsam);
        for (int m = ms; m <= mt; m++) {
          float b2 = (1.0f - fabsf(y3 - (float)m));
          size_t m1 = (size_t)jc * (wrapn(m, nsam) + n1);
          for (int l = ls; l <= lt; l++) {
            float box = (1.0f - fabsf(x3 - (float)l)) * b2 * b3;
            size_t v = (size_t)l + m1;
            float cv = ctfv2 * box;
            unsafeAtomicAdd(&num[v].x, samp.re * box);
            unsafeAtomicAdd(&num[v].y, samp.im * box);
            unsafeAtomicAdd(&c2[v], cv);
            my_n++;
            my_ttd += (double)cv;
          }
        }
      }
    }
  }
  // block reduction of NKSUM / TTD, one atomic per block
  __shared__ double sd[kThreads / 64];
  __shared__ unsigned long long sn[kThreads / 64];
  wave_reduce(my_ttd, my_n);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { sd[wid] = my_ttd; sn[wid] = my_n; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double d = 0.0;
    unsigned long long n = 0;
    for (int k = 0; k < kThreads / 64; k++) { d += sd[k]; n += sn[k]; }
    if (n) {
      atomicAdd(&nksum[t.half], n);
      unsafeAtomicAdd(&ttd[t.half], d);
    }
  }
}

}  // namespace

int launch_insert(Ctx& c, int nsam, int ntask, const InsertTask* d_tasks, const float* d_datc,
                  const float* d_ctff, float* num0, float* ctf2_0, float* num1, float* ctf2_1,
                  unsigned long long* d_nksum, double* d_ttd) {
  if (ntask <= 0) return 0;
  const int jc = nsam / 2 + 1, w = 2 * jc - 1;
  FR_CHECK(ntask <= 65535 * 64, "too many insertion tasks (%d)", ntask);
  int done = 0;
  while (done < ntask) {
    int chunk = ntask - done < 65535 ? ntask - done : 65535;
    dim3 grid((w * w + kThreads - 1) / kThreads, chunk);
    hipLaunchKernelGGL(k_insert, grid, dim3(kThreads), 0, c.stream,
                       reinterpret_cast<const float2*>(d_datc),
                       reinterpret_cast<const float2*>(d_ctff), d_tasks + done, nsam,
                       reinterpret_cast<float2*>(num0), ctf2_0, reinterpret_cast<float2*>(num1),
                       ctf2_1, d_nksum, d_ttd);
    FR_HIP(hipGetLastError());
    done += chunk;
  }
  return 0;
}

}  // namespace fr

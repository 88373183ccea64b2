// fr_prep.hip -- batched particle preparation.
//
// Replaces, for a batch of particles, the r2c FFT of the particle image in
// LMAIN (lmain.f:459) followed by CTFAPPLY (ctfapply.f:8-161): CTF table
// (ctf.f), shift phase ramp, c2r, scale, MASKCOS (maskcos.f), r2c,
// (-1)^(l+m) origin shift and PWEIGHT (pweight.f), including PWEIGHT's
// in-place double weighting of the Nyquist row and its JC+1 shell aliasing
// into bin 1's count (the reference calls PWEIGHT(NSAM,OUTC,OUTC,...)).
//
// For NSAM 64, 128 and 256 the steps run as seven launches: the CTF table;
// the r2c's row pass; one column pass that completes the r2c (written back
// into the image), applies the shift and writes DATC, and runs OUTC's
// backward column pass; the row c2r, which also sums each row block's part
// of MASKCOS's ring; the row r2c, which scales and masks as it loads; the
// forward column pass; the origin shift and PWEIGHT.  Other sizes take the
// FFT entry points and one kernel per step.
#include <hip/hip_runtime.h>

#include "fr_common.h"
#include "fr_kernels.h"
#include "fr_ctf.h"
#include "fr_fft_dev.h"

namespace fr {

namespace {

constexpr int kThreads = 256;
constexpr float kPi = 3.1415926535897f;

// ctfapply.f:64-80 at half-plane point (l, mm) of a particle (pp: its ppar
// row) with spectrum value s: the shifted value v (OUTC before its c2r) and
// DATC = v (-1)^(l+mm) (ctfapply.f:144-158's origin shift: DATC is not
// transformed again, OUTC takes it in PWEIGHT's weighting pass)
__device__ inline void shift_point(const float* pp, int l, int mm, float2 s, float2& v,
                                   float2& d) {
  const float shx = pp[0], shy = pp[1];
  float phase = shx * (float)l + shy * (float)mm;
  float sn, cs;
  sincosf(phase, &sn, &cs);
  const cf r = cmul(cf{s.x, s.y}, cf{cs, sn});
  const float sg = ((l + mm) & 1) ? -1.0f : 1.0f;
  v = make_float2(r.re, r.im);
  d = make_float2(r.re * sg, r.im * sg);
}

// ctfapply.f:64-76: the CTF at (l, mm) and (-l, -mm) (CTFR, CTFL)
__device__ inline void ctf_point(const PrepGlobals& g, const float* pp, int l, int mm,
                                 float2& cr, float2& cl) {
  const float df1 = pp[2], df2 = pp[3], ang = pp[4], amagp = pp[5];
  const float thetatrp = g.thetatr / amagp;
  cf a, b;
  ctf_pair_dev(g, df1, df2, ang, thetatrp, l, mm, a, b);
  cr = make_float2(a.re, a.im);
  cl = make_float2(b.re, b.im);
}

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
  float2 v, d, cr, cl;
  shift_point(pp, l, mm, spec[(size_t)b * ic + id], v, d);
  ctf_point(g, pp, l, mm, cr, cl);
  outc[(size_t)b * ic + id] = v;
  datc[(size_t)b * ic + id] = d;
  ctff[(size_t)b * 2 * ic + id] = cr;
  ctff[(size_t)b * 2 * ic + ic + id] = cl;
}

// ctfapply.f:64-76 alone: the CTF table, one thread per half-plane point
// (coalesced stores; the fused column pass below applies the shift)
__global__ void __launch_bounds__(kThreads)
    k_ctf_table(int nsam, int nb, const float* __restrict__ ppar, PrepGlobals g, float2* ctff) {
  const int jc = nsam / 2 + 1;
  const size_t ic = (size_t)nsam * jc;
  const size_t gid = blockIdx.x * (size_t)kThreads + threadIdx.x;
  if (gid >= ic * nb) return;
  const int b = (int)(gid / ic);
  const int id = (int)(gid - (size_t)b * ic);
  const int l = id % jc, m = id / jc;
  const int mm = m >= jc ? m - nsam : m;
  float2 cr, cl;
  ctf_point(g, ppar + 8 * b, l, mm, cr, cl);
  ctff[(size_t)b * 2 * ic + id] = cr;
  ctff[(size_t)b * 2 * ic + ic + id] = cl;
}

// The shift step between two column passes, for the box sizes whose FFT
// length is a compile-time constant: a workgroup takes CW half-plane columns
// of one particle.  kFwd: `spec` holds the row transforms only; the forward
// column transform is completed here and written back (the r2c in place of
// lmain.f:459 leaves the spectrum in the image).  Otherwise `spec` is a whole
// spectrum and is only read.  Then shift_point per point (DATC goes out) and
// the backward column transform of OUTC, which the row pass completes.  Each
// value takes the operations of the separate column passes and k_ctf_shift,
// in their order; two HBM round trips of the batch's spectra are saved.  The
// CTF table, which needs no spectrum, comes from k_ctf_table.
template <int N, bool kFwd>
__global__ void __launch_bounds__(fftdev::kFftThreads)
    k_cols_shift(float2* spec, const float* __restrict__ ppar, float2* outc, float2* datc, int nb,
                 const float2* __restrict__ tw) {
  using namespace fftdev;
  constexpr int JC = N / 2 + 1, LD = N + 1, CW = cols_cw<N>();
  constexpr int JB = (JC + CW - 1) / CW;
  constexpr size_t IC = (size_t)N * JC;
  extern __shared__ float2 cs_smem[];
  float2* buf = cs_smem;
  float2* tmp = cs_smem + CW * LD;
  // column blocks of one particle in XCD-contiguous order
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int b = lb / JB;
  if (b >= nb) return;
  const int x0 = (lb - b * JB) * CW;
  const int ncol = min(CW, JC - x0);
  float2* sp = spec + (size_t)b * IC;
  for (int e = threadIdx.x; e < N * CW; e += kFftThreads) {
    const int y = e / CW, c = e - y * CW;
    float2 v = make_float2(0.f, 0.f);
    if (c < ncol) v = sp[(size_t)y * JC + x0 + c];
    buf[c * LD + y] = v;
  }
  __syncthreads();
  float2* res = buf;
  if constexpr (kFwd) {
    res = lds_fft_t<N, CW, LD, -1>(buf, tmp, tw);
    for (int e = threadIdx.x; e < N * CW; e += kFftThreads) {
      const int y = e / CW, c = e - y * CW;
      if (c < ncol) sp[(size_t)y * JC + x0 + c] = res[c * LD + y];
    }
  }
  float2* oth = res == buf ? tmp : buf;
  const float* pp = ppar + 8 * b;
  float2* dc = datc + (size_t)b * IC;
  // each thread rewrites only the entries it read: no barrier before this loop
  for (int e = threadIdx.x; e < N * CW; e += kFftThreads) {
    const int y = e / CW, c = e - y * CW;
    if (c >= ncol) continue;
    float2 v, d;
    shift_point(pp, x0 + c, y >= JC ? y - N : y, res[c * LD + y], v, d);
    res[c * LD + y] = v;
    dc[(size_t)y * JC + x0 + c] = d;
  }
  __syncthreads();
  float2* out = lds_fft_t<N, CW, LD, +1>(res, oth, tw);
  float2* oc = outc + (size_t)b * IC;
  for (int e = threadIdx.x; e < N * CW; e += kFftThreads) {
    const int y = e / CW, c = e - y * CW;
    if (c < ncol) oc[(size_t)y * JC + x0 + c] = out[c * LD + y];
  }
}

template <int N>
int cols_shift_t(Ctx& c, int nb, float* spec, bool fwd, const float* ppar, float* outc,
                 float* datc) {
  constexpr int JC = N / 2 + 1, CW = fftdev::cols_cw<N>();
  const size_t lds = sizeof(float2) * 2 * CW * (N + 1);
  const float2* tw = c.twiddles(N);
  FR_CHECK(tw != nullptr, "twiddle table allocation failed");
  FR_CHECK(nb <= 65535, "too many images in one FFT batch (%d)", nb);
  const dim3 grid(fftdev::round8((JC + CW - 1) / CW * nb));
  auto* sp = reinterpret_cast<float2*>(spec);
  auto* oc = reinterpret_cast<float2*>(outc);
  auto* dc = reinterpret_cast<float2*>(datc);
  if (fwd)
    hipLaunchKernelGGL((k_cols_shift<N, true>), grid, dim3(fftdev::kFftThreads), lds, c.stream,
                       sp, ppar, oc, dc, nb, tw);
  else
    hipLaunchKernelGGL((k_cols_shift<N, false>), grid, dim3(fftdev::kFftThreads), lds, c.stream,
                       sp, ppar, oc, dc, nb, tw);
  FR_HIP(hipGetLastError());
  return 0;
}

// MASKCOS's squared distance of 0-based pixel (row, x) from the centre
// NSAM/2+1 of 1-based coordinates, as k_scale_mask forms it
__device__ inline float maskcos_r2(int row, int x, int nmid) {
  const float d1 = (float)(row + 1 - nmid), d2 = (float)(x + 1 - nmid);
  float crad2;
  {
#pragma clang fp contract(off)
    crad2 = d2 * d2 + d1 * d1;
  }
  return crad2;
}

// The row c2r of OUTC (as the FFT's row pass) that also sums, for MASKCOS's
// mean (ctfapply.f:107-122), the scaled pixels of the ring RI2' < r^2 <=
// RI3' of the rows it writes: part[workgroup] = (sum, count).  A workgroup's
// NSEQ row pairs lie in one image.
template <int N>
__global__ void __launch_bounds__(fftdev::kFftThreads)
    k_rows_c2r_ring(float* data, size_t img_stride, int npairs, const float* __restrict__ ppar,
                    float ri2, float ri3, float scal, float2* part,
                    const float2* __restrict__ tw) {
  using namespace fftdev;
  constexpr int NSEQ = 2048 / N, LD = N + 1, JC = N / 2 + 1, PITCH = N + 2;
  extern __shared__ float2 rr_smem[];
  __shared__ float red_s[kFftThreads / 64];
  __shared__ int red_c[kFftThreads / 64];
  float2* buf = rr_smem;
  float2* tmp = rr_smem + NSEQ * LD;
  const int pair0 = blockIdx.x * NSEQ;
  for (int e = threadIdx.x; e < NSEQ * JC; e += kFftThreads) {
    const int s = e / JC, k = e - s * JC;
    const int pr = pair0 + s;
    float2 a = make_float2(0.f, 0.f), b = make_float2(0.f, 0.f);
    if (pr < npairs) {
      const int img = (2 * pr) / N, row = (2 * pr) % N;
      const float* ra = data + img * img_stride + (size_t)row * PITCH;
      a = make_float2(ra[2 * k], ra[2 * k + 1]);
      b = make_float2(ra[PITCH + 2 * k], ra[PITCH + 2 * k + 1]);
    }
    if (k == 0 || 2 * k == N) { a.y = 0.f; b.y = 0.f; }
    buf[s * LD + k] = make_float2(a.x - b.y, a.y + b.x);
    if (k > 0 && 2 * k < N) buf[s * LD + (N - k)] = make_float2(a.x + b.y, -a.y + b.x);
  }
  __syncthreads();
  float2* res = lds_fft_t<N, NSEQ, LD, +1>(buf, tmp, tw);
  const int img0 = (2 * pair0) / N;
  const float amagp = ppar[8 * (size_t)img0 + 5];
  const float ri2p = ri2 / amagp / amagp, ri3p = ri3 / amagp / amagp;
  const int nmid = N / 2 + 1;
  float sum = 0.f;
  int cnt = 0;
  for (int e = threadIdx.x; e < NSEQ * PITCH; e += kFftThreads) {
    const int s = e / PITCH, x = e - s * PITCH;
    const int pr = pair0 + s;
    if (pr >= npairs) continue;
    const int img = (2 * pr) / N, row = (2 * pr) % N;
    float* ra = data + img * img_stride + (size_t)row * PITCH;
    if (x >= N) continue;  // the row padding is not read again
    const float2 z = res[s * LD + x];
    // pixels beyond RI3' become the ring mean in the masked r2c, which does
    // not read them: only those within RI3' are stored
    const float ca = maskcos_r2(row, x, nmid), cb = maskcos_r2(row + 1, x, nmid);
    if (ca <= ri3p) ra[x] = z.x;
    if (cb <= ri3p) ra[PITCH + x] = z.y;
    {
      if (ca > ri2p && ca <= ri3p) {
        sum += z.x * scal;
        cnt++;
      }
      if (cb > ri2p && cb <= ri3p) {
        sum += z.y * scal;
        cnt++;
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    sum += __shfl_down(sum, off, 64);
    cnt += __shfl_down(cnt, off, 64);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    red_s[wave] = sum;
    red_c[wave] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    int n = 0;
    for (int w = 0; w < kFftThreads / 64; w++) {
      t += red_s[w];
      n += red_c[w];
    }
    part[blockIdx.x] = make_float2(t, (float)n);
  }
}

// The row r2c of OUTC (as the FFT's row pass) with ctfapply.f:107-122 on its
// input: each pixel scaled by 1/NSAM^2 and MASKCOS-masked (cosine edge over
// the ring, the ring mean beyond it), the mean from the image's row-block
// partial sums of k_rows_c2r_ring, added in block order.
template <int N>
__global__ void __launch_bounds__(fftdev::kFftThreads)
    k_rows_r2c_mask(float* data, size_t img_stride, int npairs, const float* __restrict__ ppar,
                    float ri2, float ri3, float rih, float halfw, float scal,
                    const float2* __restrict__ part, const float2* __restrict__ tw) {
  using namespace fftdev;
  constexpr int NSEQ = 2048 / N, LD = N + 1, JC = N / 2 + 1, PITCH = N + 2;
  constexpr int WPI = (N / 2) / NSEQ;  // row blocks per image
  extern __shared__ float2 rm_smem[];
  float2* buf = rm_smem;
  float2* tmp = rm_smem + NSEQ * LD;
  const int pair0 = blockIdx.x * NSEQ;
  const int img0 = (2 * pair0) / N;
  float tot = 0.f;
  int itot = 0;
  for (int k = 0; k < WPI; k++) {
    const float2 q = part[(size_t)img0 * WPI + k];
    tot += q.x;
    itot += (int)q.y;
  }
  const float mean = itot ? tot / (float)itot : tot;
  const float amagp = ppar[8 * (size_t)img0 + 5];
  const float rihp = rih / amagp, ri2p = ri2 / amagp / amagp, ri3p = ri3 / amagp / amagp;
  const int nmid = N / 2 + 1;
  auto mask = [&](float v, float crad2) {
    if (crad2 > ri2p && crad2 <= ri3p) {
      const float edge = (1.0f + cosf(kPi * (sqrtf(crad2) - rihp) / halfw)) / 2.0f;
      v = v * edge + (1.0f - edge) * mean;
    } else if (crad2 > ri3p) {
      v = mean;
    }
    return v;
  };
  for (int e = threadIdx.x; e < NSEQ * N; e += kFftThreads) {
    const int s = e / N, x = e - s * N;
    const int pr = pair0 + s;
    float2 v = make_float2(0.f, 0.f);
    if (pr < npairs) {
      const int img = (2 * pr) / N, row = (2 * pr) % N;
      const float* ra = data + img * img_stride + (size_t)row * PITCH;
      // beyond RI3' the pixel is the mean: k_rows_c2r_ring did not store it
      const float ca = maskcos_r2(row, x, nmid), cb = maskcos_r2(row + 1, x, nmid);
      v = make_float2(ca > ri3p ? mean : mask(ra[x] * scal, ca),
                      cb > ri3p ? mean : mask(ra[x + PITCH] * scal, cb));
    }
    buf[s * LD + x] = v;
  }
  __syncthreads();
  float2* res = lds_fft_t<N, NSEQ, LD, -1>(buf, tmp, tw);
  for (int e = threadIdx.x; e < NSEQ * JC; e += kFftThreads) {
    const int s = e / JC, k = e - s * JC;
    const int pr = pair0 + s;
    if (pr >= npairs) continue;
    const float2 zk = res[s * LD + k];
    const float2 zm = res[s * LD + ((N - k) % N)];
    const float2 xa = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
    const float2 xb = make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
    const int img = (2 * pr) / N, row = (2 * pr) % N;
    float* ra = data + img * img_stride + (size_t)row * PITCH;
    ra[2 * k] = xa.x;
    ra[2 * k + 1] = xa.y;
    ra[PITCH + 2 * k] = xb.x;
    ra[PITCH + 2 * k + 1] = xb.y;
  }
}

// OUTC's c2r, MASKCOS and r2c rows on the fused path; `part` (NSAM^2/4096
// float2 per image) carries the ring partial sums between the two launches
template <int N>
int rows_mask_t(Ctx& c, int nb, float* outc, const float* ppar, const PrepGlobals& g,
                float2* part) {
  constexpr int NSEQ = 2048 / N;
  const int npairs = nb * (N / 2);
  const size_t lds = sizeof(float2) * 2 * NSEQ * (N + 1);
  const size_t img = (size_t)N * (N + 2);
  const float2* tw = c.twiddles(N);
  FR_CHECK(tw != nullptr, "twiddle table allocation failed");
  const float scal = 1.0f / (float)N / (float)N;
  const dim3 grid((npairs + NSEQ - 1) / NSEQ);
  hipLaunchKernelGGL(k_rows_c2r_ring<N>, grid, dim3(fftdev::kFftThreads), lds, c.stream, outc,
                     img, npairs, ppar, g.ri2, g.ri3, scal, part, tw);
  FR_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_rows_r2c_mask<N>, grid, dim3(fftdev::kFftThreads), lds, c.stream, outc,
                     img, npairs, ppar, g.ri2, g.ri3, g.rih, g.halfw, scal, part, tw);
  FR_HIP(hipGetLastError());
  return 0;
}

bool fused_cols(int nsam) { return nsam == 64 || nsam == 128 || nsam == 256; }

// the fused path's column pass with the shift (fwd: completing the forward
// transform first), then OUTC's c2r, MASKCOS and r2c rows
int fused_shift_mask(Ctx& c, int nsam, int nb, float* spec, bool fwd, const float* ppar,
                     const PrepGlobals& g, float* outc, float* datc, float2* part) {
  if (!part) {
    // the context's buffer holds 65535 images' sums
    FR_CHECK(nb <= 65535, "too many images in one batch (%d)", nb);
    part = c.prep_partials();
    FR_CHECK(part != nullptr, "preparation buffer allocation failed");
  }
  switch (nsam) {
    case 64:
      if (int r = cols_shift_t<64>(c, nb, spec, fwd, ppar, outc, datc)) return r;
      return rows_mask_t<64>(c, nb, outc, ppar, g, part);
    case 128:
      if (int r = cols_shift_t<128>(c, nb, spec, fwd, ppar, outc, datc)) return r;
      return rows_mask_t<128>(c, nb, outc, ppar, g, part);
    case 256:
      if (int r = cols_shift_t<256>(c, nb, spec, fwd, ppar, outc, datc)) return r;
      return rows_mask_t<256>(c, nb, outc, ppar, g, part);
    default: return fail("no fused column pass for NSAM %d", nsam);
  }
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

namespace {

// the origin shift and PWEIGHT of OUTC's transform (ctfapply.f:144-158)
int origin_pweight(Ctx& c, int nsam, int nb, const float* pssnr, const PrepGlobals& g,
                   const float* ctff, float* outc) {
  const size_t lds = pweight_lds(nsam);
  FR_HIP(hipFuncSetAttribute((const void*)k_origin_pweight,
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_origin_pweight, dim3(nb), dim3(kMaskThreads), lds, c.stream, nsam, pssnr, g.ictf,
                     reinterpret_cast<const float2*>(ctff), reinterpret_cast<float2*>(outc));
  FR_HIP(hipGetLastError());
  return 0;
}

// the CTF table (first on the fused path: the passes after it then read what
// the pass before them wrote, while it is still in the caches)
int ctf_table(Ctx& c, int nsam, int nb, const float* ppar, const PrepGlobals& g, float* ctff) {
  const size_t tot = (size_t)nsam * (nsam / 2 + 1) * nb;
  hipLaunchKernelGGL(k_ctf_table, dim3((unsigned)((tot + kThreads - 1) / kThreads)),
                     dim3(kThreads), 0, c.stream, nsam, nb, ppar, g,
                     reinterpret_cast<float2*>(ctff));
  FR_HIP(hipGetLastError());
  return 0;
}

// the fused path after its row passes: OUTC's forward column pass, the
// origin shift and PWEIGHT
int fused_tail(Ctx& c, int nsam, int nb, const float* pssnr, const PrepGlobals& g,
               const float* ctff, float* outc) {
  if (int r = fft_cols_2d(c, nsam, nb, outc, (size_t)nsam * (nsam + 2), -1)) return r;
  return origin_pweight(c, nsam, nb, pssnr, g, ctff, outc);
}

}  // namespace

size_t prep_partials_count(int nsam, int nb) {
  // row blocks of NSEQ = 2048 / NSAM row pairs: NSAM^2 / 4096 per image
  const size_t per = fused_cols(nsam) ? (size_t)nsam * nsam / 4096 : 1;
  return per * (size_t)(nb > 0 ? nb : 1);
}

int launch_prep(Ctx& c, int nsam, int nb, float* data, const float* ppar, const float* pssnr,
                const PrepGlobals& g, float* ctff, float* outc, float* datc, float2* part) {
  if (nb <= 0) return 0;
  FR_CHECK(nsam / 2 + 1 <= 520, "box too large for PWEIGHT bins");
  const size_t img = (size_t)nsam * (nsam + 2);
  if (!fused_cols(nsam)) {
    // lmain.f:459: r2c of the particle image, in place
    if (int r = fft_r2c_2d(c, nsam, nb, data, img)) return r;
    return launch_ctfapply(c, nsam, nb, data, ppar, pssnr, g, ctff, outc, datc, part);
  }
  // lmain.f:459's r2c: its row pass here, its column pass in the fused
  // column kernel, which continues with CTFAPPLY's shift step and OUTC's
  // backward column pass
  if (int r = ctf_table(c, nsam, nb, ppar, g, ctff)) return r;
  if (int r = fft_rows_2d(c, nsam, nb, data, img, true)) return r;
  if (int r = fused_shift_mask(c, nsam, nb, data, true, ppar, g, outc, datc, part)) return r;
  return fused_tail(c, nsam, nb, pssnr, g, ctff, outc);
}

int launch_ctfapply(Ctx& c, int nsam, int nb, const float* spec, const float* ppar,
                    const float* pssnr, const PrepGlobals& g, float* ctff, float* outc,
                    float* datc, float2* part) {
  if (nb <= 0) return 0;
  FR_CHECK(nsam / 2 + 1 <= 520, "box too large for PWEIGHT bins");
  if (fused_cols(nsam)) {
    // kFwd false: the spectrum is only read
    if (int r = ctf_table(c, nsam, nb, ppar, g, ctff)) return r;
    if (int r = fused_shift_mask(c, nsam, nb, const_cast<float*>(spec), false, ppar, g, outc,
                                 datc, part))
      return r;
    return fused_tail(c, nsam, nb, pssnr, g, ctff, outc);
  }
  const size_t img = (size_t)nsam * (nsam + 2);
  const size_t tot = (size_t)nsam * (nsam / 2 + 1) * nb;
  hipLaunchKernelGGL(k_ctf_shift, dim3((unsigned)((tot + kThreads - 1) / kThreads)),
                     dim3(kThreads), 0, c.stream, nsam, nb, reinterpret_cast<const float2*>(spec),
                     ppar, g, reinterpret_cast<float2*>(ctff), reinterpret_cast<float2*>(outc),
                     reinterpret_cast<float2*>(datc));
  FR_HIP(hipGetLastError());
  if (int r = fft_c2r_2d(c, nsam, nb, outc, img)) return r;
  // ctfapply.f:107-122; per-particle AMAGP is column 5 of ppar
  hipLaunchKernelGGL(k_scale_mask, dim3(nb), dim3(kMaskThreads), 0, c.stream, nsam, outc, img,
                     1.0f / (float)nsam / (float)nsam, ppar + 5, 8, g.ri2, g.ri3, g.rih, g.halfw,
                     0);
  FR_HIP(hipGetLastError());
  if (int r = fft_r2c_2d(c, nsam, nb, outc, img)) return r;
  return origin_pweight(c, nsam, nb, pssnr, g, ctff, outc);
}

}  // namespace fr

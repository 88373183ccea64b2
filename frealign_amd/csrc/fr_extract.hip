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

#include <cmath>
#include <vector>

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
// real-FFT buffer of each item (zero elsewhere).  flip: CC3's phase flip
// instead (cc3.f:86-89, the section negated where REAL(CTFR+CONJG(CTFL)) < 0).
__global__ void __launch_bounds__(kThreads)
    k_cc3m_fill(int nsam, int n, int maxr12, int maxr22, const float2* __restrict__ a3df,
                const float* __restrict__ dm, const int* __restrict__ pidx,
                const float2* __restrict__ ctff, float* work, int flip) {
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
    if (flip) {
      if (cr.x + cl.x < 0.0f) a = cf{-a.re, -a.im};
    } else {
      a = cmul(a, cf{cr.x + cl.x, cr.y - cl.y});
    }
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
                 const float* __restrict__ fsct, const float* __restrict__ work, float* score,
                 int flip) {
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
        if (!flip && fsct[i - 1] < 0.95f) r = fabsf(r);
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

// presb.f:95-123: per-shell Re(SAMP conj(P)) / sqrt(sum|SAMP|^2 sum|P|^2) over
// 1 <= r^2 < (NSAM/2)^2; shells with no point leave RBIN untouched.
__global__ void __launch_bounds__(kThreads)
    k_presb_bins(int nsam, const float2* __restrict__ spec, const int* __restrict__ pidx,
                 const float* __restrict__ shift, const float* __restrict__ work, float* rbin) {
  __shared__ float bins[3 * 260];
  __shared__ int cnt[260];
  const int jc = nsam / 2 + 1, nsamh = nsam / 2;
  const size_t ic = (size_t)nsam * jc;
  const int it = blockIdx.x;
  const int nb = nsamh + 2;
  for (int k = threadIdx.x; k < nb; k += blockDim.x) {
    bins[3 * k] = bins[3 * k + 1] = bins[3 * k + 2] = 0.0f;
    cnt[k] = 0;
  }
  __syncthreads();
  const float2* sp = spec + (size_t)pidx[it] * ic;
  const float2* pw = reinterpret_cast<const float2*>(work) + (size_t)it * ic;
  const float shx = shift[2 * it], shy = shift[2 * it + 1];
  const int r2 = nsamh * nsamh;
  for (int e = threadIdx.x; e < (int)ic; e += blockDim.x) {
    int i = e % jc, m = e / jc;
    int j = m >= jc ? m - nsam : m;
    int itemp = i * i + j * j;
    if (itemp < 1 || itemp >= r2) continue;
    float sn, cs;
    sincosf(shx * (float)i + shy * (float)j, &sn, &cs);
    float2 s = sp[e];
    cf samp = cmul(cf{s.x, s.y}, cf{cs, sn});
    float2 a = pw[e];
    int ibin = (int)(sqrtf((float)itemp) + 0.5f);  // 0-based shell
    atomicAdd(&bins[3 * ibin + 0], samp.re * a.x + samp.im * a.y);
    atomicAdd(&bins[3 * ibin + 1], samp.re * samp.re + samp.im * samp.im);
    atomicAdd(&bins[3 * ibin + 2], a.x * a.x + a.y * a.y);
    atomicAdd(&cnt[ibin], 1);
  }
  __syncthreads();
  float* rb = rbin + (size_t)it * nsam;
  for (int i = threadIdx.x; i < nsamh; i += blockDim.x) {
    if (cnt[i] != 0) {
      rb[2 * i] = bins[3 * i] / sqrtf(bins[3 * i + 1] * bins[3 * i + 2]);
      rb[2 * i + 1] = 1.0f;
    }
  }
}

}  // namespace

// ------------------------------------------------------------ fused CC3M
// The same CC3M score with the two full-size transforms pruned to the band
// they touch.  Only columns I < 32 and rows |J| < 32 of the half-plane are
// non-zero on the way in (the annulus of cc3m.f:74-112), and only those are
// read on the way out (cc3m.f:117-150), so with the table
//   T[n][2i] = cos(2 pi i n/N),  T[n][2i+1] = -sin(2 pi i n/N),  i < 32,
// every pass is a product with T on the matrix cores (FP32 MFMA):
//   column c2c   Y  (N x 64) = T   x B(X)       (J pairs folded: K = 64)
//   row c2r      x^T(16n x 16m) = T x (w Y)^T   (w = 1 at I = 0, else 2)
//   MASKCOS2     in registers (ring mean from a first pass over ring tiles)
//   row r2c      Z^T (64 x 16m) = T^T x x'
//   column r2c   out(64 x 64) = T^T x Z         (C +- i S' gives +-J)
// One workgroup holds one item: T, Y/Z and the small operands live in LDS.
// T is held folded (rows 0 .. N/2: row N - n of T is row n with its sine
// columns negated), and the row passes use the fold: pixel rows n and N - n
// come from the same cosine and sine partial sums (x = E +- O) and go back
// through the same table row (cosine columns take y[n] + y[N-n], sine
// columns y[n] - y[N-n]), which halves their matrix-core work.
// The row-column order and the c2r conventions (Im ignored at I = 0) are
// those of the FFT path, so the result agrees with it to FP32 rounding.
namespace {

constexpr int kFT = 512;     // threads: 8 waves (the column r2c maps waves to its 8 tile pairs)
constexpr int kFK = 64;      // reals per row: 32 frequencies
constexpr int kFP = 68;      // LDS row pitch (floats): conflict-free MFMA operand reads
constexpr int kFJ = 31;      // |J| <= 31
constexpr float kPi = 3.1415926535897f;  // MASKCOS2's PI

typedef __attribute__((__vector_size__(4 * sizeof(float)))) float fx4;

__device__ inline fx4 mfma4(float a, float b, fx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// squared wrapped distance range of the 16 coordinates v0..v0+15 from the
// corner origin (MASKCOS2 with origin 1); tiles never straddle NSAM/2
__device__ inline void tile_r2(int v0, int nsamh, int nsam, int& lo, int& hi) {
  int a = v0 < nsamh ? v0 : v0 - nsam, b = a + 15;
  int la = a * a, lb = b * b;
  lo = (a <= 0 && b >= 0) ? 0 : (la < lb ? la : lb);
  hi = la > lb ? la : lb;
}

// words of the band point table in LDS (header and at most 32 x 63 points),
// even so that the float2 band after it is 8-byte aligned
constexpr int cc3m_pts_words() { return (kCc3mPtsHdr + 32 * (2 * kFJ + 1) + 1) & ~1; }

// kWin: the score at every shift of a window around the item's shift (the
// best of them and its index), instead of at the item's shift alone
template <bool kWin>
__global__ void __launch_bounds__(kFT)
    k_cc3m_fused(int nsam, int nitem, int maxr12, int maxr22, float rbfact,
                 const float2* __restrict__ a3df,
                 const float* __restrict__ dm, const int* __restrict__ pidx,
                 const float2* __restrict__ ctff, const float2* __restrict__ spec,
                 const float* __restrict__ shift, const float* __restrict__ amag_arr,
                 const float* __restrict__ fsct, const float* __restrict__ tab,
                 const int* __restrict__ pts, const float* __restrict__ ring_tab, int ring_cnt,
                 const float* __restrict__ wmap, float ri2,
                 float ri3, float rih, float halfw, float* score, Cc3mWindow win,
                 int* __restrict__ wcode, int flip) {
  extern __shared__ __attribute__((aligned(16))) float lds[];
  const int hrow = nsam / 2 + 1;
  float* Ts = lds;                      // [nsam/2 + 1][kFP]: the folded table TF
  float* Ys = Ts + hrow * kFP;          // [nsam][kFP]: Y, then Z (columns Re | Im)
  float* Bs = Ys + nsam * kFP;          // [64][kFP]: B(X)^T, then out
  float* bins = Bs + 64 * kFP;          // [4][nb]
  const int nb = nsam / 2 + 2 > kCc3mShells ? nsam / 2 + 2 : kCc3mShells;
  int* pts_s = reinterpret_cast<int*>(bins + 4 * nb);  // the band point table
  // the particle's spectrum at the band points, [I][J + 31], read once in the
  // band fill (its loads in flight with the gathers) for the shell sums
  float2* Sp = reinterpret_cast<float2*>(pts_s + cc3m_pts_words());
  __shared__ float red_s[kFT / 64];
  __shared__ int red_c[kFT / 64];
  // e^{i SHX I} and e^{i SHY J} of the item: the shift phase of a point is
  // their product
  __shared__ float2 phx[32], phy[2 * kFJ + 1];

  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int la = lane & 15, lg = lane >> 4;
  const int jc = nsam / 2 + 1, nsamh = nsam / 2;
  const size_t ic = (size_t)nsam * jc;

  // T[n][k] (k = 2i: cos, 2i + 1: -sin) for any row n, from the folded table
  auto tval = [&](int n, int k) -> float {
    const bool up = n > nsamh;
    const float v = Ts[(up ? nsam - n : n) * kFP + (k & 1) * (kFK / 2) + (k >> 1)];
    return (up && (k & 1)) ? -v : v;
  };

  // ---- stage TF once; the workgroup then takes items blockIdx.x, +gridDim.x, ...
  for (int e = tid; e < hrow * (kFK / 4); e += kFT) {
    int n = e / (kFK / 4), q = e % (kFK / 4);
    *reinterpret_cast<float4*>(Ts + n * kFP + 4 * q) =
        *reinterpret_cast<const float4*>(tab + (size_t)n * kFK + 4 * q);
  }
  // the band point table, once (item independent)
  for (int e = tid; e < kCc3mPtsHdr + pts[kCc3mShells]; e += kFT) pts_s[e] = pts[e];
  // this thread's entries e = tid + 512 r of the ring's Q table (the same for
  // every item), held in registers across the items
  static_assert(64 * 64 % kFT == 0, "Q entries per thread");
  float qv[64 * 64 / kFT];
#pragma unroll
  for (int r = 0; r < 64 * 64 / kFT; r++) qv[r] = ring_tab ? ring_tab[tid + r * kFT] : 0.f;
  // a contiguous run of items per workgroup: consecutive items of one
  // particle (a particle's candidates) reuse its band spectrum in LDS
  const int per = (nitem + (int)gridDim.x - 1) / (int)gridDim.x;
  const int it0 = blockIdx.x * per, it1 = min(nitem, it0 + per);
  int pprev = -1;
  for (int it = it0; it < it1; it++) {
  const int p = pidx[it];
  const bool newp = p != pprev;
  pprev = p;
  const float amagp = amag_arr[it];
  const float rihp = rih / amagp, ri2p = ri2 / amagp / amagp, ri3p = ri3 / amagp / amagp;

  // ---- fill X (annulus x CTF / NSAM^2) into Ys
  float2* Xs = reinterpret_cast<float2*>(Ys);  // [32][63]: I, J + 31
  if (!kWin && tid < 32 + 2 * kFJ + 1) {
    float sn, cs;
    if (tid < 32) {
      sincosf(shift[2 * it] * (float)tid, &sn, &cs);
      phx[tid] = make_float2(cs, sn);
    } else {
      sincosf(shift[2 * it + 1] * (float)(tid - 32 - kFJ), &sn, &cs);
      phy[tid - 32] = make_float2(cs, sn);
    }
  }
  {
    const float* d = dm + 9 * (size_t)it;
    const float2* cp = ctff + (size_t)p * 2 * ic;
    const float2* sps = spec + (size_t)p * ic;
    const float scal = 1.0f / (float)nsam / (float)nsam;
    // unrolled: the gathers of a thread's points are in flight together
#pragma unroll 4
    for (int e = tid; e < 32 * (2 * kFJ + 1); e += kFT) {
      int i = e / (2 * kFJ + 1), j = e % (2 * kFJ + 1) - kFJ;
      int itemp = i * i + j * j;
      float2 out = make_float2(0.f, 0.f);
      if (itemp >= maxr12 && itemp < maxr22) {
        size_t id = (size_t)i + (size_t)jc * wrapn(j, nsam);
        cf a = ainterpo_dev(nsam, a3df, d, i, j);
        float2 cr = cp[id], cl = cp[id + ic];
        if (newp) Sp[e] = sps[id];
        if (flip) {  // cc3.f:86-89
          if (cr.x + cl.x < 0.0f) a = cf{-a.re, -a.im};
        } else {
          a = cmul(a, cf{cr.x + cl.x, cr.y - cl.y});
        }
        out = make_float2(a.re * scal, a.im * scal);
      }
      Xs[e] = out;
    }
  }
  __syncthreads();
  // B(X)^T[c][k]: column c = (I, re|im), k = (|J|, cos|sin).  At AMAG = 1
  // the MASKCOS2 ring sum is linear in B: sum_{c,k} Q[c][k] B^T[c][k] with the
  // precomputed Q (the ring's rows of T through both transforms)
  const bool ringq = ring_tab != nullptr && amagp == 1.0f;
  float rsum = 0.f;
  int rcnt = 0;
#pragma unroll
  for (int r = 0; r < 64 * 64 / kFT; r++) {
    const int e = tid + r * kFT;
    int c = e >> 6, k = e & 63;
    int i = c >> 1, jj = k >> 1;
    float2 xp = Xs[i * (2 * kFJ + 1) + kFJ + jj];
    float2 xm = Xs[i * (2 * kFJ + 1) + kFJ - jj];
    float2 P = jj ? make_float2(xp.x + xm.x, xp.y + xm.y) : xp;
    float2 Q = jj ? make_float2(xp.x - xm.x, xp.y - xm.y) : make_float2(0.f, 0.f);
    float v;
    if (!(c & 1)) v = (k & 1) ? Q.y : P.x;   // Y_re = P_re T_re + Q_im T_im
    else v = (k & 1) ? -Q.x : P.y;           // Y_im = P_im T_re - Q_re T_im
    Bs[c * kFP + k] = v;
    if (ringq) rsum += qv[r] * v;
  }
  __syncthreads();

  // ---- column c2c: Y[m][c] = sum_k T[m][k] B[k][c], times the c2r weight,
  // folded like the rows: for m < N/2, E = TF[m][0:32] against the cosine
  // rows of B and O = TF[m][32:64] against its sine rows give Y[m] = E + O
  // and Y[N-m] = E - O (row 0 once); row N/2, which has no partner, on the
  // vector ALU.  Folded tiles (16 m x 16 c) go two at a time (four chains).
  const int mt = nsam / 16;
  {
    const int nct = (nsamh / 16) * 4;
    auto ctile = [&](int t, fx4& e, fx4& o) {
      const int m0 = (t >> 2) * 16, c0 = (t & 3) * 16;
#pragma unroll
      for (int ks = 0; ks < kFK / 8; ks++) {
        e = mfma4(Ts[(m0 + la) * kFP + 4 * ks + lg], Bs[(c0 + la) * kFP + 8 * ks + 2 * lg], e);
        o = mfma4(Ts[(m0 + la) * kFP + kFK / 2 + 4 * ks + lg],
                  Bs[(c0 + la) * kFP + 8 * ks + 2 * lg + 1], o);
      }
    };
    auto cstore = [&](int t, const fx4& e, const fx4& o) {
      const int m0 = (t >> 2) * 16, c0 = (t & 3) * 16;
      const float w = (c0 + la) < 2 ? 1.0f : 2.0f;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int m = m0 + 4 * lg + i;
        Ys[m * kFP + c0 + la] = (e[i] + o[i]) * w;
        if (m) Ys[(nsam - m) * kFP + c0 + la] = (e[i] - o[i]) * w;
      }
    };
    for (int t = wave; t < nct; t += 2 * (kFT / 64)) {
      const int t1 = t + kFT / 64;
      fx4 e0 = {0.f, 0.f, 0.f, 0.f}, o0 = e0, e1 = e0, o1 = e0;
      if (t1 < nct) {
        const int m0 = (t >> 2) * 16, c0 = (t & 3) * 16;
        const int m1 = (t1 >> 2) * 16, c1 = (t1 & 3) * 16;
#pragma unroll
        for (int ks = 0; ks < kFK / 8; ks++) {
          e0 = mfma4(Ts[(m0 + la) * kFP + 4 * ks + lg], Bs[(c0 + la) * kFP + 8 * ks + 2 * lg],
                     e0);
          o0 = mfma4(Ts[(m0 + la) * kFP + kFK / 2 + 4 * ks + lg],
                     Bs[(c0 + la) * kFP + 8 * ks + 2 * lg + 1], o0);
          e1 = mfma4(Ts[(m1 + la) * kFP + 4 * ks + lg], Bs[(c1 + la) * kFP + 8 * ks + 2 * lg],
                     e1);
          o1 = mfma4(Ts[(m1 + la) * kFP + kFK / 2 + 4 * ks + lg],
                     Bs[(c1 + la) * kFP + 8 * ks + 2 * lg + 1], o1);
        }
        cstore(t, e0, o0);
        cstore(t1, e1, o1);
      } else {
        ctile(t, e0, o0);
        cstore(t, e0, o0);
      }
    }
    // row N/2: Y[N/2][c] = sum_k T[N/2][k] B[k][c]
    for (int c = tid; c < kFK; c += kFT) {
      float y = 0.f;
      for (int j = 0; j < kFK / 2; j++)
        y += Ts[nsamh * kFP + j] * Bs[c * kFP + 2 * j] +
             Ts[nsamh * kFP + kFK / 2 + j] * Bs[c * kFP + 2 * j + 1];
      Ys[nsamh * kFP + c] = y * (c < 2 ? 1.0f : 2.0f);
    }
  }
  __syncthreads();

  const bool use_wmap = wmap != nullptr && amagp == 1.0f;
  // ---- MASKCOS2 ring mean: x over ri2' < r^2 <= ri3'.  At AMAG = 1 the
  // ring sum was taken from B above; otherwise a pass over the ring tiles on
  // the matrix cores.
  if (ringq) {
    rcnt = tid == 0 ? ring_cnt : 0;
  } else
  for (int rb = wave; rb < mt; rb += kFT / 64) {
    const int m0 = rb * 16;
    int mlo, mhi;
    tile_r2(m0, nsamh, nsam, mlo, mhi);
    float yv[kFK / 4];
#pragma unroll
    for (int ks = 0; ks < kFK / 4; ks++) yv[ks] = Ys[(m0 + la) * kFP + 4 * ks + lg];
    const int m = m0 + la;
    const int dy = m < nsamh ? m : m - nsam;
    for (int nt = 0; nt < mt; nt++) {
      const int n0 = nt * 16;
      int nlo, nhi;
      tile_r2(n0, nsamh, nsam, nlo, nhi);
      if ((float)(nlo + mlo) > ri3p || (float)(nhi + mhi) <= ri2p) continue;
      fx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < kFK / 4; ks++)
        acc = mfma4(tval(n0 + la, 4 * ks + lg), yv[ks], acc);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        int n = n0 + 4 * lg + i, dx = n < nsamh ? n : n - nsam;
        float crad2 = (float)(dx * dx + dy * dy);
        if (crad2 > ri2p && crad2 <= ri3p) {
          rsum += acc[i];
          rcnt++;
        }
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    rsum += __shfl_down(rsum, off, 64);
    rcnt += __shfl_down(rcnt, off, 64);
  }
  if (lane == 0) {
    red_s[wave] = rsum;
    red_c[wave] = rcnt;
  }
  __syncthreads();
  float tot = 0.f;
  int itot = 0;
#pragma unroll
  for (int k = 0; k < kFT / 64; k++) {
    tot += red_s[k];
    itot += red_c[k];
  }
  const float mean = itot ? tot / (float)itot : tot;

  // ---- rows: x = T (wY)^T, MASKCOS2, Z^T = T^T x'.  The mask gives
  // x' = mean + (x - mean) w, with w = 1 inside RI2', the cosine edge in the
  // ring and 0 outside RI3'; since sum_n T[n][k] is N for the DC term (k = 0)
  // and 0 for every other k, Z^T = N mean e_0 + T^T ((x - mean) w), and tiles
  // wholly outside the mask add nothing.  Folded: a tile of table rows
  // r = 16t .. 16t+15 (r < N/2) gives E = TF[r][0:32] wYre and
  // O = TF[r][32:64] wYim, x[r] = E + O, x[N-r] = E - O; y = (x - mean) w
  // (the same w at r and N - r: equal wrapped distances); then
  // Z[m][i] += TF[r][i] (y[r] + y[N-r]), Z[m][32+i] += TF[r][32+i] (y[r] -
  // y[N-r]).  Row 0 is its own partner (counted once); row N/2 has none and
  // is added on the vector ALU.  Z goes back with its columns blocked (Re |
  // Im).  The active tiles of a row block go through the matrix cores two at
  // a time, their four chains interleaved.
  for (int rb = wave; rb < mt; rb += kFT / 64) {
    const int m0 = rb * 16;
    int mlo, mhi;
    tile_r2(m0, nsamh, nsam, mlo, mhi);
    const int m = m0 + la;
    const int dy = m < nsamh ? m : m - nsam;
    // wY of row m in the E/O operand layout: I = 4 ks + lg
    float yre[kFK / 8], yim[kFK / 8];
#pragma unroll
    for (int ks = 0; ks < kFK / 8; ks++) {
      yre[ks] = Ys[m * kFP + 8 * ks + 2 * lg];
      yim[ks] = Ys[m * kFP + 8 * ks + 2 * lg + 1];
    }
    fx4 z[4];
#pragma unroll
    for (int kt = 0; kt < 4; kt++) z[kt] = fx4{0.f, 0.f, 0.f, 0.f};
    // active folded tiles of this row block as a bit set (wave-uniform)
    unsigned act = 0u;
    for (int nt = 0; nt < nsamh / 16; nt++) {
      int nlo, nhi;
      tile_r2(nt * 16, nsamh, nsam, nlo, nhi);
      if (!((float)(nlo + mlo) > ri3p)) act |= 1u << nt;
    }
    // MASKCOS2 weights w of this lane's pixels (n0 + 4 lg + i, m) of a tile:
    // 1 inside RI2', from the precomputed map at AMAG = 1, else per pixel
    auto weights = [&](int n0) -> fx4 {
      int nlo, nhi;
      tile_r2(n0, nsamh, nsam, nlo, nhi);
      if (!((float)(nhi + mhi) > ri2p)) return fx4{1.f, 1.f, 1.f, 1.f};
      if (use_wmap) return *reinterpret_cast<const fx4*>(wmap + (size_t)m * nsam + n0 + 4 * lg);
      fx4 w;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        int n = n0 + 4 * lg + i, dx = n < nsamh ? n : n - nsam;
        float crad2 = (float)(dx * dx + dy * dy);
        float e = 1.0f;
        if (crad2 > ri3p) e = 0.0f;
        else if (crad2 > ri2p) e = (1.0f + cosf(kPi * (sqrtf(crad2) - rihp) / halfw)) / 2.0f;
        w[i] = e;
      }
      return w;
    };
    // E and O of folded tile t on the matrix cores (two independent chains)
    auto eo = [&](int t, fx4& e, fx4& o) {
      e = fx4{0.f, 0.f, 0.f, 0.f};
      o = e;
#pragma unroll
      for (int ks = 0; ks < kFK / 8; ks++) {
        e = mfma4(Ts[(16 * t + la) * kFP + 4 * ks + lg], yre[ks], e);
        o = mfma4(Ts[(16 * t + la) * kFP + kFK / 2 + 4 * ks + lg], yim[ks], o);
      }
    };
    // masked sums and differences of the pixel pairs (r, N - r), r = 16t+4lg+i
    auto fold = [&](int t, const fx4& w, const fx4& e, const fx4& o, fx4& sp, fx4& sm) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const float ya = (e[i] + o[i] - mean) * w[i];
        float yb = (e[i] - o[i] - mean) * w[i];
        if (t == 0 && lg == 0 && i == 0) yb = 0.f;  // r = 0: its own partner
        sp[i] = ya + yb;
        sm[i] = ya - yb;
      }
    };
    // k-step i takes r = 16t + 4 lg + i: the C layout of E, O is the B layout
    auto zacc = [&](int t, const fx4& sp, const fx4& sm) {
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int kt = 0; kt < 4; kt++)
          z[kt] = mfma4(Ts[(16 * t + 4 * lg + i) * kFP + 16 * kt + la], kt < 2 ? sp[i] : sm[i],
                        z[kt]);
    };
    // the table operands of a tile's Z products, loaded ahead of the
    // products and the folds they wait for
    auto zload = [&](int t, float (&ta)[16]) {
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int kt = 0; kt < 4; kt++) ta[4 * i + kt] = Ts[(16 * t + 4 * lg + i) * kFP + 16 * kt + la];
    };
    auto zacc_r = [&](const float (&ta)[16], const fx4& sp, const fx4& sm) {
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int kt = 0; kt < 4; kt++) z[kt] = mfma4(ta[4 * i + kt], kt < 2 ? sp[i] : sm[i], z[kt]);
    };
    while (__builtin_popcount(act) >= 2) {
      const int t0 = __builtin_ctz(act);
      act &= act - 1u;
      const int t1 = __builtin_ctz(act);
      act &= act - 1u;
      fx4 e0, o0, e1, o1, p0, q0, p1, q1;
      float za0[16], za1[16];
      zload(t0, za0);
      zload(t1, za1);
      // the weights first: their loads are in flight under the products
      const fx4 w0 = weights(16 * t0), w1 = weights(16 * t1);
      e0 = fx4{0.f, 0.f, 0.f, 0.f};
      o0 = e0;
      e1 = e0;
      o1 = e0;
#pragma unroll
      for (int ks = 0; ks < kFK / 8; ks++) {
        e0 = mfma4(Ts[(16 * t0 + la) * kFP + 4 * ks + lg], yre[ks], e0);
        o0 = mfma4(Ts[(16 * t0 + la) * kFP + kFK / 2 + 4 * ks + lg], yim[ks], o0);
        e1 = mfma4(Ts[(16 * t1 + la) * kFP + 4 * ks + lg], yre[ks], e1);
        o1 = mfma4(Ts[(16 * t1 + la) * kFP + kFK / 2 + 4 * ks + lg], yim[ks], o1);
      }
      fold(t0, w0, e0, o0, p0, q0);
      fold(t1, w1, e1, o1, p1, q1);
      zacc_r(za0, p0, q0);
      zacc_r(za1, p1, q1);
    }
    if (act) {
      const int t0 = __builtin_ctz(act);
      fx4 e0, o0, p0, q0;
      const fx4 w0 = weights(16 * t0);
      eo(t0, e0, o0);
      fold(t0, w0, e0, o0, p0, q0);
      zacc(t0, p0, q0);
    }
    // row N/2: x = sum_I (-1)^I wYre[m][I] (its sine terms vanish), added to
    // the cosine rows of Z as (-1)^i y
    {
      const float crad2 = (float)(nsamh * nsamh + dy * dy);
      if (__ballot(crad2 <= ri3p)) {
        float xs = 0.f;
#pragma unroll
        for (int ks = 0; ks < kFK / 8; ks++) xs += yre[ks];
        xs = (lg & 1) ? -xs : xs;  // I = 4 ks + lg: the sign of lane group lg
        xs += __shfl_xor(xs, 16, 64);
        xs += __shfl_xor(xs, 32, 64);
        float w = 1.0f;
        if (crad2 > ri3p) w = 0.0f;
        else if (use_wmap) w = wmap[(size_t)m * nsam + nsamh];
        else if (crad2 > ri2p) w = (1.0f + cosf(kPi * (sqrtf(crad2) - rihp) / halfw)) / 2.0f;
        const float y = (xs - mean) * w;
#pragma unroll
        for (int kt = 0; kt < 2; kt++)
#pragma unroll
          for (int i = 0; i < 4; i++) z[kt][i] += (i & 1) ? -y : y;
      }
    }
    if (lg == 0) z[0][0] += mean * (float)nsam;  // row k = 0, the DC term
    // Z[m][k]: this wave alone read rows m0..m0+15 of Ys
#pragma unroll
    for (int kt = 0; kt < 4; kt++)
#pragma unroll
      for (int i = 0; i < 4; i++) Ys[(m0 + la) * kFP + 16 * kt + 4 * lg + i] = z[kt][i];
  }
  __syncthreads();

  // ---- column r2c: out[r][c] = sum_m T[m][r] Z[m][c]  (into Bs), folded over
  // m: the cosine rows r = j take P[m] = Z[m] + Z[N-m] and the sine rows
  // r = 32 + j take M[m] = Z[m] - Z[N-m] against TF rows m = 0 .. N/2 (P = Z
  // and M = 0 at m = 0 and N/2, their own partners).  out is written with its
  // rows blocked (cos | sin) like its columns (Re | Im).  Wave w takes the
  // cosine and sine tiles of rows j0 = 16 (w >> 2) and columns 16 (w & 3).
  static_assert(kFT == 512, "eight waves: one (cosine, sine) tile pair each");
  {
    const int c0 = (wave & 3) * 16, j0 = (wave >> 2) * 16;
    fx4 ac0 = {0.f, 0.f, 0.f, 0.f}, ac1 = ac0, as0 = ac0, as1 = ac0;
    for (int s4 = 0; s4 <= nsamh / 4; s4 += 2) {
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int m = 4 * (s4 + h) + lg;
        float p = 0.f, q = 0.f, tc = 0.f, ts = 0.f;
        if (m <= nsamh) {
          const float za = Ys[m * kFP + c0 + la];
          if (m == 0 || m == nsamh) {
            p = za;
          } else {
            const float zb = Ys[(nsam - m) * kFP + c0 + la];
            p = za + zb;
            q = za - zb;
          }
          tc = Ts[m * kFP + j0 + la];
          ts = Ts[m * kFP + kFK / 2 + j0 + la];
        }
        if (h) {
          ac1 = mfma4(tc, p, ac1);
          as1 = mfma4(ts, q, as1);
        } else {
          ac0 = mfma4(tc, p, ac0);
          as0 = mfma4(ts, q, as0);
        }
      }
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      Bs[(j0 + 4 * lg + i) * kFP + c0 + la] = ac0[i] + ac1[i];
      Bs[(kFK / 2 + j0 + 4 * lg + i) * kFP + c0 + la] = as0[i] + as1[i];
    }
  }
  __syncthreads();

  if constexpr (!kWin) {
  // ---- cc3m.f:117-150: shifted particle x conj(masked projection), shells.
  // Each wave owns a fixed set of shells (zig-zag, so the point counts
  // balance), walks their points from the sorted table and reduces them with
  // a fixed butterfly: no atomics, the same sums on every run.
  {
    for (int sh = 0; sh < kCc3mShells; sh++) {
      const int z = sh & 15;
      if ((z < 8 ? z : 15 - z) != wave) continue;
      const int b0 = pts_s[sh], b1 = pts_s[sh + 1];
      float q = 0.f, s2 = 0.f, a2 = 0.f;
#pragma unroll 4
      for (int e = b0 + lane; e < b1; e += 64) {
        const int pk = pts_s[kCc3mPtsHdr + e];
        const int i = pk & 0xff, j = (pk >> 8) - 64;
        int jj = j < 0 ? -j : j;
        float cre = Bs[jj * kFP + i], cim = Bs[jj * kFP + kFK / 2 + i];
        float sre = Bs[(kFK / 2 + jj) * kFP + i], sim = Bs[(kFK / 2 + jj) * kFP + kFK / 2 + i];
        float2 a = j >= 0 ? make_float2(cre - sim, cim + sre) : make_float2(cre + sim, cim - sre);
        const float2 ex = phx[i], ey = phy[j + kFJ];
        const float cs = ex.x * ey.x - ex.y * ey.y, sn = ex.x * ey.y + ex.y * ey.x;
        const float2 sv = Sp[i * (2 * kFJ + 1) + j + kFJ];
        cf samp = cmul(cf{sv.x, sv.y}, cf{cs, sn});
        q += samp.re * a.x + samp.im * a.y;
        s2 += samp.re * samp.re + samp.im * samp.im;
        a2 += a.x * a.x + a.y * a.y;
      }
      for (int off = 32; off > 0; off >>= 1) {
        q += __shfl_xor(q, off, 64);
        s2 += __shfl_xor(s2, off, 64);
        a2 += __shfl_xor(a2, off, 64);
      }
      if (lane == 0) {
        bins[4 * sh + 0] = q;
        bins[4 * sh + 1] = s2;
        bins[4 * sh + 2] = a2;
        bins[4 * sh + 3] = (float)(b1 - b0);
      }
    }
  }
  __syncthreads();
  // cc3m.f:139-167 over shells 1 .. NSAM/2-1 (the band holds shells below
  // kCc3mShells): one wave, FP64 sums reduced by a fixed butterfly
  if (wave == 0) {
    double s1 = 0.0, s2 = 0.0, s3 = 0.0;
    const int top = nsam / 2 < kCc3mShells ? nsam / 2 : kCc3mShells;
    for (int i = 2 + lane; i <= top; i += 64) {
      const float* b = bins + 4 * (i - 1);
      if (b[3] != 0.0f) {
        float r = b[0] / sqrtf(b[2]);
        if (!flip && fsct[i - 1] < 0.95f) r = fabsf(r);
        float wgt = expf(-rbfact * (float)((i - 1) * (i - 1)));
        s3 += (double)(r * wgt);
        s2 += (double)(wgt * wgt * b[1]);
        s1 += 1.0;
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      s1 += __shfl_xor(s1, off, 64);
      s2 += __shfl_xor(s2, off, 64);
      s3 += __shfl_xor(s3, off, 64);
    }
    if (lane == 0) {
      if (s1 * s2 != 0.0) s3 = s3 / sqrt(s1 * s2);
      score[it] = (float)s3;
    }
  }
  } else {
  // ---- cc3m.f:117-167 at every shift s of the window.  The shell sums of
  // SAMP conj(AINPO) are Q_k(s) = Re sum C e^{i s.(I,J)} with
  // C = SPEC conj(AINPO); the other sums do not depend on s (|SAMP| = |SPEC|).
  // C goes to LDS (the Y/Z buffer is free), the shell sums of |SPEC|^2 and
  // |AINPO|^2 to bins; then each wave takes its shells with one lane per shift
  // (two for windows above 64 shifts) and the phase e^{i sx I} e^{i sy J} from
  // two small tables, and adds R_k WGT into its FP64 partial per shift.
  float2* Cs = reinterpret_cast<float2*>(Ys);
  {
    for (int sh = 0; sh < kCc3mShells; sh++) {
      const int z = sh & 15;
      if ((z < 8 ? z : 15 - z) != wave) continue;
      const int b0 = pts_s[sh], b1 = pts_s[sh + 1];
      float s2 = 0.f, a2 = 0.f;
      for (int e = b0 + lane; e < b1; e += 64) {
        const int pk = pts_s[kCc3mPtsHdr + e];
        const int i = pk & 0xff, j = (pk >> 8) - 64;
        int jj = j < 0 ? -j : j;
        float cre = Bs[jj * kFP + i], cim = Bs[jj * kFP + kFK / 2 + i];
        float sre = Bs[(kFK / 2 + jj) * kFP + i], sim = Bs[(kFK / 2 + jj) * kFP + kFK / 2 + i];
        float2 a = j >= 0 ? make_float2(cre - sim, cim + sre) : make_float2(cre + sim, cim - sre);
        const float2 sv = Sp[i * (2 * kFJ + 1) + j + kFJ];
        Cs[e] = make_float2(sv.x * a.x + sv.y * a.y, sv.y * a.x - sv.x * a.y);
        s2 += sv.x * sv.x + sv.y * sv.y;
        a2 += a.x * a.x + a.y * a.y;
      }
      for (int off = 32; off > 0; off >>= 1) {
        s2 += __shfl_xor(s2, off, 64);
        a2 += __shfl_xor(a2, off, 64);
      }
      if (lane == 0) {
        bins[4 * sh + 1] = s2;
        bins[4 * sh + 2] = a2;
        bins[4 * sh + 3] = (float)(b1 - b0);
      }
    }
  }
  __syncthreads();
  const int nsx = win.nsx, nsy = win.nsy, ns = nsx * nsy;
  float2* Ex = reinterpret_cast<float2*>(Bs);          // [I][ix]
  float2* Ey = Ex + 32 * kCc3mWinMax;                   // [J + 31][iy]
  double* part = reinterpret_cast<double*>(Ey + (2 * kFJ + 1) * kCc3mWinMax);  // [wave][s]
  {
    const float shx0 = shift[2 * it], shy0 = shift[2 * it + 1];
    for (int e = tid; e < 32 * nsx; e += kFT) {
      const int i = e / nsx, ix = e % nsx;
      const float sx = shx0 + (float)(ix - nsx / 2) * win.dsx;
      float sn, cs;
      sincosf(sx * (float)i, &sn, &cs);
      Ex[i * kCc3mWinMax + ix] = make_float2(cs, sn);
    }
    for (int e = tid; e < (2 * kFJ + 1) * nsy; e += kFT) {
      const int jj = e / nsy, iy = e % nsy;
      const float sy = shy0 + (float)(iy - nsy / 2) * win.dsy;
      float sn, cs;
      sincosf(sy * (float)(jj - kFJ), &sn, &cs);
      Ey[jj * kCc3mWinMax + iy] = make_float2(cs, sn);
    }
  }
  __syncthreads();
  const int top = nsam / 2 < kCc3mShells ? nsam / 2 : kCc3mShells;
  {
    const int s0 = lane, s1 = lane + 64;
    const bool v0 = s0 < ns, v1 = s1 < ns;
    const int ix0 = v0 ? s0 / nsy : 0, iy0 = v0 ? s0 % nsy : 0;
    const int ix1 = v1 ? s1 / nsy : 0, iy1 = v1 ? s1 % nsy : 0;
    double acc0 = 0.0, acc1 = 0.0;
    for (int sh = 1; sh < top; sh++) {
      const int z = sh & 15;
      if ((z < 8 ? z : 15 - z) != wave) continue;
      const float cnt = bins[4 * sh + 3];
      if (cnt == 0.0f) continue;
      const int b0 = pts_s[sh], b1 = pts_s[sh + 1];
      float q0 = 0.f, q1 = 0.f;
      for (int e = b0; e < b1; e++) {
        const int pk = pts_s[kCc3mPtsHdr + e];
        const int i = pk & 0xff, jj = (pk >> 8) - 64 + kFJ;
        const float2 c = Cs[e];
        {
          const float2 ex = Ex[i * kCc3mWinMax + ix0], ey = Ey[jj * kCc3mWinMax + iy0];
          const float er = ex.x * ey.x - ex.y * ey.y, ei = ex.x * ey.y + ex.y * ey.x;
          q0 += c.x * er - c.y * ei;
        }
        if (v1) {
          const float2 ex = Ex[i * kCc3mWinMax + ix1], ey = Ey[jj * kCc3mWinMax + iy1];
          const float er = ex.x * ey.x - ex.y * ey.y, ei = ex.x * ey.y + ex.y * ey.x;
          q1 += c.x * er - c.y * ei;
        }
      }
      const float ra = sqrtf(bins[4 * sh + 2]);
      float r0 = q0 / ra, r1 = q1 / ra;
      if (!flip && fsct[sh] < 0.95f) {
        r0 = fabsf(r0);
        r1 = fabsf(r1);
      }
      const float wgt = expf(-rbfact * (float)(sh * sh));
      acc0 += (double)(r0 * wgt);
      acc1 += (double)(r1 * wgt);
    }
    if (v0) part[wave * 128 + s0] = acc0;
    if (v1) part[wave * 128 + s1] = acc1;
  }
  __syncthreads();
  if (wave == 0) {
    double t1 = 0.0, t2 = 0.0;
    for (int i = 2 + lane; i <= top; i += 64) {
      const float* b = bins + 4 * (i - 1);
      if (b[3] != 0.0f) {
        const float wgt = expf(-rbfact * (float)((i - 1) * (i - 1)));
        t2 += (double)(wgt * wgt * b[1]);
        t1 += 1.0;
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      t1 += __shfl_xor(t1, off, 64);
      t2 += __shfl_xor(t2, off, 64);
    }
    float bv = -3.0e38f;
    int bs = 0x7fffffff;
    for (int s = lane; s < ns; s += 64) {
      double s3 = 0.0;
      for (int w = 0; w < kFT / 64; w++) s3 += part[w * 128 + s];
      if (t1 * t2 != 0.0) s3 = s3 / sqrt(t1 * t2);
      const float v = (float)s3;
      if (v > bv) {
        bv = v;
        bs = s;
      }
    }
    // the first maximum in window order (IX outer, IY inner)
    for (int off = 32; off > 0; off >>= 1) {
      const float ov = __shfl_xor(bv, off, 64);
      const int os = __shfl_xor(bs, off, 64);
      if (ov > bv || (ov == bv && os < bs)) {
        bv = ov;
        bs = os;
      }
    }
    if (lane == 0) {
      score[it] = bv;
      wcode[it] = bs < ns ? bs : 0;
    }
  }
  }
  __syncthreads();  // bins and LDS buffers are reused by the next item
  }
}

}  // namespace

bool cc3m_fused_ok(int nsam, int maxr2) {
  return nsam % 32 == 0 && nsam >= 64 && nsam <= 256 && maxr2 <= 32;
}

size_t cc3m_fused_lds(int nsam) {
  const int nb = nsam / 2 + 2 > kCc3mShells ? nsam / 2 + 2 : kCc3mShells;
  // TF, Y/Z, B/out, the shell bins, the band point table (header and at
  // most 32 x 63 points, padded to 8 bytes) and the particle's band
  return sizeof(float) * ((size_t)(nsam / 2 + 1) * kFP + (size_t)nsam * kFP + 64 * kFP + 4 * nb) +
         sizeof(int) * cc3m_pts_words() + sizeof(float2) * 32 * (2 * kFJ + 1);
}

// T[n][2i] = cos(2 pi i n/N), T[n][2i+1] = -sin(2 pi i n/N) for i < 32, from
// the exact residue (i n) mod N in double precision.
void cc3m_fused_table(int nsam, std::vector<float>& tab) {
  tab.assign((size_t)nsam * kFK, 0.f);
  for (int n = 0; n < nsam; n++)
    for (int i = 0; i < kFK / 2; i++) {
      double ph = 2.0 * 3.14159265358979323846 * (double)((i * n) % nsam) / (double)nsam;
      tab[(size_t)n * kFK + 2 * i] = (float)std::cos(ph);
      tab[(size_t)n * kFK + 2 * i + 1] = (float)-std::sin(ph);
    }
}

// The folded table of the kernel: TF[r][i] = cos(2 pi i r/N) and
// TF[r][32 + i] = -sin(2 pi i r/N) for r = 0 .. N/2, i < 32 (rows N/2+1 .. N-1
// of T are rows N - r of it, the sine columns negated).
void cc3m_fused_table_folded(int nsam, std::vector<float>& tab) {
  const int h = nsam / 2;
  tab.assign((size_t)(h + 1) * kFK, 0.f);
  for (int r = 0; r <= h; r++)
    for (int i = 0; i < kFK / 2; i++) {
      double ph = 2.0 * 3.14159265358979323846 * (double)((i * r) % nsam) / (double)nsam;
      tab[(size_t)r * kFK + i] = (float)std::cos(ph);
      tab[(size_t)r * kFK + kFK / 2 + i] = (float)-std::sin(ph);
    }
}

// The band points of cc3m.f:117-150 (MAXR12 <= I*I+J*J < MAXR22, I < 32,
// |J| <= 31) grouped by shell INT(SQRT(FLOAT(ITEMP))+0.5), in the (I, J)
// loop order within a shell.
void cc3m_fused_points(int maxr12, int maxr22, std::vector<int>& pts) {
  std::vector<std::vector<int>> by(kCc3mShells);
  for (int i = 0; i < 32; i++)
    for (int j = -kFJ; j <= kFJ; j++) {
      int itemp = i * i + j * j;
      if (itemp < maxr12 || itemp >= maxr22) continue;
      int sh = (int)(sqrtf((float)itemp) + 0.5f);
      if (sh < kCc3mShells) by[sh].push_back(i | ((j + 64) << 8));
    }
  pts.assign(kCc3mPtsHdr, 0);
  for (int sh = 0; sh < kCc3mShells; sh++) {
    pts[sh] = (int)(pts.size() - kCc3mPtsHdr);
    pts.insert(pts.end(), by[sh].begin(), by[sh].end());
  }
  pts[kCc3mShells] = (int)(pts.size() - kCc3mPtsHdr);
}

// R[m][k] = sum of T[n][k] over the pixels (n, m) of the MASKCOS2 ring
// RI2 < r^2 <= RI3 (wrapped distances from the origin, as the kernel tests
// them at AMAG = 1), summed in double; returns the ring's pixel count.
int cc3m_fused_ring(int nsam, float ri2, float ri3, const std::vector<float>& dft,
                    std::vector<float>& ring) {
  const int nsamh = nsam / 2;
  std::vector<double> acc((size_t)nsam * kFK, 0.0);
  int cnt = 0;
  for (int m = 0; m < nsam; m++) {
    const int dy = m < nsamh ? m : m - nsam;
    for (int n = 0; n < nsam; n++) {
      const int dx = n < nsamh ? n : n - nsam;
      const float crad2 = (float)(dx * dx + dy * dy);
      if (!(crad2 > ri2 && crad2 <= ri3)) continue;
      cnt++;
      for (int k = 0; k < kFK; k++) acc[(size_t)m * kFK + k] += dft[(size_t)n * kFK + k];
    }
  }
  ring.resize(acc.size());
  for (size_t e = 0; e < acc.size(); e++) ring[e] = (float)acc[e];
  return cnt;
}

// The ring sum of the kernel as a function of B (the column transform's
// input): sum_{(n,m) in ring} x[n][m] = sum_{c,k} Q[c][k] B^T[c][k] with
// Q[c][k] = w_c sum_m T[m][k] R[m][c] (w_c = 1 for c < 2, else 2: the c2r
// weight on Y), R from cc3m_fused_ring; in double, q is 64 x 64 [c][k].
int cc3m_fused_ring_q(int nsam, float ri2, float ri3, std::vector<float>& q) {
  std::vector<float> dft, ring;
  cc3m_fused_table(nsam, dft);
  const int cnt = cc3m_fused_ring(nsam, ri2, ri3, dft, ring);
  std::vector<double> t((size_t)nsam * kFK);
  for (int m = 0; m < nsam; m++)
    for (int i = 0; i < kFK / 2; i++) {
      double ph = 2.0 * 3.14159265358979323846 * (double)((i * m) % nsam) / (double)nsam;
      t[(size_t)m * kFK + 2 * i] = std::cos(ph);
      t[(size_t)m * kFK + 2 * i + 1] = -std::sin(ph);
    }
  q.assign((size_t)kFK * kFK, 0.f);
  for (int c = 0; c < kFK; c++)
    for (int k = 0; k < kFK; k++) {
      double acc = 0.0;
      for (int m = 0; m < nsam; m++) acc += t[(size_t)m * kFK + k] * (double)ring[(size_t)m * kFK + c];
      q[(size_t)c * kFK + k] = (float)((c < 2 ? 1.0 : 2.0) * acc);
    }
  return cnt;
}

// MASKCOS2 (maskcos2.f) weights at AMAG = 1 for every pixel (n, m) of the
// NSAM x NSAM image, indexed [m][n] with wrapped distances from the origin:
// 1 for r2 <= RI2, (1 + COS(PI (SQRT(r2) - RIH) / HALFW)) / 2 for
// RI2 < r2 <= RI3, 0 beyond.
void cc3m_fused_wmap(int nsam, float ri2, float ri3, float rih, float halfw,
                     std::vector<float>& w) {
  const int nsamh = nsam / 2;
  w.assign((size_t)nsam * nsam, 0.f);
  for (int m = 0; m < nsam; m++) {
    const int dy = m < nsamh ? m : m - nsam;
    for (int n = 0; n < nsam; n++) {
      const int dx = n < nsamh ? n : n - nsam;
      const float crad2 = (float)(dx * dx + dy * dy);
      float e = 1.0f;
      if (crad2 > ri3) e = 0.0f;
      else if (crad2 > ri2) e = (1.0f + cosf(kPi * (sqrtf(crad2) - rih) / halfw)) / 2.0f;
      w[(size_t)m * nsam + n] = e;
    }
  }
}

int launch_cc3m_fused(Ctx& c, const Cc3mArgs& a, const float* a3df, const float* spec,
                      const float* ctff, const int* pidx, const float* dm, const float* shift,
                      const float* amag, const float* fsct, const Cc3mTables& t, float* score,
                      int blocks, const Cc3mWindow* win, int* wcode) {
  if (a.n <= 0) return 0;
  FR_CHECK(cc3m_fused_ok(a.nsam, a.maxr2), "fused CC3M needs NSAM %% 32 == 0, NSAM <= 256, MAXR2 <= 32");
  FR_CHECK(t.dft && t.pts, "fused CC3M tables missing");
  const size_t lds = cc3m_fused_lds(a.nsam);
  Cc3mWindow w{1, 1, 0.f, 0.f};
  if (win) {
    FR_CHECK(wcode != nullptr, "shift window without an index output");
    FR_CHECK(win->nsx >= 1 && win->nsy >= 1 && win->nsx <= kCc3mWinMax &&
                 win->nsy <= kCc3mWinMax && win->nsx % 2 == 1 && win->nsy % 2 == 1,
             "shift window %d x %d: odd sizes up to %d", win->nsx, win->nsy, kCc3mWinMax);
    w = *win;
  }
  const void* fn = win ? (const void*)k_cc3m_fused<true> : (const void*)k_cc3m_fused<false>;
  FR_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  // persistent grid: `blocks` workgroups (0: one per item; below 0: one full
  // wave of workgroups), each looping over items, so that the tables are
  // staged once per workgroup rather than once per item
  int grid = (blocks > 0 && blocks < a.n) ? blocks : a.n;
  if (blocks < 0) {
    if (c.cus <= 0) {
      hipDeviceProp_t prop;
      FR_HIP(hipGetDeviceProperties(&prop, c.device));
      c.cus = prop.multiProcessorCount;
    }
    int per_cu = 0;
    FR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kFT, lds));
    const long long full = (long long)c.cus * (per_cu > 0 ? per_cu : 1);
    if (full < a.n) grid = (int)full;
  }
  if (win)
    hipLaunchKernelGGL(k_cc3m_fused<true>, dim3((unsigned)grid), dim3(kFT), lds, c.stream, a.nsam,
                       a.n, a.maxr1 * a.maxr1, a.maxr2 * a.maxr2, a.rbfact,
                       reinterpret_cast<const float2*>(a3df), dm, pidx,
                       reinterpret_cast<const float2*>(ctff),
                       reinterpret_cast<const float2*>(spec), shift, amag, fsct, t.dft, t.pts,
                       t.ring, t.ring_cnt, t.wmap, a.ri2, a.ri3, a.rih, a.halfw, score, w, wcode,
                       a.cc3);
  else
    hipLaunchKernelGGL(k_cc3m_fused<false>, dim3((unsigned)grid), dim3(kFT), lds, c.stream,
                       a.nsam, a.n, a.maxr1 * a.maxr1, a.maxr2 * a.maxr2, a.rbfact,
                       reinterpret_cast<const float2*>(a3df), dm, pidx,
                       reinterpret_cast<const float2*>(ctff),
                       reinterpret_cast<const float2*>(spec), shift, amag, fsct, t.dft, t.pts,
                       t.ring, t.ring_cnt, t.wmap, a.ri2, a.ri3, a.rih, a.halfw, score, w,
                       (int*)nullptr, a.cc3);
  FR_HIP(hipGetLastError());
  return 0;
}

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
                     reinterpret_cast<const float2*>(ctff), work, a.cc3);
  FR_HIP(hipGetLastError());
  if (int r = fft_c2r_2d(c, nsam, a.n, work, img)) return r;
  // cc3m.f:114: MASKCOS2 with AMAG of each item (scal 1: already scaled)
  hipLaunchKernelGGL(k_scale_mask, dim3(a.n), dim3(kMaskThreads), 0, c.stream, nsam, work, img, 1.0f,
                     amag, 1, a.ri2, a.ri3, a.rih, a.halfw, 1);
  FR_HIP(hipGetLastError());
  if (int r = fft_r2c_2d(c, nsam, a.n, work, img)) return r;
  hipLaunchKernelGGL(k_cc3m_score, dim3(a.n), dim3(kThreads), 0, c.stream, nsam,
                     a.maxr1 * a.maxr1, a.maxr2 * a.maxr2, a.rbfact,
                     reinterpret_cast<const float2*>(spec), pidx, shift, fsct, work, score, a.cc3);
  FR_HIP(hipGetLastError());
  return 0;
}

// PRESB (presb.f:8-160) for a batch: the CC3M pipeline over 1 <= r^2 < (NSAM/2)^2
int launch_presb(Ctx& c, int nsam, int n, const float* a3df, const float* spec, const float* ctff,
                 const int* pidx, const float* dm, const float* shift, const float* amag,
                 float ri2, float ri3, float rih, float halfw, float* work, float* rbin) {
  if (n <= 0) return 0;
  FR_CHECK(nsam / 2 + 2 <= 260, "box too large for PRESB bins");
  const int nsamh = nsam / 2;
  const size_t ic = (size_t)nsam * (nsam / 2 + 1);
  const size_t img = (size_t)nsam * (nsam + 2);
  size_t tot = ic * n;
  hipLaunchKernelGGL(k_cc3m_fill, dim3((unsigned)((tot + kThreads - 1) / kThreads)), dim3(kThreads),
                     0, c.stream, nsam, n, 1, nsamh * nsamh, reinterpret_cast<const float2*>(a3df),
                     dm, pidx, reinterpret_cast<const float2*>(ctff), work, 0);
  FR_HIP(hipGetLastError());
  if (int r = fft_c2r_2d(c, nsam, n, work, img)) return r;
  hipLaunchKernelGGL(k_scale_mask, dim3(n), dim3(kMaskThreads), 0, c.stream, nsam, work, img, 1.0f,
                     amag, 1, ri2, ri3, rih, halfw, 1);
  FR_HIP(hipGetLastError());
  if (int r = fft_r2c_2d(c, nsam, n, work, img)) return r;
  hipLaunchKernelGGL(k_presb_bins, dim3(n), dim3(kThreads), 0, c.stream, nsam,
                     reinterpret_cast<const float2*>(spec), pidx, shift, work, rbin);
  FR_HIP(hipGetLastError());
  return 0;
}

}  // namespace fr

// fr_ccp.hip -- PSEARCH's exhaustive search computed the way the reference
// computes it: CCP (ccp.f:8-275) with PAD (pad.f:8-68) and ROTATE2D
// (rotate2d.f:8-77), for many (particle, view) items at once.
//
// Per item (particle p, view v):
//   1. the CTF-multiplied central section on MAXR1^2 <= r^2 < MAXR2^2
//      (ccp.f:105-131), times SCAL = 1/NSAM^2 (pad.f:28-31);
//   2. PAD's padded 4 NSAM transform, evaluated only where ROTATE2D can read
//      it for the correlation annulus (|L|, |M| <= 4 MAXR2).  The padded image
//      is CIRC (the mean of the image's central cross, pad.f:36-45) except at
//      the NSAM^2 pixels, so its transform is
//        F(L,M) = sum_k H(k) D(4kx-L) D(4ky-M) - CIRC (D(-L) D(-M) + B(L,M))
//      with H the section's Hermitian extension (the c2r's view of it),
//      D(m) = sum_{q=-N/2}^{N/2-1} e^{2 pi i q m/(4N)} and B the transform of
//      the pixels that pad.f:124-140 moves with -CIRC (their source slot sits
//      in a row-padding column of the 4 NSAM array, which the +CIRC loop skips).
//      Two small complex products per item (the section's band is 63 x 63 at
//      MAXR2 = 32), FP32;
//   3. for each in-plane step PSIS (ccp.f:144-153) the nearest-neighbour
//      rotation of ROTATE2D (NINT of the FP32 rotated coordinates, host cos/sin),
//      then for each quadrant the index permutation of ccp.f:175-209, the cross
//      spectrum with the B-factor weighted particle, and the NSAMR-grid
//      correlation map at the 15 x 15 shifts CCP reads (-7..7: the -6..6 argmax
//      window and its 3x3 box), as a separable direct DFT (the J phase, then
//      the I phase, with the c2r weights 1 / 2);
//   4. CCP's epilogue: argmax over IX (outer) x IY, first maximum; 3x3 box sum
//      / sqrt(SUM) / sqrt(SUM2) / 9 (ccp.f:223-252); the first best (IPSI,
//      IQUAD) in the reference's loop order.
// PSEARCH's top-IPMAX (strict '>', CC > 0) is the shared k_topk.
#include <math.h>

#include <vector>

#include "fr_ccp.h"
#include "fr_extract.h"

namespace fr {
namespace {

constexpr int kThreads = 256;
constexpr int kWin = 15;  // shifts -7..7 on the NSAMR grid
constexpr int kMc = 64;   // M rows per chunk of the padded-transform products

__device__ __forceinline__ float2 cmulf(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

// ------------------------------------------------------------------ PAD band
// One workgroup per item.  LDS: H (2R-1)^2, G (2R-1) x kMc, D (16R+1).
__global__ void __launch_bounds__(kThreads)
    k_ccp_pad(CcpArgs a, const float2* __restrict__ a3df, const float2* __restrict__ ctff,
              const float* __restrict__ dm_all, int item0, const float2* __restrict__ dtab,
              const float2* __restrict__ ktab, const float2* __restrict__ wtab,
              float2* __restrict__ fband) {
  extern __shared__ float2 lds[];
  const int R = a.maxr2, W = 2 * R - 1, R4 = 4 * R;
  const int nL = R4 + 1, nM = 8 * R + 1, nD = 16 * R + 1;
  float2* H = lds;                       // [kx + R-1][ky + R-1]
  float2* G = H + (size_t)W * W;         // [kx + R-1][m]
  float2* D = G + (size_t)W * kMc;       // D(m), m = -8R..8R
  float2* S0 = D + nD;                   // section column I = 0, by J + R-1
  __shared__ float red[kThreads / 64];
  __shared__ float circ_s;
  const int item = item0 + blockIdx.x;
  const int p = item / a.nview;
  const int tid = threadIdx.x;
  const int nsam = a.nsam, jc = nsam / 2 + 1;
  const size_t ic = (size_t)nsam * jc;
  float dm[6];
#pragma unroll
  for (int k = 0; k < 6; k++) dm[k] = dm_all[(size_t)item * 9 + k];
  for (int e = tid; e < W * W; e += kThreads) H[e] = make_float2(0.f, 0.f);
  for (int e = tid; e < nD; e += kThreads) D[e] = dtab[e];
  for (int e = tid; e < W; e += kThreads) S0[e] = make_float2(0.f, 0.f);
  __syncthreads();
  // the section (ccp.f:108-131) times SCAL; H(I,J) = S, H(-I,-J) = conj S for I > 0
  const float2* cr = ctff + (size_t)p * 2 * ic;
  const float2* cl = cr + ic;
  const int lo2 = a.maxr1 * a.maxr1, hi2 = R * R;
  for (int e = tid; e < R * W; e += kThreads) {
    const int i = e / W, j = e % W - (R - 1);
    const int it = i * i + j * j;
    if (it < lo2 || it >= hi2 || i > jc - 2 || abs(j) == jc - 1) continue;
    const size_t id = (size_t)i + (size_t)jc * wrapn(j, nsam);
    cf v = ainterpo_dev(nsam, a3df, dm, i, j);
    float2 c = make_float2(cr[id].x + cl[id].x, cr[id].y - cl[id].y);
    float2 s = cmulf(make_float2(v.re, v.im), c);
    s.x *= a.scal;
    s.y *= a.scal;
    if (i > 0) {
      H[(size_t)(i + R - 1) * W + (j + R - 1)] = s;
      H[(size_t)(-i + R - 1) * W + (-j + R - 1)] = conjf2(s);
    } else {
      S0[j + R - 1] = s;
    }
  }
  __syncthreads();
  // the c2r reads column I = 0 as its Hermitian part
  for (int e = tid; e < W; e += kThreads) {
    const float2 u = S0[e], w = S0[W - 1 - e];
    H[(size_t)(R - 1) * W + e] = make_float2(0.5f * (u.x + w.x), 0.5f * (u.y - w.y));
  }
  __syncthreads();
  // CIRC = Re sum_k H(k) W(k) / (4 NSAM - 4)   (pad.f:36-45)
  float part = 0.f;
  for (int e = tid; e < W * W; e += kThreads) {
    const float2 h = H[e], w = wtab[e];
    part += h.x * w.x - h.y * w.y;
  }
  for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
  if ((tid & 63) == 0) red[tid >> 6] = part;
  __syncthreads();
  if (tid == 0) {
    float s = 0.f;
    for (int k = 0; k < kThreads / 64; k++) s += red[k];
    circ_s = s / (float)(4 * nsam - 4);
  }
  __syncthreads();
  const float circ = circ_s;
  float2* out = fband + (size_t)blockIdx.x * nM * nL;
  for (int m0 = 0; m0 < nM; m0 += kMc) {
    const int mc = min(kMc, nM - m0);
    // G(kx, M) = sum_ky H(kx, ky) D(4 ky - M), M = m0 + m - 4R; 4 kx x 4 m per thread
    for (int t = tid; t < ((W + 3) / 4) * (kMc / 4); t += kThreads) {
      const int kb = (t / (kMc / 4)) * 4, mb = (t % (kMc / 4)) * 4;
      float2 acc[4][4];
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int q = 0; q < 4; q++) acc[u][q] = make_float2(0.f, 0.f);
      for (int ky = 0; ky < W; ky++) {
        float2 h[4], d[4];
#pragma unroll
        for (int u = 0; u < 4; u++) h[u] = (kb + u < W) ? H[(size_t)(kb + u) * W + ky] : make_float2(0.f, 0.f);
#pragma unroll
        for (int q = 0; q < 4; q++) {
          // D index: 4 (ky - (R-1)) - (m0 + mb + q - 4R) + 8R
          int di = 4 * (ky - (R - 1)) - (m0 + mb + q - R4) + 8 * R;
          di = min(max(di, 0), nD - 1);
          d[q] = D[di];
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
          for (int q = 0; q < 4; q++) {
            acc[u][q].x += h[u].x * d[q].x - h[u].y * d[q].y;
            acc[u][q].y += h[u].x * d[q].y + h[u].y * d[q].x;
          }
      }
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int q = 0; q < 4; q++)
          if (kb + u < W) G[(size_t)(kb + u) * kMc + mb + q] = acc[u][q];
    }
    __syncthreads();
    // F(L, M) = sum_kx D(4 kx - L) G(kx, M) - CIRC K(L, M); 4 L x 8 m per thread
    for (int t = tid; t < ((nL + 3) / 4) * (kMc / 8); t += kThreads) {
      const int lb = (t % ((nL + 3) / 4)) * 4, mb = (t / ((nL + 3) / 4)) * 8;
      float2 acc[4][8];
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int q = 0; q < 8; q++) acc[u][q] = make_float2(0.f, 0.f);
      for (int kx = 0; kx < W; kx++) {
        float2 d[4], g[8];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          int di = 4 * (kx - (R - 1)) - (lb + u) + 8 * R;
          di = min(max(di, 0), nD - 1);
          d[u] = D[di];
        }
#pragma unroll
        for (int q = 0; q < 8; q++) g[q] = G[(size_t)kx * kMc + mb + q];
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
          for (int q = 0; q < 8; q++) {
            acc[u][q].x += d[u].x * g[q].x - d[u].y * g[q].y;
            acc[u][q].y += d[u].x * g[q].y + d[u].y * g[q].x;
          }
      }
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const int mrow = m0 + mb + q;
        if (mb + q >= mc) continue;
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int l = lb + u;
          if (l >= nL) continue;
          const float2 k = ktab[(size_t)mrow * nL + l];
          out[(size_t)mrow * nL + l] =
              make_float2(acc[u][q].x - circ * k.x, acc[u][q].y - circ * k.y);
        }
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------ rotate + scan
// One workgroup per item; wave q scores quadrant q+1 of every in-plane step.
__global__ void __launch_bounds__(kThreads)
    k_ccp_scan(CcpArgs a, const float2* __restrict__ spec, const float* __restrict__ bw,
               int item0, const float2* __restrict__ fband, const float2* __restrict__ etab,
               const float2* __restrict__ e1tab, const unsigned long long* __restrict__ region,
               const float* __restrict__ psitab, float* __restrict__ score,
               int* __restrict__ code) {
  extern __shared__ float2 lds[];
  const int R = a.maxr2, W = 2 * R - 1, R4 = 4 * R;
  const int nL = R4 + 1;
  float2* X = lds;                           // [I][J + R-1] weighted particle on the region
  float2* U = X + (size_t)R * W;             // [a][b + R-1] rotated section (ROTATE2D)
  float2* E = U + (size_t)R * W;             // [J + R-1][n2]
  float2* E1 = E + (size_t)W * kWin;         // [I][n1]
  float2* Z = E1 + (size_t)R * kWin;         // [wave][I][n2]
  float* M = (float*)(Z + (size_t)4 * R * kWin);  // [wave][n1 * 15 + n2]
  __shared__ unsigned long long reg[64];
  __shared__ float red[4];
  __shared__ float wbest[4];
  __shared__ int wcode[4];
  const int item = item0 + blockIdx.x;
  const int p = item / a.nview;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int nsam = a.nsam, jc = nsam / 2 + 1;
  const float2* fb = fband + (size_t)blockIdx.x * (8 * R + 1) * nL;
  const float2* sp = spec + (size_t)p * nsam * jc;
  for (int e = tid; e < R; e += kThreads) reg[e] = region[e];
  for (int e = tid; e < W * kWin; e += kThreads) E[e] = etab[e];
  for (int e = tid; e < R * kWin; e += kThreads) E1[e] = e1tab[e];
  __syncthreads();
  // CBUF = SPEC * EXP(-RBFACT r^2) (psearch.f:50-59) on the correlation region
  float s2 = 0.f;
  for (int e = tid; e < R * W; e += kThreads) {
    const int i = e / W, jj = e % W, j = jj - (R - 1);
    float2 x = make_float2(0.f, 0.f);
    if ((reg[i] >> jj) & 1ull) {
      const float2 v = sp[(size_t)i + (size_t)jc * wrapn(j, nsam)];
      const float w = bw[i * i + j * j];
      x = make_float2(v.x * w, v.y * w);
      s2 += x.x * x.x + x.y * x.y;
    }
    X[e] = x;
  }
  for (int off = 32; off > 0; off >>= 1) s2 += __shfl_xor(s2, off, 64);
  if (lane == 0) red[wv] = s2;
  __syncthreads();
  const float sum2 = red[0] + red[1] + red[2] + red[3];
  const int iq = a.iq;
  float best = -1e30f;
  int bcode = 0;
  const int nx = 12 * a.mask4 + 1, ny = 12 * a.mask5 + 1;
  for (int ipsi = 0; ipsi < a.np; ipsi++) {
    const float cps = psitab[2 * ipsi], sps = psitab[2 * ipsi + 1];
    // ROTATE2D (rotate2d.f:47-74) on the half disc r < MAXR2
    for (int e = tid; e < R * W; e += kThreads) {
      const int ia = e / W, jb = e % W - (R - 1);
      float2 u = make_float2(0.f, 0.f);
      if (ia * ia + jb * jb < R * R) {
        float x, y;
        {
#pragma clang fp contract(off)
          const float fi = (float)ia, fj = (float)jb;
          x = cps * fi + (-sps) * fj;
          y = sps * fi + cps * fj;
        }
        // NINT(IPAD*X): |l|, |m| <= 4R for r < R; the clamp only guards
        // non-finite input
        const int l = min(max((int)roundf(4.0f * x), -R4), R4);
        const int m = min(max((int)roundf(4.0f * y), -R4), R4);
        if (l >= 0)
          u = fb[(size_t)(m + R4) * nL + l];
        else
          u = conjf2(fb[(size_t)(-m + R4) * nL + (-l)]);
      }
      U[e] = u;
    }
    __syncthreads();
    float ssum = 0.f;
    if (wv < iq) {
      // stage 1: Z(I, n2) = sum_J w_I X(I,J) conj(T_q(I,J)) e^{2 pi i J n2 / NSAMR}
      const int I = lane & 31, h = lane >> 5;
      float2 z[8];
#pragma unroll
      for (int k = 0; k < 8; k++) z[k] = make_float2(0.f, 0.f);
      if (I < R) {
        const unsigned long long rb = reg[I];
        const float wI = I == 0 ? 1.0f : 2.0f;
        for (int jj = 0; jj < W; jj++) {
          const int J = jj - (R - 1);
          // quadrant permutations of ccp.f:175-209
          int ua, ub;
          bool cj;
          if (wv == 0) { ua = I; ub = J; cj = false; }
          else if (wv == 1) {
            if (J <= 0) { ua = -J; ub = I; cj = false; } else { ua = J; ub = -I; cj = true; }
          } else if (wv == 2) { ua = I; ub = J; cj = true; }
          else {
            if (J > 0) { ua = J; ub = -I; cj = false; } else { ua = -J; ub = I; cj = true; }
          }
          float2 t = ((rb >> jj) & 1ull) ? U[(size_t)ua * W + (ub + R - 1)] : make_float2(0.f, 0.f);
          if (cj) t.y = -t.y;
          ssum += t.x * t.x + t.y * t.y;
          const float2 x = X[(size_t)I * W + jj];
          float2 c = make_float2(x.x * t.x + x.y * t.y, x.y * t.x - x.x * t.y);  // x conj(t)
          c.x *= wI;
          c.y *= wI;
#pragma unroll
          for (int k = 0; k < 8; k++) {
            const int n2 = h * 8 + k;
            if (n2 < kWin) {
              const float2 e = E[(size_t)jj * kWin + n2];
              z[k].x += c.x * e.x - c.y * e.y;
              z[k].y += c.x * e.y + c.y * e.x;
            }
          }
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
          const int n2 = h * 8 + k;
          if (n2 < kWin) Z[((size_t)wv * R + I) * kWin + n2] = z[k];
        }
      }
      if (h != 0) ssum = 0.f;
      for (int off = 32; off > 0; off >>= 1) ssum += __shfl_xor(ssum, off, 64);
    }
    __syncthreads();
    if (wv < iq) {
      // stage 2: map(n1, n2) = sum_I Re(Z(I, n2) e^{2 pi i I n1 / NSAMR})
      for (int o = lane; o < kWin * kWin; o += 64) {
        const int n1 = o / kWin, n2 = o % kWin;
        float v = 0.f;
        for (int i = 0; i < R; i++) {
          const float2 zz = Z[((size_t)wv * R + i) * kWin + n2];
          const float2 e = E1[(size_t)i * kWin + n1];
          v += zz.x * e.x - zz.y * e.y;
        }
        M[wv * kWin * kWin + o] = v;
      }
    }
    __syncthreads();
    if (wv < iq) {
      // argmax over IX (outer) x IY, first maximum (ccp.f:229-241)
      float bv = -1e30f;
      int bo = 0x7fffffff;
      for (int o = lane; o < nx * ny; o += 64) {
        const int ix = o / ny - 6 * a.mask4, iy = o % ny - 6 * a.mask5;
        const float v = M[wv * kWin * kWin + (ix + 7) * kWin + (iy + 7)];
        if (v > bv) { bv = v; bo = o; }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const float ov = __shfl_xor(bv, off, 64);
        const int oo = __shfl_xor(bo, off, 64);
        if (ov > bv || (ov == bv && oo < bo)) { bv = ov; bo = oo; }
      }
      // a map without a finite value leaves bo unset: take the first shift
      if (bo >= nx * ny) bo = 0;
      const int ixm = bo / ny - 6 * a.mask4, iym = bo % ny - 6 * a.mask5;
      // 3x3 box in IX-outer order, then the normalisation (ccp.f:242-252)
      float box = 0.f;
      for (int dx = -1; dx <= 1; dx++)
        for (int dy = -1; dy <= 1; dy++)
          box += M[wv * kWin * kWin + (ixm + dx + 7) * kWin + (iym + dy + 7)];
      const float cc = box / sqrtf(ssum) / sqrtf(sum2) / 9.0f;
      if (cc > best) {
        best = cc;
        bcode = ((ipsi * 4 + wv) * 225) + (ixm + 7) * 15 + (iym + 7);
      }
    }
    __syncthreads();
  }
  if (lane == 0) {
    wbest[wv] = best;
    wcode[wv] = bcode;
  }
  __syncthreads();
  if (tid == 0) {
    // the first best in the reference's (IPSI outer, IQUAD inner) order
    float b = -1e30f;
    int bc = 0, bp = 0x7fffffff;
    for (int q = 0; q < iq; q++) {
      const int ip = wcode[q] / 900;
      if (wbest[q] > b || (wbest[q] == b && ip < bp)) {
        b = wbest[q];
        bc = wcode[q];
        bp = ip;
      }
    }
    score[item] = b;
    code[item] = bc;
  }
}

}  // namespace

// ------------------------------------------------------------------- tables
void ccp_tables(int nsam, int maxr2, CcpTablesHost* t) {
  const int R = maxr2, W = 2 * R - 1, R4 = 4 * R, nL = R4 + 1, nM = 8 * R + 1;
  const int n4 = 4 * nsam, nsamh = nsam / 2, jc = nsam / 2 + 1;
  int nsamrh = nsamh / 2;
  if (2 * nsamrh != nsamh) nsamrh = (nsamh + 1) / 2;
  const int nsamr = 2 * nsamrh;
  const double tw = 2.0 * M_PI;
  // D(m), m = -8R..8R
  t->d.assign(2 * (16 * R + 1), 0.f);
  std::vector<double> dre(16 * R + 1), dim(16 * R + 1);
  for (int m = -8 * R; m <= 8 * R; m++) {
    double sr = 0, si = 0;
    for (int q = -nsamh; q < nsamh; q++) {
      const double ang = tw * (double)q * (double)m / (double)n4;
      sr += cos(ang);
      si += sin(ang);
    }
    dre[m + 8 * R] = sr;
    dim[m + 8 * R] = si;
    t->d[2 * (m + 8 * R)] = (float)sr;
    t->d[2 * (m + 8 * R) + 1] = (float)si;
  }
  // pixels that pad.f:129-140 moves with -CIRC: source slot in a padding column
  std::vector<std::pair<int, int>> bug;
  for (int j = 0; j < nsam; j++)
    for (int i = 0; i < nsam; i++) {
      const long src = (long)i + (long)(nsam + 2) * j;
      if (src % (n4 + 2) >= n4) bug.emplace_back(i < nsamh ? i : i - nsam, j < nsamh ? j : j - nsam);
    }
  // K(L, M) = D(-L) D(-M) + B(L, M), M = -4R..4R (rows), L = 0..4R
  t->k.assign(2 * (size_t)nM * nL, 0.f);
  for (int mm = 0; mm < nM; mm++) {
    const int M = mm - R4;
    for (int L = 0; L <= R4; L++) {
      const double ar = dre[-L + 8 * R], ai = dim[-L + 8 * R];
      const double br = dre[-M + 8 * R], bi = dim[-M + 8 * R];
      double kr = ar * br - ai * bi, ki = ar * bi + ai * br;
      for (auto& q : bug) {
        const double ang = -tw * ((double)L * q.first + (double)M * q.second) / (double)n4;
        kr += cos(ang);
        ki += sin(ang);
      }
      t->k[2 * ((size_t)mm * nL + L)] = (float)kr;
      t->k[2 * ((size_t)mm * nL + L) + 1] = (float)ki;
    }
  }
  // W(kx, ky) = sum over the central cross of e^{2 pi i k.q/N} (rows/columns
  // NSAM/2 and NSAM/2+1, 1-based)
  t->w.assign(2 * (size_t)W * W, 0.f);
  const int ca = nsamh - 1, cb = nsamh;
  for (int kx = -R + 1; kx <= R - 1; kx++)
    for (int ky = -R + 1; ky <= R - 1; ky++) {
      double sr = 0, si = 0;
      for (int i = 0; i < nsam; i++)
        for (int j = 0; j < nsam; j++) {
          if (!(i == ca || i == cb || j == ca || j == cb)) continue;
          const double ang = tw * ((double)kx * i + (double)ky * j) / (double)nsam;
          sr += cos(ang);
          si += sin(ang);
        }
      t->w[2 * ((size_t)(kx + R - 1) * W + (ky + R - 1))] = (float)sr;
      t->w[2 * ((size_t)(kx + R - 1) * W + (ky + R - 1)) + 1] = (float)si;
    }
  // phase tables of the NSAMR-grid map
  t->e.assign(2 * (size_t)W * kWin, 0.f);
  for (int jj = 0; jj < W; jj++)
    for (int n = 0; n < kWin; n++) {
      const double ang = tw * (double)(jj - (R - 1)) * (double)(n - 7) / (double)nsamr;
      t->e[2 * ((size_t)jj * kWin + n)] = (float)cos(ang);
      t->e[2 * ((size_t)jj * kWin + n) + 1] = (float)sin(ang);
    }
  t->e1.assign(2 * (size_t)R * kWin, 0.f);
  for (int i = 0; i < R; i++)
    for (int n = 0; n < kWin; n++) {
      const double ang = tw * (double)i * (double)(n - 7) / (double)nsamr;
      t->e1[2 * ((size_t)i * kWin + n)] = (float)cos(ang);
      t->e1[2 * ((size_t)i * kWin + n) + 1] = (float)sin(ang);
    }
  // correlation region of ccp.f:160-174 on the NSAMR grid
  t->region.assign(R, 0ull);
  for (int i = 0; i < R; i++)
    for (int jj = 0; jj < W; jj++) {
      const int j = jj - (R - 1), it = i * i + j * j;
      if (it >= 1 && it < R * R && abs(j) != jc - 1 && i <= nsamrh && j >= -nsamrh &&
          j <= nsamrh - 1)
        t->region[i] |= 1ull << jj;
    }
}

size_t ccp_band_bytes(int maxr2) {
  return sizeof(float2) * (size_t)(8 * maxr2 + 1) * (4 * maxr2 + 1);
}

static size_t pad_lds(int R) {
  const int W = 2 * R - 1;
  return sizeof(float2) * ((size_t)W * W + (size_t)W * kMc + (16 * R + 1) + W);
}
static size_t scan_lds(int R) {
  const int W = 2 * R - 1;
  return sizeof(float2) * ((size_t)2 * R * W + (size_t)W * kWin + (size_t)R * kWin +
                           (size_t)4 * R * kWin) +
         sizeof(float) * 4 * kWin * kWin;
}

int launch_ccp(Ctx& c, const CcpArgs& a, int item0, int nitems, const float* a3df,
               const float* spec, const float* ctff, const float* dm, const float* bw,
               const CcpTablesDev& t, const float* psitab, float* fband, float* score,
               int* code) {
  if (nitems <= 0) return 0;
  FR_CHECK(a.maxr2 >= 2 && a.maxr2 <= kCcpMaxR, "exact CCP search: MAXR2 must be 2..%d",
           kCcpMaxR);
  const size_t l1 = pad_lds(a.maxr2), l2 = scan_lds(a.maxr2);
  FR_CHECK(l1 <= 160 * 1024 && l2 <= 160 * 1024, "exact CCP search: LDS");
  FR_HIP(hipFuncSetAttribute((const void*)k_ccp_pad, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)l1));
  FR_HIP(hipFuncSetAttribute((const void*)k_ccp_scan, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)l2));
  hipLaunchKernelGGL(k_ccp_pad, dim3(nitems), dim3(kThreads), l1, c.stream, a,
                     reinterpret_cast<const float2*>(a3df), reinterpret_cast<const float2*>(ctff),
                     dm, item0, reinterpret_cast<const float2*>(t.d),
                     reinterpret_cast<const float2*>(t.k), reinterpret_cast<const float2*>(t.w),
                     reinterpret_cast<float2*>(fband));
  FR_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_ccp_scan, dim3(nitems), dim3(kThreads), l2, c.stream, a,
                     reinterpret_cast<const float2*>(spec), bw, item0,
                     reinterpret_cast<const float2*>(fband), reinterpret_cast<const float2*>(t.e),
                     reinterpret_cast<const float2*>(t.e1), t.region, psitab, score, code);
  FR_HIP(hipGetLastError());
  return 0;
}

}  // namespace fr

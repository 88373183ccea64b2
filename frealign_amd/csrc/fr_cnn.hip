// fr_cnn.hip -- the learned orientation proposal, trained and run on the GPU
// with this file's kernels only.
//
// Network: nconv 3x3 / stride 2 / pad 1 convolutions with ReLU, then one
// fully connected layer to K orientation-class logits.  Activations are
// channel-major, [C][B][H][W], so every convolution is one GEMM over the
// batch: out[C_out][B*P] = W[C_out][C_in*9] x cols[C_in*9][B*P], with the
// im2col operand gathered inside the GEMM -- on the way forward, for the
// weight gradients (positions as the reduction axis, split over many
// workgroups) and for the input gradients (the transposed convolution over
// K = C_out*9, the ReLU derivative applied in the epilogue): no column
// matrix is ever written.
// The GEMM runs on the matrix cores (v_mfma_f32_16x16x4f32).  Weights are
// held as BF16-rounded values (wq) with an FP32 master copy for Adam.
//
// Data: the network input is the particle spectrum (DATC) phase-flipped by
// the sign of its CTF, cropped to S x S in Fourier space and brought back to
// real space (S-point c2r), then normalised.  Training images are made the
// same way from central sections of the current reference (AINTERPO3DS),
// times the CTF of ctf.f, a shift ramp and white noise at a given SNR.
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "fr_cnn.h"
#include "fr_ctf.h"
#include "fr_extract.h"

namespace fr {

namespace {

constexpr int kT = 256;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float fx4;

// ------------------------------------------------------------------ GEMM
// C[m][n] (+)= sum_k A[m][k] B[k][n] over k in this split's range, with
// arbitrary element strides (so transposed operands cost nothing), an
// optional per-row bias and ReLU.  64x64 tiles, K staged 16 at a time.
struct GemmArgs {
  int M, N, K;
  const float* A;
  long long sam, sak;
  const float* B;
  long long sbk, sbn;
  float* C;
  long long scm, scn;
  int kchunk;        // K range per blockIdx.z (multiple of 16)
  long long csplit;  // C offset between splits
  const float* bias;
  int relu;
  int accumulate;
  // implicit convolution operands (B is then the activation or gradient
  // tensor, [C][batch][H][H], never an explicit column matrix):
  //   kGemmConv  (forward)        B[k][n] = in[ci][b][2oy+ky-1][2ox+kx-1],
  //              k = ci*9 + ky*3 + kx, n = b*Ho*Ho + oy*Ho + ox;
  //   kGemmConvW (weight grad.)   B[k][n] = the same element with the roles
  //              swapped: k = b*Ho*Ho + oy*Ho + ox, n = ci*9 + ky*3 + kx;
  //   kGemmConvX (input grad.)    one parity class (py, px) = (conv_cls >> 1,
  //              conv_cls & 1) of input positions y = 2t+py, x = 2u+px,
  //              n = b*Ho*Ho + t*Ho + u; its taps are the ky with
  //              y+1-ky even (ky = 1 for py = 0, ky = 0, 2 for py = 1), kx
  //              likewise, ntap = |ky set| |kx set|, k = co*ntap + j, and
  //              B[k][n] = g[co][b][(y+1-ky)/2][(x+1-kx)/2];
  // zero outside the image.  conv_B is the batch, conv_H the input size H.
  int conv_B, conv_H, conv_cls;
  // epilogue (kGemmConvX): C element (ci, n) goes to gin[ci][b][y][x]
  // (row stride scm), zeroed unless mask at the same place is > 0 (the ReLU
  // derivative of the layer below)
  const float* mask;
};

enum { kGemmPlain = 0, kGemmConv = 1, kGemmConvW = 2, kGemmConvX = 3 };

constexpr int kGP = 80;  // LDS pitch: 16*g + lane -> distinct banks

template <int kMode>
__global__ void __launch_bounds__(kT) k_gemm_t(GemmArgs g) {
  constexpr bool kConv = kMode == kGemmConv;
  // two LDS stages: the next K tile is fetched into registers while the
  // matrix cores work on the current one (one barrier per K tile)
  __shared__ float As[2][16][kGP];
  __shared__ float Bs[2][16][kGP];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int kbeg = blockIdx.z * g.kchunk;
  const int kend = min(g.K, kbeg + g.kchunk);
  const bool a_kfast = g.sak == 1;
  const bool b_nfast = g.sbn == 1;
  fx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = fx4{0.f, 0.f, 0.f, 0.f};
  // this thread's four A and four B elements of a K tile
  int am[4], ak[4], bn[4], bk[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int e = tid + kT * q;
    if (a_kfast) {
      ak[q] = e & 15;
      am[q] = e >> 4;
    } else {
      am[q] = e & 63;
      ak[q] = e >> 6;
    }
    if (b_nfast) {
      bn[q] = e & 63;
      bk[q] = e >> 6;
    } else {
      bk[q] = e & 15;
      bn[q] = e >> 4;
    }
  }
  float ra[4], rb[4];
  // implicit convolution: this thread's output positions, fixed over K
  long long cb[4];
  int cy[4], cx[4];
  if (kConv || kMode == kGemmConvX) {
    // forward: n -> output position (b, oy, ox), taps add (ky-1, kx-1) to
    // (2oy, 2ox); input gradient: n -> input position (b, y, x)
    const int H = g.conv_H, Ho = H / 2, P = Ho * Ho;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int n = n0 + bn[q];
      if (kConv) {
        const int b = n / P, pp = n - b * P, oy = pp / Ho, ox = pp - oy * Ho;
        cb[q] = (long long)b * H * H;
        cy[q] = 2 * oy - 1;
        cx[q] = 2 * ox - 1;
      } else {
        const int b = n / P, pp = n - b * P, t = pp / Ho, u = pp - t * Ho;
        cb[q] = (long long)b * P;
        cy[q] = 2 * t + (g.conv_cls >> 1) + 1;  // y + 1 = 2oy + ky
        cx[q] = 2 * u + (g.conv_cls & 1) + 1;
      }
    }
  }
  // weight gradient: this thread's K position (b, oy, ox) walks with the K
  // tiles (bk fixed per thread, +16 per tile); its four n are fixed taps
  int wb = 0, woy = 0, wox = 0;
  int wci[4], wky[4], wkx[4];
  if (kMode == kGemmConvW) {
    const int Ho = g.conv_H / 2, P = Ho * Ho;
    const int k = kbeg + bk[0];
    wb = k / P;
    const int pp = k - wb * P;
    woy = pp / Ho;
    wox = pp - woy * Ho;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int n = n0 + bn[q];
      wci[q] = n / 9;
      const int kk = n - 9 * wci[q];
      wky[q] = kk / 3;
      wkx[q] = kk - 3 * wky[q];
    }
  }
  auto fetch = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int m = m0 + am[q], k = k0 + ak[q];
      ra[q] = (m < g.M && k < kend) ? g.A[m * g.sam + k * g.sak] : 0.f;
      const int n = n0 + bn[q], k2 = k0 + bk[q];
      if (kMode == kGemmConvW) {
        // the K position (wb, woy, wox) is this thread's, shared by its four n
        float v = 0.f;
        const int H = g.conv_H;
        if (n < g.N && k2 < kend) {
          const int y = 2 * woy + wky[q] - 1, x = 2 * wox + wkx[q] - 1;
          if (y >= 0 && y < H && x >= 0 && x < H)
            v = g.B[((long long)wci[q] * g.conv_B + wb) * H * H + (long long)y * H + x];
        }
        rb[q] = v;
      } else if (kMode == kGemmConvX) {
        float v = 0.f;
        if (n < g.N && k2 < kend) {
          const int py = g.conv_cls >> 1, px = g.conv_cls & 1;
          const int nty = py + 1, ntx = px + 1, ntap = nty * ntx;
          const int co = k2 / ntap, j = k2 - ntap * co, jy = j / ntx, jx = j - jy * ntx;
          const int ky = py ? 2 * jy : 1, kx = px ? 2 * jx : 1;
          const int oy = (cy[q] - ky) >> 1, ox = (cx[q] - kx) >> 1, Ho = g.conv_H / 2;
          if (oy < Ho && ox < Ho)
            v = g.B[(long long)co * g.conv_B * Ho * Ho + cb[q] + (long long)oy * Ho + ox];
        }
        rb[q] = v;
      } else if (kConv) {
        float v = 0.f;
        if (n < g.N && k2 < kend) {
          const int ci = k2 / 9, kk = k2 - 9 * ci, ky = kk / 3, kx = kk - 3 * ky;
          const int y = cy[q] + ky, x = cx[q] + kx, H = g.conv_H;
          if (y >= 0 && y < H && x >= 0 && x < H)
            v = g.B[(long long)ci * g.conv_B * H * H + cb[q] + (long long)y * H + x];
        }
        rb[q] = v;
      } else {
        rb[q] = (n < g.N && k2 < kend) ? g.B[k2 * g.sbk + n * g.sbn] : 0.f;
      }
    }
    if (kMode == kGemmConvW) {  // the next K tile: 16 positions on
      const int Ho = g.conv_H / 2;
      wox += 16;
      while (wox >= Ho) {
        wox -= Ho;
        if (++woy == Ho) {
          woy = 0;
          wb++;
        }
      }
    }
  };
  auto put = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      As[buf][ak[q]][am[q]] = ra[q];
      Bs[buf][bk[q]][bn[q]] = rb[q];
    }
  };
  int buf = 0;
  if (kbeg < kend) {
    fetch(kbeg);
    put(0);
  }
  __syncthreads();
  for (int k0 = kbeg; k0 < kend; k0 += 16) {
    const bool more = k0 + 16 < kend;
    if (more) fetch(k0 + 16);
#pragma unroll
    for (int kk = 0; kk < 16; kk += 4) {
      float a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; i++)
        a[i] = As[buf][kk + (lane >> 4)][wm * 32 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 2; j++)
        b[j] = Bs[buf][kk + (lane >> 4)][wn * 32 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    if (more) put(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  float* C = g.C + blockIdx.z * g.csplit;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int m = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + r;
        const int n = n0 + wn * 32 + j * 16 + (lane & 15);
        if (m < g.M && n < g.N) {
          float v = acc[i][j][r];
          if (g.bias) v += g.bias[m];
          if (g.relu) v = fmaxf(v, 0.f);
          long long at = m * g.scm + n * g.scn;
          if (kMode == kGemmConvX) {
            const int H = g.conv_H, Ho = H / 2, P = Ho * Ho;
            const int b = n / P, pp = n - b * P, t = pp / Ho, u = pp - t * Ho;
            at = m * g.scm + (long long)b * H * H + (long long)(2 * t + (g.conv_cls >> 1)) * H +
                 (2 * u + (g.conv_cls & 1));
            if (!(g.mask[at] > 0.f)) v = 0.f;
          }
          float* p = C + at;
          if (g.accumulate) v += *p;
          *p = v;
        }
      }
}

// The same GEMM with BF16 operands (each FP32 element rounded to nearest
// even as it is staged) on v_mfma_f32_16x16x32_bf16, FP32 accumulation and
// epilogue: the training step's mixed-precision form.  K is staged 32 at a
// time; LDS holds the tiles k-contiguous per row (A[m][k], B[n][k]) so that
// each lane's 8-element operand is one 16-byte read.  Same modes, arguments
// and epilogue as k_gemm_t.
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 gbf16x8;
constexpr int kBP = 40;  // LDS row pitch (bf16): 80 bytes

template <int kMode>
__global__ void __launch_bounds__(kT) k_gemm_b(GemmArgs g) {
  constexpr bool kConv = kMode == kGemmConv;
  __shared__ __attribute__((aligned(16))) __bf16 As[2][64][kBP];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[2][64][kBP];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int la = lane & 15, lg = lane >> 4;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int kbeg = blockIdx.z * g.kchunk;
  const int kend = min(g.K, kbeg + g.kchunk);
  const bool a_kfast = g.sak == 1;
  const bool b_nfast = g.sbn == 1;
  fx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = fx4{0.f, 0.f, 0.f, 0.f};
  // this thread's eight A and eight B elements of a K tile of 32
  int am[8], ak[8], bn[8], bk[8];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int e = tid + kT * q;
    if (a_kfast) {
      ak[q] = e & 31;
      am[q] = e >> 5;
    } else {
      am[q] = e & 63;
      ak[q] = e >> 6;
    }
    if (b_nfast) {
      bn[q] = e & 63;
      bk[q] = e >> 6;
    } else {
      bk[q] = e & 31;
      bn[q] = e >> 5;
    }
  }
  // implicit convolution: the n of an n-fast B operand is fixed per thread
  long long cb = 0;
  int cy = 0, cx = 0;
  const int H = g.conv_H, Ho = H / 2, P = Ho * Ho;
  if (kConv || kMode == kGemmConvX) {
    const int n = n0 + bn[0];
    if (kConv) {
      const int b = n / P, pp = n - b * P, oy = pp / Ho, ox = pp - oy * Ho;
      cb = (long long)b * H * H;
      cy = 2 * oy - 1;
      cx = 2 * ox - 1;
    } else {
      const int b = n / P, pp = n - b * P, t = pp / Ho, u = pp - t * Ho;
      cb = (long long)b * P;
      cy = 2 * t + (g.conv_cls >> 1) + 1;
      cx = 2 * u + (g.conv_cls & 1) + 1;
    }
  }
  // weight gradient: the k of a k-fast B operand is fixed per thread; its
  // eight n are taps (ci, ky, kx)
  int wci[8], wky[8], wkx[8];
  if (kMode == kGemmConvW) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int n = n0 + bn[q];
      wci[q] = n / 9;
      const int kk = n - 9 * wci[q];
      wky[q] = kk / 3;
      wkx[q] = kk - 3 * wky[q];
    }
  }
  float ra[8], rb[8];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int m = m0 + am[q], k = k0 + ak[q];
      ra[q] = (m < g.M && k < kend) ? g.A[m * g.sam + k * g.sak] : 0.f;
    }
    if (kMode == kGemmConvW) {
      const int k2 = k0 + bk[0];
      int wb = 0, woy = 0, wox = 0;
      if (k2 < kend) {
        wb = k2 / P;
        const int pp = k2 - wb * P;
        woy = pp / Ho;
        wox = pp - woy * Ho;
      }
#pragma unroll
      for (int q = 0; q < 8; q++) {
        float v = 0.f;
        if (n0 + bn[q] < g.N && k2 < kend) {
          const int y = 2 * woy + wky[q] - 1, x = 2 * wox + wkx[q] - 1;
          if (y >= 0 && y < H && x >= 0 && x < H)
            v = g.B[((long long)wci[q] * g.conv_B + wb) * H * H + (long long)y * H + x];
        }
        rb[q] = v;
      }
    } else {
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const int n = n0 + bn[q], k2 = k0 + bk[q];
        float v = 0.f;
        if (n < g.N && k2 < kend) {
          if (kMode == kGemmConvX) {
            const int py = g.conv_cls >> 1, px = g.conv_cls & 1;
            const int ntx = px + 1, ntap = (py + 1) * ntx;
            const int co = k2 / ntap, j = k2 - ntap * co, jy = j / ntx, jx = j - jy * ntx;
            const int ky = py ? 2 * jy : 1, kx = px ? 2 * jx : 1;
            const int oy = (cy - ky) >> 1, ox = (cx - kx) >> 1;
            if (oy < Ho && ox < Ho)
              v = g.B[(long long)co * g.conv_B * P + cb + (long long)oy * Ho + ox];
          } else if (kConv) {
            const int ci = k2 / 9, kk = k2 - 9 * ci, ky = kk / 3, kx = kk - 3 * ky;
            const int y = cy + ky, x = cx + kx;
            if (y >= 0 && y < H && x >= 0 && x < H)
              v = g.B[(long long)ci * g.conv_B * H * H + cb + (long long)y * H + x];
          } else {
            v = g.B[k2 * g.sbk + n * g.sbn];
          }
        }
        rb[q] = v;
      }
    }
  };
  auto put = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      As[buf][am[q]][ak[q]] = (__bf16)ra[q];
      Bs[buf][bn[q]][bk[q]] = (__bf16)rb[q];
    }
  };
  int buf = 0;
  if (kbeg < kend) {
    fetch(kbeg);
    put(0);
  }
  __syncthreads();
  for (int k0 = kbeg; k0 < kend; k0 += 32) {
    const bool more = k0 + 32 < kend;
    if (more) fetch(k0 + 32);
    gbf16x8 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; i++)
      a[i] = *reinterpret_cast<const gbf16x8*>(&As[buf][wm * 32 + i * 16 + la][8 * lg]);
#pragma unroll
    for (int j = 0; j < 2; j++)
      b[j] = *reinterpret_cast<const gbf16x8*>(&Bs[buf][wn * 32 + j * 16 + la][8 * lg]);
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    if (more) put(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  float* C = g.C + blockIdx.z * g.csplit;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int m = m0 + wm * 32 + i * 16 + 4 * lg + r;
        const int n = n0 + wn * 32 + j * 16 + la;
        if (m < g.M && n < g.N) {
          float v = acc[i][j][r];
          if (g.bias) v += g.bias[m];
          if (g.relu) v = fmaxf(v, 0.f);
          long long at = m * g.scm + n * g.scn;
          if (kMode == kGemmConvX) {
            const int b = n / P, pp = n - b * P, t = pp / Ho, u = pp - t * Ho;
            at = m * g.scm + (long long)b * H * H + (long long)(2 * t + (g.conv_cls >> 1)) * H +
                 (2 * u + (g.conv_cls & 1));
            if (!(g.mask[at] > 0.f)) v = 0.f;
          }
          float* p = C + at;
          if (g.accumulate) v += *p;
          *p = v;
        }
      }
}

// out[i] = sum over s of part[s][i] for many splits and few outputs: one
// block per output, thread t summing splits t, t+256, ... in order, then a
// fixed tree (deterministic)
__global__ void __launch_bounds__(kT)
    k_split_sum_wide(long long n, int nsplit, const float* __restrict__ part, float* out) {
  __shared__ float red[kT];
  const long long i = blockIdx.x;
  float s = 0.f;
  for (int q = threadIdx.x; q < nsplit; q += kT) s += part[q * n + i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[i] = red[0];
}

// out[i] = sum over s of part[s][i], in split order (deterministic)
__global__ void __launch_bounds__(kT)
    k_split_sum(long long n, int nsplit, const float* __restrict__ part, float* out) {
  const long long i = blockIdx.x * (long long)kT + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int q = 0; q < nsplit; q++) s += part[q * n + i];
  out[i] = s;
}

// X[c*P + p][b] = act[c][b*P + p]
__global__ void __launch_bounds__(kT)
    k_flatten(int C, int B, int P, const float* __restrict__ act, float* X) {
  const long long total = (long long)C * B * P;
  const long long idx = blockIdx.x * (long long)kT + threadIdx.x;
  if (idx >= total) return;
  const int p = (int)(idx % P);
  const int b = (int)((idx / P) % B);
  const int c = (int)(idx / ((long long)P * B));
  X[((long long)c * P + p) * B + b] = act[idx];
}

// g[c][b*P + p] = (act > 0) * gX[c*P + p][b]
__global__ void __launch_bounds__(kT)
    k_unflatten_relu(int C, int B, int P, const float* __restrict__ gX,
                     const float* __restrict__ act, float* g) {
  const long long total = (long long)C * B * P;
  const long long idx = blockIdx.x * (long long)kT + threadIdx.x;
  if (idx >= total) return;
  const int p = (int)(idx % P);
  const int b = (int)((idx / P) % B);
  const int c = (int)(idx / ((long long)P * B));
  g[idx] = act[idx] > 0.f ? gX[((long long)c * P + p) * B + b] : 0.f;
}

// out[m] = sum_n g[m*ld + n*sn] over n < N; one block per row, fixed order
__global__ void __launch_bounds__(kT)
    k_rowsum(int N, long long ld, long long sn, const float* __restrict__ g, float* out) {
  __shared__ float red[kT / 64];
  const float* row = g + blockIdx.x * ld;
  float s = 0.f;
  for (int n = threadIdx.x; n < N; n += kT) s += row[n * sn];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// part[m * nsplit + s] = sum of g[m*ld + n] over split s of 0..N-1 (splits of
// `chunk` elements); one block per (row, split), fixed order per thread
__global__ void __launch_bounds__(kT)
    k_rowsum_part(int N, long long ld, int chunk, const float* __restrict__ g, float* part) {
  __shared__ float red[kT / 64];
  const int m = blockIdx.x, sp = blockIdx.y;
  const float* row = g + m * ld;
  const int n0 = sp * chunk, n1 = min(N, n0 + chunk);
  float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
  int n = n0 + threadIdx.x;
  for (; n + 3 * kT < n1; n += 4 * kT) {
    s0 += row[n];
    s1 += row[n + kT];
    s2 += row[n + 2 * kT];
    s3 += row[n + 3 * kT];
  }
  for (; n < n1; n += kT) s0 += row[n];
  float s = (s0 + s1) + (s2 + s3);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[m * gridDim.y + sp] = red[0] + red[1] + red[2] + red[3];
}

__device__ inline float block_max(float v, float* red) {
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_down(v, off, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

__device__ inline float block_sum(float v, float* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

// softmax cross-entropy of logits[b][*] against label[b]: loss[b] and
// grad[b][k] = (p_k - [k == label]) / B; one block per sample
__global__ void __launch_bounds__(kT)
    k_softmax_xent(int K, int B, const float* __restrict__ logits, const int* __restrict__ label,
                   float* loss, float* grad) {
  __shared__ float red[kT / 64];
  const int b = blockIdx.x;
  const float* z = logits + (long long)b * K;
  float mx = -3.0e38f;
  for (int k = threadIdx.x; k < K; k += kT) mx = fmaxf(mx, z[k]);
  mx = block_max(mx, red);
  float s = 0.f;
  for (int k = threadIdx.x; k < K; k += kT) s += expf(z[k] - mx);
  s = block_sum(s, red);
  int y = label[b];
  y = y < 0 ? 0 : (y >= K ? K - 1 : y);  // labels are validated by the caller; stay in bounds
  for (int k = threadIdx.x; k < K; k += kT) {
    float p = expf(z[k] - mx) / s;
    grad[(long long)b * K + k] = (p - (k == y ? 1.f : 0.f)) / (float)B;
  }
  if (threadIdx.x == 0) loss[b] = logf(s) + mx - z[y];
}

// top-k classes by probability (ties: lower class) and their probabilities
__global__ void __launch_bounds__(kT)
    k_propose(int K, const float* __restrict__ logits, int topk, int* top, float* prob) {
  extern __shared__ float pz[];  // K probabilities
  __shared__ float red[kT / 64];
  __shared__ float bv[kT];
  __shared__ int bi[kT];
  const int b = blockIdx.x;
  const float* z = logits + (long long)b * K;
  float mx = -3.0e38f;
  for (int k = threadIdx.x; k < K; k += kT) mx = fmaxf(mx, z[k]);
  mx = block_max(mx, red);
  float s = 0.f;
  for (int k = threadIdx.x; k < K; k += kT) {
    float e = expf(z[k] - mx);
    pz[k] = e;
    s += e;
  }
  s = block_sum(s, red);
  for (int k = threadIdx.x; k < K; k += kT) pz[k] /= s;
  __syncthreads();
  for (int t = 0; t < topk; t++) {
    float v = -1.f;
    int vi = K;
    for (int k = threadIdx.x; k < K; k += kT)
      if (pz[k] > v) {
        v = pz[k];
        vi = k;
      }
    bv[threadIdx.x] = v;
    bi[threadIdx.x] = vi;
    __syncthreads();
    for (int off = kT / 2; off > 0; off >>= 1) {
      if (threadIdx.x < off) {
        float v2 = bv[threadIdx.x + off];
        int i2 = bi[threadIdx.x + off];
        if (v2 > bv[threadIdx.x] || (v2 == bv[threadIdx.x] && i2 < bi[threadIdx.x])) {
          bv[threadIdx.x] = v2;
          bi[threadIdx.x] = i2;
        }
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      top[(long long)b * topk + t] = bi[0] < K ? bi[0] : -1;
      prob[(long long)b * topk + t] = bv[0];
      if (bi[0] < K) pz[bi[0]] = -1.f;
    }
    __syncthreads();
  }
}

// Adam on one parameter segment; wq gets the BF16-rounded value for weights
// and the exact value for biases
__device__ inline float bf16_round(float x) {
  unsigned u = __float_as_uint(x);
  u = (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u;
  return __uint_as_float(u);
}

__global__ void __launch_bounds__(kT)
    k_adam(long long n, float* master, float* wq, const float* __restrict__ grad, float* m,
           float* v, float lr, float c1, float c2, int round) {
  const long long i = blockIdx.x * (long long)kT + threadIdx.x;
  if (i >= n) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float g = grad[i];
  float mi = b1 * m[i] + (1.f - b1) * g;
  float vi = b2 * v[i] + (1.f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  float w = master[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  master[i] = w;
  wq[i] = round ? bf16_round(w) : w;
}

// splitmix64 finaliser: a counter-based stream (seed, counter) -> 64 bits
__device__ inline unsigned long long mix64(unsigned long long z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// two standard normals (Box-Muller) from counter `ctr` of stream `seed`
__device__ inline float2 normal2(unsigned long long seed, unsigned long long ctr) {
  unsigned long long r = mix64(seed ^ mix64(ctr));
  float u1 = ((float)(r >> 40) + 1.0f) * (1.0f / 16777217.0f);  // (0, 1]
  float u2 = (float)((r >> 16) & 0xffffffull) * (1.0f / 16777216.0f);
  float rad = sqrtf(-2.0f * logf(u1));
  float s, c;
  sincosf(6.2831853071796f * u2, &s, &c);
  return make_float2(rad * c, rad * s);
}

// He-normal weights (std sqrt(2/fan_in)), zero biases
__global__ void __launch_bounds__(kT)
    k_init(long long n, unsigned long long seed, long long off, float stdv, float* master,
           float* wq, int round) {
  const long long i = blockIdx.x * (long long)kT + threadIdx.x;
  if (i >= n) return;
  float w = stdv > 0.f ? normal2(seed, (unsigned long long)(off + i)).x * stdv : 0.f;
  if (round) w = bf16_round(w);  // the master starts on the BF16 grid
  master[i] = w;
  wq[i] = w;
}

// ------------------------------------------------------------ network input
// From DATC (origin-centred spectrum) and CTFF: the S x S crop, phase-flipped
// by sign(Re(CTFR + conj CTFL)) and multiplied by (-1)^(l+j) so that the
// S-point c2r puts the particle centre at S/2.  The Nyquist row and column of
// the crop are zero.  Output: in-place c2r layout, pitch S+2 floats.
__global__ void __launch_bounds__(kT)
    k_cnn_crop(int nsam, int S, int B, const float2* __restrict__ datc,
               const float2* __restrict__ ctff, float2* out) {
  const int jc = nsam / 2 + 1, js = S / 2 + 1;
  const size_t ic = (size_t)nsam * jc, is = (size_t)S * js;
  const size_t gid = blockIdx.x * (size_t)kT + threadIdx.x;
  if (gid >= is * B) return;
  const int b = (int)(gid / is);
  const int e = (int)(gid - (size_t)b * is);
  const int l = e % js, m = e / js;
  const int j = m >= js ? m - S : m;
  float2 v = make_float2(0.f, 0.f);
  if (l < S / 2 && j > -S / 2 && j < S / 2) {
    const size_t id = (size_t)l + (size_t)jc * wrapn(j, nsam);
    const float2 d = datc[(size_t)b * ic + id];
    const float2 cr = ctff[(size_t)b * 2 * ic + id], cl = ctff[(size_t)b * 2 * ic + ic + id];
    float sg = (cr.x + cl.x) < 0.f ? -1.f : 1.f;
    if ((l + j) & 1) sg = -sg;
    v = make_float2(d.x * sg, d.y * sg);
  }
  out[gid] = v;
}

// Training images in the same form: section x CTF x shift ramp + white noise
// at SNR, cropped and flipped as k_cnn_crop.  One block per image.
//   spar[b]: SHX, SHY (radians per frequency unit, the ramp is
//   exp(-i(SHX l + SHY j)) as in a particle shifted by (SHX, SHY)), DF1, DF2,
//   ANGAST, SNR, 0, 0.
// The SNR is the signal variance over `disc` pixels of the NSAM^2 image
// (Parseval on the crop) over the per-pixel noise variance.
__global__ void __launch_bounds__(kT)
    k_cnn_synth(int nsam, int S, const float2* __restrict__ a3df, const float* __restrict__ dm,
                const float* __restrict__ spar, PrepGlobals g, float disc,
                unsigned long long seed, float2* out) {
  __shared__ float red[kT / 64];
  const int js = S / 2 + 1;
  const size_t is = (size_t)S * js;
  const int b = blockIdx.x;
  const float* sp = spar + 8 * (size_t)b;
  const float shx = sp[0], shy = sp[1], df1 = sp[2], df2 = sp[3], ang = sp[4], snr = sp[5];
  const float* d = dm + 9 * (size_t)b;
  float2* o = out + (size_t)b * is;
  float pw = 0.f, dc = 0.f;
  for (int e = threadIdx.x; e < (int)is; e += kT) {
    const int l = e % js, m = e / js;
    const int j = m >= js ? m - S : m;
    float2 v = make_float2(0.f, 0.f);
    if (l < S / 2 && j > -S / 2 && j < S / 2) {
      cf a = ainterpo_dev(nsam, a3df, d, l, j);
      cf cr = ctf_dev(g, df1, df2, ang, g.thetatr, l, j);
      cf cl = ctf_dev(g, df1, df2, ang, g.thetatr, -l, -j);
      cf c = cf{cr.re + cl.re, cr.im - cl.im};
      float sn, cs;
      sincosf(-(shx * (float)l + shy * (float)j), &sn, &cs);
      cf s = cmul(cmul(a, c), cf{cs, sn});
      v = make_float2(s.re, s.im);
      const float w = l > 0 ? 2.f : 1.f;
      pw += w * (s.re * s.re + s.im * s.im);
      if (l == 0 && j == 0) dc = s.re;
    }
    o[e] = v;
  }
  pw = block_sum(pw, red);
  dc = block_sum(dc, red);
  const float n2 = (float)nsam * (float)nsam;
  const float sumsq = pw / n2, mean = dc / disc;
  float var = sumsq / disc - mean * mean;
  var = var > 0.f ? var : 0.f;
  const float sig = sqrtf(n2 * var / snr * 0.5f);  // per component of a coefficient
  for (int e = threadIdx.x; e < (int)is; e += kT) {
    const int l = e % js, m = e / js;
    const int j = m >= js ? m - S : m;
    if (!(l < S / 2 && j > -S / 2 && j < S / 2)) continue;
    float2 v = o[e];
    float2 z = normal2(seed, ((unsigned long long)b << 32) | (unsigned)e);
    v.x += sig * z.x;
    v.y += sig * z.y;
    cf cr = ctf_dev(g, df1, df2, ang, g.thetatr, l, j);
    cf cl = ctf_dev(g, df1, df2, ang, g.thetatr, -l, -j);
    float sg = (cr.re + cl.re) < 0.f ? -1.f : 1.f;
    if ((l + j) & 1) sg = -sg;
    o[e] = make_float2(v.x * sg, v.y * sg);
  }
}

// (x - mean) / std over each S x S image (pitch S+2 in, dense out)
// the same with the image held in registers (S^2 <= 16 x 1024): one read
__global__ void __launch_bounds__(1024)
    k_cnn_normalise_reg(int S, const float* __restrict__ buf, float* img) {
  __shared__ float red[16];
  const int b = blockIdx.x, t = threadIdx.x, n = S * S;
  const float* x = buf + (size_t)b * S * (S + 2);
  float v[16];
  float s = 0.f;
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const int e = t + 1024 * j;
    v[j] = e < n ? x[(e / S) * (S + 2) + e % S] : 0.f;
    s += v[j];
  }
  auto bsum = [&](float a) {
    for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
    __syncthreads();
    if ((t & 63) == 0) red[t >> 6] = a;
    __syncthreads();
    float r = 0.f;
    for (int k = 0; k < 16; k++) r += red[k];
    return r;
  };
  const float mean = bsum(s) / (float)n;
  float q = 0.f;
#pragma unroll
  for (int j = 0; j < 16; j++)
    if (t + 1024 * j < n) q += (v[j] - mean) * (v[j] - mean);
  const float var = bsum(q) / (float)n;
  const float inv = var > 0.f ? 1.0f / sqrtf(var) : 1.0f;
  float* y = img + (size_t)b * n;
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const int e = t + 1024 * j;
    if (e < n) y[e] = (v[j] - mean) * inv;
  }
}

__global__ void __launch_bounds__(kT)
    k_cnn_normalise(int S, const float* __restrict__ buf, float* img) {
  __shared__ float red[kT / 64];
  const int b = blockIdx.x;
  const float* x = buf + (size_t)b * S * (S + 2);
  float s = 0.f;
  for (int e = threadIdx.x; e < S * S; e += kT) s += x[(e / S) * (S + 2) + e % S];
  const float mean = block_sum(s, red) / (float)(S * S);
  float q = 0.f;
  for (int e = threadIdx.x; e < S * S; e += kT) {
    float t = x[(e / S) * (S + 2) + e % S] - mean;
    q += t * t;
  }
  const float var = block_sum(q, red) / (float)(S * S);
  const float inv = var > 0.f ? 1.0f / sqrtf(var) : 1.0f;
  float* y = img + (size_t)b * S * S;
  for (int e = threadIdx.x; e < S * S; e += kT) y[e] = (x[(e / S) * (S + 2) + e % S] - mean) * inv;
}

// ------------------------------------------------------ orientation classes
// Orientation matrices are A[k][j] = dm[3j + k] (the transpose of the DM
// rows); two orientations are equivalent under symmetry S when S A = B.  The
// cosine relation trace((S A)^T B) = 1 + 2 cos(angle) gives the distance.
__device__ inline float sym_trace(const float* s, const float* a, const float* b) {
  float t = 0.f;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      float sa = s[3 * i + 0] * a[3 * j + 0] + s[3 * i + 1] * a[3 * j + 1] +
                 s[3 * i + 2] * a[3 * j + 2];  // (S A)[i][j], A[k][j] = a[3j + k]
      t += sa * b[3 * j + i];                     // B[i][j] = b[3j + i]
    }
  return t;
}

__global__ void __launch_bounds__(kT)
    k_orient_nearest(int n, const float* __restrict__ dm, int ncls, const float* __restrict__ dmc,
                     int nsym, const float* __restrict__ sym, int* label) {
  // one block per orientation; thread t scans classes t, t + kT, ...; the
  // block keeps the largest trace, ties going to the lowest class (the first
  // maximum of a scan over classes in order)
  __shared__ float ss[60 * 9];
  __shared__ float rv[kT / 64];
  __shared__ int ri[kT / 64];
  for (int e = threadIdx.x; e < nsym * 9; e += kT) ss[e] = sym[e];
  __syncthreads();
  const int i = blockIdx.x;
  float a[9];
  for (int q = 0; q < 9; q++) a[q] = dm[9 * (size_t)i + q];
  float best = -3.0e38f;
  int bi = -1;
  for (int c = threadIdx.x; c < ncls; c += kT) {
    float b[9];
    for (int q = 0; q < 9; q++) b[q] = dmc[9 * (size_t)c + q];
    for (int q = 0; q < nsym; q++) {
      float t = sym_trace(ss + 9 * q, a, b);
      if (t > best) {
        best = t;
        bi = c;
      }
    }
  }
  auto better = [](float v, int k, float bv, int bk) {
    return k >= 0 && (bk < 0 || v > bv || (v == bv && k < bk));
  };
  for (int off = 32; off > 0; off >>= 1) {
    float v = __shfl_down(best, off, 64);
    int k = __shfl_down(bi, off, 64);
    if (better(v, k, best, bi)) {
      best = v;
      bi = k;
    }
  }
  if ((threadIdx.x & 63) == 0) {
    rv[threadIdx.x >> 6] = best;
    ri[threadIdx.x >> 6] = bi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kT / 64; w++)
      if (better(rv[w], ri[w], best, bi)) {
        best = rv[w];
        bi = ri[w];
      }
    label[i] = bi;
  }
}

// mask[b][w] bit i: orientation a = 32 w + i lies within the angle of
// orientation b (max over symmetry of the trace >= tmin)
__global__ void __launch_bounds__(kT)
    k_orient_within(int na, const float* __restrict__ dma, int nb, const float* __restrict__ dmb,
                    int nsym, const float* __restrict__ sym, float tmin, unsigned* mask) {
  __shared__ float ss[60 * 9];
  for (int e = threadIdx.x; e < nsym * 9; e += kT) ss[e] = sym[e];
  __syncthreads();
  const int nw = (na + 31) / 32;
  const long long gid = blockIdx.x * (long long)kT + threadIdx.x;
  if (gid >= (long long)nw * nb) return;
  const int w = (int)(gid % nw), b = (int)(gid / nw);
  float bm[9];
  for (int q = 0; q < 9; q++) bm[q] = dmb[9 * (size_t)b + q];
  unsigned bits = 0;
  for (int i = 0; i < 32; i++) {
    const int a = 32 * w + i;
    if (a >= na) break;
    float am[9];
    for (int q = 0; q < 9; q++) am[q] = dma[9 * (size_t)a + q];
    float t = -3.0e38f;
    for (int q = 0; q < nsym; q++) t = fmaxf(t, sym_trace(ss + 9 * q, am, bm));
    if (t >= tmin) bits |= 1u << i;
  }
  mask[gid] = bits;
}

inline unsigned nblk(long long n) { return (unsigned)((n + kT - 1) / kT); }

// the FP32 (k_gemm_t) or BF16-operand (k_gemm_b) form of one GEMM launch
template <int kMode>
void launch_gemm(Ctx& c, dim3 grid, const GemmArgs& g, bool bf16) {
  if (bf16)
    hipLaunchKernelGGL(k_gemm_b<kMode>, grid, dim3(kT), 0, c.stream, g);
  else
    hipLaunchKernelGGL(k_gemm_t<kMode>, grid, dim3(kT), 0, c.stream, g);
}

int gemm(Ctx& c, int M, int N, int K, const float* A, long long sam, long long sak, const float* B,
         long long sbk, long long sbn, float* C, long long scm, long long scn,
         const float* bias = nullptr, bool relu = false, bool accumulate = false,
         bool bf16 = false) {
  GemmArgs g{M, N, K, A, sam, sak, B, sbk, sbn, C, scm, scn, ((K + 15) / 16) * 16, 0, bias,
             relu ? 1 : 0, accumulate ? 1 : 0, 0, 0, 0, nullptr};
  dim3 grid((N + 63) / 64, (M + 63) / 64, 1);
  FR_CHECK(grid.y < 65536u, "GEMM too tall");
  launch_gemm<kGemmPlain>(c, grid, g, bf16);
  FR_HIP(hipGetLastError());
  return 0;
}

// out[C][B*Ho*Ho] = W[C][cin*9] x im2col(in) + bias, ReLU, with the im2col
// operand gathered inside the GEMM (the same K order as the explicit form)
int conv_forward(Ctx& c, int C, int cin, int B, int H, const float* W, const float* in,
                 const float* bias, float* out, bool bf16) {
  const int Ho = H / 2;
  const long long NB = (long long)B * Ho * Ho;
  FR_CHECK(NB < (1ll << 31), "batch too large");
  const int K = cin * 9, N = (int)NB;
  GemmArgs g{C, N, K, W, K, 1, in, NB, 1, out, NB, 1, ((K + 15) / 16) * 16, 0, bias, 1, 0, B, H,
             0, nullptr};
  dim3 grid((N + 63) / 64, (C + 63) / 64, 1);
  launch_gemm<kGemmConv>(c, grid, g, bf16);
  FR_HIP(hipGetLastError());
  return 0;
}

// split-K partial products of the weight gradients: at most this many floats
constexpr long long kPartFloats = 16ll << 20;
// workgroups the split-K weight gradient aims for (the GPU's 256 CUs, 8 each)
constexpr int kWgradWorkgroups = 2048;

// gW[C][cin*9] = sum over the B*Ho*Ho output positions n of
// g[C][n] x in[ci][b][2oy+ky-1][2ox+kx-1] (the im2col operand gathered in the
// GEMM), with the positions split over enough workgroups to fill the GPU;
// the splits' partial products (in `part`) are summed in split order
int conv_wgrad(Ctx& c, int C, int cin, int B, int H, const float* g_, const float* in,
               float* gw, float* part, bool bf16) {
  const int Ho = H / 2, N = cin * 9;
  const long long NB = (long long)B * Ho * Ho;
  const int tiles = ((N + 63) / 64) * ((C + 63) / 64);
  long long nsplit = (kWgradWorkgroups + tiles - 1) / tiles;
  nsplit = std::min(nsplit, std::max(1ll, NB / 256));
  nsplit = std::min(nsplit, std::max(1ll, kPartFloats / ((long long)C * N)));
  // whole K tiles per split (16 for FP32, 32 for BF16 operands)
  const int kt = bf16 ? 32 : 16;
  int kchunk = (int)(((NB + nsplit - 1) / nsplit + kt - 1) / kt * kt);
  nsplit = (NB + kchunk - 1) / kchunk;
  FR_CHECK(nsplit < 65536 && nsplit * C * N <= kPartFloats, "weight-gradient split too large");
  GemmArgs g{C, N, (int)NB, g_, NB, 1, in, 1, NB, part, N, 1, kchunk, (long long)C * N,
             nullptr, 0, 0, B, H, 0, nullptr};
  dim3 grid((N + 63) / 64, (C + 63) / 64, (unsigned)nsplit);
  launch_gemm<kGemmConvW>(c, grid, g, bf16);
  FR_HIP(hipGetLastError());
  if ((long long)C * N < 16384 && nsplit > 64)
    hipLaunchKernelGGL(k_split_sum_wide, dim3((unsigned)(C * N)), dim3(kT), 0, c.stream,
                       (long long)C * N, (int)nsplit, part, gw);
  else
    hipLaunchKernelGGL(k_split_sum, dim3(nblk((long long)C * N)), dim3(kT), 0, c.stream,
                       (long long)C * N, (int)nsplit, part, gw);
  FR_HIP(hipGetLastError());
  return 0;
}

// the layer's weights ordered for the input gradient's GEMMs, one block per
// parity class (py, px) (classes 0..3 at offsets 0, 1, 3, 5 times C*cin):
// wt_cls[ci][co*ntap + j] = w[co][ci*9 + ky*3 + kx] for the class's j-th tap
__global__ void __launch_bounds__(kT)
    k_wgrad_order(int C, int cin, const float* __restrict__ w, float* wt) {
  const int idx = blockIdx.x * kT + threadIdx.x;
  if (idx >= C * cin * 9) return;
  // idx -> class and element: classes hold 1, 2, 2, 4 taps
  const int per = C * cin;
  int cls, base;
  if (idx < per) { cls = 0; base = 0; }
  else if (idx < 3 * per) { cls = 1; base = per; }
  else if (idx < 5 * per) { cls = 2; base = 3 * per; }
  else { cls = 3; base = 5 * per; }
  const int py = cls >> 1, px = cls & 1, ntx = px + 1, ntap = (py + 1) * ntx;
  const int e = idx - base;
  const int ci = e / (C * ntap), r = e - ci * C * ntap, co = r / ntap, j = r - co * ntap;
  const int jy = j / ntx, jx = j - jy * ntx;
  const int ky = py ? 2 * jy : 1, kx = px ? 2 * jx : 1;
  wt[idx] = w[(co * cin + ci) * 9 + ky * 3 + kx];
}

// gin[ci][b][y][x] = (act[ci][b][y][x] > 0) x sum over co, ky, kx of
// w[co][ci][ky][kx] g[co][b][oy][ox] with 2oy+ky-1 = y, 2ox+kx-1 = x: the
// transposed convolution as four GEMMs, one per parity class of (y, x), each
// over only the taps that reach it (K = C x 1, 2, 2, 4), with the gradient
// gathered in place (no column matrix) and the ReLU derivative applied on
// the way out
int conv_xgrad(Ctx& c, int C, int cin, int B, int H, const float* w, const float* g_,
               const float* act, float* wt, float* gin, bool bf16) {
  const int Ho = H / 2;
  const long long N = (long long)B * Ho * Ho;
  FR_CHECK((long long)B * H * H < (1ll << 31), "batch too large");
  hipLaunchKernelGGL(k_wgrad_order, dim3(nblk((long long)C * cin * 9)), dim3(kT), 0, c.stream, C,
                     cin, w, wt);
  FR_HIP(hipGetLastError());
  static const int off[4] = {0, 1, 3, 5};
  for (int cls = 0; cls < 4; cls++) {
    const int ntap = ((cls >> 1) + 1) * ((cls & 1) + 1), K = C * ntap;
    GemmArgs g{cin, (int)N, K, wt + (size_t)off[cls] * C * cin, K, 1, g_, 0, 1, gin,
               (long long)B * H * H, 1, ((K + 15) / 16) * 16, 0, nullptr, 0, 0, B, H, cls, act};
    dim3 grid((unsigned)((N + 63) / 64), (cin + 63) / 64, 1);
    launch_gemm<kGemmConvX>(c, grid, g, bf16);
    FR_HIP(hipGetLastError());
  }
  return 0;
}

// workspace offsets (floats) for a batch of B
struct Ws {
  size_t act[8];  // act[l]: output of convolution l, [C][B][S>>(l+1)]^2
  size_t X, logits, wt, gA, gB, gX, gL, part, total;
};

constexpr int kMaxSplit = 64;  // splits of the bias-gradient row sums (x9)

Ws ws_layout(const CnnLayout& lo, int B, bool backward) {
  Ws w{};
  size_t o = 0;
  for (int l = 0; l < lo.L; l++) {
    w.act[l] = o;
    o += (size_t)lo.C * B * lo.hout(l) * lo.hout(l);
  }
  w.X = o;
  o += (size_t)lo.F() * B;
  w.logits = o;
  o += (size_t)lo.K * B;
  size_t gmax = 0;
  for (int l = 0; l < lo.L; l++) {
    size_t ga = (size_t)lo.C * B * lo.hout(l) * lo.hout(l);
    gmax = ga > gmax ? ga : gmax;
  }
  // every convolution GEMM, forward and backward, gathers its operand from
  // the activation or gradient tensor in place: no column matrices
  if (backward) {
    w.wt = o;
    o += (size_t)lo.C * lo.C * 9;
    w.gA = o;
    o += gmax;
    w.gB = o;
    o += gmax;
    w.gX = o;
    o += (size_t)lo.F() * B;
    w.gL = o;
    o += (size_t)lo.K * B;
    w.part = o;
    // split-K partials of the weight gradients, or the bias row sums' partials
    size_t pc = (size_t)lo.C * 9 * kMaxSplit;
    o += (size_t)kPartFloats > pc ? (size_t)kPartFloats : pc;
  }
  w.total = o;
  return w;
}

}  // namespace

int cnn_layout(int S, int C, int L, int K, CnnLayout* out) {
  FR_CHECK(L >= 1 && L <= 7, "nconv must be 1..7 (got %d)", L);
  FR_CHECK(S >= 2 && S <= 1024 && (S % (1 << L)) == 0, "input size %d must be a multiple of 2^%d",
           S, L);
  FR_CHECK(S % 2 == 0 && (S == 0 || fft_supported(S)), "input size %d not supported by the FFT", S);
  FR_CHECK(C >= 1 && C <= 256, "channels must be 1..256 (got %d)", C);
  FR_CHECK(K >= 1 && K <= 65536, "classes must be 1..65536 (got %d)", K);
  CnnLayout lo;
  lo.S = S;
  lo.C = C;
  lo.L = L;
  lo.K = K;
  long long o = 0;
  for (int l = 0; l < L; l++) {
    lo.w_off[l] = o;
    o += (long long)C * lo.cin(l) * 9;
    lo.b_off[l] = o;
    o += C;
  }
  lo.fw_off = o;
  o += (long long)K * lo.F();
  lo.fb_off = o;
  o += K;
  lo.total = o;
  *out = lo;
  return 0;
}

size_t cnn_workspace(const CnnLayout& lo, int B, bool backward) {
  return ws_layout(lo, B, backward).total;
}

int cnn_init(Ctx& c, const CnnLayout& lo, unsigned long long seed, float* master, float* wq) {
  for (int l = 0; l <= lo.L; l++) {
    long long wo, bo, nw, nbias;
    float stdv;
    if (l < lo.L) {
      wo = lo.w_off[l];
      bo = lo.b_off[l];
      nw = (long long)lo.C * lo.cin(l) * 9;
      nbias = lo.C;
      stdv = std::sqrt(2.0f / (float)(lo.cin(l) * 9));
    } else {
      wo = lo.fw_off;
      bo = lo.fb_off;
      nw = (long long)lo.K * lo.F();
      nbias = lo.K;
      stdv = std::sqrt(1.0f / (float)lo.F());
    }
    hipLaunchKernelGGL(k_init, dim3(nblk(nw)), dim3(kT), 0, c.stream, nw, seed, wo, stdv,
                       master + wo, wq + wo, 1);
    FR_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_init, dim3(nblk(nbias)), dim3(kT), 0, c.stream, nbias, seed, bo, 0.0f,
                       master + bo, wq + bo, 0);
    FR_HIP(hipGetLastError());
  }
  return 0;
}

namespace {

int forward_impl(Ctx& c, const CnnLayout& lo, const float* wq, int B, const float* img,
                 float* ws, const Ws& w, bool bf16) {
  const float* in = img;
  for (int l = 0; l < lo.L; l++) {
    const int H = lo.hin(l), cin = lo.cin(l);
    // out[C][NB] = W[C][cin*9] x cols[cin*9][NB] + b, ReLU
    if (int r = conv_forward(c, lo.C, cin, B, H, wq + lo.w_off[l], in, wq + lo.b_off[l],
                             ws + w.act[l], bf16))
      return r;
    in = ws + w.act[l];
  }
  const int P = lo.hout(lo.L - 1) * lo.hout(lo.L - 1);
  const long long nx = (long long)lo.C * B * P;
  hipLaunchKernelGGL(k_flatten, dim3(nblk(nx)), dim3(kT), 0, c.stream, lo.C, B, P, in, ws + w.X);
  FR_HIP(hipGetLastError());
  // logits[b][k] = sum_f Wfc[k][f] X[f][b] + bfc[k]  (C element (k, b) at b*K + k)
  return gemm(c, lo.K, B, lo.F(), wq + lo.fw_off, lo.F(), 1, ws + w.X, B, 1, ws + w.logits, 1,
              lo.K, wq + lo.fb_off, false, false, bf16);
}

}  // namespace

int cnn_forward(Ctx& c, const CnnLayout& lo, const float* wq, int B, const float* img, float* ws,
                float* logits) {
  Ws w = ws_layout(lo, B, false);
  if (int r = forward_impl(c, lo, wq, B, img, ws, w, false)) return r;
  FR_HIP(hipMemcpyAsync(logits, ws + w.logits, sizeof(float) * (size_t)lo.K * B,
                        hipMemcpyDeviceToDevice, c.stream));
  return 0;
}

int cnn_backward(Ctx& c, const CnnLayout& lo, const float* wq, int B, const float* img,
                 const int* label, float* ws, float* loss, float* grad, bool bf16) {
  Ws w = ws_layout(lo, B, true);
  if (int r = forward_impl(c, lo, wq, B, img, ws, w, bf16)) return r;
  const int K = lo.K, F = lo.F();
  hipLaunchKernelGGL(k_softmax_xent, dim3(B), dim3(kT), 0, c.stream, K, B, ws + w.logits, label,
                     loss, ws + w.gL);
  FR_HIP(hipGetLastError());
  // gWfc[k][f] = sum_b gL[b][k] X[f][b]
  if (int r = gemm(c, K, F, B, ws + w.gL, 1, K, ws + w.X, 1, B, grad + lo.fw_off, F, 1, nullptr,
                   false, false, bf16))
    return r;
  // gbfc[k] = sum_b gL[b][k]
  hipLaunchKernelGGL(k_rowsum, dim3(K), dim3(kT), 0, c.stream, B, 1ll, (long long)K, ws + w.gL,
                     grad + lo.fb_off);
  FR_HIP(hipGetLastError());
  // gX[f][b] = sum_k Wfc[k][f] gL[b][k]
  if (int r = gemm(c, F, B, K, wq + lo.fw_off, 1, F, ws + w.gL, 1, K, ws + w.gX, B, 1, nullptr,
                   false, false, bf16))
    return r;
  const int PL = lo.hout(lo.L - 1) * lo.hout(lo.L - 1);
  const long long nl = (long long)lo.C * B * PL;
  float* gcur = ws + w.gA;
  float* gnext = ws + w.gB;
  hipLaunchKernelGGL(k_unflatten_relu, dim3(nblk(nl)), dim3(kT), 0, c.stream, lo.C, B, PL,
                     ws + w.gX, ws + w.act[lo.L - 1], gcur);
  FR_HIP(hipGetLastError());
  for (int l = lo.L - 1; l >= 0; l--) {
    const int H = lo.hin(l), Ho = lo.hout(l), cin = lo.cin(l);
    const long long NB = (long long)B * Ho * Ho;
    const float* in = l == 0 ? img : ws + w.act[l - 1];
    // gW[C][cin*9] = sum_n g[C][n] im2col(in)[cin*9][n]
    if (int r = conv_wgrad(c, lo.C, cin, B, H, gcur, in, grad + lo.w_off[l], ws + w.part, bf16))
      return r;
    // gb[C] = sum_n g[C][n]: per-split partials in ws.part (free again once
    // the split-K sum above has run), then the splits in order
    const int bchunk = 16384;
    const int bsplit = (int)((NB + bchunk - 1) / bchunk);
    FR_CHECK(bsplit <= kMaxSplit * 9, "batch too large for the bias reduction");
    hipLaunchKernelGGL(k_rowsum_part, dim3(lo.C, bsplit), dim3(kT), 0, c.stream, (int)NB, NB,
                       bchunk, gcur, ws + w.part);
    FR_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_rowsum, dim3(lo.C), dim3(kT), 0, c.stream, bsplit, (long long)bsplit, 1ll,
                       ws + w.part, grad + lo.b_off[l]);
    FR_HIP(hipGetLastError());
    if (l > 0) {
      // the input gradient through the layer below's ReLU
      if (int r = conv_xgrad(c, lo.C, cin, B, H, wq + lo.w_off[l], gcur, ws + w.act[l - 1],
                             ws + w.wt, gnext, bf16))
        return r;
      float* t = gcur;
      gcur = gnext;
      gnext = t;
    }
  }
  return 0;
}

int cnn_adam(Ctx& c, const CnnLayout& lo, float* master, float* wq, const float* grad, float* m,
             float* v, int step, float lr) {
  FR_CHECK(step >= 1, "Adam step counts from 1");
  const float c1 = 1.0f - std::pow(0.9f, (float)step);
  const float c2 = 1.0f - std::pow(0.999f, (float)step);
  for (int l = 0; l <= lo.L; l++) {
    long long wo = l < lo.L ? lo.w_off[l] : lo.fw_off;
    long long bo = l < lo.L ? lo.b_off[l] : lo.fb_off;
    long long nw = bo - wo;
    long long nbias = l < lo.L ? lo.C : lo.K;
    hipLaunchKernelGGL(k_adam, dim3(nblk(nw)), dim3(kT), 0, c.stream, nw, master + wo, wq + wo,
                       grad + wo, m + wo, v + wo, lr, c1, c2, 1);
    FR_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_adam, dim3(nblk(nbias)), dim3(kT), 0, c.stream, nbias, master + bo,
                       wq + bo, grad + bo, m + bo, v + bo, lr, c1, c2, 0);
    FR_HIP(hipGetLastError());
  }
  return 0;
}

int cnn_propose(Ctx& c, int K, int B, const float* logits, int topk, int* top, float* prob) {
  FR_CHECK(topk >= 1 && topk <= K && topk <= 64, "topk must be 1..min(K, 64)");
  FR_CHECK((size_t)K * sizeof(float) <= 48 * 1024, "too many classes for the proposal kernel");
  hipLaunchKernelGGL(k_propose, dim3(B), dim3(kT), sizeof(float) * K, c.stream, K, logits, topk,
                     top, prob);
  FR_HIP(hipGetLastError());
  return 0;
}

int cnn_images(Ctx& c, int nsam, int S, int B, const float* datc, const float* ctff, float* buf,
               float* img) {
  const size_t is = (size_t)S * (S / 2 + 1);
  hipLaunchKernelGGL(k_cnn_crop, dim3(nblk((long long)(is * B))), dim3(kT), 0, c.stream, nsam, S,
                     B, reinterpret_cast<const float2*>(datc), reinterpret_cast<const float2*>(ctff),
                     reinterpret_cast<float2*>(buf));
  FR_HIP(hipGetLastError());
  if (int r = fft_c2r_2d(c, S, B, buf, (size_t)S * (S + 2))) return r;
  if (S * S <= 16 * 1024)
    hipLaunchKernelGGL(k_cnn_normalise_reg, dim3(B), dim3(1024), 0, c.stream, S, buf, img);
  else
    hipLaunchKernelGGL(k_cnn_normalise, dim3(B), dim3(kT), 0, c.stream, S, buf, img);
  FR_HIP(hipGetLastError());
  return 0;
}

int cnn_synth(Ctx& c, int nsam, int S, int B, const float* a3df, const float* dm,
              const float* spar, const PrepGlobals& g, float disc, unsigned long long seed,
              float* buf, float* img) {
  hipLaunchKernelGGL(k_cnn_synth, dim3(B), dim3(kT), 0, c.stream, nsam, S,
                     reinterpret_cast<const float2*>(a3df), dm, spar, g, disc, seed,
                     reinterpret_cast<float2*>(buf));
  FR_HIP(hipGetLastError());
  if (int r = fft_c2r_2d(c, S, B, buf, (size_t)S * (S + 2))) return r;
  if (S * S <= 16 * 1024)
    hipLaunchKernelGGL(k_cnn_normalise_reg, dim3(B), dim3(1024), 0, c.stream, S, buf, img);
  else
    hipLaunchKernelGGL(k_cnn_normalise, dim3(B), dim3(kT), 0, c.stream, S, buf, img);
  FR_HIP(hipGetLastError());
  return 0;
}

int orient_nearest(Ctx& c, int n, const float* dm, int ncls, const float* dmc, int nsym,
                   const float* sym, int* label) {
  FR_CHECK(nsym >= 1 && nsym <= 60, "1..60 symmetry operators");
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_orient_nearest, dim3(n), dim3(kT), 0, c.stream, n, dm, ncls, dmc, nsym,
                     sym, label);
  FR_HIP(hipGetLastError());
  return 0;
}

int orient_within(Ctx& c, int na, const float* dma, int nb, const float* dmb, int nsym,
                  const float* sym, float tmin, unsigned* mask) {
  FR_CHECK(nsym >= 1 && nsym <= 60, "1..60 symmetry operators");
  const long long nw = (na + 31) / 32;
  hipLaunchKernelGGL(k_orient_within, dim3(nblk(nw * nb)), dim3(kT), 0, c.stream, na, dma, nb,
                     dmb, nsym, sym, tmin, mask);
  FR_HIP(hipGetLastError());
  return 0;
}


// ------------------------------------------------------------------------
// Inference with BF16 activations (the proposal's forward).  The weights are
// the same BF16 values; activations are stored as BF16, channels innermost
// ([B][H][W][C]), and the convolutions after the first run on the BF16
// matrix cores with FP32 accumulation.  K is ordered (tap, channel) so that
// a fragment's 8 consecutive K values are 8 channels of one input pixel: one
// 16-byte load.  Bias and ReLU are applied in FP32, then rounded to BF16.
// The fully connected layer reads the last BF16 activation directly.
namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
constexpr int kInferC = 64;  // output channels per tile (the whole layer)

struct WOff {
  long long o[8];
};

// weights of convolutions 1..L-1 as BF16 [co][tap][ci]
__global__ void __launch_bounds__(kT)
    k_cnn_wbf16(int C, int L, const float* __restrict__ wq, WOff woff, __bf16* wb) {
  const long long per = (long long)C * 9 * C;
  const long long gid = blockIdx.x * (long long)kT + threadIdx.x;
  if (gid >= per * (L - 1)) return;
  const int l = 1 + (int)(gid / per);
  const long long e = gid % per;
  const int ci = (int)(e % C), t = (int)((e / C) % 9), co = (int)(e / (9LL * C));
  wb[gid] = (__bf16)wq[woff.o[l] + ((long long)co * C + ci) * 9 + t];
}

// convolution 0 (one input channel): one thread per output pixel, all C
// channels, FP32 arithmetic in the im2col order of the training forward
__global__ void __launch_bounds__(kT)
    k_cnn_conv0_bf16(int B, int H, int C, const float* __restrict__ img,
                     const float* __restrict__ w0, const float* __restrict__ b0, __bf16* out) {
  __shared__ float ws[kInferC * 9 + kInferC];
  for (int e = threadIdx.x; e < C * 9; e += kT) ws[e] = w0[e];
  for (int e = threadIdx.x; e < C; e += kT) ws[C * 9 + e] = b0[e];
  __syncthreads();
  const int Ho = H / 2;
  const long long pos = blockIdx.x * (long long)kT + threadIdx.x;
  if (pos >= (long long)B * Ho * Ho) return;
  const int b = (int)(pos / (Ho * Ho)), r = (int)(pos % (Ho * Ho)), oy = r / Ho, ox = r % Ho;
  float x[9];
#pragma unroll
  for (int t = 0; t < 9; t++) {
    const int y = 2 * oy + t / 3 - 1, xx = 2 * ox + t % 3 - 1;
    x[t] = (y >= 0 && y < H && xx >= 0 && xx < H) ? img[((long long)b * H + y) * H + xx] : 0.f;
  }
  uint4* dst = reinterpret_cast<uint4*>(out + pos * C);
  for (int c0 = 0; c0 < C; c0 += 8) {
    bf16x8_t v;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      float z = 0.f;
#pragma unroll
      for (int t = 0; t < 9; t++) z += ws[(c0 + j) * 9 + t] * x[t];
      z += ws[C * 9 + c0 + j];
      v[j] = (__bf16)(z > 0.f ? z : 0.f);
    }
    dst[c0 / 8] = *reinterpret_cast<uint4*>(&v);
  }
}

// convolution l >= 1 with C = 64: a workgroup computes all 64 output
// channels of 256 output pixels (the weights are read once per 256 pixels);
// wave w takes pixels 64 w .. 64 w + 63 as four 16-pixel blocks
__global__ void __launch_bounds__(kT)
    k_cnn_conv_bf16(int B, int H, const __bf16* __restrict__ in, const __bf16* __restrict__ wb,
                    const float* __restrict__ bias, __bf16* out) {
  constexpr int C = kInferC;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, n = lane & 15;
  const int Ho = H / 2;
  const long long npos = (long long)B * Ho * Ho;
  long long pos[4];
  int pb[4], py[4], px[4];
#pragma unroll
  for (int nb = 0; nb < 4; nb++) {
    pos[nb] = blockIdx.x * 256LL + wave * 64 + nb * 16 + n;
    pb[nb] = py[nb] = px[nb] = -1;
    if (pos[nb] < npos) {
      pb[nb] = (int)(pos[nb] / (Ho * Ho));
      const int r = (int)(pos[nb] % (Ho * Ho));
      py[nb] = 2 * (r / Ho) - 1;
      px[nb] = 2 * (r % Ho) - 1;
    }
  }
  fx4 acc[4][4];
#pragma unroll
  for (int rb = 0; rb < 4; rb++)
#pragma unroll
    for (int nb = 0; nb < 4; nb++) acc[rb][nb] = fx4{0.f, 0.f, 0.f, 0.f};
  for (int s = 0; s < 18; s++) {
    const int tap = s >> 1, ci0 = (s & 1) * 32 + 8 * g;
    const int ky = tap / 3, kx = tap % 3;
    bf16x8_t bf[4];
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
      const int y = py[nb] + ky, x = px[nb] + kx;
      uint4 bv = make_uint4(0u, 0u, 0u, 0u);
      if (pb[nb] >= 0 && y >= 0 && y < H && x >= 0 && x < H)
        bv = *reinterpret_cast<const uint4*>(in + (((long long)pb[nb] * H + y) * H + x) * C + ci0);
      bf[nb] = *reinterpret_cast<bf16x8_t*>(&bv);
    }
#pragma unroll
    for (int rb = 0; rb < 4; rb++) {
      const uint4 av =
          *reinterpret_cast<const uint4*>(wb + ((long long)(rb * 16 + n) * 9 + tap) * C + ci0);
      const bf16x8_t af = *reinterpret_cast<const bf16x8_t*>(&av);
#pragma unroll
      for (int nb = 0; nb < 4; nb++)
        acc[rb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bf[nb], acc[rb][nb], 0, 0, 0);
    }
  }
  // lane holds output channels rb*16 + 4g + i of pixel block nb, pixel n
#pragma unroll
  for (int nb = 0; nb < 4; nb++) {
    if (pos[nb] >= npos) continue;
#pragma unroll
    for (int rb = 0; rb < 4; rb++) {
      const int co = rb * 16 + 4 * g;
      __bf16 v[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const float z = acc[rb][nb][i] + bias[co + i];
        v[i] = (__bf16)(z > 0.f ? z : 0.f);
      }
      *reinterpret_cast<uint2*>(out + pos[nb] * C + co) = *reinterpret_cast<uint2*>(v);
    }
  }
}

// fully connected weights W[k][c P + p] as BF16 [k][p][c]
__global__ void __launch_bounds__(kT)
    k_cnn_fcw_bf16(int K, int P, int C, const float* __restrict__ W, __bf16* wf) {
  const long long gid = blockIdx.x * (long long)kT + threadIdx.x;
  if (gid >= (long long)K * P * C) return;
  const int c = (int)(gid % C), p = (int)((gid / C) % P);
  const long long k = gid / ((long long)C * P);
  wf[gid] = (__bf16)W[k * C * P + (long long)c * P + p];
}

// logits[b][k] = sum_f Wfc[k][f] X_f(b) + bfc[k] from the last activation
// [B][P][C] (BF16), f = c P + p; K ordered (p, c) so that a fragment is 8
// channels of one pixel (weights from k_cnn_fcw_bf16).  A workgroup: 16
// classes x 64 images, wave w the images 16 w .. 16 w + 15.
__global__ void __launch_bounds__(kT)
    k_cnn_fc_bf16(int K, int B, int P, int C, const __bf16* __restrict__ wf,
                  const float* __restrict__ bias, const __bf16* __restrict__ in, float* logits) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, n = lane & 15;
  const int k0 = blockIdx.x * 16, kr = k0 + n;
  const int b = blockIdx.y * 64 + wave * 16 + n;
  fx4 acc = fx4{0.f, 0.f, 0.f, 0.f};
  for (int p = 0; p < P; p++)
    for (int c0 = 0; c0 < C; c0 += 32) {
      const int c = c0 + 8 * g;
      bf16x8_t af, bfv;
      uint4 av = make_uint4(0u, 0u, 0u, 0u);
      if (kr < K) av = *reinterpret_cast<const uint4*>(wf + ((long long)kr * P + p) * C + c);
      af = *reinterpret_cast<bf16x8_t*>(&av);
      uint4 bv = make_uint4(0u, 0u, 0u, 0u);
      if (b < B) bv = *reinterpret_cast<const uint4*>(in + ((long long)b * P + p) * C + c);
      bfv = *reinterpret_cast<bf16x8_t*>(&bv);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bfv, acc, 0, 0, 0);
    }
  // lane holds classes k0 + 4g + i of image b
  if (b >= B) return;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int k = k0 + 4 * g + i;
    if (k < K) logits[(long long)b * K + k] = acc[i] + bias[k];
  }
}

}  // namespace

bool cnn_bf16_ok(const CnnLayout& lo) { return lo.C == kInferC && lo.L >= 2 && lo.L <= 8; }

size_t cnn_workspace_bf16(const CnnLayout& lo, int B) {
  const size_t act = (size_t)B * lo.hout(0) * lo.hout(0) * lo.C;  // largest activation
  const size_t wb = (size_t)(lo.L - 1) * lo.C * 9 * lo.C;
  // bytes: two BF16 activation buffers, the BF16 convolution and fully
  // connected weights
  return 2 * act * 2 + wb * 2 + (size_t)lo.K * lo.F() * 2 + 256;
}

int cnn_forward_bf16(Ctx& c, const CnnLayout& lo, const float* wq, int B, const float* img,
                     void* ws, float* logits) {
  FR_CHECK(cnn_bf16_ok(lo), "BF16 inference needs C = %d channels and 2..8 convolutions",
           kInferC);
  unsigned char* p = reinterpret_cast<unsigned char*>(ws);
  const size_t act = (size_t)B * lo.hout(0) * lo.hout(0) * lo.C;
  __bf16* a0 = reinterpret_cast<__bf16*>(p);
  __bf16* a1 = a0 + act;
  __bf16* wb = a1 + act;
  const size_t nwb = (size_t)(lo.L - 1) * lo.C * 9 * lo.C;
  WOff woff;
  for (int l = 0; l < 8; l++) woff.o[l] = lo.w_off[l];
  hipLaunchKernelGGL(k_cnn_wbf16, dim3(nblk((long long)nwb)), dim3(kT), 0, c.stream, lo.C, lo.L,
                     wq, woff, wb);
  FR_HIP(hipGetLastError());
  const int H0 = lo.hin(0), Ho0 = lo.hout(0);
  hipLaunchKernelGGL(k_cnn_conv0_bf16, dim3(nblk((long long)B * Ho0 * Ho0)), dim3(kT), 0,
                     c.stream, B, H0, lo.C, img, wq + lo.w_off[0], wq + lo.b_off[0], a0);
  FR_HIP(hipGetLastError());
  __bf16* cur = a0;
  __bf16* nxt = a1;
  for (int l = 1; l < lo.L; l++) {
    const int H = lo.hin(l), Ho = lo.hout(l);
    const long long npos = (long long)B * Ho * Ho;
    hipLaunchKernelGGL(k_cnn_conv_bf16, dim3((unsigned)((npos + 255) / 256)), dim3(kT), 0, c.stream,
                       B, H, cur, wb + (size_t)(l - 1) * lo.C * 9 * lo.C, wq + lo.b_off[l], nxt);
    FR_HIP(hipGetLastError());
    __bf16* t = cur;
    cur = nxt;
    nxt = t;
  }
  const int P = lo.hout(lo.L - 1) * lo.hout(lo.L - 1);
  __bf16* wf = wb + nwb;
  hipLaunchKernelGGL(k_cnn_fcw_bf16, dim3(nblk((long long)lo.K * lo.F())), dim3(kT), 0, c.stream,
                     lo.K, P, lo.C, wq + lo.fw_off, wf);
  FR_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_cnn_fc_bf16, dim3((lo.K + 15) / 16, (B + 63) / 64), dim3(kT), 0, c.stream,
                     lo.K, B, P, lo.C, wf, wq + lo.fb_off, cur, logits);
  FR_HIP(hipGetLastError());
  return 0;
}

}  // namespace fr

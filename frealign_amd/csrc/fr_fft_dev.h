// fr_fft_dev.h -- device building blocks of the batched FFTs: complex
// helpers, the radix-2/3/4/5 butterflies and the Stockham autosort passes in
// LDS with the length known at compile time.  Used by the FFT kernels and by
// kernels that run a transform between other work on the same tile.
#pragma once
#include <hip/hip_runtime.h>

namespace fr {
namespace fftdev {

constexpr int kFftThreads = 256;  // threads of every workgroup that runs these passes

// XCD-aware block order.  The hardware deals a launch's workgroups to the 8
// XCDs round robin (workgroup p to XCD p mod 8); with a grid of g8 blocks (a
// multiple of 8), logical block (p mod 8) (g8 / 8) + p / 8 gives each XCD a
// contiguous range of logical blocks, so neighbouring blocks whose data share
// cache lines (the 8-column blocks of one half-plane row) share one L2.
__device__ inline int xcd_block(int p, int g8) { return (p & 7) * (g8 >> 3) + (p >> 3); }
inline int round8(int g) { return (g + 7) & ~7; }

__device__ inline float2 f2(float a, float b) { return make_float2(a, b); }
__device__ inline float2 add2(float2 a, float2 b) { return f2(a.x + b.x, a.y + b.y); }
__device__ inline float2 sub2(float2 a, float2 b) { return f2(a.x - b.x, a.y - b.y); }
__device__ inline float2 mul2(float2 a, float2 b) {
  return f2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline float2 scl2(float2 a, float s) { return f2(a.x * s, a.y * s); }
// multiply by i*sign
__device__ inline float2 muli(float2 a, int sign) { return sign > 0 ? f2(-a.y, a.x) : f2(a.y, -a.x); }

template <int R>
__device__ inline void butterfly(float2* u, int sign) {
  if constexpr (R == 2) {
    float2 a = u[0], b = u[1];
    u[0] = add2(a, b);
    u[1] = sub2(a, b);
  } else if constexpr (R == 4) {
    float2 a0 = add2(u[0], u[2]), a1 = sub2(u[0], u[2]);
    float2 b0 = add2(u[1], u[3]), jt = muli(sub2(u[1], u[3]), sign);
    u[0] = add2(a0, b0);
    u[2] = sub2(a0, b0);
    u[1] = add2(a1, jt);
    u[3] = sub2(a1, jt);
  } else if constexpr (R == 3) {
    const float h = 0.86602540378443864676f;
    float2 s = add2(u[1], u[2]), d = sub2(u[1], u[2]);
    float2 v0 = add2(u[0], s);
    float2 m = sub2(u[0], scl2(s, 0.5f));
    float2 nn = scl2(muli(d, sign), h);
    u[0] = v0;
    u[1] = add2(m, nn);
    u[2] = sub2(m, nn);
  } else if constexpr (R == 5) {
    const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
    const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
    float2 a1 = add2(u[1], u[4]), b1 = sub2(u[1], u[4]);
    float2 a2 = add2(u[2], u[3]), b2 = sub2(u[2], u[3]);
    float2 v0 = add2(add2(u[0], a1), a2);
    float2 p1 = add2(u[0], add2(scl2(a1, c1), scl2(a2, c2)));
    float2 q1 = muli(add2(scl2(b1, s1), scl2(b2, s2)), sign);
    float2 p2 = add2(u[0], add2(scl2(a1, c2), scl2(a2, c1)));
    float2 q2 = muli(sub2(scl2(b1, s2), scl2(b2, s1)), sign);
    u[0] = v0;
    u[1] = add2(p1, q1);
    u[4] = sub2(p1, q1);
    u[2] = add2(p2, q2);
    u[3] = sub2(p2, q2);
  }
}

// Radix of pass i in make_plan's order (0 past the last pass).
constexpr int radix_at(int n, int i) {
  int m = n, k = 0;
  while (m % 4 == 0) { if (k == i) return 4; m /= 4; k++; }
  while (m % 2 == 0) { if (k == i) return 2; m /= 2; k++; }
  while (m % 3 == 0) { if (k == i) return 3; m /= 3; k++; }
  while (m % 5 == 0) { if (k == i) return 5; m /= 5; k++; }
  return 0;
}

template <int R, int N, int P, int NSEQ, int LD, int SIGN>
__device__ inline void stockham_pass_t(const float2* in, float2* out,
                                       const float2* __restrict__ tw) {
  constexpr int T = N / R;
  constexpr int stride = N / (P * R);
  for (int b = threadIdx.x; b < NSEQ * T; b += kFftThreads) {
    const int s = b / T, i = b - s * T;
    const int k = i % P;
    const float2* x = in + s * LD;
    float2 u[R];
#pragma unroll
    for (int r = 0; r < R; r++) u[r] = x[i + r * T];
#pragma unroll
    for (int r = 1; r < R; r++) {
      float2 w = tw[(r * k * stride) % N];
      if (SIGN > 0) w.y = -w.y;
      u[r] = mul2(u[r], w);
    }
    butterfly<R>(u, SIGN);
    float2* y = out + s * LD;
    const int base = (i - k) * R + k;
#pragma unroll
    for (int q = 0; q < R; q++) y[base + q * P] = u[q];
  }
}

template <int N, int NSEQ, int LD, int SIGN, int I = 0, int P = 1>
__device__ inline float2* lds_fft_t(float2* buf, float2* tmp, const float2* __restrict__ tw) {
  constexpr int R = radix_at(N, I);
  if constexpr (R == 0) {
    return buf;
  } else {
    stockham_pass_t<R, N, P, NSEQ, LD, SIGN>(buf, tmp, tw);
    __syncthreads();
    return lds_fft_t<N, NSEQ, LD, SIGN, I + 1, P * R>(tmp, buf, tw);
  }
}


// columns per workgroup of the column pass (16 at 256 measured no faster)
template <int N>
constexpr int cols_cw() {
  constexpr int JC = N / 2 + 1;
  return (2048 / N) < JC ? (2048 / N) : JC;
}

}  // namespace fftdev
}  // namespace fr

// fr_fft.hip -- batched 2D real FFTs for the particle and projection planes.
//
// Replaces the FFTW 3.3 plans of fftw_plans.f:12-13 (fftwf_plan_dft_r2c_2d /
// fftwf_plan_dft_c2r_2d, in place, unnormalised) for the box sizes the hot
// path uses: N = 2^a 3^b 5^c, N <= 1024 (64, 96, 128, 160, 256 and the 4x
// padded 1024 of pad.f).
//
// Structure: row pass (two real rows packed into one complex sequence),
// column pass over the JC half-plane columns, both as Stockham autosort FFTs
// in LDS with radix-4/2/3/5 butterflies and an accurate twiddle table
// (W_N^k computed on the host in FP64).
#include <hip/hip_runtime.h>

#include <vector>

#include "fr_common.h"
#include "fr_fft_dev.h"

namespace fr {

namespace {

using namespace fftdev;
constexpr int kThreads = kFftThreads;
constexpr int kMaxPass = 12;

struct Plan {
  int n;
  int npass;
  int radix[kMaxPass];
};

bool make_plan(int n, Plan* p) {
  if (n < 2 || n > 1024) return false;
  p->n = n;
  p->npass = 0;
  int m = n;
  while (m % 4 == 0 && p->npass < kMaxPass) { p->radix[p->npass++] = 4; m /= 4; }
  while (m % 2 == 0 && p->npass < kMaxPass) { p->radix[p->npass++] = 2; m /= 2; }
  while (m % 3 == 0 && p->npass < kMaxPass) { p->radix[p->npass++] = 3; m /= 3; }
  while (m % 5 == 0 && p->npass < kMaxPass) { p->radix[p->npass++] = 5; m /= 5; }
  return m == 1;
}

template <int R>
__device__ inline void stockham_pass(const float2* in, float2* out, int n, int ld, int nseq, int p,
                                     int sign, const float2* __restrict__ tw) {
  const int T = n / R;
  const int stride = n / (p * R);
  for (int b = threadIdx.x; b < nseq * T; b += blockDim.x) {
    int s = b / T, i = b - s * T;
    int k = i % p;
    const float2* x = in + s * ld;
    float2 u[R];
#pragma unroll
    for (int r = 0; r < R; r++) u[r] = x[i + r * T];
#pragma unroll
    for (int r = 1; r < R; r++) {
      int e = (r * k * stride) % n;
      float2 w = tw[e];
      if (sign > 0) w.y = -w.y;
      u[r] = mul2(u[r], w);
    }
    butterfly<R>(u, sign);
    float2* y = out + s * ld;
    int base = (i - k) * R + k;
#pragma unroll
    for (int q = 0; q < R; q++) y[base + q * p] = u[q];
  }
}

// In-LDS FFT of nseq sequences (stride ld); returns the buffer holding the result.
__device__ float2* lds_fft(float2* buf, float2* tmp, const Plan& pl, int ld, int nseq, int sign,
                           const float2* __restrict__ tw) {
  int p = 1;
  for (int ps = 0; ps < pl.npass; ps++) {
    int R = pl.radix[ps];
    switch (R) {
      case 4: stockham_pass<4>(buf, tmp, pl.n, ld, nseq, p, sign, tw); break;
      case 2: stockham_pass<2>(buf, tmp, pl.n, ld, nseq, p, sign, tw); break;
      case 3: stockham_pass<3>(buf, tmp, pl.n, ld, nseq, p, sign, tw); break;
      default: stockham_pass<5>(buf, tmp, pl.n, ld, nseq, p, sign, tw); break;
    }
    __syncthreads();
    float2* t = buf;
    buf = tmp;
    tmp = t;
    p *= R;
  }
  return buf;
}

// ---- row pass, forward: rows (2q, 2q+1) -> z = a + i b -> FFT -> unpack to half spectra
__global__ void __launch_bounds__(kThreads) k_rows_r2c(float* data, Plan pl, int nrows_per_img,
                                                       size_t img_stride, int npairs, int nseq,
                                                       const float2* __restrict__ tw) {
  extern __shared__ float2 smem[];
  const int n = pl.n, ld = n + 1;
  float2* buf = smem;
  float2* tmp = smem + nseq * ld;
  const int pair0 = blockIdx.x * nseq;
  const int pitch = n + 2;
  for (int e = threadIdx.x; e < nseq * n; e += blockDim.x) {
    int s = e / n, x = e - s * n;
    int pr = pair0 + s;
    float2 v = f2(0.f, 0.f);
    if (pr < npairs) {
      int img = (2 * pr) / nrows_per_img, row = (2 * pr) % nrows_per_img;
      const float* ra = data + img * img_stride + (size_t)row * pitch;
      v = f2(ra[x], ra[x + pitch]);
    }
    buf[s * ld + x] = v;
  }
  __syncthreads();
  float2* res = lds_fft(buf, tmp, pl, ld, nseq, -1, tw);
  const int jc = n / 2 + 1;
  for (int e = threadIdx.x; e < nseq * jc; e += blockDim.x) {
    int s = e / jc, k = e - s * jc;
    int pr = pair0 + s;
    if (pr >= npairs) continue;
    float2 zk = res[s * ld + k];
    float2 zm = res[s * ld + ((n - k) % n)];
    // Xa = (Z_k + conj Z_-k)/2 ; Xb = (Z_k - conj Z_-k)/(2i)
    float2 xa = f2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
    float2 xb = f2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
    int img = (2 * pr) / nrows_per_img, row = (2 * pr) % nrows_per_img;
    float* ra = data + img * img_stride + (size_t)row * pitch;
    ra[2 * k] = xa.x;
    ra[2 * k + 1] = xa.y;
    ra[pitch + 2 * k] = xb.x;
    ra[pitch + 2 * k + 1] = xb.y;
  }
}

// ---- row pass, backward: half spectra of rows (2q, 2q+1) -> real rows
__global__ void __launch_bounds__(kThreads) k_rows_c2r(float* data, Plan pl, int nrows_per_img,
                                                       size_t img_stride, int npairs, int nseq,
                                                       const float2* __restrict__ tw) {
  extern __shared__ float2 smem[];
  const int n = pl.n, ld = n + 1, jc = n / 2 + 1, pitch = n + 2;
  float2* buf = smem;
  float2* tmp = smem + nseq * ld;
  const int pair0 = blockIdx.x * nseq;
  for (int e = threadIdx.x; e < nseq * jc; e += blockDim.x) {
    int s = e / jc, k = e - s * jc;
    int pr = pair0 + s;
    float2 a = f2(0.f, 0.f), b = f2(0.f, 0.f);
    if (pr < npairs) {
      int img = (2 * pr) / nrows_per_img, row = (2 * pr) % nrows_per_img;
      const float* ra = data + img * img_stride + (size_t)row * pitch;
      a = f2(ra[2 * k], ra[2 * k + 1]);
      b = f2(ra[pitch + 2 * k], ra[pitch + 2 * k + 1]);
    }
    if (k == 0 || 2 * k == n) { a.y = 0.f; b.y = 0.f; }
    // Z_k = A_k + i B_k ; Z_{n-k} = conj(A_k) + i conj(B_k)
    buf[s * ld + k] = f2(a.x - b.y, a.y + b.x);
    if (k > 0 && 2 * k < n) buf[s * ld + (n - k)] = f2(a.x + b.y, -a.y + b.x);
  }
  __syncthreads();
  float2* res = lds_fft(buf, tmp, pl, ld, nseq, +1, tw);
  for (int e = threadIdx.x; e < nseq * pitch; e += blockDim.x) {
    int s = e / pitch, x = e - s * pitch;
    int pr = pair0 + s;
    if (pr >= npairs) continue;
    int img = (2 * pr) / nrows_per_img, row = (2 * pr) % nrows_per_img;
    float* ra = data + img * img_stride + (size_t)row * pitch;
    float2 z = x < n ? res[s * ld + x] : f2(0.f, 0.f);
    ra[x] = z.x;
    ra[pitch + x] = z.y;
  }
}

// ---- column pass over the half-plane: JC columns of length n (complex, stride JC)
__global__ void __launch_bounds__(kThreads) k_cols(float* data, Plan pl, size_t img_stride, int cw,
                                                   int nimg, int sign,
                                                   const float2* __restrict__ tw) {
  extern __shared__ float2 smem[];
  const int n = pl.n, ld = n + 1, jc = n / 2 + 1;
  float2* buf = smem;
  float2* tmp = smem + cw * ld;
  // column blocks of one image in XCD-contiguous order
  const int jb = (jc + cw - 1) / cw;
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int img = lb / jb;
  if (img >= nimg) return;
  const int x0 = (lb - img * jb) * cw;
  const int ncol = min(cw, jc - x0);
  float2* base = reinterpret_cast<float2*>(data + img * img_stride);
  for (int e = threadIdx.x; e < n * cw; e += blockDim.x) {
    int y = e / cw, c = e - y * cw;
    float2 v = f2(0.f, 0.f);
    if (c < ncol) v = base[(size_t)y * jc + x0 + c];
    buf[c * ld + y] = v;
  }
  __syncthreads();
  float2* res = lds_fft(buf, tmp, pl, ld, cw, sign, tw);
  for (int e = threadIdx.x; e < n * cw; e += blockDim.x) {
    int y = e / cw, c = e - y * cw;
    if (c < ncol) base[(size_t)y * jc + x0 + c] = res[c * ld + y];
  }
}

// the third axis of a 3D transform: ncols columns of length n, element y of
// column q at complex index y * rs + q (rs = JC*NSAM for a padded volume)
__global__ void __launch_bounds__(kThreads) k_cols_z(float* data, Plan pl, int ncols, size_t rs,
                                                     int cw, int sign,
                                                     const float2* __restrict__ tw) {
  extern __shared__ float2 smem[];
  const int n = pl.n, ld = n + 1;
  float2* buf = smem;
  float2* tmp = smem + cw * ld;
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int x0 = lb * cw;
  if (x0 >= ncols) return;
  const int ncol = min(cw, ncols - x0);
  float2* base = reinterpret_cast<float2*>(data);
  for (int e = threadIdx.x; e < n * cw; e += blockDim.x) {
    int y = e / cw, c = e - y * cw;
    float2 v = f2(0.f, 0.f);
    if (c < ncol) v = base[(size_t)y * rs + x0 + c];
    buf[c * ld + y] = v;
  }
  __syncthreads();
  float2* res = lds_fft(buf, tmp, pl, ld, cw, sign, tw);
  for (int e = threadIdx.x; e < n * cw; e += blockDim.x) {
    int y = e / cw, c = e - y * cw;
    if (c < ncol) base[(size_t)y * rs + x0 + c] = res[c * ld + y];
  }
}

int row_seqs(int n) {
  int s = 2048 / n;
  return s < 1 ? 1 : s;
}

// ---- the same three kernels with the length known at compile time (the
// box sizes of the bench and the proposal: 64, 128, 256): every index
// division and modulus becomes a shift or a constant multiply.
template <int N>
__global__ void __launch_bounds__(kThreads) k_rows_r2c_t(float* data, size_t img_stride,
                                                         int npairs, const float2* __restrict__ tw) {
  constexpr int NSEQ = 2048 / N, LD = N + 1, JC = N / 2 + 1, PITCH = N + 2;
  extern __shared__ float2 smem[];
  float2* buf = smem;
  float2* tmp = smem + NSEQ * LD;
  const int pair0 = blockIdx.x * NSEQ;
  for (int e = threadIdx.x; e < NSEQ * N; e += kThreads) {
    const int s = e / N, x = e - s * N;
    const int pr = pair0 + s;
    float2 v = f2(0.f, 0.f);
    if (pr < npairs) {
      const int img = (2 * pr) / N, row = (2 * pr) % N;
      const float* ra = data + img * img_stride + (size_t)row * PITCH;
      v = f2(ra[x], ra[x + PITCH]);
    }
    buf[s * LD + x] = v;
  }
  __syncthreads();
  float2* res = lds_fft_t<N, NSEQ, LD, -1>(buf, tmp, tw);
  for (int e = threadIdx.x; e < NSEQ * JC; e += kThreads) {
    const int s = e / JC, k = e - s * JC;
    const int pr = pair0 + s;
    if (pr >= npairs) continue;
    const float2 zk = res[s * LD + k];
    const float2 zm = res[s * LD + ((N - k) % N)];
    const float2 xa = f2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
    const float2 xb = f2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
    const int img = (2 * pr) / N, row = (2 * pr) % N;
    float* ra = data + img * img_stride + (size_t)row * PITCH;
    ra[2 * k] = xa.x;
    ra[2 * k + 1] = xa.y;
    ra[PITCH + 2 * k] = xb.x;
    ra[PITCH + 2 * k + 1] = xb.y;
  }
}

template <int N>
__global__ void __launch_bounds__(kThreads) k_rows_c2r_t(float* data, size_t img_stride,
                                                         int npairs, const float2* __restrict__ tw) {
  constexpr int NSEQ = 2048 / N, LD = N + 1, JC = N / 2 + 1, PITCH = N + 2;
  extern __shared__ float2 smem[];
  float2* buf = smem;
  float2* tmp = smem + NSEQ * LD;
  const int pair0 = blockIdx.x * NSEQ;
  for (int e = threadIdx.x; e < NSEQ * JC; e += kThreads) {
    const int s = e / JC, k = e - s * JC;
    const int pr = pair0 + s;
    float2 a = f2(0.f, 0.f), b = f2(0.f, 0.f);
    if (pr < npairs) {
      const int img = (2 * pr) / N, row = (2 * pr) % N;
      const float* ra = data + img * img_stride + (size_t)row * PITCH;
      a = f2(ra[2 * k], ra[2 * k + 1]);
      b = f2(ra[PITCH + 2 * k], ra[PITCH + 2 * k + 1]);
    }
    if (k == 0 || 2 * k == N) { a.y = 0.f; b.y = 0.f; }
    buf[s * LD + k] = f2(a.x - b.y, a.y + b.x);
    if (k > 0 && 2 * k < N) buf[s * LD + (N - k)] = f2(a.x + b.y, -a.y + b.x);
  }
  __syncthreads();
  float2* res = lds_fft_t<N, NSEQ, LD, +1>(buf, tmp, tw);
  for (int e = threadIdx.x; e < NSEQ * PITCH; e += kThreads) {
    const int s = e / PITCH, x = e - s * PITCH;
    const int pr = pair0 + s;
    if (pr >= npairs) continue;
    const int img = (2 * pr) / N, row = (2 * pr) % N;
    float* ra = data + img * img_stride + (size_t)row * PITCH;
    const float2 z = x < N ? res[s * LD + x] : f2(0.f, 0.f);
    ra[x] = z.x;
    ra[PITCH + x] = z.y;
  }
}


template <int N, int SIGN>
__global__ void __launch_bounds__(kThreads) k_cols_t(float* data, size_t img_stride, int nimg,
                                                     const float2* __restrict__ tw) {
  constexpr int JC = N / 2 + 1, LD = N + 1;
  constexpr int CW = cols_cw<N>();
  constexpr int JB = (JC + CW - 1) / CW;
  extern __shared__ float2 smem[];
  float2* buf = smem;
  float2* tmp = smem + CW * LD;
  // column blocks of one image in XCD-contiguous order
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int img = lb / JB;
  if (img >= nimg) return;
  const int x0 = (lb - img * JB) * CW;
  const int ncol = min(CW, JC - x0);
  float2* base = reinterpret_cast<float2*>(data + img * img_stride);
  for (int e = threadIdx.x; e < N * CW; e += kThreads) {
    const int y = e / CW, c = e - y * CW;
    float2 v = f2(0.f, 0.f);
    if (c < ncol) v = base[(size_t)y * JC + x0 + c];
    buf[c * LD + y] = v;
  }
  __syncthreads();
  float2* res = lds_fft_t<N, CW, LD, SIGN>(buf, tmp, tw);
  for (int e = threadIdx.x; e < N * CW; e += kThreads) {
    const int y = e / CW, c = e - y * CW;
    if (c < ncol) base[(size_t)y * JC + x0 + c] = res[c * LD + y];
  }
}

template <int N>
int rows_pass_t(Ctx& c, int nimg, float* data, size_t img_stride, bool fwd, const float2* tw) {
  constexpr int NSEQ = 2048 / N;
  const int npairs = nimg * (N / 2);
  const size_t lds = sizeof(float2) * 2 * NSEQ * (N + 1);
  dim3 grid((npairs + NSEQ - 1) / NSEQ);
  if (fwd)
    hipLaunchKernelGGL(k_rows_r2c_t<N>, grid, dim3(kThreads), lds, c.stream, data, img_stride,
                       npairs, tw);
  else
    hipLaunchKernelGGL(k_rows_c2r_t<N>, grid, dim3(kThreads), lds, c.stream, data, img_stride,
                       npairs, tw);
  FR_HIP(hipGetLastError());
  return 0;
}

template <int N>
int cols_pass_t(Ctx& c, int nimg, float* data, size_t img_stride, int sign, const float2* tw) {
  constexpr int JC = N / 2 + 1;
  constexpr int CW = cols_cw<N>();
  const size_t lds = sizeof(float2) * 2 * CW * (N + 1);
  const long long nblk = (long long)((JC + CW - 1) / CW) * nimg;
  FR_CHECK(nblk <= (1ll << 30), "too many images in one FFT batch (%d)", nimg);
  dim3 grid(round8((int)nblk));
  if (sign < 0)
    hipLaunchKernelGGL((k_cols_t<N, -1>), grid, dim3(kThreads), lds, c.stream, data, img_stride,
                       nimg, tw);
  else
    hipLaunchKernelGGL((k_cols_t<N, +1>), grid, dim3(kThreads), lds, c.stream, data, img_stride,
                       nimg, tw);
  FR_HIP(hipGetLastError());
  return 0;
}

}  // namespace

bool fft_supported(int nsam) {
  Plan p;
  return (nsam % 2 == 0) && make_plan(nsam, &p);
}

static int rows_pass(Ctx& c, int nsam, int nimg, float* data, size_t img_stride, bool fwd) {
  Plan pl;
  FR_CHECK(make_plan(nsam, &pl) && nsam % 2 == 0, "FFT length %d not supported", nsam);
  const float2* tw = c.twiddles(nsam);
  FR_CHECK(tw != nullptr, "twiddle table allocation failed");
  switch (nsam) {
    case 64: return rows_pass_t<64>(c, nimg, data, img_stride, fwd, tw);
    case 128: return rows_pass_t<128>(c, nimg, data, img_stride, fwd, tw);
    case 256: return rows_pass_t<256>(c, nimg, data, img_stride, fwd, tw);
    default: break;
  }
  int npairs = nimg * (nsam / 2);
  int nseq = row_seqs(nsam);
  size_t lds = sizeof(float2) * 2 * nseq * (nsam + 1);
  dim3 grid((npairs + nseq - 1) / nseq);
  if (fwd)
    hipLaunchKernelGGL(k_rows_r2c, grid, dim3(kThreads), lds, c.stream, data, pl, nsam, img_stride,
                       npairs, nseq, tw);
  else
    hipLaunchKernelGGL(k_rows_c2r, grid, dim3(kThreads), lds, c.stream, data, pl, nsam, img_stride,
                       npairs, nseq, tw);
  FR_HIP(hipGetLastError());
  return 0;
}

static int cols_pass(Ctx& c, int nsam, int nimg, float* data, size_t img_stride, int sign) {
  Plan pl;
  FR_CHECK(make_plan(nsam, &pl), "FFT length %d not supported", nsam);
  const float2* tw = c.twiddles(nsam);
  FR_CHECK(tw != nullptr, "twiddle table allocation failed");
  FR_CHECK(nimg <= 65535, "too many images in one FFT batch (%d)", nimg);
  switch (nsam) {
    case 64: return cols_pass_t<64>(c, nimg, data, img_stride, sign, tw);
    case 128: return cols_pass_t<128>(c, nimg, data, img_stride, sign, tw);
    case 256: return cols_pass_t<256>(c, nimg, data, img_stride, sign, tw);
    default: break;
  }
  int jc = nsam / 2 + 1;
  int cw = 2048 / nsam;
  if (cw < 1) cw = 1;
  if (cw > jc) cw = jc;
  size_t lds = sizeof(float2) * 2 * cw * (nsam + 1);
  const long long nblk = (long long)((jc + cw - 1) / cw) * nimg;
  FR_CHECK(nblk <= (1ll << 30), "too many images in one FFT batch (%d)", nimg);
  dim3 grid(round8((int)nblk));
  hipLaunchKernelGGL(k_cols, grid, dim3(kThreads), lds, c.stream, data, pl, img_stride, cw, nimg,
                     sign, tw);
  FR_HIP(hipGetLastError());
  return 0;
}

int fft_rows_2d(Ctx& c, int nsam, int nimg, float* data, size_t img_stride, bool fwd) {
  if (nimg <= 0) return 0;
  return rows_pass(c, nsam, nimg, data, img_stride, fwd);
}

int fft_cols_2d(Ctx& c, int nsam, int nimg, float* data, size_t img_stride, int sign) {
  if (nimg <= 0) return 0;
  return cols_pass(c, nsam, nimg, data, img_stride, sign);
}

int fft_r2c_2d(Ctx& c, int nsam, int nimg, float* data, size_t img_stride) {
  if (nimg <= 0) return 0;
  if (int r = rows_pass(c, nsam, nimg, data, img_stride, true)) return r;
  return cols_pass(c, nsam, nimg, data, img_stride, -1);
}

int fft_c2r_2d(Ctx& c, int nsam, int nimg, float* data, size_t img_stride) {
  if (nimg <= 0) return 0;
  if (int r = cols_pass(c, nsam, nimg, data, img_stride, +1)) return r;
  return rows_pass(c, nsam, nimg, data, img_stride, false);
}

static int z_pass(Ctx& c, int nsam, float* data, int sign) {
  Plan pl;
  FR_CHECK(make_plan(nsam, &pl), "FFT length %d not supported", nsam);
  const float2* tw = c.twiddles(nsam);
  FR_CHECK(tw != nullptr, "twiddle table allocation failed");
  const int jc = nsam / 2 + 1;
  const int ncols = jc * nsam;
  int cw = 2048 / nsam;
  if (cw < 1) cw = 1;
  const size_t lds = sizeof(float2) * 2 * cw * (nsam + 1);
  dim3 grid(round8((ncols + cw - 1) / cw));
  hipLaunchKernelGGL(k_cols_z, grid, dim3(kThreads), lds, c.stream, data, pl, ncols,
                     (size_t)ncols, cw, sign, tw);
  FR_HIP(hipGetLastError());
  return 0;
}

// 3D transforms of a padded volume (NSAM+2 row pitch, NSAM^2 rows): the 2D
// transform of every z-slice, then the z axis
int fft_r2c_3d(Ctx& c, int nsam, float* data) {
  const size_t img = (size_t)nsam * (nsam + 2);
  if (int r = fft_r2c_2d(c, nsam, nsam, data, img)) return r;
  return z_pass(c, nsam, data, -1);
}

int fft_c2r_3d(Ctx& c, int nsam, float* data) {
  const size_t img = (size_t)nsam * (nsam + 2);
  if (int r = z_pass(c, nsam, data, +1)) return r;
  return fft_c2r_2d(c, nsam, nsam, data, img);
}

}  // namespace fr

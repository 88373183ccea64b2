// fr_search.hip -- exhaustive orientation x translation search as an MFMA GEMM.
//
// The reference's PSEARCH/CCP (psearch.f:8-121, ccp.f:8-275) correlates,
// for every view and in-plane angle, the B-factor-weighted particle spectrum
// with the CTF-multiplied projection and reads the cross-correlation map on
// the NSAMR = NSAM/2 grid inside +-6 pixels (ccp.f:229-252).  Here the same
// quantity is written as a dot product over the resolution annulus k:
//
//   cc(p, o, s) = sum_k w_k Re( X_p(k) conj(c_p(k)) e^{i 2pi k.s/NSAMR} conj(T_o(k)) )
//
// with X_p = CBUF of psearch.f:50-59, c_p = CTFR + conj(CTFL), w_k = 2 for
// I > 0 and 1 for I = 0 (the c2r convention of the NSAMR transform), and
// T_o(k) the CTF-free central section at orientation o.  The CTF goes to the
// particle side, so the templates T_o are shared by every particle: rows are
// the 15x15 shift hypotheses s of one particle (generated on the fly in LDS
// from the particle vector and separable phase tables), columns are
// templates, K is the real-interleaved annulus.  The epilogue does what
// ccp.f:225-252 does with the map: argmax over the inner 13x13 (first max
// wins, x outer), the 3x3 box sum around it, and the normalisation by
// sqrt(sum |c T|^2) sqrt(sum |X|^2) / 9.
//
// The element type is bf16 (v_mfma_f32_16x16x32_bf16, FP32 accumulation) or
// fp32 (v_mfma_f32_16x16x4f32, the plain FP32 path).
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include "fr_common.h"
#include "fr_extract.h"
#include "fr_search.h"

namespace fr {

namespace {

constexpr int kThreads = 256;
constexpr int kRows = 256;     // 15x15 = 225 shift rows, padded
constexpr int kCols = 256;     // templates per block
constexpr int kKStep = 32;     // reals per K step (16 complex coefficients)
constexpr int kShift = 15;     // shift grid edge (-7..7 on the NSAMR grid)
constexpr int kInner = 6;      // argmax window (-6..6), ccp.f:229-241
#ifndef FR_SEARCH_BDEPTH
#define FR_SEARCH_BDEPTH 1
#endif
constexpr int kBDepth = FR_SEARCH_BDEPTH;  // template K chunks in flight per thread

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

template <typename T>
struct Elem;
template <>
struct Elem<__bf16> {
  static constexpr int kPad = 40;  // LDS row pitch in elements (80 B: conflict-free b128 reads)
  static __device__ inline __bf16 cvt(float v) { return (__bf16)v; }
  static __device__ inline float f(__bf16 v) { return (float)v; }
};
template <>
struct Elem<float> {
  static constexpr int kPad = 36;
  static __device__ inline float cvt(float v) { return v; }
  static __device__ inline float f(float v) { return v; }
};

}  // namespace

// ---------------------------------------------------------------- bank
// T_o(k) for every orientation o and annulus point k (AINTERPO3DS of the
// plane DM_o*(I,J)), stored real-interleaved with row pitch kpad.
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_bank_build(int nsam, int norient, int kc, int kpad, const float2* __restrict__ a3df,
                 const float* __restrict__ dm, const int2* __restrict__ kidx, T* bank) {
  const size_t gid = blockIdx.x * (size_t)kThreads + threadIdx.x;
  const int half = kpad / 2;
  if (gid >= (size_t)norient * half) return;
  const int o = (int)(gid / half), k = (int)(gid - (size_t)o * half);
  T* row = bank + (size_t)o * kpad;
  if (k >= kc) {
    row[2 * k] = Elem<T>::cvt(0.f);
    row[2 * k + 1] = Elem<T>::cvt(0.f);
    return;
  }
  int2 ij = kidx[k];
  cf v = ainterpo_dev(nsam, a3df, dm + 9 * (size_t)o, ij.x, ij.y);
  row[2 * k] = Elem<T>::cvt(v.re);
  row[2 * k + 1] = Elem<T>::cvt(v.im);
}

// ------------------------------------------------------- particle features
// psearch.f:50-59 (CBUF), the CTF of ccp.f:125-129 folded onto the particle,
// the c2r weights, |c|^2 for the normalisation and SUM2 of ccp.f:212.
__global__ void __launch_bounds__(kThreads)
    k_search_feats(int nsam, int kc, int kpad, int maxr22, float rbfact,
                   const float2* __restrict__ outc, const float2* __restrict__ ctff,
                   const int2* __restrict__ kidx, float2* ap, float* c2, float* sum2) {
  const int jc = nsam / 2 + 1, p = blockIdx.x, half = kpad / 2;
  const size_t ic = (size_t)nsam * jc;
  const float2* sp = outc + (size_t)p * ic;
  const float2* cp = ctff + (size_t)p * 2 * ic;
  // ap and c2 null: SUM2 only (the separable search forms its own terms)
  for (int k = threadIdx.x; ap && k < half; k += blockDim.x) {
    float2 a = make_float2(0.f, 0.f);
    float cc = 0.f;
    if (k < kc) {
      int2 ij = kidx[k];
      size_t id = (size_t)ij.x + (size_t)jc * wrapn(ij.y, nsam);
      float w = expf(-rbfact * (float)(ij.x * ij.x + ij.y * ij.y));
      float2 s = sp[id];
      float2 cr = cp[id], cl = cp[id + ic];
      float cre = cr.x + cl.x, cim = cr.y - cl.y;  // c = CTFR + conj(CTFL)
      float wk = ij.x > 0 ? 2.0f : 1.0f;
      float xr = s.x * w, xi = s.y * w;
      // w_k * X * conj(c)
      a = make_float2(wk * (xr * cre + xi * cim), wk * (xi * cre - xr * cim));
      cc = cre * cre + cim * cim;
    }
    ap[(size_t)p * half + k] = a;
    c2[(size_t)p * half + k] = cc;
  }
  // SUM2: |CBUF|^2 over 1 <= r^2 < MAXR2^2 on the half-plane (ccp.f:194-216)
  __shared__ float red[kThreads / 64];
  float acc = 0.f;
  const int r = (int)sqrtf((float)maxr22) + 1;
  for (int e = threadIdx.x; e < (r + 1) * (2 * r + 1); e += blockDim.x) {
    int i = e / (2 * r + 1), j = e % (2 * r + 1) - r;
    int it = i * i + j * j;
    if (it < 1 || it >= maxr22) continue;
    size_t id = (size_t)i + (size_t)jc * wrapn(j, nsam);
    float w = expf(-rbfact * (float)it);
    float2 s = sp[id];
    acc += (s.x * w) * (s.x * w) + (s.y * w) * (s.y * w);
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) sum2[p] = red[0] + red[1] + red[2] + red[3];
}

// ---------------------------------------------------------------- norms
// sum_k |c_p(k)|^2 |T_o(k)|^2 for every (particle, orientation): the
// normalisation SUM of ccp.f:211-214 with the CTF folded on the particle.
// Tiles of 64 particles x 128 templates; K staged through LDS.
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_search_norm(int norient, int kpad, int np, const T* __restrict__ bank,
                  const float* __restrict__ c2, float* nrm) {
  constexpr int TP = 64, TO = 128, KS = 32;  // particles, templates, complex per step
  __shared__ float cs[TP][KS + 1];
  __shared__ float ts[TO][KS + 1];
  const int half = kpad / 2;
  const int o0 = blockIdx.x * TO, p0 = blockIdx.y * TP;
  const int tid = threadIdx.x;
  // thread -> (4 particles) x (8 templates)
  const int tp = (tid & 15) * 4, to = (tid >> 4) * 8;
  float acc[4][8] = {};
  for (int k0 = 0; k0 < half; k0 += KS) {
    for (int e = tid; e < TP * KS; e += kThreads) {
      int pp = e / KS, kk = e % KS;
      int p = p0 + pp, k = k0 + kk;
      cs[pp][kk] = (p < np && k < half) ? c2[(size_t)p * half + k] : 0.f;
    }
    for (int e = tid; e < TO * KS; e += kThreads) {
      int oo = e / KS, kk = e % KS;
      int o = o0 + oo, k = k0 + kk;
      float v = 0.f;
      if (o < norient && k < half) {
        float re = Elem<T>::f(bank[(size_t)o * kpad + 2 * k]);
        float im = Elem<T>::f(bank[(size_t)o * kpad + 2 * k + 1]);
        v = re * re + im * im;
      }
      ts[oo][kk] = v;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < KS; kk++) {
      float a[4], b[8];
#pragma unroll
      for (int i = 0; i < 4; i++) a[i] = cs[tp + i][kk];
#pragma unroll
      for (int j = 0; j < 8; j++) b[j] = ts[to + j][kk];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 8; j++) acc[i][j] += a[i] * b[j];
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 8; j++) {
      int p = p0 + tp + i, o = o0 + to + j;
      if (p < np && o < norient) nrm[(size_t)p * norient + o] = acc[i][j];
    }
}

// ---------------------------------------------------------------- GEMM
constexpr int kGemmThreads = 512;  // 8 waves: 4 row groups x 2 column groups

template <typename T>
__device__ inline void mma_step(const T* As, const T* Bs, floatx4 (&acc)[4][8], int wr, int wc,
                                int lane);

template <>
__device__ inline void mma_step<__bf16>(const __bf16* As, const __bf16* Bs,
                                        floatx4 (&acc)[4][8], int wr, int wc, int lane) {
  constexpr int P = Elem<__bf16>::kPad;
  const int lr = lane & 15, kq = (lane >> 4) * 8;
  bf16x8 a[4], b[8];
#pragma unroll
  for (int rt = 0; rt < 4; rt++)
    a[rt] = *reinterpret_cast<const bf16x8*>(As + (wr * 64 + rt * 16 + lr) * P + kq);
#pragma unroll
  for (int ct = 0; ct < 8; ct++)
    b[ct] = *reinterpret_cast<const bf16x8*>(Bs + (wc * 128 + ct * 16 + lr) * P + kq);
#pragma unroll
  for (int rt = 0; rt < 4; rt++)
#pragma unroll
    for (int ct = 0; ct < 8; ct++)
      acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[rt], b[ct], acc[rt][ct], 0, 0, 0);
}

template <>
__device__ inline void mma_step<float>(const float* As, const float* Bs, floatx4 (&acc)[4][8],
                                       int wr, int wc, int lane) {
  constexpr int P = Elem<float>::kPad;
  const int lr = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int kk = 0; kk < kKStep; kk += 4) {
    float a[4], b[8];
#pragma unroll
    for (int rt = 0; rt < 4; rt++) a[rt] = As[(wr * 64 + rt * 16 + lr) * P + kk + kq];
#pragma unroll
    for (int ct = 0; ct < 8; ct++) b[ct] = Bs[(wc * 128 + ct * 16 + lr) * P + kk + kq];
#pragma unroll
    for (int rt = 0; rt < 4; rt++)
#pragma unroll
      for (int ct = 0; ct < 8; ct++)
        acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[rt], b[ct], acc[rt][ct], 0, 0, 0);
  }
}

// LDS layout of one GEMM tile: two operand buffers (A rows x P, B cols x P)
// and three table buffers, each holding the chunk's 16x15 X'(k) e^{i I sx}
// and 16x15 e^{i J sy} phase entries and its 16 weights |c_p(k)|^2.  The
// epilogue map [kRows][65] floats aliases the operand buffers; the column
// norms' partial sums alias the table buffers.
template <typename T>
struct Lds {
  static constexpr size_t a_bytes = (size_t)kRows * Elem<T>::kPad * sizeof(T);
  static constexpr size_t b_bytes = (size_t)kCols * Elem<T>::kPad * sizeof(T);
  static constexpr size_t op_bytes = a_bytes + b_bytes;  // one buffer
  static constexpr int tab_entries = 2 * 16 * kShift + 16;
  static constexpr size_t tab_one = tab_entries * sizeof(float2);
  static constexpr size_t tab_off = 2 * op_bytes;
  static constexpr size_t ep_bytes = (size_t)kRows * 65 * sizeof(float);
  static constexpr size_t end_tab = tab_off + 3 * tab_one;
  static constexpr size_t total = end_tab > ep_bytes ? end_tab : ep_bytes;
  static_assert(ep_bytes <= tab_off, "epilogue map overlaps the tables");
  static_assert(3 * tab_one >= kGemmThreads * sizeof(float), "norm partials exceed the tables");
};

template <typename T>
constexpr size_t lds_total() {
  return Lds<T>::total;
}

// B chunk transfer: kCols templates x 32 reals, in 16-byte pieces.
template <typename T>
struct BPieces {
  static constexpr int per_row = 32 * sizeof(T) / 16;      // 4 (bf16) or 8 (fp32)
  static constexpr int total = kCols * per_row;             // 1024 or 2048
  static constexpr int per_thread = total / kGemmThreads;   // 2 or 4
  static constexpr int elems = 16 / sizeof(T);              // elements per piece
};

// One tile: the 225 shift rows of particle p against columns o0..o0+255 of
// its template list (list entry cand_p[o], or o itself when cand_p is null;
// entries outside 0..nbank-1 are empty columns).  The normalisation
// sum_k |c_p(k)|^2 |T_o(k)|^2 (SUM of ccp.f:211-214 with the CTF folded on
// the particle) comes from nrm_p[o], or, with kFused, is accumulated
// alongside the MFMA work from the template chunks already in LDS: that
// pays where the template rows are gathered (candidate lists) and the K
// loop waits on memory, not where it streams the bank (the exhaustive
// search, whose norms k_search_norm computes).  Outputs are indexed by list
// position.
template <typename T, bool kFused, bool kLoop>
__device__ __forceinline__ void gemm_tile(unsigned char* smem_raw, int o0, int ncol, int nbank,
                                          int kpad, const T* __restrict__ bank,
                                          const int* __restrict__ cand_p,
                                          const float2* __restrict__ apv,
                                          const float* __restrict__ c2p,
                                          const float* __restrict__ nrm_p, float s2,
                                          const float2* __restrict__ shx_tab,
                                          const float2* __restrict__ shy_tab, float* score_p,
                                          int* brow_p) {
  constexpr int P = Elem<T>::kPad;
  using L = Lds<T>;
  using BP = BPieces<T>;
  auto As = [&](int buf) { return reinterpret_cast<T*>(smem_raw + buf * L::op_bytes); };
  auto Bs = [&](int buf) { return reinterpret_cast<T*>(smem_raw + buf * L::op_bytes + L::a_bytes); };
  auto tabx = [&](int buf) {
    return reinterpret_cast<float2*>(smem_raw + L::tab_off + buf * L::tab_one);
  };
  auto taby = [&](int buf) { return tabx(buf) + 16 * kShift; };

  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wr = wave & 3, wc = wave >> 2;
  const int half = kpad / 2;
  const int nchunk = half / 16;

  floatx4 acc[4][8];
#pragma unroll
  for (int rt = 0; rt < 4; rt++)
#pragma unroll
    for (int ct = 0; ct < 8; ct++) acc[rt][ct] = floatx4{0.f, 0.f, 0.f, 0.f};
  float nacc = 0.f;  // norm partial: column tid/2, complex 8*(tid%2)..+7 of each chunk

  // this thread's B pieces: piece bq of templates bc0 + u * bstep, for the
  // kBDepth chunks in flight (chunk c in register slot c % kBDepth)
  const int bq = tid % BP::per_row;
  const int bc0 = tid / BP::per_row;
  constexpr int bstep = kGemmThreads / BP::per_row;
  uint4 breg[kBDepth][BP::per_thread];
  float2 treg;

  int orow[BP::per_thread];
#pragma unroll
  for (int u = 0; u < BP::per_thread; u++) {
    int col = o0 + bc0 + u * bstep;
    int o = col < ncol ? (cand_p ? cand_p[col] : col) : -1;
    orow[u] = (o >= 0 && o < nbank) ? o : -1;
  }
  auto load_b = [&](uint4 (&r)[BP::per_thread], int chunk) {
#pragma unroll
    for (int u = 0; u < BP::per_thread; u++) {
      if (orow[u] >= 0)
        r[u] = *reinterpret_cast<const uint4*>(bank + (size_t)orow[u] * kpad + 32 * chunk +
                                                bq * BP::elems);
      else
        r[u] = make_uint4(0u, 0u, 0u, 0u);
    }
  };
  auto store_b = [&](int buf, const uint4 (&r)[BP::per_thread]) {
    T* dst = Bs(buf);
#pragma unroll
    for (int u = 0; u < BP::per_thread; u++)
      *reinterpret_cast<uint4*>(dst + (bc0 + u * bstep) * P + bq * BP::elems) = r[u];
  };
  // one table entry per thread: 240 X'(k) e^{i I sx}, 240 e^{i J sy}, 16 |c(k)|^2
  auto load_tab = [&](int chunk) {
    const int k0 = 16 * chunk;
    float2 v = make_float2(0.f, 0.f);
    if (tid < 16 * kShift) {
      int kk = tid / kShift, sx = tid % kShift;
      float2 a = apv[k0 + kk];
      float2 ex = shx_tab[(size_t)(k0 + kk) * kShift + sx];
      v = make_float2(a.x * ex.x - a.y * ex.y, a.x * ex.y + a.y * ex.x);
    } else if (tid < 32 * kShift) {
      int e2 = tid - 16 * kShift;
      v = shy_tab[(size_t)(k0 + e2 / kShift) * kShift + e2 % kShift];
    } else if (kFused && tid < L::tab_entries) {
      v = make_float2(c2p[k0 + tid - 32 * kShift], 0.f);
    }
    treg = v;
  };
  auto store_tab = [&](int tbuf) {
    if (tid < (kFused ? L::tab_entries : 32 * kShift)) tabx(tbuf)[tid] = treg;
  };
  // A chunk: row (sx, sy) = X'(k) e^{i k.s}; two threads per row, 8 complex each
  const int row = tid >> 1, hk = (tid & 1) * 8;
  const int rsx = row / kShift, rsy = row % kShift;
  auto gen_a = [&](int buf, int tbuf) {
    T* dst = As(buf) + row * P + 2 * hk;
    const float2* tx = tabx(tbuf);
    const float2* ty = taby(tbuf);
    constexpr int E = 16 / sizeof(T);  // elements per 16-byte store
#pragma unroll
    for (int g = 0; g < 16 / E; g++) {
      T v[E];
#pragma unroll
      for (int q = 0; q < E / 2; q++) {
        int kk = hk + g * (E / 2) + q;
        float re = 0.f, im = 0.f;
        if (row < kShift * kShift) {
          float2 a = tx[kk * kShift + rsx], e = ty[kk * kShift + rsy];
          re = a.x * e.x - a.y * e.y;
          im = a.x * e.y + a.y * e.x;
        }
        v[2 * q] = Elem<T>::cvt(re);
        v[2 * q + 1] = Elem<T>::cvt(im);
      }
      *reinterpret_cast<uint4*>(dst + g * E) = *reinterpret_cast<const uint4*>(v);
    }
  };
  // norm: |T|^2 of this thread's 8 complex of the chunk in buf, weighted by |c|^2
  auto norm_step = [&](int buf, int tbuf) {
    constexpr int E = 16 / sizeof(T);
    const T* src = Bs(buf) + row * P + 2 * hk;
    const float2* w = tabx(tbuf) + 32 * kShift + hk;
    T v[16];
#pragma unroll
    for (int g = 0; g < 16 / E; g++)
      *reinterpret_cast<uint4*>(v + g * E) = *reinterpret_cast<const uint4*>(src + g * E);
#pragma unroll
    for (int q = 0; q < 8; q++) {
      float re = Elem<T>::f(v[2 * q]), im = Elem<T>::f(v[2 * q + 1]);
      nacc += w[q].x * (re * re + im * im);
    }
  };

  // prologue: tables for chunks 0 and 1, template loads for chunks
  // 0..kBDepth-1 in flight, operands for chunk 0 in LDS
  load_tab(0);
  store_tab(0);
  if (nchunk > 1) {
    load_tab(1);
    store_tab(1);
  }
#pragma unroll
  for (int d = 0; d < kBDepth; d++)
    if (d < nchunk) load_b(breg[d], d);
  store_b(0, breg[0]);
  if (kBDepth < nchunk) load_b(breg[0], kBDepth);
  __syncthreads();
  gen_a(0, 0);
  __syncthreads();

  // chunk kc's templates were loaded kBDepth chunks earlier; the loop is
  // unrolled by kBDepth so that every register slot index is static
  for (int kc0 = 0; kc0 < nchunk; kc0 += kBDepth) {
#pragma unroll
    for (int d = 0; d < kBDepth; d++) {
      const int kc = kc0 + d;
      if (kc >= nchunk) break;
      const int cur = kc & 1, nxt = cur ^ 1;
      const bool more = kc + 1 < nchunk;
      if (kc + 2 < nchunk) load_tab(kc + 2);
      mma_step<T>(As(cur), Bs(cur), acc, wr, wc, lane);
      if (kFused) norm_step(cur, kc % 3);
      if (more) {
        gen_a(nxt, (kc + 1) % 3);
        store_b(nxt, breg[(d + 1) % kBDepth]);
        if (kc + 1 + kBDepth < nchunk) load_b(breg[(d + 1) % kBDepth], kc + 1 + kBDepth);
      }
      if (kc + 2 < nchunk) store_tab((kc + 2) % 3);
      __syncthreads();
    }
  }

  // ------------------------------------------------------------ epilogue
  float* ep = reinterpret_cast<float*>(smem_raw);  // [kRows][65] floats, reuses A/B space
  float* nl = reinterpret_cast<float*>(smem_raw + L::tab_off);
  if (kFused) nl[tid] = nacc;
#pragma unroll
  for (int q4 = 0; q4 < 4; q4++) {  // 64-column quarters of the 256 templates
    if (q4) __syncthreads();
    if (wc == (q4 >> 1)) {
#pragma unroll
      for (int rt = 0; rt < 4; rt++)
#pragma unroll
        for (int ct = 0; ct < 4; ct++)
#pragma unroll
          for (int i = 0; i < 4; i++) {
            int r = wr * 64 + rt * 16 + (lane >> 4) * 4 + i;
            int c = ct * 16 + (lane & 15);
            ep[r * 65 + c] = acc[rt][ct + 4 * (q4 & 1)][i];
          }
    }
    __syncthreads();
    if (tid < 64) {
      const int c = tid, cc = q4 * 64 + c, o = o0 + cc;
      // ccp.f:229-241: argmax over IX (outer) and IY (inner) in -6..6, first max wins
      float best = -3.0e38f;
      int bx = 0, by = 0;
      for (int ix = -kInner; ix <= kInner; ix++)
        for (int iy = -kInner; iy <= kInner; iy++) {
          float v = ep[((ix + 7) * kShift + (iy + 7)) * 65 + c];
          if (v > best) { best = v; bx = ix; by = iy; }
        }
      // ccp.f:242-252: 3x3 box sum
      float box = 0.f;
      for (int dx = -1; dx <= 1; dx++)
        for (int dy = -1; dy <= 1; dy++) box += ep[((bx + dx + 7) * kShift + (by + dy + 7)) * 65 + c];
      if (o < ncol) {
        float n = kFused ? nl[2 * cc] + nl[2 * cc + 1] : nrm_p[o];
        float v = (n > 0.f && s2 > 0.f) ? box / sqrtf(n) / sqrtf(s2) / 9.0f : -3.0e38f;
        score_p[o] = v;
        brow_p[o] = (bx + 7) * kShift + (by + 7);
      }
    }
  }
  if (kLoop) __syncthreads();  // LDS is reused by the next tile
}

// Uniform lists: nrow columns per particle (the whole bank when cand is
// null, else the particle's row of cand); norms from nrm (p x nrow) unless
// kFused.  XCD-aware mapping: blocks are
// dealt round-robin to the 8 XCDs; each XCD walks its own template tiles with
// the particles innermost, so a template tile stays resident in that XCD's L2
// while every particle uses it.
template <typename T, bool kFused>
__global__ void __launch_bounds__(kGemmThreads)
    k_search_gemm(int nrow, int nbank, int kpad, int np, int ntile, const T* __restrict__ bank,
                  const int* __restrict__ cand, const float2* __restrict__ ap,
                  const float* __restrict__ c2, const float* __restrict__ nrm,
                  const float* __restrict__ sum2,
                  const float2* __restrict__ shx_tab, const float2* __restrict__ shy_tab,
                  float* score, int* best_row) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const int b = blockIdx.x;
  const int xcd = b & 7, j = b >> 3;
  const int tile = (j / np) * 8 + xcd;
  const int p = j % np;
  if (tile >= ntile) return;
  const int half = kpad / 2;
  const size_t base = (size_t)p * nrow;
  gemm_tile<T, kFused, false>(smem_raw, tile * kCols, nrow, nbank, kpad, bank,
                       cand ? cand + base : nullptr, ap + (size_t)p * half, c2 + (size_t)p * half,
                       kFused ? nullptr : nrm + base, sum2[p], shx_tab, shy_tab, score + base,
                       best_row + base);
}

// Ragged lists: particle p's list is cand[loff[p] .. loff[p+1]) and covers
// tiles toff[p] .. toff[p+1]) of the launch; list positions at or beyond
// `cap` (the capacity of cand, score and best_row) are neither read nor
// written.  Persistent: each workgroup walks tiles blockIdx.x,
// blockIdx.x + gridDim.x, ... up to toff[np].
template <typename T>
__global__ void __launch_bounds__(kGemmThreads)
    k_search_gemm_ragged(int np, int nbank, int kpad, const T* __restrict__ bank,
                         const int* __restrict__ loff, const int* __restrict__ toff,
                         const int* __restrict__ cand, int cap, const float2* __restrict__ ap,
                         const float* __restrict__ c2, const float* __restrict__ sum2,
                         const float2* __restrict__ shx_tab, const float2* __restrict__ shy_tab,
                         float* score, int* best_row) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  const int total = toff[np];
  const int half = kpad / 2;
  for (int t = blockIdx.x; t < total; t += gridDim.x) {
    // the particle whose tile range holds t: the largest p with toff[p] <= t
    int lo = 0, hi = np - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (toff[mid] <= t) lo = mid;
      else hi = mid - 1;
    }
    const int p = lo;
    const int base = loff[p];
    const int end = loff[p + 1] < cap ? loff[p + 1] : cap;
    if (base < 0 || end <= base) continue;  // uniform across the workgroup
    gemm_tile<T, true, true>(smem_raw, (t - toff[p]) * kCols, end - base, nbank, kpad, bank,
                       cand + base, ap + (size_t)p * half, c2 + (size_t)p * half, nullptr,
                       sum2[p], shx_tab, shy_tab, score + base, best_row + base);
  }
}

// ---------------------------------------------------------------- top-K
// CCP returns one (psi, shift) per view: the first maximum over its in-plane
// angles in (IPSI outer, IQUAD inner) order (ccp.f:144-264); PSEARCH then
// keeps the IPMAX best views by strict '>' in grid order (psearch.f:89-102),
// so ties keep the earlier view.  A particle's list comes in groups of
// `group` consecutive entries per view.  One workgroup per particle: each
// thread keeps a sorted top-K of its views, then a bitonic merge.  Uniform
// lists: n entries per particle, group `group`; ragged lists (loff non-null):
// particle p's entries loff[p]..min(loff[p+1], cap)) in groups of grp[p].  Output
// indices are list positions, mapped to bank orientations through cand when
// it is non-null; an orientation outside 0..nbank-1 is reported as -1 (an
// empty column, as its -3e38 score says).
__global__ void __launch_bounds__(kThreads)
    k_topk(int nbank, int n, int group, const int* __restrict__ loff, const int* __restrict__ grp,
           const int* __restrict__ cand, int cap, int topk, const float* __restrict__ score,
           const int* __restrict__ best_row, int* top_o, int* top_row, float* top_s) {
  constexpr int kK = 32;
  __shared__ float sv[kThreads * kK];
  __shared__ int si[kThreads * kK];
  const int p = blockIdx.x, tid = threadIdx.x;
  const size_t base = loff ? (size_t)loff[p] : (size_t)p * n;
  int len = n;
  if (loff) {
    const int end = loff[p + 1] < cap ? loff[p + 1] : cap;
    len = (loff[p] >= 0 && end > loff[p]) ? end - loff[p] : 0;
  }
  const int gsz = grp ? grp[p] : group;
  const float* sc = score + base;
  const int ngroup = gsz > 0 ? len / gsz : 0;
  // few groups (candidate lists, a bank of up to 2048 views): sort the group
  // maxima themselves, padded to a power of two
  if (ngroup <= 2048) {
    int np2 = 1;
    while (np2 < ngroup || np2 < topk) np2 <<= 1;
    for (int g = tid; g < np2; g += kThreads) {
      float x = -3.4e38f;
      int xo = 0x7fffffff;
      if (g < ngroup) {
        xo = g * gsz;
        for (int u = 0; u < gsz; u++) {
          float y = sc[g * gsz + u];
          if (y > x) { x = y; xo = g * gsz + u; }
        }
      }
      sv[g] = x;
      si[g] = xo;
    }
    __syncthreads();
    for (int size = 2; size <= np2; size <<= 1)
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        for (int e = tid; e < np2 / 2; e += kThreads) {
          int lo = 2 * e - (e & (stride - 1));
          int hi = lo + stride;
          bool desc = ((lo & size) == 0);
          float a = sv[lo], bb = sv[hi];
          int ia = si[lo], ib = si[hi];
          bool a_first = (a > bb) || (a == bb && ia < ib);
          if (a_first != desc) {
            sv[lo] = bb; sv[hi] = a;
            si[lo] = ib; si[hi] = ia;
          }
        }
        __syncthreads();
      }
  } else {
  float v[kK];
  int ix[kK];
#pragma unroll
  for (int q = 0; q < kK; q++) { v[q] = -3.4e38f; ix[q] = 0x7fffffff; }
  for (int g = tid; g < ngroup; g += kThreads) {
    float x = -3.4e38f;
    int xo = g * gsz;
    for (int u = 0; u < gsz; u++) {
      float y = sc[g * gsz + u];
      if (y > x) { x = y; xo = g * gsz + u; }
    }
    if (!(x > v[kK - 1])) continue;
    int q = kK - 1;
    while (q > 0 && x > v[q - 1]) { v[q] = v[q - 1]; ix[q] = ix[q - 1]; q--; }
    v[q] = x;
    ix[q] = xo;
  }
#pragma unroll
  for (int q = 0; q < kK; q++) { sv[tid * kK + q] = v[q]; si[tid * kK + q] = ix[q]; }
  __syncthreads();
  // bitonic sort of kThreads*kK entries, descending score, ascending index on ties
  const int ns = kThreads * kK;
  for (int size = 2; size <= ns; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int e = tid; e < ns / 2; e += kThreads) {
        int lo = 2 * e - (e & (stride - 1));
        int hi = lo + stride;
        bool desc = ((lo & size) == 0);
        float a = sv[lo], bb = sv[hi];
        int ia = si[lo], ib = si[hi];
        bool a_first = (a > bb) || (a == bb && ia < ib);
        if (a_first != desc) {
          sv[lo] = bb; sv[hi] = a;
          si[lo] = ib; si[hi] = ia;
        }
      }
      __syncthreads();
    }
  }
  // PSEARCH starts every TESTPAR(6,K) at 0.0 and admits a view only when
  // CC > TESTPAR(6,K) (psearch.f:42-45,89-90): a view scoring <= 0 never
  // enters the list.  Such slots, like empty columns, are reported as
  // orientation -1, row -1 and TESTPAR's start value 0.0.
  for (int q = tid; q < topk; q += kThreads) {
    int t = si[q];
    bool ok = t >= 0 && t < len && sv[q] > 0.0f;
    int o = ok ? (cand ? cand[base + t] : t) : -1;
    if (o >= nbank) o = -1;  // a list entry outside the bank is an empty column
    top_o[(size_t)p * topk + q] = o;
    top_row[(size_t)p * topk + q] = o >= 0 ? best_row[base + t] : -1;
    top_s[(size_t)p * topk + q] = o >= 0 ? sv[q] : 0.0f;
  }
}

// ------------------------------------------------------------- launchers
size_t search_gemm_lds(bool bf16) { return bf16 ? lds_total<__bf16>() : lds_total<float>(); }

int launch_bank_build(Ctx& c, int nsam, int norient, int kc, int kpad, const float* a3df,
                      const float* dm, const int* kidx, void* bank, bool bf16) {
  size_t tot = (size_t)norient * (kpad / 2);
  dim3 grid((unsigned)((tot + kThreads - 1) / kThreads));
  if (bf16)
    hipLaunchKernelGGL(k_bank_build<__bf16>, grid, dim3(kThreads), 0, c.stream, nsam, norient, kc,
                       kpad, reinterpret_cast<const float2*>(a3df), dm,
                       reinterpret_cast<const int2*>(kidx), reinterpret_cast<__bf16*>(bank));
  else
    hipLaunchKernelGGL(k_bank_build<float>, grid, dim3(kThreads), 0, c.stream, nsam, norient, kc,
                       kpad, reinterpret_cast<const float2*>(a3df), dm,
                       reinterpret_cast<const int2*>(kidx), reinterpret_cast<float*>(bank));
  FR_HIP(hipGetLastError());
  return 0;
}

int launch_search_feats(Ctx& c, int nsam, int np, int kc, int kpad, int maxr22, float rbfact,
                        const float* outc, const float* ctff, const int* kidx, float* ap,
                        float* c2, float* sum2) {
  hipLaunchKernelGGL(k_search_feats, dim3(np), dim3(kThreads), 0, c.stream, nsam, kc, kpad, maxr22,
                     rbfact, reinterpret_cast<const float2*>(outc),
                     reinterpret_cast<const float2*>(ctff), reinterpret_cast<const int2*>(kidx),
                     reinterpret_cast<float2*>(ap), c2, sum2);
  FR_HIP(hipGetLastError());
  return 0;
}

namespace {
template <typename T>
int set_gemm_lds(const void* fn, size_t lds) {
  FR_CHECK(lds <= 160 * 1024, "LDS budget exceeded (%zu)", lds);
  FR_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  return 0;
}
}  // namespace

int launch_search_norm(Ctx& c, int norient, int kpad, int np, const void* bank, bool bf16,
                       const float* c2, float* nrm) {
  dim3 grid((norient + 127) / 128, (np + 63) / 64);
  if (bf16)
    hipLaunchKernelGGL(k_search_norm<__bf16>, grid, dim3(kThreads), 0, c.stream, norient, kpad, np,
                       reinterpret_cast<const __bf16*>(bank), c2, nrm);
  else
    hipLaunchKernelGGL(k_search_norm<float>, grid, dim3(kThreads), 0, c.stream, norient, kpad, np,
                       reinterpret_cast<const float*>(bank), c2, nrm);
  FR_HIP(hipGetLastError());
  return 0;
}

namespace {
template <typename T, bool kFused>
int launch_gemm_t(Ctx& c, size_t nblk, size_t lds, int nrow, int nbank, int kpad, int np,
                  int ntile, const void* bank, const int* cand, const float2* ap, const float* c2,
                  const float* nrm, const float* sum2, const float2* sx, const float2* sy,
                  float* score, int* best_row) {
  const void* fn = (const void*)k_search_gemm<T, kFused>;
  if (int r = set_gemm_lds<T>(fn, lds)) return r;
  hipLaunchKernelGGL((k_search_gemm<T, kFused>), dim3((unsigned)nblk), dim3(kGemmThreads), lds,
                     c.stream, nrow, nbank, kpad, np, ntile, reinterpret_cast<const T*>(bank), cand,
                     ap, c2, nrm, sum2, sx, sy, score, best_row);
  FR_HIP(hipGetLastError());
  return 0;
}
}  // namespace

int launch_search_gemm(Ctx& c, int nrow, int nbank, int kpad, int np, const void* bank,
                       bool bf16, const int* cand, const float* ap, const float* c2,
                       const float* nrm, const float* sum2, const float* shx_tab,
                       const float* shy_tab, float* score, int* best_row) {
  const int ntile = (nrow + kCols - 1) / kCols;
  const int ntile8 = (ntile + 7) / 8 * 8;
  const size_t nblk = (size_t)ntile8 * np;
  FR_CHECK(nblk < (1u << 31), "search grid too large");
  const size_t lds = search_gemm_lds(bf16);
  const float2* apv = reinterpret_cast<const float2*>(ap);
  const float2* sx = reinterpret_cast<const float2*>(shx_tab);
  const float2* sy = reinterpret_cast<const float2*>(shy_tab);
  if (bf16)
    return nrm ? launch_gemm_t<__bf16, false>(c, nblk, lds, nrow, nbank, kpad, np, ntile, bank,
                                              cand, apv, c2, nrm, sum2, sx, sy, score, best_row)
               : launch_gemm_t<__bf16, true>(c, nblk, lds, nrow, nbank, kpad, np, ntile, bank,
                                             cand, apv, c2, nrm, sum2, sx, sy, score, best_row);
  return nrm ? launch_gemm_t<float, false>(c, nblk, lds, nrow, nbank, kpad, np, ntile, bank, cand,
                                           apv, c2, nrm, sum2, sx, sy, score, best_row)
             : launch_gemm_t<float, true>(c, nblk, lds, nrow, nbank, kpad, np, ntile, bank, cand,
                                          apv, c2, nrm, sum2, sx, sy, score, best_row);
}

int launch_search_gemm_ragged(Ctx& c, int np, int nbank, int kpad, const void* bank, bool bf16,
                              const int* loff, const int* toff, const int* cand, int cap,
                              int max_tiles, const float* ap, const float* c2, const float* sum2,
                              const float* shx_tab, const float* shy_tab, float* score,
                              int* best_row) {
  const size_t lds = search_gemm_lds(bf16);
  const void* fn = bf16 ? (const void*)k_search_gemm_ragged<__bf16>
                        : (const void*)k_search_gemm_ragged<float>;
  if (int r = bf16 ? set_gemm_lds<__bf16>(fn, lds) : set_gemm_lds<float>(fn, lds)) return r;
  // persistent grid: as many workgroups as the device holds at once
  int per_cu = 0;
  FR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kGemmThreads, lds));
  if (per_cu < 1) per_cu = 1;
  if (c.cus <= 0) {
    hipDeviceProp_t prop;
    FR_HIP(hipGetDeviceProperties(&prop, c.device));
    c.cus = prop.multiProcessorCount;
  }
  int grid = c.cus * per_cu;
  if (max_tiles < grid) grid = max_tiles;
  if (grid < 1) return 0;
  const float2* apv = reinterpret_cast<const float2*>(ap);
  const float2* sx = reinterpret_cast<const float2*>(shx_tab);
  const float2* sy = reinterpret_cast<const float2*>(shy_tab);
  if (bf16)
    hipLaunchKernelGGL(k_search_gemm_ragged<__bf16>, dim3(grid), dim3(kGemmThreads), lds,
                       c.stream, np, nbank, kpad, reinterpret_cast<const __bf16*>(bank), loff, toff,
                       cand, cap, apv, c2, sum2, sx, sy, score, best_row);
  else
    hipLaunchKernelGGL(k_search_gemm_ragged<float>, dim3(grid), dim3(kGemmThreads), lds,
                       c.stream, np, nbank, kpad, reinterpret_cast<const float*>(bank), loff, toff,
                       cand, cap, apv, c2, sum2, sx, sy, score, best_row);
  FR_HIP(hipGetLastError());
  return 0;
}

int launch_topk(Ctx& c, int nbank, int np, int n, int group, const int* loff, const int* grp,
                const int* cand, int cap, int topk, const float* score, const int* best_row,
                int* top_o, int* top_row, float* top_s) {
  FR_CHECK(topk >= 1 && topk <= 32, "topk must be in 1..32");
  if (!loff)
    FR_CHECK(group >= 1 && n % group == 0, "list length %d is not a multiple of the per-view "
             "group %d", n, group);
  hipLaunchKernelGGL(k_topk, dim3(np), dim3(kThreads), 0, c.stream, nbank, n, group, loff, grp,
                     cand, cap, topk, score, best_row, top_o, top_row, top_s);
  FR_HIP(hipGetLastError());
  return 0;
}

}  // namespace fr

// fr_search_sep.hip -- the BF16 orientation x translation search with the
// translations applied in two separable matrix-core stages.
//
// The score of ccp.f:194-252 for particle p, template o and shift
// s = (sx, sy) on the NSAMR grid is, with the CTF and the c2r weights folded
// onto the particle (X'_p, as in fr_search.hip):
//
//   cc(p, o, s) = Re sum_{I,J} Y(I, J) e^{i I phi(sx)} e^{i J phi(sy)},
//   Y(I, J)     = X'_p(I, J) conj(T_o(I, J)),  phi(s) = 2 pi (s - 7) / NSAMR.
//
// The phase factor is separable, so the 15 x 15 map comes from
//   stage 1:  Z(I, sy)  = sum_J Y(I, J) e^{i J phi(sy)}       (32 x 64 x 15 complex)
//   stage 2:  cc(sx, sy) = Re sum_I Z(I, sy) e^{i I phi(sx)}  (15 x 32 x 15 complex)
// which is 18 v_mfma_f32_16x16x32_bf16 per (particle, template) instead of
// the 225-row GEMM's 104: the elementwise product Y is formed on the vector
// ALU (v_dot2_f32_bf16), stage 1 consumes it as the A operand, and stage 2
// takes stage 1's accumulators as its B operand lane for lane, with no data
// exchange between lanes.  The epilogue (argmax over the inner 13 x 13,
// first maximum with x outer, 3 x 3 box sum, normalisation) is ccp.f's, as
// in the GEMM search.
//
// Layout ("separable" bank): the annulus I in 0..31, J in -32..31 on a grid;
// grid point (I, J) of fragment q (row block I / 16, K step (J + 32) / 16)
// sits in lane (J + 32) % 16 / 4 * 16 + I % 16, slot (J + 32) % 4.  A
// template is 8 fragments x 64 lanes x 4 complex bf16 = 8 KiB, points
// outside the annulus are zero.  The particle terms X'_p (as bf16 pairs
// (re, im) and (im, -re)) and |c_p|^2 (FP32) use the same order.
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include "fr_common.h"
#include "fr_extract.h"
#include "fr_search.h"

namespace fr {

namespace {

constexpr int kSepThreads = 256;  // 4 waves; a wave takes one template at a time
constexpr int kSepTile = 256;     // list entries per work item
constexpr int kSepFbTile = 64;    // templates per work item of the fallback form (few
                                  // particles: short items keep the GPU full)
constexpr int kSepElems = 2048;   // grid points per template (8 x 64 x 4)

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

// grid point of element e = (q * 64 + lane) * 4 + t of the separable layout
__device__ inline void sep_point(int e, int& I, int& J) {
  const int q = e >> 8, lane = (e >> 2) & 63, t = e & 3;
  I = (q >> 2) * 16 + (lane & 15);
  J = (q & 3) * 16 + 4 * (lane >> 4) + t - 32;
}

// the search annulus of search_annulus() (fr_api.cpp): I >= 0,
// lo2 <= I^2 + J^2 < hi2, |J| != JC - 1
__device__ inline bool sep_in(int I, int J, int lo2, int hi2, int jc) {
  const int it = I * I + J * J;
  return it >= lo2 && it < hi2 && abs(J) != jc - 1 && I <= jc - 2;
}

__device__ inline unsigned pack_bf16(float lo, float hi) {
  bf16x2 v = {(__bf16)lo, (__bf16)hi};
  return *reinterpret_cast<unsigned*>(&v);
}

__device__ inline bf16x2 as_bf2(unsigned u) { return *reinterpret_cast<bf16x2*>(&u); }

// Wave-wide maximum and sum through DPP row operations (fixed order), the
// result read from lane 63: a uniform value.
template <int kCtrl, int kRowMask>
__device__ inline float dpp_f(float old, float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), kCtrl,
                                                    kRowMask, 0xF, false));
}

__device__ inline float wave_max(float v) {
  const float lo = -3.4e38f;
  v = fmaxf(v, dpp_f<0xB1, 0xF>(lo, v));   // quad_perm [1,0,3,2]
  v = fmaxf(v, dpp_f<0x4E, 0xF>(lo, v));   // quad_perm [2,3,0,1]
  v = fmaxf(v, dpp_f<0x141, 0xF>(lo, v));  // row_half_mirror
  v = fmaxf(v, dpp_f<0x140, 0xF>(lo, v));  // row_mirror
  v = fmaxf(v, dpp_f<0x142, 0xA>(lo, v));  // row_bcast15
  v = fmaxf(v, dpp_f<0x143, 0xC>(lo, v));  // row_bcast31
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

__device__ inline float wave_sum(float v) {
  v += dpp_f<0xB1, 0xF>(0.f, v);
  v += dpp_f<0x4E, 0xF>(0.f, v);
  v += dpp_f<0x141, 0xF>(0.f, v);
  v += dpp_f<0x140, 0xF>(0.f, v);
  v += dpp_f<0x142, 0xA>(0.f, v);
  v += dpp_f<0x143, 0xC>(0.f, v);
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// cos and sin of 2 pi (a mod n) / n: the argument reduced exactly first
__device__ inline void phase(int a, int n, float& c, float& s) {
  int r = a % n;
  if (r < 0) r += n;
  sincosf(6.283185307179586f * (float)r / (float)n, &s, &c);
}

// ------------------------------------------------------------------ bank
__global__ void __launch_bounds__(256)
    k_bank_build_sep(int nsam, int norient, int lo2, int hi2, const float2* __restrict__ a3df,
                     const float* __restrict__ dm, unsigned* bank) {
  const size_t gid = blockIdx.x * (size_t)256 + threadIdx.x;
  if (gid >= (size_t)norient * kSepElems) return;
  const int o = (int)(gid / kSepElems), e = (int)(gid % kSepElems);
  int I, J;
  sep_point(e, I, J);
  float re = 0.f, im = 0.f;
  if (sep_in(I, J, lo2, hi2, nsam / 2 + 1)) {
    cf v = ainterpo_dev(nsam, a3df, dm + 9 * (size_t)o, I, J);
    re = v.re;
    im = v.im;
  }
  bank[gid] = pack_bf16(re, im);
}

// ------------------------------------------------------- particle terms
// X'_p = w_k X conj(c) and |c|^2 at every grid point, the arithmetic of
// k_search_feats (psearch.f:50-59 and the CTF of ccp.f:125-129).
__global__ void __launch_bounds__(256)
    k_sep_feats(int nsam, int lo2, int hi2, float rbfact, const float2* __restrict__ outc,
                const float2* __restrict__ ctff, unsigned* xa, unsigned* xb, float* c2s) {
  const int jc = nsam / 2 + 1, p = blockIdx.x;
  const size_t ic = (size_t)nsam * jc;
  const float2* sp = outc + (size_t)p * ic;
  const float2* cp = ctff + (size_t)p * 2 * ic;
  for (int e = threadIdx.x; e < kSepElems; e += 256) {
    int I, J;
    sep_point(e, I, J);
    float ar = 0.f, ai = 0.f, cc = 0.f;
    if (sep_in(I, J, lo2, hi2, jc)) {
      size_t id = (size_t)I + (size_t)jc * wrapn(J, nsam);
      float w = expf(-rbfact * (float)(I * I + J * J));
      float2 s = sp[id];
      float2 cr = cp[id], cl = cp[id + ic];
      float cre = cr.x + cl.x, cim = cr.y - cl.y;
      float wk = I > 0 ? 2.0f : 1.0f;
      float xr = s.x * w, xi = s.y * w;
      ar = wk * (xr * cre + xi * cim);
      ai = wk * (xi * cre - xr * cim);
      cc = cre * cre + cim * cim;
    }
    const size_t o = (size_t)p * kSepElems + e;
    xa[o] = pack_bf16(ar, ai);
    xb[o] = pack_bf16(ai, -ar);
    c2s[o] = cc;
  }
}

// ------------------------------------------------------------------ search
// One work item, seen from one wave: particle p against positions lo .. hi-1
// of its list (entries base .. base + len).  Uniform lists: the workgroup's
// four waves take four consecutive particles against the same 256 list
// positions (each template is fetched once into the CU's cache and read by
// all four); ragged lists: the four waves split one particle's 256-entry
// tile.
struct SepItem {
  int p, base, lo, hi;
  bool listed;  // the entries are list positions (cand), else bank rows in order
};

// kSepMixed: the ragged tiles of the listed particles, then the fallback
// items, in one persistent launch
enum { kSepUniform = 0, kSepRagged = 1, kSepFallback = 2, kSepMixed = 3 };

template <int kMode>
__device__ inline bool sep_item(int idx, int wave, int np, int nrow, int ntile,
                                const int* __restrict__ loff, const int* __restrict__ toff,
                                int cap, const int* __restrict__ fb,
                                const int* __restrict__ ord, SepItem& it) {
  if (kMode == kSepFallback) {
    // the uniform mapping over the fallback particles fb[1 .. fb[0]], whose
    // lists are the bank in order at loff[p]
    const int cnt = fb[0], npq = (cnt + 3) / 4;
    const int xcd = idx & 7, j = idx >> 3;
    const int tile = (j / npq) * 8 + xcd;
    const int q = (j % npq) * 4 + wave;
    if (tile >= ntile || q >= cnt) return false;
    it.p = fb[1 + q];
    it.base = loff[it.p];
    const int end = loff[it.p + 1] < cap ? loff[it.p + 1] : cap;
    it.lo = tile * kSepFbTile;
    it.hi = min(min(nrow, end - it.base), it.lo + kSepFbTile);
    it.listed = false;
    return it.base >= 0 && it.lo < it.hi;
  }
  if (kMode == kSepUniform) {
    // XCD-aware: blocks are dealt round-robin to the 8 XCDs; each XCD walks
    // its template tiles with the particle quads innermost
    const int npq = (np + 3) / 4;
    const int xcd = idx & 7, j = idx >> 3;
    const int tile = (j / npq) * 8 + xcd;
    it.p = (j % npq) * 4 + wave;
    if (tile >= ntile || it.p >= np) return false;
    it.base = it.p * nrow;
    it.lo = tile * kSepTile;
    it.hi = min(nrow, it.lo + kSepTile);
    it.listed = true;
    return true;
  }
  // the slot whose tile range holds idx: the largest q with toff[q] <= idx;
  // its particle is ord[q] (q itself without an order)
  int lo = 0, hi = np - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (toff[mid] <= idx) lo = mid;
    else hi = mid - 1;
  }
  it.p = ord ? ord[lo] : lo;
  it.base = loff[it.p];
  const int end = loff[it.p + 1] < cap ? loff[it.p + 1] : cap;
  it.lo = (idx - toff[lo]) * kSepTile + wave * 64;
  it.hi = min(end - it.base, it.lo + 64);
  it.listed = true;
  return it.base >= 0 && it.lo < it.hi;
}

template <int kMode>
__global__ void __launch_bounds__(kSepThreads, 2)
    k_search_sep(int np, int nbank, int nsamr, int nrow, int ntile, const int* __restrict__ loff,
                 const int* __restrict__ toff, int cap, const int* __restrict__ cand,
                 const uint4* __restrict__ bank, const uint4* __restrict__ xa,
                 const uint4* __restrict__ xb, const float4* __restrict__ c2s,
                 const float* __restrict__ sum2, float* score, int* best_row,
                 const int* __restrict__ fb, const int* __restrict__ ord) {
  constexpr bool kRagged = kMode == kSepRagged;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int g = lane >> 4, n = lane & 15;

  // stage 1 B operand: rows (J, re|im) of this lane's K slots, column sy = n:
  // Z_re gets (cos, -sin), Z_im (sin, cos) of J phi(sy)
  bf16x8 ey[4][2];
#pragma unroll
  for (int ks = 0; ks < 4; ks++)
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int J = ks * 16 + 4 * g + (j >> 1) - 32;
      float c = 0.f, s = 0.f;
      if (n < 15) phase(J * (n - 7), nsamr, c, s);
      ey[ks][0][j] = (__bf16)((j & 1) ? -s : c);
      ey[ks][1][j] = (__bf16)((j & 1) ? c : s);
    }
  // stage 2 A operand: row sx = n, K slots (I, re|im): (cos, -sin) of I phi(sx)
  bf16x8 ex[2];
#pragma unroll
  for (int s2i = 0; s2i < 2; s2i++)
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int I = s2i * 16 + 4 * g + (j & 3);
      float c = 0.f, s = 0.f;
      if (n < 15) phase(I * (n - 7), nsamr, c, s);
      ex[s2i][j] = (__bf16)((j >> 2) ? -s : c);
    }

  uint4 pa[8], pb[8];
  float4 pc[8];
  int pcur = -1;
  float s2 = 0.f;

  // an entry outside the bank reads row 0 and is scored as an empty column
  auto load_t = [&](uint4 (&tv)[8], int o) {
    const uint4* src = bank + (size_t)(o >= 0 ? o : 0) * 512 + lane;
#pragma unroll
    for (int q = 0; q < 8; q++) tv[q] = src[q * 64];
  };

  // one (particle, template) pair: the 15 x 15 map, then ccp.f's epilogue
  auto pair = [&](const uint4 (&tv)[8], float& sc, int& row) {
    floatx4 acc1[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++) acc1[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
    float nacc = 0.f;
    // fragments in (K step, row block) order: four independent accumulator
    // chains between dependent matrix-core operations
#pragma unroll
    for (int qq = 0; qq < 8; qq++) {
      const int q = (qq & 1) * 4 + (qq >> 1);
      const unsigned tw[4] = {tv[q].x, tv[q].y, tv[q].z, tv[q].w};
      const unsigned aw[4] = {pa[q].x, pa[q].y, pa[q].z, pa[q].w};
      const unsigned bw[4] = {pb[q].x, pb[q].y, pb[q].z, pb[q].w};
      const float cw[4] = {pc[q].x, pc[q].y, pc[q].z, pc[q].w};
      bf16x8 y;
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const bf16x2 T = as_bf2(tw[t]);
        const float yr = __builtin_amdgcn_fdot2_f32_bf16(as_bf2(aw[t]), T, 0.f, false);
        const float yi = __builtin_amdgcn_fdot2_f32_bf16(as_bf2(bw[t]), T, 0.f, false);
        nacc = fmaf(cw[t], __builtin_amdgcn_fdot2_f32_bf16(T, T, 0.f, false), nacc);
        y[2 * t] = (__bf16)yr;
        y[2 * t + 1] = (__bf16)yi;
      }
      acc1[q >> 2][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(y, ey[q & 3][0], acc1[q >> 2][0],
                                                                0, 0, 0);
      acc1[q >> 2][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(y, ey[q & 3][1], acc1[q >> 2][1],
                                                                0, 0, 0);
    }
    nacc = wave_sum(nacc);
    // stage 2: this lane's accumulators (rows I = 16 s + 4 g + i, column sy)
    // are its B operand for K step s (two chains, added)
    floatx4 acc2h[2];
#pragma unroll
    for (int s = 0; s < 2; s++) {
      bf16x8 z;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        z[i] = (__bf16)acc1[s][0][i];
        z[4 + i] = (__bf16)acc1[s][1][i];
      }
      acc2h[s] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ex[s], z, floatx4{0.f, 0.f, 0.f, 0.f},
                                                         0, 0, 0);
    }
    floatx4 acc2;
#pragma unroll
    for (int i = 0; i < 4; i++) acc2[i] = acc2h[0][i] + acc2h[1][i];
    // lane (g, n) holds map[sx = 4 g + i][sy = n]; ccp.f:229-241: argmax over
    // the inner 13 x 13, x outer, first maximum: the maximum value across the
    // wave, then the first (sx, sy) holding it from the lanes' ballots
    bool ok[4];
    float m = -3.4e38f;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int sx = 4 * g + i;
      ok[i] = sx >= 1 && sx <= 13 && n >= 1 && n <= 13;
      if (ok[i]) m = fmaxf(m, acc2[i]);
    }
    m = wave_max(m);
    unsigned long long hit[4];
#pragma unroll
    for (int i = 0; i < 4; i++) hit[i] = __ballot(ok[i] && acc2[i] == m);
    int key = 7 * 16 + 7;  // no ordered value (NaN map): the centre
#pragma unroll
    for (int sx = 13; sx >= 1; sx--) {
      const unsigned bits = (unsigned)(hit[sx & 3] >> (16 * (sx >> 2))) & 0xffffu;
      if (bits) key = sx * 16 + __builtin_ctz(bits);
    }
    const int bx = key >> 4, by = key & 15;
    // ccp.f:242-252: 3 x 3 box sum around the maximum
    float box = 0.f;
#pragma unroll
    for (int dx = -1; dx <= 1; dx++) {
      const int sxn = bx + dx, ii = sxn & 3;
      const float v = ii == 0 ? acc2[0] : ii == 1 ? acc2[1] : ii == 2 ? acc2[2] : acc2[3];
#pragma unroll
      for (int dy = -1; dy <= 1; dy++) box += __shfl(v, (sxn >> 2) * 16 + by + dy, 64);
    }
    sc = (nacc > 0.f && s2 > 0.f) ? box / sqrtf(nacc) / sqrtf(s2) / 9.0f : -3.0e38f;
    row = bx * 15 + by;
  };

  // persistent forms loop over their items; the uniform form is one item
  // per workgroup
  const int nfb = kMode == kSepFallback || kMode == kSepMixed
                      ? (ntile + 7) / 8 * 8 * ((fb[0] + 3) / 4) : 0;
  const int nlist = kRagged || kMode == kSepMixed ? toff[np] : 0;
  const int nitem = kMode == kSepUniform ? (int)gridDim.x : nlist + nfb;
  const int stride = kMode == kSepUniform ? nitem : (int)gridDim.x;
  for (int idx = blockIdx.x; idx < nitem; idx += stride) {
    SepItem it;
    bool ok;
    if (kMode == kSepMixed) {
      ok = idx < nlist
               ? sep_item<kSepRagged>(idx, wave, np, nrow, ntile, loff, toff, cap, fb, ord, it)
               : sep_item<kSepFallback>(idx - nlist, wave, np, nrow, ntile, loff, toff, cap, fb,
                                        ord, it);
    } else {
      ok = sep_item<kMode>(idx, wave, np, nrow, ntile, loff, toff, cap, fb, ord, it);
    }
    if (!ok) continue;
    if (it.p != pcur) {
      const size_t b = (size_t)it.p * 512 + lane;
#pragma unroll
      for (int q = 0; q < 8; q++) {
        pa[q] = xa[b + q * 64];
        pb[q] = xb[b + q * 64];
        pc[q] = c2s[b + q * 64];
      }
      s2 = sum2[it.p];
      pcur = it.p;
    }
    const int* cl = cand && it.listed ? cand + it.base : nullptr;
    // the score of an empty column (zero map, no norm): the GEMM form's
    auto put = [&](int pos, int o, float sc, int row) {
      if (lane == 0) {
        score[it.base + pos] = o >= 0 ? sc : -3.0e38f;
        best_row[it.base + pos] = o >= 0 ? row : 16;
      }
    };
    for (int c0 = it.lo; c0 < it.hi; c0 += 64) {
      const int c1 = min(it.hi, c0 + 64);
      // the chunk's bank rows, one per lane, read ahead
      int mine = -1;
      if (c0 + lane < c1) {
        const int o = cl ? cl[c0 + lane] : c0 + lane;
        mine = (o >= 0 && o < nbank) ? o : -1;
      }
      auto bank_row = [&](int pos) { return __builtin_amdgcn_readlane(mine, pos - c0); };
      uint4 t0[8], t1[8];
      int o0 = bank_row(c0), o1 = -1;
      load_t(t0, o0);
      // two positions per trip, the next template's loads in flight meanwhile
      for (int pos = c0; pos < c1; pos += 2) {
        float sc;
        int row;
        const bool two = pos + 1 < c1;
        if (two) {
          o1 = bank_row(pos + 1);
          load_t(t1, o1);
        }
        pair(t0, sc, row);
        put(pos, o0, sc, row);
        if (!two) break;
        if (pos + 2 < c1) {
          o0 = bank_row(pos + 2);
          load_t(t0, o0);
        }
        pair(t1, sc, row);
        put(pos + 1, o1, sc, row);
      }
    }
  }
}

}  // namespace

// ------------------------------------------------------------- launchers
int launch_bank_build_sep(Ctx& c, int nsam, int norient, int lo2, int hi2, const float* a3df,
                          const float* dm, void* bank) {
  const size_t tot = (size_t)norient * kSepElems;
  if (tot == 0) return 0;
  hipLaunchKernelGGL(k_bank_build_sep, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                     c.stream, nsam, norient, lo2, hi2, reinterpret_cast<const float2*>(a3df), dm,
                     reinterpret_cast<unsigned*>(bank));
  FR_HIP(hipGetLastError());
  return 0;
}

int launch_sep_feats(Ctx& c, int nsam, int np, int lo2, int hi2, float rbfact, const float* outc,
                     const float* ctff, void* xa, void* xb, float* c2s) {
  if (np <= 0) return 0;
  hipLaunchKernelGGL(k_sep_feats, dim3(np), dim3(256), 0, c.stream, nsam, lo2, hi2, rbfact,
                     reinterpret_cast<const float2*>(outc), reinterpret_cast<const float2*>(ctff),
                     reinterpret_cast<unsigned*>(xa), reinterpret_cast<unsigned*>(xb), c2s);
  FR_HIP(hipGetLastError());
  return 0;
}

int launch_search_sep(Ctx& c, int np, int nbank, int nsamr, int nrow, const int* loff,
                      const int* toff, int cap, int max_tiles, const int* cand, const void* bank,
                      const void* xa, const void* xb, const float* c2s, const float* sum2,
                      float* score, int* best_row, const int* fb, const int* ord) {
  const uint4* bk = reinterpret_cast<const uint4*>(bank);
  const uint4* a = reinterpret_cast<const uint4*>(xa);
  const uint4* b = reinterpret_cast<const uint4*>(xb);
  const float4* cc = reinterpret_cast<const float4*>(c2s);
  if (!loff) {
    const int ntile = (nrow + kSepTile - 1) / kSepTile;
    const size_t nblk = (size_t)((ntile + 7) / 8 * 8) * ((np + 3) / 4);
    FR_CHECK(nblk < (1u << 31), "search grid too large");
    if (nblk == 0) return 0;
    hipLaunchKernelGGL(k_search_sep<kSepUniform>, dim3((unsigned)nblk), dim3(kSepThreads), 0,
                       c.stream, np, nbank, nsamr, nrow, ntile, nullptr, nullptr, 0, cand, bk, a,
                       b, cc, sum2, score, best_row, nullptr, nullptr);
    FR_HIP(hipGetLastError());
    return 0;
  }
  // persistent grid over the ragged tiles
  int per_cu = 0;
  FR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, (const void*)k_search_sep<kSepRagged>, kSepThreads, 0));
  if (per_cu < 1) per_cu = 1;
  if (c.cus <= 0) {
    hipDeviceProp_t prop;
    FR_HIP(hipGetDeviceProperties(&prop, c.device));
    c.cus = prop.multiProcessorCount;
  }
  const int full = c.cus * per_cu;
  if (fb) {
    // the listed particles' tiles and the fallback particles' items (four
    // particles against shared 64-template tiles, the uniform mapping) in one
    // persistent launch, so that neither pass leaves a tail of idle CUs
    const int ntile = (nbank + kSepFbTile - 1) / kSepFbTile;
    hipLaunchKernelGGL(k_search_sep<kSepMixed>, dim3(full), dim3(kSepThreads), 0, c.stream, np,
                       nbank, nsamr, nbank, ntile, loff, toff, cap, cand, bk, a, b, cc, sum2,
                       score, best_row, fb, ord);
    FR_HIP(hipGetLastError());
    return 0;
  }
  int grid = full;
  if (max_tiles < grid) grid = max_tiles;
  if (grid >= 1) {
    hipLaunchKernelGGL(k_search_sep<kSepRagged>, dim3(grid), dim3(kSepThreads), 0, c.stream, np,
                       nbank, nsamr, 0, 0, loff, toff, cap, cand, bk, a, b, cc, sum2, score,
                       best_row, nullptr, ord);
    FR_HIP(hipGetLastError());
  }
  return 0;
}

}  // namespace fr

// fr_local.hip -- the local orientation and shift search of a mode-1
// refinement: for every particle, the grid of (2N+1)^3 Euler angles
// PHI0 + a*DANG, THETA0 + b*DANG, PSI0 + c*DANG (a, b, c = -N..N) around its
// current orientation, each scored by CC3M (cc3m.f:8-169, the score PREFINE
// maximises, prefine.f:51-67) at every shift of a (2M+1)^2 window around its
// current shift, and the best (orientation, shift) kept.  PREFINE then starts
// from that point (fr_refine_batch).
//
//   * The grid angles are summed on the host in FP32 and their COS/SIN taken
//     there with the C library, as for every other rotation matrix on the
//     path; k_local_dm forms DM from them with the products of cc3m.f:48-65 in
//     the same order without contraction, so each DM is the one the host's
//     rotation gives for the same angles, bit for bit.
//   * The scores come from the fused CC3M kernel with a shift window: one
//     extraction, masking and transform pair per orientation, the shell sums
//     at every shift from one set of cross-spectrum products.
//   * k_local_best takes each particle's first maximum over its orientations
//     in grid order (PHI outer, then THETA, then PSI), strict '>' as PREFINE's
//     and PSEARCH's comparisons.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/frealign_hip.h"
#include "fr_common.h"
#include "fr_ctx.h"
#include "fr_host.h"
#include "fr_kernels.h"

using namespace fr;

#define CTX_CHECK(c) FR_CHECK((c) != nullptr, "null fr_ctx")

namespace {

// Items i = q * norient + o of particles p0 + q: o = (a * n + b) * n + c.
// trig: per particle, [angle 0..2][offset 0..n-1] (COS, SIN).  Writes DM,
// the particle index, the window centre and the magnification per item.
__global__ void k_local_dm(int nitem, int n, int p0, const float* __restrict__ trig,
                           const float* __restrict__ amag, const float* __restrict__ centre,
                           const float* __restrict__ sm, float* __restrict__ dm,
                           int* __restrict__ pidx, float* __restrict__ shift,
                           float* __restrict__ amag_it) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nitem) return;
  const int norient = n * n * n;
  const int q = i / norient, o = i % norient;
  const int a = o / (n * n), b = (o / n) % n, c = o % n;
  const float* t = trig + (size_t)(p0 + q) * 6 * n;
  const float cphi = t[2 * a], sphi = t[2 * a + 1];
  const float cthe = t[2 * (n + b)], sthe = t[2 * (n + b) + 1];
  const float cpsi = t[2 * (2 * n + c)], spsi = t[2 * (2 * n + c) + 1];
  float d[9];
  d[0] = cphi * cthe * cpsi - sphi * spsi;
  d[1] = sphi * cthe * cpsi + cphi * spsi;
  d[2] = -sthe * cpsi;
  d[3] = -cphi * cthe * spsi - sphi * cpsi;
  d[4] = -sphi * cthe * spsi + cphi * cpsi;
  d[5] = sthe * spsi;
  d[6] = sthe * cphi;
  d[7] = sthe * sphi;
  d[8] = cthe;
  const float am = amag[p0 + q];
  if (am != 0.0f) {
    const float ab = fabsf(am);
    for (int k = 0; k < 9; k++) d[k] = d[k] / ab;
  }
  float* out = dm + 9 * (size_t)i;
  if (sm) {  // matmul_t.f:17-24: DM <- DM^T-style product with SM
    for (int r = 0; r < 3; r++)
      for (int s = 0; s < 3; s++) {
        float v = d[3 * r] * sm[3 * s];
        v = v + d[1 + 3 * r] * sm[1 + 3 * s];
        v = v + d[2 + 3 * r] * sm[2 + 3 * s];
        out[s + 3 * r] = v;
      }
  } else {
    for (int k = 0; k < 9; k++) out[k] = d[k];
  }
  pidx[i] = p0 + q;
  shift[2 * (size_t)i] = centre[2 * (size_t)(p0 + q)];
  shift[2 * (size_t)i + 1] = centre[2 * (size_t)(p0 + q) + 1];
  amag_it[i] = fabsf(am);
}

// one wave per particle: the first maximum of its norient scores; writes
// (orientation, shift index, score) to out[3 q ..]
__global__ void __launch_bounds__(64)
    k_local_best(int norient, const float* __restrict__ score, const int* __restrict__ code,
                 int* __restrict__ out) {
  const int q = blockIdx.x, lane = threadIdx.x;
  const float* s = score + (size_t)q * norient;
  float bv = -3.0e38f;
  int bo = 0x7fffffff;
  for (int o = lane; o < norient; o += 64) {
    const float v = s[o];
    if (v > bv) {
      bv = v;
      bo = o;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const float ov = __shfl_xor(bv, off, 64);
    const int oo = __shfl_xor(bo, off, 64);
    if (ov > bv || (ov == bv && oo < bo)) {
      bv = ov;
      bo = oo;
    }
  }
  if (lane == 0) {
    const int o = bo < norient ? bo : 0;
    out[3 * q] = o;
    out[3 * q + 1] = code[(size_t)q * norient + o];
    out[3 * q + 2] = __float_as_int(s[o]);
  }
}

constexpr int kItemBudget = 1 << 16;  // items per launch

}  // namespace

extern "C" {

int fr_local_search_batch(fr_ctx* c, int nsam, int np, const float* d_a3df, const float* d_spec,
                          const float* d_ctff, const float* amag, float* par, int nang,
                          float dang, int nsh, float dsh, int maxr1, int maxr2, float rbfact,
                          const float* d_fsct, const float* ring, const float* sm,
                          float* d_scores, int* d_codes) {
  CTX_CHECK(c);
  FR_CHECK(d_a3df && d_spec && d_ctff && amag && par && d_fsct && ring, "null argument");
  FR_CHECK(np >= 0, "bad particle count");
  FR_CHECK(nang >= 0 && nang <= 12, "angular half-width %d: 0 .. 12 steps", nang);
  FR_CHECK(nsh >= 0 && 2 * nsh + 1 <= kCc3mWinMax, "shift half-width %d: 0 .. %d steps", nsh,
           (kCc3mWinMax - 1) / 2);
  FR_CHECK(maxr1 >= 0 && maxr2 > maxr1, "bad resolution band %d..%d", maxr1, maxr2);
  FR_CHECK(cc3m_fused_ok(nsam, maxr2),
           "local search: NSAM must be a multiple of 32 in 64..256 and MAXR2 <= 32");
  if (np == 0) return 0;
  const int n = 2 * nang + 1, norient = n * n * n;
  // host: grid angles in FP32 and their COS/SIN (the host rotation's functions)
  std::vector<float> trig((size_t)np * 6 * n), centre(2 * (size_t)np);
  for (int p = 0; p < np; p++) {
    const float* x = par + 6 * (size_t)p;
    for (int ang = 0; ang < 3; ang++)
      for (int k = 0; k < n; k++) {
        const float v = x[ang] + (float)(k - nang) * dang;
        trig[(size_t)p * 6 * n + 2 * (ang * n + k)] = cosf(v);
        trig[(size_t)p * 6 * n + 2 * (ang * n + k) + 1] = sinf(v);
      }
    centre[2 * p] = x[3];
    centre[2 * p + 1] = x[4];
  }
  std::vector<float> am(amag, amag + np);
  const int chunk_p = std::max(1, std::min(np, kItemBudget / norient));
  const size_t chunk_items = (size_t)chunk_p * norient;
  int err = 0;
  float* d_trig = (float*)c->slot(6, sizeof(float) * (trig.size() + centre.size() + np + 9), &err);
  if (!d_trig) return err;
  float* d_centre = d_trig + trig.size();
  float* d_am = d_centre + centre.size();
  float* d_sm = d_am + np;
  FR_HIP(hipMemcpyAsync(d_trig, trig.data(), sizeof(float) * trig.size(), hipMemcpyHostToDevice,
                        c->stream));
  FR_HIP(hipMemcpyAsync(d_centre, centre.data(), sizeof(float) * centre.size(),
                        hipMemcpyHostToDevice, c->stream));
  FR_HIP(hipMemcpyAsync(d_am, am.data(), sizeof(float) * np, hipMemcpyHostToDevice, c->stream));
  if (sm)
    FR_HIP(hipMemcpyAsync(d_sm, sm, sizeof(float) * 9, hipMemcpyHostToDevice, c->stream));
  // per-chunk item arrays: DM, particle, window centre, magnification, score, index
  const size_t per_item = 9 * sizeof(float) + sizeof(int) + 2 * sizeof(float) + sizeof(float) +
                          sizeof(float) + sizeof(int);
  char* base = (char*)c->slot(7, per_item * chunk_items, &err);
  if (!base) return err;
  float* d_dm = (float*)base;
  int* d_pi = (int*)(d_dm + 9 * chunk_items);
  float* d_sh = (float*)(d_pi + chunk_items);
  float* d_ai = d_sh + 2 * chunk_items;
  float* d_sc = d_ai + chunk_items;
  int* d_cd = (int*)(d_sc + chunk_items);
  int* d_best = (int*)c->slot(8, 3 * sizeof(int) * (size_t)np, &err);
  if (!d_best) return err;
  Cc3mArgs a{nsam, 0, maxr1, maxr2, rbfact, ring[0], ring[1], ring[2], ring[3]};
  Cc3mTables tabs;
  if (int r = c->cc3m_tables(a, &tabs)) return r;
  const Cc3mWindow win{2 * nsh + 1, 2 * nsh + 1, dsh, dsh};
  const fr_ctx::Tune tn = c->tune_for(nsam);
  std::vector<int> best(3 * (size_t)np);
  for (int p0 = 0; p0 < np; p0 += chunk_p) {
    const int pc = std::min(chunk_p, np - p0);
    const int ni = pc * norient;
    hipLaunchKernelGGL(k_local_dm, dim3((ni + 255) / 256), dim3(256), 0, c->stream, ni, n, p0,
                       d_trig, d_am, d_centre, sm ? d_sm : nullptr, d_dm, d_pi, d_sh, d_ai);
    FR_HIP(hipGetLastError());
    a.n = ni;
    float* sc = d_scores ? d_scores + (size_t)p0 * norient : d_sc;
    int* cd = d_codes ? d_codes + (size_t)p0 * norient : d_cd;
    int pslot = c->prof_begin(kProfCc3m);
    if (int r = launch_cc3m_fused(*c, a, d_a3df, d_spec, d_ctff, d_pi, d_dm, d_sh, d_ai, d_fsct,
                                  tabs, sc, tn.cc3m_blocks, &win, cd))
      return r;
    c->prof_end(pslot);
    hipLaunchKernelGGL(k_local_best, dim3(pc), dim3(64), 0, c->stream, norient, sc, cd,
                       d_best + 3 * (size_t)p0);
    FR_HIP(hipGetLastError());
  }
  FR_HIP(hipMemcpyAsync(best.data(), d_best, sizeof(int) * best.size(), hipMemcpyDeviceToHost,
                        c->stream));
  FR_HIP(hipStreamSynchronize(c->stream));
  const int ns = 2 * nsh + 1;
  for (int p = 0; p < np; p++) {
    float* x = par + 6 * (size_t)p;
    const int o = best[3 * p], wc = best[3 * p + 1];
    float bs;
    memcpy(&bs, &best[3 * p + 2], sizeof bs);
    const int ia = o / (n * n), ib = (o / n) % n, ic = o % n;
    const int ix = wc / ns, iy = wc % ns;
    const float phi = x[0] + (float)(ia - nang) * dang;
    const float the = x[1] + (float)(ib - nang) * dang;
    const float psi = x[2] + (float)(ic - nang) * dang;
    const float shx = x[3] + (float)(ix - nsh) * dsh;
    const float shy = x[4] + (float)(iy - nsh) * dsh;
    x[0] = phi;
    x[1] = the;
    x[2] = psi;
    x[3] = shx;
    x[4] = shy;
    x[5] = bs;
  }
  return 0;
}

}  // extern "C"

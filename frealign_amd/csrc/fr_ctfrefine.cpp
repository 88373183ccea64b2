// fr_ctfrefine.cpp -- the defocus and magnification refinements of a mode-1/3
// run, batched: CTFREFINE (ctfrefine.f:8-100) and MAGREFINE (magrefine.f:8-73),
// each a VA04A (va04.f) over CALCFX's CTF and magnification sections
// (calcfx.f:63-140).
//
// A group is what one reference call refines: the particles of one film
// (consecutive in the batch), or one particle when defocus is refined per
// particle (FPART, CI in ctfrefine.f).  Every group's minimiser runs as the
// resumable fr::Powell state machine; each round gathers the active groups'
// particles, re-applies the CTF at the parameters each group's minimiser
// asks for (CTFAPPLY on the device, from the particles' stored spectra) and
// scores every particle with CC3 at its fixed orientation in one launch:
//
//   defocus (IVFLAG 1):  F = sum_K [-CC3(K; DF1+x1, DF2+x2, AST+x3)
//                                   - FD(SIG2, x1, x2, DFSTD)] / n
//                        (fd.f; x2 = x1 and x3 = 0 unless FASTIG)
//   magnification (2):   F = sum_K -CC3(K; AMAG = x) / n
//
// with CTFAPPLY's ICTF = 0 and CC3 at zero shift (the particle shift is in
// CTFAPPLY's phase ramp), MAXR = INT(NSAM RMAX |AMAG|) per particle, as
// calcfx.f computes them.  The sums run on the host in the reference's FP32
// order over the device scores.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <vector>

#include "../../include/frealign_hip.h"
#include "fr_common.h"
#include "fr_ctx.h"
#include "fr_host.h"
#include "fr_kernels.h"
#include "fr_powell.h"

using namespace fr;

#define CTX_CHECK(c) FR_CHECK((c) != nullptr, "null fr_ctx")

namespace {

#pragma clang fp contract(off)

// fd.f: SIG2 * (-D1**2/2/DFSTD**2 - D2**2/2/DFSTD**2)
float fd(float sig2, float d1, float d2, float dfstd) {
  return sig2 * (-(d1 * d1) / 2.0f / (dfstd * dfstd) - (d2 * d2) / 2.0f / (dfstd * dfstd));
}

struct Job {
  int mode;            // 1: defocus, 2: magnification
  int nsam, np, ng;
  const float* d_spec;
  const float* ppar;   // host np x 8
  const float* euler;  // host np x 3
  const int* gstart;   // host ng + 1
  int fastig;
  float dfstd;
  const float* sig2;   // host ng (defocus)
  const float* d_a3df;
  const float* d_pssnr;
  PrepGlobals g;
  float rmax1, rmax2, rbfact;
  const float* ring;
};

int run(fr_ctx* c, const Job& j, float* out, int* nevals) {
  const int nsam = j.nsam, jc = nsam / 2 + 1;
  const size_t ic = (size_t)nsam * jc;
  const int npar = j.mode == 1 ? (j.fastig ? 3 : 1) : 1;
  // ctfrefine.f:32-33 / magrefine.f:34-35
  const float e_def[3] = {1.0f, 1.0f, 1.0f}, e_mag[1] = {0.01f};
  std::vector<Powell> pw(j.ng);
  for (int g = 0; g < j.ng; g++) {
    float x0[3] = {0.0f, 0.0f, 0.0f};
    if (j.mode == 2) x0[0] = j.ppar[8 * (size_t)j.gstart[g] + 5];  // AMAGP(IFIRSTF)
    pw[g].start(x0, j.mode == 1 ? e_def : e_mag, npar, j.mode == 1 ? 1000.0f : 100.0f, 1, 50);
  }
  std::vector<int> active(j.ng);
  for (int g = 0; g < j.ng; g++) active[g] = g;
  std::vector<float> pp, dm, am, sc;
  std::vector<int> item_p, groupof;
  std::vector<float> fsct_one((size_t)jc + 1, 1.0f);
  int err = 0;
  float* d_fs = (float*)c->slot(16, sizeof(float) * fsct_one.size(), &err);
  if (!d_fs) return err;
  FR_HIP(hipMemcpy(d_fs, fsct_one.data(), sizeof(float) * fsct_one.size(),
                   hipMemcpyHostToDevice));
  std::vector<int> pending;
  while (true) {
    pending.clear();
    for (int g : active)
      if (pw[g].run()) pending.push_back(g);
    if (pending.empty()) break;
    // the particles of the groups that want an evaluation, in group order
    item_p.clear();
    groupof.clear();
    for (int g : pending)
      for (int p = j.gstart[g]; p < j.gstart[g + 1]; p++) {
        item_p.push_back(p);
        groupof.push_back(g);
      }
    const int m = (int)item_p.size();
    pp.assign(8 * (size_t)m, 0.0f);
    dm.assign(9 * (size_t)m, 0.0f);
    am.assign(m, 0.0f);
    std::vector<int> mr1(m), mr2(m);
    for (int q = 0; q < m; q++) {
      const int p = item_p[q], g = groupof[q];
      const float* src = j.ppar + 8 * (size_t)p;
      const float* x = pw[g].x();
      float* d = &pp[8 * (size_t)q];
      memcpy(d, src, 8 * sizeof(float));
      float amag = src[5];
      if (j.mode == 1) {  // calcfx.f:75-80
        const float x1 = x[0], x2 = npar == 1 ? x[0] : x[1], x3 = npar == 1 ? 0.0f : x[2];
        d[2] = src[2] + x1;
        d[3] = src[3] + x2;
        d[4] = src[4] + x3;
      } else {  // calcfx.f:118-124
        amag = x[0];
        d[5] = amag;
      }
      am[q] = fabsf(amag);
      rotation(j.euler[3 * (size_t)p], j.euler[3 * (size_t)p + 1], j.euler[3 * (size_t)p + 2],
               amag, nullptr, &dm[9 * (size_t)q]);
      mr1[q] = (int)((float)nsam * j.rmax1 * fabsf(amag));
      mr2[q] = (int)((float)nsam * j.rmax2 * fabsf(amag));
    }
    // device work: spectra gathered, CTFAPPLY, then CC3 per (MAXR1, MAXR2)
    float* d_sp = (float*)c->slot(17, sizeof(float) * 2 * ic * (size_t)m, &err);
    if (!d_sp) return err;
    float* d_ctff = (float*)c->slot(18, sizeof(float) * 4 * ic * (size_t)m, &err);
    if (!d_ctff) return err;
    float* d_outc = (float*)c->slot(19, sizeof(float) * 2 * ic * (size_t)m, &err);
    if (!d_outc) return err;
    float* d_datc = (float*)c->slot(20, sizeof(float) * 2 * ic * (size_t)m, &err);
    if (!d_datc) return err;
    float* d_pp = (float*)c->slot(21, sizeof(float) * 8 * (size_t)m, &err);
    if (!d_pp) return err;
    float* d_it = (float*)c->slot(22, sizeof(float) * 13 * (size_t)m, &err);
    if (!d_it) return err;
    int* d_pi = (int*)c->slot(23, sizeof(int) * (size_t)m, &err);
    if (!d_pi) return err;
    for (int q = 0; q < m;) {  // consecutive particles copy as one run
      int r = q + 1;
      while (r < m && item_p[r] == item_p[r - 1] + 1) r++;
      FR_HIP(hipMemcpyAsync(d_sp + 2 * ic * (size_t)q, j.d_spec + 2 * ic * (size_t)item_p[q],
                            sizeof(float) * 2 * ic * (size_t)(r - q), hipMemcpyDeviceToDevice,
                            c->stream));
      q = r;
    }
    FR_HIP(hipMemcpyAsync(d_pp, pp.data(), sizeof(float) * 8 * m, hipMemcpyHostToDevice,
                          c->stream));
    if (int r = launch_ctfapply(*c, nsam, m, d_sp, d_pp, j.d_pssnr, j.g, d_ctff, d_outc, d_datc))
      return r;
    // item arrays: DM (9), shift (2, zero), AMAG (1), score (1)
    std::vector<float> items(13 * (size_t)m, 0.0f);
    std::vector<int> pidx(m);
    std::map<std::pair<int, int>, std::vector<int>> by_band;
    for (int q = 0; q < m; q++) by_band[{mr1[q], mr2[q]}].push_back(q);
    // lay the items out band by band so each launch reads one contiguous range
    std::vector<int> order;
    for (auto& kv : by_band) order.insert(order.end(), kv.second.begin(), kv.second.end());
    for (int k = 0; k < m; k++) {
      const int q = order[k];
      memcpy(&items[9 * (size_t)k], &dm[9 * (size_t)q], 9 * sizeof(float));
      items[9 * (size_t)m + 2 * (size_t)k] = 0.0f;
      items[9 * (size_t)m + 2 * (size_t)k + 1] = 0.0f;
      items[11 * (size_t)m + k] = am[q];
      pidx[k] = q;
    }
    FR_HIP(hipMemcpyAsync(d_it, items.data(), sizeof(float) * 12 * m, hipMemcpyHostToDevice,
                          c->stream));
    FR_HIP(hipMemcpyAsync(d_pi, pidx.data(), sizeof(int) * m, hipMemcpyHostToDevice,
                          c->stream));
    float* d_sc = d_it + 12 * (size_t)m;
    int k0 = 0;
    for (auto& kv : by_band) {
      const int cnt = (int)kv.second.size();
      FR_CHECK(kv.first.second > kv.first.first && kv.first.first >= 0,
               "resolution band %d..%d empty (RMAX1, RMAX2, AMAG)", kv.first.first,
               kv.first.second);
      if (int r = cc3m_batch(c, nsam, cnt, j.d_a3df, d_outc, d_ctff, d_pi + k0,
                             d_it + 9 * (size_t)k0, d_it + 9 * (size_t)m + 2 * (size_t)k0,
                             d_it + 11 * (size_t)m + k0, kv.first.first, kv.first.second,
                             j.rbfact, d_fs, j.ring, d_sc + k0, 1))
        return r;
      k0 += cnt;
    }
    sc.assign(m, 0.0f);
    std::vector<float> sc_k(m);
    FR_HIP(hipMemcpyAsync(sc_k.data(), d_sc, sizeof(float) * m, hipMemcpyDeviceToHost,
                          c->stream));
    FR_HIP(hipStreamSynchronize(c->stream));
    for (int k = 0; k < m; k++) sc[order[k]] = sc_k[k];
    // CALCFX's sums in particle order, FP32
    size_t q = 0;
    for (int g : pending) {
      const int n = j.gstart[g + 1] - j.gstart[g];
      const float* x = pw[g].x();
      float rft = 0.0f;
      for (int k = 0; k < n; k++, q++) {
        if (j.mode == 1) {
          const float x1 = x[0], x2 = npar == 1 ? x[0] : x[1];
          rft = rft - sc[q] - fd(j.sig2[g], x1, x2, j.dfstd);
        } else {
          const float ptemp = -sc[q];
          rft = rft + ptemp;
        }
      }
      pw[g].supply(rft / (float)n);
    }
    active = pending;
  }
  for (int g = 0; g < j.ng; g++) {
    const float* x = pw[g].x();
    float* o = out + 4 * (size_t)g;
    if (j.mode == 1) {
      o[0] = x[0];
      o[1] = npar == 1 ? x[0] : x[1];  // ctfrefine.f:62
      o[2] = npar == 1 ? 0.0f : x[2];
    } else {
      o[0] = x[0];
      o[1] = o[2] = 0.0f;
    }
    o[3] = pw[g].value();
    if (nevals) nevals[g] = pw[g].evaluations();
  }
  return 0;
}

int check_common(fr_ctx* c, int nsam, int np, const float* d_spec, const float* ppar,
                 const float* euler, int ng, const int* gstart, const float* d_a3df,
                 const float* d_pssnr, const float* gpar, const float* ring, float* out) {
  CTX_CHECK(c);
  FR_CHECK(d_spec && ppar && euler && gstart && d_a3df && d_pssnr && gpar && ring && out,
           "null argument");
  FR_CHECK(nsam >= 8 && nsam <= 1024 && nsam % 2 == 0, "unsupported box size %d", nsam);
  FR_CHECK(np >= 0 && ng >= 0, "bad sizes");
  FR_CHECK(ng == 0 || (gstart[0] == 0 && gstart[ng] == np), "groups must cover the particles");
  for (int g = 0; g < ng; g++)
    FR_CHECK(gstart[g + 1] > gstart[g], "group %d is empty", g);
  return 0;
}

PrepGlobals globals(const float* gpar) {
  return PrepGlobals{gpar[0], gpar[1], gpar[2], gpar[3], gpar[4], gpar[5],
                     gpar[6], gpar[7], gpar[8], gpar[9], gpar[10], 0};
}

}  // namespace

extern "C" {

int fr_ctfrefine_batch(fr_ctx* c, int nsam, int np, const float* d_spec, const float* ppar,
                       const float* euler, int ng, const int* gstart, int fastig, float dfstd,
                       const float* sig2, const float* d_a3df, const float* d_pssnr,
                       const float* gpar, float rmax1, float rmax2, float rbfact,
                       const float* ring, float* out, int* nevals) {
  if (int r = check_common(c, nsam, np, d_spec, ppar, euler, ng, gstart, d_a3df, d_pssnr, gpar,
                           ring, out))
    return r;
  FR_CHECK(sig2 != nullptr, "null argument");
  FR_CHECK(dfstd != 0.0f, "DFSTD must be non-zero (fd.f divides by it)");
  if (ng == 0) return 0;
  Job j{1, nsam, np, ng, d_spec, ppar, euler, gstart, fastig, dfstd, sig2, d_a3df, d_pssnr,
        globals(gpar), rmax1, rmax2, rbfact, ring};
  return run(c, j, out, nevals);
}

int fr_magrefine_batch(fr_ctx* c, int nsam, int np, const float* d_spec, const float* ppar,
                       const float* euler, int ng, const int* gstart, const float* d_a3df,
                       const float* d_pssnr, const float* gpar, float rmax1, float rmax2,
                       float rbfact, const float* ring, float* out, int* nevals) {
  if (int r = check_common(c, nsam, np, d_spec, ppar, euler, ng, gstart, d_a3df, d_pssnr, gpar,
                           ring, out))
    return r;
  if (ng == 0) return 0;
  Job j{2, nsam, np, ng, d_spec, ppar, euler, gstart, 0, 1.0f, nullptr, d_a3df, d_pssnr,
        globals(gpar), rmax1, rmax2, rbfact, ring};
  return run(c, j, out, nevals);
}

}  // extern "C"

// fr_psearch.cpp -- C-ABI of the reference-exact global search: PSEARCH
// (psearch.f:8-121) over CCP with PAD and ROTATE2D (ccp.f:8-275, pad.f:8-68,
// rotate2d.f:8-77).  The per-view scan runs on the GPU (fr_ccp.hip); the
// bookkeeping PSEARCH does around it -- the RANDOM jitter, the view angles
// XPAR, TESTPAR's strict '>' insertion with CC > 0 admission and the running
// best -- runs on the host in the reference's FP32 order, since it is
// sequential per particle and touches a few hundred numbers.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <vector>

#include "../../include/frealign_hip.h"
#include "fr_ccp.h"
#include "fr_common.h"
#include "fr_ctx.h"
#include "fr_host.h"

using namespace fr;

#define CTX_CHECK(c) FR_CHECK((c) != nullptr, "null fr_ctx")

namespace {

constexpr float kPi = 3.1415926535897f;  // PI of psearch.f:35, ccp.f:34
constexpr size_t kBandBudget = (size_t)1 << 30;  // bytes of padded-transform bands per launch

template <typename T>
int upload(const std::vector<T>& h, const T** d) {
  T* p = nullptr;
  FR_HIP(hipMalloc(&p, sizeof(T) * h.size()));
  FR_HIP(hipMemcpy(p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  *d = p;
  return 0;
}

int nsamr_of(int nsam) {
  const int nsamh = nsam / 2;
  int nsamrh = nsamh / 2;
  if (2 * nsamrh != nsamh) nsamrh = (nsamh + 1) / 2;
  return 2 * nsamrh;
}

// NINT (round half away from zero) of an FP32 value
int nint(float v) { return (int)lroundf(v); }

// NPSI of ccp.f:62-64 for CCP's DANGIN (degrees)
int npsi_of(float dang) {
  float d = 0.25f;
  if (dang != 0.0f) d = dang;
  return nint(90.0f / d);
}

}  // namespace

int fr_ctx::ccp_tables(int nsam, int maxr2, CcpTablesDev* t) {
  const std::pair<int, int> key(nsam, maxr2);
  auto it = ccp_tabs.find(key);
  if (it == ccp_tabs.end()) {
    CcpTablesHost h;
    fr::ccp_tables(nsam, maxr2, &h);
    CcpTablesDev d;
    if (int r = upload(h.d, &d.d)) return r;
    if (int r = upload(h.k, &d.k)) return r;
    if (int r = upload(h.w, &d.w)) return r;
    if (int r = upload(h.e, &d.e)) return r;
    if (int r = upload(h.e1, &d.e1)) return r;
    if (int r = upload(h.region, &d.region)) return r;
    it = ccp_tabs.emplace(key, d).first;
  }
  *t = it->second;
  return 0;
}

extern "C" {

int fr_ccp_view_batch(fr_ctx* c, int nsam, int np, int nview, const float* d_a3df,
                      const float* d_spec, const float* d_ctff, const float* d_dm, int maxr1,
                      int maxr2, float rbfact, float dang, int iquadmax, const int* mask,
                      float* d_score, int* d_code) {
  CTX_CHECK(c);
  FR_CHECK(nsam >= 8 && nsam <= 1024 && nsam % 2 == 0, "unsupported box size %d", nsam);
  FR_CHECK(d_a3df && d_spec && d_ctff && d_dm && mask && d_score && d_code, "null argument");
  FR_CHECK(np >= 0 && nview >= 0, "bad sizes");
  const int nsamr = nsamr_of(nsam);
  FR_CHECK(maxr2 >= 2 && maxr2 <= kCcpMaxR && maxr2 <= nsamr / 2,
           "exact CCP search: MAXR2 %d must be 2..min(%d, NSAMR/2 = %d)", maxr2, kCcpMaxR,
           nsamr / 2);
  FR_CHECK(maxr1 >= 0 && maxr1 < maxr2, "MAXR1 must be 0..MAXR2-1");
  FR_CHECK(iquadmax >= 1 && iquadmax <= 4, "IQUADMAX must be 1..4");
  if (np == 0 || nview == 0) return 0;
  const int npsi = npsi_of(dang);
  FR_CHECK(npsi >= 1 && npsi <= 3600, "NPSI %d out of range (DANG %g)", npsi, (double)dang);
  CcpArgs a;
  a.nsam = nsam;
  a.nview = nview;
  a.maxr1 = maxr1;
  a.maxr2 = maxr2;
  a.np = mask[2] ? npsi : 1;  // ccp.f:137-142
  a.iq = mask[2] ? iquadmax : 1;
  a.mask4 = mask[3] ? 1 : 0;
  a.mask5 = mask[4] ? 1 : 0;
  a.scal = 1.0f / (float)nsam / (float)nsam;  // pad.f:28
  CcpTablesDev t;
  if (int r = c->ccp_tables(nsam, maxr2, &t)) return r;
  // in-plane steps PSIS = (IPSI-1)*PI/2/NPSI, their COS and SIN (ccp.f:146,
  // rotate2d.f:36-37), and the B-factor weights EXP(-RBFACT*r^2) (psearch.f:57)
  std::vector<float> tabs(2 * (size_t)a.np + (size_t)maxr2 * maxr2);
  for (int ip = 0; ip < a.np; ip++) {
    const float psis = (float)ip * kPi / 2.0f / (float)npsi;
    tabs[2 * ip] = cosf(psis);
    tabs[2 * ip + 1] = sinf(psis);
  }
  for (int r2 = 0; r2 < maxr2 * maxr2; r2++)
    tabs[2 * a.np + r2] = expf(-rbfact * (float)r2);
  int err = 0;
  float* d_tabs = (float*)c->slot(33, sizeof(float) * tabs.size(), &err);
  if (!d_tabs) return err;
  FR_HIP(hipMemcpyAsync(d_tabs, tabs.data(), sizeof(float) * tabs.size(), hipMemcpyHostToDevice,
                        c->stream));
  const size_t band = ccp_band_bytes(maxr2);
  const long long total = (long long)np * nview;
  const int chunk = (int)std::min<long long>(total, (long long)std::max<size_t>(1, kBandBudget / band));
  float* d_band = (float*)c->slot(29, band * (size_t)chunk, &err);
  if (!d_band) return err;
  for (long long i0 = 0; i0 < total; i0 += chunk) {
    const int n = (int)std::min<long long>(chunk, total - i0);
    if (int r = launch_ccp(*c, a, (int)i0, n, d_a3df, d_spec, d_ctff, d_dm, d_tabs + 2 * a.np, t,
                           d_tabs, d_band, d_score, d_code))
      return r;
  }
  // the host table must outlive the asynchronous copy
  FR_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int fr_psearch(fr_ctx* c, int nsam, int np, const float* amag, const float* d_a3df,
               const float* d_spec, const float* d_ctff, float dangin, int maxr1, int maxr2,
               float* par, const float* sang, int nsang, int iquadmax, const int* mask,
               float rbfact, float* testpar, int ipmax, int* iran, const float* sm, char asym) {
  CTX_CHECK(c);
  FR_CHECK(amag && par && sang && mask && testpar && iran, "null argument");
  FR_CHECK(np >= 0 && nsang >= 1 && ipmax >= 1, "bad sizes");
  FR_CHECK(asym != 'H' && asym != 'h', "helical PSEARCH (ASYM = H) is not supported");
  if (np == 0) return 0;
  const int nsamr = nsamr_of(nsam);
  // psearch.f:61-67: the jitter of each particle's call, in call order
  const float dang_r = dangin / 180.0f * kPi;
  float dccp = dangin / 5.0f;
  if (dccp < 1.0f) dccp = 1.0f;
  const int npsi = npsi_of(dccp);
  std::vector<float> xpar((size_t)np * nsang * 3), dm((size_t)np * nsang * 9);
  for (int k = 0; k < np; k++) {
    const float phir = (lcg_random(iran) - 0.5f) * dang_r * (float)mask[0];
    const float thetar = (lcg_random(iran) - 0.5f) * dang_r * (float)mask[1];
    const float psir = (lcg_random(iran) - 0.5f) * dang_r * (float)mask[2];
    const float* p = par + 5 * (size_t)k;
    for (int v = 0; v < nsang; v++) {
      float* x = &xpar[((size_t)k * nsang + v) * 3];
      x[0] = (float)mask[0] * (sang[3 * v] + phir) + p[0];  // psearch.f:71-77
      x[1] = (float)mask[1] * (sang[3 * v + 1] + thetar) + p[1];
      x[2] = (float)mask[2] * (sang[3 * v + 2] + psir) + p[2];
      for (int q = 0; q < 3; q++)
        if (x[q] > kPi) x[q] = x[q] - 2.0f * kPi;  // psearch.f:78-80
      rotation(x[0], x[1], x[2], amag[k], sm, &dm[((size_t)k * nsang + v) * 9]);
    }
  }
  int err = 0;
  const size_t nit = (size_t)np * nsang;
  float* d_dm = (float*)c->slot(30, sizeof(float) * 9 * nit, &err);
  if (!d_dm) return err;
  float* d_sc = (float*)c->slot(31, sizeof(float) * nit, &err);
  if (!d_sc) return err;
  int* d_cd = (int*)c->slot(32, sizeof(int) * nit, &err);
  if (!d_cd) return err;
  FR_HIP(hipMemcpy(d_dm, dm.data(), sizeof(float) * 9 * nit, hipMemcpyHostToDevice));
  if (int r = fr_ccp_view_batch(c, nsam, np, nsang, d_a3df, d_spec, d_ctff, d_dm, maxr1, maxr2,
                                rbfact, dccp, iquadmax, mask, d_sc, d_cd))
    return r;
  std::vector<float> sc(nit);
  std::vector<int> cd(nit);
  FR_HIP(hipMemcpy(sc.data(), d_sc, sizeof(float) * nit, hipMemcpyDeviceToHost));
  FR_HIP(hipMemcpy(cd.data(), d_cd, sizeof(int) * nit, hipMemcpyDeviceToHost));
  for (int k = 0; k < np; k++) {
    float* tp = testpar + (size_t)6 * ipmax * k;
    for (int q = 0; q < ipmax; q++) tp[6 * q + 5] = 0.0f;  // psearch.f:43-45
    float ccmax = -1e30f;
    float* p = par + 5 * (size_t)k;
    float best[5] = {p[0], p[1], p[2], p[3], p[4]};
    for (int v = 0; v < nsang; v++) {
      const size_t it = (size_t)k * nsang + v;
      const float cc = sc[it];
      const int code = cd[it];
      const int ipsi = code / 900, iquad = (code / 225) % 4, row = code % 225;
      const int ixm = row / 15 - 7, iym = row % 15 - 7;
      const float* x = &xpar[it * 3];
      float xp[5];
      xp[0] = x[0];
      xp[1] = x[1];
      // ccp.f:146,256-258: PSIS, PSIM = PSI + (IQUAD-1)*PI/2 + PSIS, SX = IXM/NSAMR*2PI
      const float psis = (float)ipsi * kPi / 2.0f / (float)npsi;
      xp[2] = x[2] + (float)iquad * kPi / 2.0f + psis;
      if (xp[2] > kPi) xp[2] = xp[2] - 2.0f * kPi;  // psearch.f:88
      xp[3] = (float)ixm / (float)nsamr * kPi * 2.0f;
      xp[4] = (float)iym / (float)nsamr * kPi * 2.0f;
      // psearch.f:89-102: first slot the score beats, strict '>'
      for (int q = 0; q < ipmax; q++) {
        if (cc > tp[6 * q + 5]) {
          for (int qq = ipmax - 1; qq >= q + 1; qq--)
            for (int j = 0; j < 6; j++) tp[6 * qq + j] = tp[6 * (qq - 1) + j];
          for (int j = 0; j < 5; j++) tp[6 * q + j] = xp[j];
          tp[6 * q + 5] = cc;
          break;
        }
      }
      if (cc > ccmax) {  // psearch.f:105-113
        ccmax = cc;
        for (int j = 0; j < 5; j++) best[j] = xp[j];
      }
    }
    for (int j = 0; j < 5; j++) p[j] = best[j];
  }
  return 0;
}

}  // extern "C"

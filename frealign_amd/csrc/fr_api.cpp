// fr_api.cpp -- the C-ABI entry points (declared in include/frealign_hip.h).
//
// Every entry point validates its shapes on the host before any launch,
// returns 0 on success or a negative status with a message retrievable from
// fr_last_error(); nothing falls back to a CPU path.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/frealign_hip.h"
#include "fr_common.h"
#include "fr_host.h"
#include "fr_kernels.h"
#include "fr_cnn.h"
#include "fr_powell.h"
#include "fr_lists.h"
#include "fr_search.h"
#include "fr_ctx.h"

namespace fr {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int fail(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return -1;
}

int Ctx::ensure_scratch(size_t bytes) {
  if (bytes <= scratch_bytes) return 0;
  if (scratch) {
    FR_HIP(hipStreamSynchronize(stream));
    FR_HIP(hipFree(scratch));
    scratch = nullptr;
    scratch_bytes = 0;
  }
  FR_HIP(hipMalloc(&scratch, bytes));
  scratch_bytes = bytes;
  return 0;
}

const float2* Ctx::twiddles(int n) {
  if (n < 1 || n > 1024) return nullptr;
  if (tw[n]) return tw[n];
  std::vector<float2> h(n);
  for (int k = 0; k < n; k++) {
    double a = -2.0 * M_PI * (double)k / (double)n;
    h[k] = make_float2((float)cos(a), (float)sin(a));
  }
  float2* d = nullptr;
  if (hipMalloc(&d, sizeof(float2) * n) != hipSuccess) return nullptr;
  if (hipMemcpy(d, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    return nullptr;
  }
  tw[n] = d;
  return d;
}

float2* Ctx::prep_partials() {
  if (!prep_part && hipMalloc(&prep_part, sizeof(float2) * 65535 * 16) != hipSuccess)
    prep_part = nullptr;
  return prep_part;
}

int Ctx::prof_begin(int name) {
  if (!prof_on || nprof >= 4096) return -1;
  ProfRec& r = prof[nprof];
  if (hipEventCreate(&r.a) != hipSuccess) return -1;
  if (hipEventCreate(&r.b) != hipSuccess) {
    (void)hipEventDestroy(r.a);
    return -1;
  }
  r.name = name;
  (void)hipEventRecord(r.a, stream);
  return nprof++;
}

void Ctx::prof_end(int slot) {
  if (slot >= 0) (void)hipEventRecord(prof[slot].b, stream);
}

Ctx::~Ctx() {
  for (int i = 0; i < nprof; i++) {
    (void)hipEventDestroy(prof[i].a);
    (void)hipEventDestroy(prof[i].b);
  }
  if (scratch) (void)hipFree(scratch);
  if (prep_part) (void)hipFree(prep_part);
  for (auto* p : tw)
    if (p) (void)hipFree(p);
}

}  // namespace fr

void* fr_ctx::slot(int k, size_t bytes, int* err) {
  if (bytes <= slot_bytes[k]) return slots[k];
  if (slots[k]) {
    if (hipStreamSynchronize(stream) != hipSuccess || hipFree(slots[k]) != hipSuccess) {
      *err = fr::fail("scratch release failed");
      return nullptr;
    }
    slots[k] = nullptr;
    slot_bytes[k] = 0;
  }
  if (hipMalloc(&slots[k], bytes) != hipSuccess) {
    *err = fr::fail("hipMalloc of %zu scratch bytes failed", bytes);
    return nullptr;
  }
  slot_bytes[k] = bytes;
  return slots[k];
}

void* fr_ctx::host_pinned(size_t bytes, int* err) {
  if (bytes <= hpin_bytes) return hpin;
  if (hpin) {
    if (hipStreamSynchronize(stream) != hipSuccess || hipHostFree(hpin) != hipSuccess) {
      *err = fr::fail("pinned staging release failed");
      return nullptr;
    }
    hpin = nullptr;
    hpin_bytes = 0;
  }
  if (hipHostMalloc(&hpin, bytes, hipHostMallocDefault) != hipSuccess) {
    *err = fr::fail("hipHostMalloc of %zu staging bytes failed", bytes);
    return nullptr;
  }
  hpin_bytes = bytes;
  return hpin;
}

namespace {
void sheltest_finish(const double* h, int ns, float fmask, int* npix, int* npixt, double* dfsc,
                     float* pssnr);

template <typename T>
int upload_table(const std::vector<T>& h, T** d) {
  FR_HIP(hipMalloc(d, sizeof(T) * h.size()));
  FR_HIP(hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return 0;
}
}  // namespace

int fr_ctx::cc3m_tables(const fr::Cc3mArgs& a, fr::Cc3mTables* t) {
  auto di = dft_tabs.find(a.nsam);
  if (di == dft_tabs.end()) {
    std::vector<float> folded;
    fr::cc3m_fused_table_folded(a.nsam, folded);
    float* d = nullptr;
    if (int r = upload_table(folded, &d)) return r;
    di = dft_tabs.emplace(a.nsam, d).first;
  }
  t->dft = di->second;
  const std::pair<int, int> pk(a.maxr1 * a.maxr1, a.maxr2 * a.maxr2);
  auto pi = pts_tabs.find(pk);
  if (pi == pts_tabs.end()) {
    std::vector<int> pts;
    fr::cc3m_fused_points(pk.first, pk.second, pts);
    int* d = nullptr;
    if (int r = upload_table(pts, &d)) return r;
    pi = pts_tabs.emplace(pk, d).first;
  }
  t->pts = pi->second;
  const auto rk = std::make_tuple(a.nsam, a.ri2, a.ri3);
  auto ri = ring_tabs.find(rk);
  if (ri == ring_tabs.end()) {
    std::vector<float> ring;
    RingTab rt;
    rt.cnt = fr::cc3m_fused_ring_q(a.nsam, a.ri2, a.ri3, ring);
    if (int r = upload_table(ring, &rt.d)) return r;
    ri = ring_tabs.emplace(rk, rt).first;
  }
  t->ring = ri->second.d;
  t->ring_cnt = ri->second.cnt;
  const auto wk = std::make_tuple(a.nsam, a.ri2, a.ri3, a.rih, a.halfw);
  auto wi = wmap_tabs.find(wk);
  if (wi == wmap_tabs.end()) {
    std::vector<float> w;
    fr::cc3m_fused_wmap(a.nsam, a.ri2, a.ri3, a.rih, a.halfw, w);
    float* d = nullptr;
    if (int r = upload_table(w, &d)) return r;
    wi = wmap_tabs.emplace(wk, d).first;
  }
  t->wmap = wi->second;
  return 0;
}

fr_ctx::Tune fr_ctx::tune_for(int nsam) const {
  auto it = tune.find(nsam);
  return it == tune.end() ? Tune{} : it->second;
}

using namespace fr;

#define CTX_CHECK(c) FR_CHECK((c) != nullptr, "null fr_ctx")
#define SIZE_CHECK(nsam) \
  FR_CHECK(fft_supported(nsam) && (nsam) >= 8 && (nsam) <= 1024, "unsupported box size %d", nsam)

extern "C" {

const char* fr_version(void) { return "frealign-mi355x 0.1 (gfx950)"; }

int fr_profile_enable(fr_ctx* c, int on) {
  CTX_CHECK(c);
  c->prof_on = on != 0;
  return 0;
}

int fr_profile_read(fr_ctx* c, int which, double* total_ms, long long* count) {
  CTX_CHECK(c);
  FR_CHECK(which >= 0 && which < kProfN && total_ms && count, "bad profile query");
  FR_HIP(hipStreamSynchronize(c->stream));
  double t = 0.0;
  long long n = 0;
  int keep = 0;
  for (int i = 0; i < c->nprof; i++) {
    auto& r = c->prof[i];
    if (r.name == which) {
      float ms = 0.f;
      FR_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      t += ms;
      n++;
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    } else {
      c->prof[keep++] = r;
    }
  }
  c->nprof = keep;
  *total_ms = t;
  *count = n;
  return 0;
}
const char* fr_last_error(void) { return g_err.c_str(); }

int fr_create(int device, fr_ctx** out) {
  FR_CHECK(out != nullptr, "null output pointer");
  int n = 0;
  FR_HIP(hipGetDeviceCount(&n));
  FR_CHECK(device >= 0 && device < n, "device %d out of range (%d devices)", device, n);
  FR_HIP(hipSetDevice(device));
  // the no-allocation entry points' device scratch, once per device
  if (int r = fr::noalloc_scratch_init(device)) return r;
  fr_ctx* c = new fr_ctx();
  c->device = device;
  c->stream = nullptr;
  *out = c;
  return 0;
}

int fr_destroy(fr_ctx* c) {
  if (!c) return 0;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (int k = 0; k < fr_ctx::kSlots; k++)
    if (c->slots[k]) (void)hipFree(c->slots[k]);
  if (c->hpin) (void)hipHostFree(c->hpin);
  if (c->d_kidx) (void)hipFree(c->d_kidx);
  for (auto& e : c->dft_tabs) (void)hipFree(e.second);
  for (auto& e : c->pts_tabs) (void)hipFree(e.second);
  for (auto& e : c->ring_tabs) (void)hipFree(e.second.d);
  for (auto& e : c->wmap_tabs) (void)hipFree(e.second);
  for (auto& e : c->ccp_tabs) {
    (void)hipFree((void*)e.second.d);
    (void)hipFree((void*)e.second.k);
    (void)hipFree((void*)e.second.w);
    (void)hipFree((void*)e.second.e);
    (void)hipFree((void*)e.second.e1);
    (void)hipFree((void*)e.second.region);
  }
  if (c->d_shx) (void)hipFree(c->d_shx);
  if (c->d_shy) (void)hipFree(c->d_shy);
  delete c;
  return 0;
}

int fr_set_stream(fr_ctx* c, void* stream) {
  CTX_CHECK(c);
  c->stream = reinterpret_cast<hipStream_t>(stream);
  return 0;
}

int fr_synchronize(fr_ctx* c) {
  CTX_CHECK(c);
  FR_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

// ------------------------------------------------------------ host geometry
int fr_search_grid(float rmm, float ri, float dangin, char asym, int nasym, const int* pmask,
                   int asymstore, float* sang, int cap, int* nsang, int* iquadmax) {
  FR_CHECK(pmask && nsang && iquadmax, "null argument");
  int n = search_grid(rmm, ri, dangin, asym, nasym, pmask, asymstore, sang, cap, iquadmax);
  FR_CHECK(n >= 0, "invalid symmetry '%c'", asym);
  *nsang = n;
  return 0;
}

int fr_point_group(char asym, int nasym, float* symop, int* jsym, int* nsym, int* iredun) {
  FR_CHECK(symop && jsym && nsym && iredun, "null argument");
  FR_CHECK(point_group(asym, nasym, symop, jsym, nsym, iredun) == 0,
           "invalid point group %c%d", asym, nasym);
  return 0;
}

int fr_rotation(int n, const float* euler, float amag, const float* sm, float* dm) {
  FR_CHECK(n >= 0 && euler && dm, "null argument");
  for (int i = 0; i < n; i++)
    rotation(euler[3 * i], euler[3 * i + 1], euler[3 * i + 2], amag, sm, dm + 9 * i);
  return 0;
}

float fr_lcg_random(int* iran) { return lcg_random(iran); }

// ------------------------------------------------------------------- FFTs
int fr_fft2d_r2c(fr_ctx* c, int nsam, int nimg, float* d_data) {
  CTX_CHECK(c);
  SIZE_CHECK(nsam);
  return fft_r2c_2d(*c, nsam, nimg, d_data, (size_t)nsam * (nsam + 2));
}

int fr_fft2d_c2r(fr_ctx* c, int nsam, int nimg, float* d_data) {
  CTX_CHECK(c);
  SIZE_CHECK(nsam);
  return fft_c2r_2d(*c, nsam, nimg, d_data, (size_t)nsam * (nsam + 2));
}

int fr_fft3d_r2c(fr_ctx* c, int nsam, float* d_data) {
  CTX_CHECK(c);
  SIZE_CHECK(nsam);
  FR_CHECK(d_data != nullptr, "null argument");
  return fft_r2c_3d(*c, nsam, d_data);
}

int fr_fft3d_c2r(fr_ctx* c, int nsam, float* d_data) {
  CTX_CHECK(c);
  SIZE_CHECK(nsam);
  FR_CHECK(d_data != nullptr, "null argument");
  return fft_c2r_3d(*c, nsam, d_data);
}

// ------------------------------------------------------- particle preparation
int fr_prep_batch(fr_ctx* c, int nsam, int nb, float* d_data, const float* d_ppar,
                  const float* d_pssnr, const float* gpar, int ictf, float* d_ctff,
                  float* d_outc, float* d_datc) {
  CTX_CHECK(c);
  SIZE_CHECK(nsam);
  FR_CHECK(nb >= 0 && gpar && d_data && d_ppar && d_pssnr && d_ctff && d_outc && d_datc,
           "null argument");
  FR_CHECK(d_outc != d_data, "OUTC must not alias the image buffer");
  PrepGlobals g{gpar[0], gpar[1], gpar[2], gpar[3], gpar[4], gpar[5], gpar[6],
                gpar[7], gpar[8], gpar[9], gpar[10], ictf};
  FR_CHECK(g.halfw != 0.0f, "HALFW must be non-zero");
  int pslot = c->prof_begin(kProfPrep);
  int r = launch_prep(*c, nsam, nb, d_data, d_ppar, d_pssnr, g, d_ctff, d_outc, d_datc);
  c->prof_end(pslot);
  return r;
}

// ------------------------------------------------------ extraction / scoring
int fr_pextract_batch(fr_ctx* c, int nsam, int n, const float* d_a3df, const float* d_dm,
                      const int* d_pidx, const float* d_ctff, float* d_outc) {
  CTX_CHECK(c);
  SIZE_CHECK(nsam);
  FR_CHECK(n >= 0 && d_a3df && d_dm && d_ctff && d_outc, "null argument");
  return launch_pextract(*c, nsam, n, d_a3df, d_dm, d_pidx, d_ctff, d_outc);
}

int fr_cc3m_batch(fr_ctx* c, int nsam, int n, const float* d_a3df, const float* d_spec,
                  const float* d_ctff, const int* d_pidx, const float* d_dm,
                  const float* d_shift, const float* d_amag, int maxr1, int maxr2, float rbfact,
                  const float* d_fsct, const float* ring, float* d_score) {
  return fr::cc3m_batch(c, nsam, n, d_a3df, d_spec, d_ctff, d_pidx, d_dm, d_shift, d_amag, maxr1,
                        maxr2, rbfact, d_fsct, ring, d_score, 0);
}

int fr_cc3_batch(fr_ctx* c, int nsam, int n, const float* d_a3df, const float* d_spec,
                 const float* d_ctff, const int* d_pidx, const float* d_dm, const float* d_shift,
                 const float* d_amag, int maxr1, int maxr2, float rbfact, const float* d_fsct,
                 const float* ring, float* d_score) {
  return fr::cc3m_batch(c, nsam, n, d_a3df, d_spec, d_ctff, d_pidx, d_dm, d_shift, d_amag, maxr1,
                        maxr2, rbfact, d_fsct, ring, d_score, 1);
}

}  // extern "C"

namespace fr {
int cc3m_batch(fr_ctx* c, int nsam, int n, const float* d_a3df, const float* d_spec,
               const float* d_ctff, const int* d_pidx, const float* d_dm, const float* d_shift,
               const float* d_amag, int maxr1, int maxr2, float rbfact, const float* d_fsct,
               const float* ring, float* d_score, int cc3) {
  CTX_CHECK(c);
  SIZE_CHECK(nsam);
  FR_CHECK(n >= 0 && d_a3df && d_spec && d_ctff && d_pidx && d_dm && d_shift && d_amag &&
               d_fsct && ring && d_score,
           "null argument");
  FR_CHECK(maxr1 >= 0 && maxr2 > maxr1, "bad resolution band %d..%d", maxr1, maxr2);
  if (n == 0) return 0;
  Cc3mArgs a{nsam, 0, maxr1, maxr2, rbfact, ring[0], ring[1], ring[2], ring[3]};
  a.cc3 = cc3;
  const fr_ctx::Tune tn = c->tune_for(nsam);
  if (fr::cc3m_fused_ok(nsam, maxr2) && tn.cc3m_fused) {
    Cc3mTables tabs;
    if (int r = c->cc3m_tables(a, &tabs)) return r;
    int pslot = c->prof_begin(kProfCc3m);
    for (int o = 0; o < n; o += 65535) {
      a.n = n - o < 65535 ? n - o : 65535;
      if (int r = launch_cc3m_fused(*c, a, d_a3df, d_spec, d_ctff, d_pidx + o,
                                    d_dm + 9 * (size_t)o, d_shift + 2 * (size_t)o, d_amag + o,
                                    d_fsct, tabs, d_score + o, tn.cc3m_blocks))
        return r;
    }
    c->prof_end(pslot);
    return 0;
  }
  const size_t img = (size_t)nsam * (nsam + 2);
  // process in chunks so the FFT work buffer stays bounded
  const int chunk = (int)((size_t)512 * 1024 * 1024 / (img * sizeof(float)));
  const int step = chunk < 1 ? 1 : (chunk > 65535 ? 65535 : chunk);
  int err = 0;
  float* work = (float*)c->slot(0, img * sizeof(float) * (size_t)(n < step ? n : step), &err);
  if (!work) return err;
  int pslot = c->prof_begin(kProfCc3m);
  for (int o = 0; o < n; o += step) {
    a.n = n - o < step ? n - o : step;
    if (int r = launch_cc3m(*c, a, d_a3df, d_spec, d_ctff, d_pidx + o, d_dm + 9 * (size_t)o,
                            d_shift + 2 * (size_t)o, d_amag + o, d_fsct, work, d_score + o))
      return r;
  }
  c->prof_end(pslot);
  return 0;
}
}  // namespace fr

extern "C" {

int fr_presb_batch(fr_ctx* c, int nsam, int n, const float* d_a3df, const float* d_spec,
                   const float* d_ctff, const int* d_pidx, const float* d_dm,
                   const float* d_shift, const float* d_amag, const float* ring, float* d_rbin) {
  CTX_CHECK(c);
  SIZE_CHECK(nsam);
  FR_CHECK(n >= 0 && d_a3df && d_spec && d_ctff && d_pidx && d_dm && d_shift && d_amag && ring &&
               d_rbin,
           "null argument");
  FR_CHECK(ring[3] != 0.0f, "HALFW must be non-zero");
  if (n == 0) return 0;
  const size_t img = (size_t)nsam * (nsam + 2);
  const int chunk = (int)((size_t)512 * 1024 * 1024 / (img * sizeof(float)));
  const int step = chunk < 1 ? 1 : (chunk > 65535 ? 65535 : chunk);
  int err = 0;
  float* work = (float*)c->slot(0, img * sizeof(float) * (size_t)(n < step ? n : step), &err);
  if (!work) return err;
  for (int o = 0; o < n; o += step) {
    int m = n - o < step ? n - o : step;
    if (int r = launch_presb(*c, nsam, m, d_a3df, d_spec, d_ctff, d_pidx + o, d_dm + 9 * (size_t)o,
                             d_shift + 2 * (size_t)o, d_amag + o, ring[0], ring[1], ring[2],
                             ring[3], work, d_rbin + (size_t)o * nsam))
      return r;
  }
  return 0;
}

// -------------------------------------------------------------- insertion
namespace {
int insert_batch(fr_ctx* c, int nsam, int nb, const float* d_datc, const float* d_ctff,
                 const float* recb, const int* half, int nsym, const int* jsym,
                 const float* symop, float pbc, float boff, float psize, float* d_num0,
                 float* d_ctf2_0, float* d_num1, float* d_ctf2_1, long long* d_nksum,
                 double* d_ttd, const fr::InsertStats* stats);
}  // namespace

int fr_insert_batch(fr_ctx* c, int nsam, int nb, const float* d_datc, const float* d_ctff,
                    const float* recb, const int* half, int nsym, const int* jsym,
                    const float* symop, float pbc, float boff, float psize, float* d_num0,
                    float* d_ctf2_0, float* d_num1, float* d_ctf2_1, long long* d_nksum,
                    double* d_ttd) {
  return insert_batch(c, nsam, nb, d_datc, d_ctff, recb, half, nsym, jsym, symop, pbc, boff, psize,
                      d_num0, d_ctf2_0, d_num1, d_ctf2_1, d_nksum, d_ttd, nullptr);
}

int fr_insert_batch_se(fr_ctx* c, int nsam, int nb, const float* d_datc, const float* d_ctff,
                       const float* recb, const int* half, int nsym, const int* jsym,
                       const float* symop, float pbc, float boff, float psize, float* d_num0,
                       float* d_ctf2_0, float* d_num1, float* d_ctf2_1, long long* d_nksum,
                       double* d_ttd, const float* d_c3df, int* d_ksum, float* d_vsum,
                       float* d_psum, float* d_asum) {
  FR_CHECK(d_c3df && d_ksum && d_vsum && d_psum && d_asum, "null statistics argument");
  const fr::InsertStats st{d_c3df, d_ksum, d_vsum, d_psum, d_asum};
  return insert_batch(c, nsam, nb, d_datc, d_ctff, recb, half, nsym, jsym, symop, pbc, boff, psize,
                      d_num0, d_ctf2_0, d_num1, d_ctf2_1, d_nksum, d_ttd, &st);
}

namespace {
int insert_batch(fr_ctx* c, int nsam, int nb, const float* d_datc, const float* d_ctff,
                 const float* recb, const int* half, int nsym, const int* jsym,
                 const float* symop, float pbc, float boff, float psize, float* d_num0,
                 float* d_ctf2_0, float* d_num1, float* d_ctf2_1, long long* d_nksum,
                 double* d_ttd, const fr::InsertStats* stats) {
  CTX_CHECK(c);
  SIZE_CHECK(nsam);
  FR_CHECK(nb >= 0 && d_datc && d_ctff && recb && half && jsym && symop && d_num0 && d_ctf2_0 &&
               d_num1 && d_ctf2_1 && d_nksum && d_ttd,
           "null argument");
  FR_CHECK(nsym >= 1 && nsym <= 8, "nsym %d out of range", nsym);
  FR_CHECK(pbc != 0.0f && psize != 0.0f, "PBC and PSIZE must be non-zero");
  int ncopy = 1;
  for (int k = 0; k < nsym; k++) {
    FR_CHECK(jsym[k] >= 1 && jsym[k] <= 60, "bad rotational order %d", jsym[k]);
    ncopy *= jsym[k];
  }
  FR_HIP(hipStreamSynchronize(c->stream));  // host task buffer reuse
  auto& tasks = c->host_tasks;
  tasks.clear();
  const int nsamh = nsam / 2;
  std::vector<float> copies(9 * (size_t)ncopy);
  for (int b = 0; b < nb; b++) {
    const float* r = recb + 10 * (size_t)b;
    FR_CHECK(half[b] == 0 || half[b] == 1, "half-map flag must be 0 or 1");
    FR_CHECK(r[7] != 0.0f, "SIG2N of particle %d is zero", b);
    float on = r[6] / r[7];  // a3d3_s.f:44
    if (on == 0.0f) continue;  // pinsert.f:40
    FR_CHECK(r[5] != 0.0f, "AMAG of particle %d is zero", b);
    // pinsert.f:42-62
    float pref = pbc / 100.0f;
    int irrec = (int)((float)nsam * r[9]);
    FR_CHECK(irrec >= 0 && irrec <= nsam, "RREC of particle %d out of range", b);
    float tmp = (r[8] / pref - boff / pbc) / (psize * psize);
    if (tmp > 2.0f) tmp = 2.0f;
    tmp = 2.0f * tmp / ((float)nsamh * (float)nsamh);
    float tm[9];
    rotation(r[0], r[1], r[2], r[5], nullptr, tm);
    int nc = symmetry_copies(tm, nsym, jsym, symop, copies.data());
    FR_CHECK(nc == ncopy, "symmetry enumeration mismatch");
    for (int k = 0; k < nc; k++) {
      InsertTask t;
      memcpy(t.dm, copies.data() + 9 * k, sizeof t.dm);
      t.shx = r[3];
      t.shy = r[4];
      t.tmp = tmp;
      t.occ = on;
      t.part = b;
      t.half = half[b];
      t.irrec2 = irrec * irrec;
      t.ilo = (short)-nsamh;
      t.ihi = (short)nsamh;
      tasks.push_back(t);
    }
  }
  if (tasks.empty()) return 0;
  int err = 0;
  InsertTask* d_tasks = (InsertTask*)c->slot(1, sizeof(InsertTask) * tasks.size(), &err);
  if (!d_tasks) return err;
  FR_HIP(hipMemcpyAsync(d_tasks, tasks.data(), sizeof(InsertTask) * tasks.size(),
                        hipMemcpyHostToDevice, c->stream));
  void* scratch = c->slot(22, insert_scratch_bytes(nsam, nb, (int)tasks.size()), &err);
  if (!scratch) return err;
  int pslot = c->prof_begin(kProfInsert);
  int r = launch_insert(*c, nsam, nb, (int)tasks.size(), d_tasks, d_datc, d_ctff, d_num0,
                        d_ctf2_0, d_num1, d_ctf2_1, reinterpret_cast<unsigned long long*>(d_nksum),
                        d_ttd, scratch, stats);
  c->prof_end(pslot);
  return r;
}
}  // namespace

// ------------------------------------------------------------- statistics
int fr_sheltest(fr_ctx* c, int nsam, int istp, const float* d_a, const float* d_d,
                const float* d_s, const float* d_v, float fmask, int* npix, int* npixt,
                double* dfsc, float* pssnr, int* nstp) {
  CTX_CHECK(c);
  SIZE_CHECK(nsam);
  FR_CHECK(istp >= 1 && d_a && d_d && d_s && d_v && npix && npixt && dfsc && pssnr && nstp,
           "null argument");
  const int jc = nsam / 2 + 1;
  int ns = jc / istp;
  if (ns * istp < jc) ns++;
  int err = 0;
  double* acc = (double*)c->slot(2, sizeof(double) * 6 * ns, &err);
  if (!acc) return err;
  if (int r = launch_shell_sums(*c, nsam, istp, ns, d_a, d_d, d_s, d_v, acc)) return r;
  std::vector<double> h(6 * (size_t)ns);
  FR_HIP(hipMemcpyAsync(h.data(), acc, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                        c->stream));
  FR_HIP(hipStreamSynchronize(c->stream));
  sheltest_finish(h.data(), ns, fmask, npix, npixt, dfsc, pssnr);
  *nstp = ns;
  return 0;
}

int fr_sheltest_noalloc(fr_ctx* c, int nsam, int istp, const float* d_a, const float* d_d,
                        const float* d_s, const float* d_v, float fmask, int* npix, int* npixt,
                        double* dfsc, float* pssnr, int* nstp) {
  CTX_CHECK(c);
  SIZE_CHECK(nsam);
  FR_CHECK(istp >= 1 && d_a && d_d && d_s && d_v && npix && npixt && dfsc && pssnr && nstp,
           "null argument");
  const int jc = nsam / 2 + 1;
  int ns = jc / istp;
  if (ns * istp < jc) ns++;
  FR_CHECK(ns <= fr::kNoAllocShells,
           "fr_sheltest_noalloc: %d shells exceed its fixed scratch of %d shells "
           "(NSAM/2+1 over ISTP)", ns, fr::kNoAllocShells);
  double* acc = fr::noalloc_scratch(c->device);
  FR_CHECK(acc != nullptr, "fr_sheltest_noalloc: no scratch address for device %d", c->device);
  if (int r = launch_shell_sums(*c, nsam, istp, ns, d_a, d_d, d_s, d_v, acc, true)) return r;
  double h[6 * fr::kNoAllocShells];
  FR_HIP(hipMemcpyAsync(h, acc, sizeof(double) * 6 * ns, hipMemcpyDeviceToHost, c->stream));
  FR_HIP(hipStreamSynchronize(c->stream));
  sheltest_finish(h, ns, fmask, npix, npixt, dfsc, pssnr);
  *nstp = ns;
  return 0;
}

}  // extern "C"

namespace {
// sheltest.f:90-112: FSC, NPIX, NPIXT and PSSNR of each shell from its FP64 sums
void sheltest_finish(const double* h, int ns, float fmask, int* npix, int* npixt, double* dfsc,
                     float* pssnr) {
  for (int i = 0; i < ns; i++) {
    const double* a = h + 6 * i;
    npix[i] = (int)a[4];
    npixt[i] = (int)a[5];
    double d = a[0];
    float ps = (float)a[3];
    if (npix[i] != 0) {
      d = d / sqrt(a[1] * a[2]);
      float f = (float)d;
      float a1 = fabsf(f) < 1.0f ? (float)(2.0 * fabs(d) / (1.0 - fabs(d))) : 1000.0f;
      if (ps != 0.0f) ps = a1 * (float)npix[i] / ps * fmask;
    }
    dfsc[i] = d;
    pssnr[i] = ps;
  }
}
}  // namespace

extern "C" {

int fr_storeshift(fr_ctx* c, int nsam, double std_, double vtd, int iredun, const float* d_a,
                  const float* d_s, const float* d_d, const float* d_v, float* d_b, float* d_c) {
  CTX_CHECK(c);
  SIZE_CHECK(nsam);
  FR_CHECK(d_a && d_s && d_d && d_v && d_b && d_c, "null argument");
  FR_CHECK(std_ != 0.0 && vtd != 0.0, "STD or VTD zero: no acceptable particles");
  return launch_storeshift(*c, nsam, std_, vtd, iredun, d_a, d_s, d_d, d_v, d_b, d_c);
}

// ------------------------------------------------------ volume finalisation
int fr_combine_arrays(fr_ctx* c, int nsam, long long nn3, int nvol1, int nvol2, int ifsc,
                      float* d_a3df, float* d_s3df, float* d_d3df, float* d_v3df, int imp,
                      const long long* nksum1, const double* ttd1, const double* ttd2,
                      long long* nksum, double* std_, double* vtd) {
  CTX_CHECK(c);
  SIZE_CHECK(nsam);
  FR_CHECK(d_a3df && d_s3df && nksum1 && ttd1 && nksum && std_, "null argument");
  FR_CHECK(ifsc != 0 || (d_d3df && d_v3df && ttd2 && vtd), "null second half-map argument");
  FR_CHECK(nvol1 >= 1 && nvol2 >= 1 && imp >= 1, "bad NVOL1/NVOL2/IMP");
  FR_CHECK(nn3 >= (long long)(nsam / 2 + 1) * nsam * nsam, "NN3 smaller than the half-volume");
  if (int r = launch_combine(*c, nsam, (size_t)nn3, nvol1, d_a3df, d_s3df)) return r;
  if (ifsc == 0)
    if (int r = launch_combine(*c, nsam, (size_t)nn3, nvol2, d_d3df, d_v3df)) return r;
  // combine_arrays.f:53-67
  for (int i = 0; i < imp; i++) *nksum += nksum1[i];
  const int n = nvol1 != 1 ? nvol1 : imp;
  for (int i = 0; i < n; i++) {
    *std_ = *std_ + ttd1[i];
    if (ifsc == 0) *vtd = *vtd + ttd2[i];
  }
  return 0;
}

int fr_shiftvol(fr_ctx* c, int nsam, double* std_, double vtd, float rlim, int iredun,
                float* d_a3df, float* d_d3df, float* d_s3df, const float* d_v3df, int ifsc,
                const float* pssnr, float fpart, const float* fsc, float fmask, int nshell) {
  CTX_CHECK(c);
  SIZE_CHECK(nsam);
  FR_CHECK(std_ && d_a3df && d_s3df, "null argument");
  FR_CHECK(ifsc != 0 || (d_d3df && d_v3df), "null second half-map argument");
  const int ir = (int)((float)nsam * rlim);
  FR_CHECK(ir >= 0, "negative RLIM");
  if (fpart != 0.0f) {
    FR_CHECK(pssnr && fsc, "null PSSNR/FSC with FPART != 0");
    FR_CHECK(fmask != 0.0f, "FMASK must be non-zero with FPART != 0");
    // shell index INT(SQRT(JRAD2)+0.5)+1 <= IR+1 inside the RLIM sphere
    FR_CHECK(nshell >= ir + 1, "PSSNR/FSC need %d shells (got %d)", ir + 1, nshell);
  }
  *std_ = *std_ + vtd;  // shiftvol.f:23
  int err = 0;
  float* tabs = nullptr;
  if (fpart != 0.0f) {
    tabs = (float*)c->slot(27, sizeof(float) * 2 * (size_t)(ir + 1), &err);
    if (!tabs) return err;
    FR_HIP(hipMemcpyAsync(tabs, pssnr, sizeof(float) * (ir + 1), hipMemcpyHostToDevice,
                          c->stream));
    FR_HIP(hipMemcpyAsync(tabs + ir + 1, fsc, sizeof(float) * (ir + 1), hipMemcpyHostToDevice,
                          c->stream));
  }
  int r = launch_shiftvol(*c, nsam, *std_ * iredun, ir * ir, ifsc, fpart, fmask, d_a3df, d_d3df,
                          d_s3df, d_v3df, tabs, tabs ? tabs + ir + 1 : nullptr);
  if (r) return r;
  // the host tables are read by the copies above: finish before returning
  FR_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int fr_correct3d(fr_ctx* c, int nsam, float* d_a3dv, int interp, int ipad, const float* sinclut,
                 int nlut) {
  CTX_CHECK(c);
  SIZE_CHECK(nsam);
  FR_CHECK(d_a3dv && sinclut, "null argument");
  FR_CHECK(nlut >= kSincLut, "SINCLUT needs %d entries (got %d)", kSincLut, nlut);
  float scal = 1.0f / (float)nsam;  // correct3d.f:23-24
  if (ipad != 0) scal = scal / (float)ipad;
  return launch_correct3d(*c, nsam, scal, interp, sinclut, d_a3dv);
}

int fr_mask3d(fr_ctx* c, int nsam, float* d_a3dv, float rio, float ric, float halfw,
              float* fmask, float* ave, float* std_) {
  CTX_CHECK(c);
  SIZE_CHECK(nsam);
  FR_CHECK(d_a3dv && fmask && ave && std_, "null argument");
  FR_CHECK(halfw != 0.0f, "HALFW must be non-zero");
  // mask3d.f:24-36: radii in FP32 as the reference forms them
  float ri = rio, ri4 = rio * rio, ri5 = ric * ric;
  float rm = fminf(ri, (float)(nsam / 2) - 1.0f - halfw / 2.0f);
  float rii = rm + halfw / 2.0f, rih = rm - halfw / 2.0f;
  if (rih < 0.0f) rih = 0.0f;
  float ri2 = rih * rih, ri3 = rii * rii;
  const float xc = 1.0f + 0.5f * (float)nsam;
  int err = 0;
  double* part = (double*)c->slot(28, sizeof(double) * (5 * (size_t)kMask3dBlocks + 8), &err);
  if (!part) return err;
  double* out = part + 5 * (size_t)kMask3dBlocks;
  double h[5];
  if (int r = launch_mask3d_sums(*c, nsam, d_a3dv, xc, ri2, ri3, ri4, ri5, 1, part, out)) return r;
  FR_HIP(hipMemcpyAsync(h, out, sizeof h, hipMemcpyDeviceToHost, c->stream));
  FR_HIP(hipStreamSynchronize(c->stream));
  // mask3d.f:49-55 (counts are exact in double)
  double bg = h[1] != 0.0 ? h[0] / h[1] : h[0];
  double sum1 = h[2], sum2 = h[3];
  if (h[4] != 0.0) {
    sum1 = sum1 / h[4];
    sum2 = sum2 / h[4];
  }
  *ave = (float)sum1;
  *std_ = (float)sqrt(fabs(sum2 - sum1 * sum1));
  double sum3 = 0.0;
  if (int r = launch_mask3d_apply(*c, nsam, d_a3dv, xc, ri2, ri3, rih, rii, halfw, bg, 0, part, out))
    return r;
  FR_HIP(hipMemcpyAsync(h, out, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  FR_HIP(hipStreamSynchronize(c->stream));
  sum3 = h[0];
  if (ric > 0.0f) {  // mask3d.f:74-118: the inner mask, on the masked volume
    ri = ric;
    rm = fmaxf(ri, halfw / 2.0f);
    rii = rm + halfw / 2.0f;
    rih = rm - halfw / 2.0f;
    ri2 = rih * rih;
    ri3 = rii * rii;
    if (int r = launch_mask3d_sums(*c, nsam, d_a3dv, xc, ri2, ri3, 0.f, 0.f, 0, part, out))
      return r;
    FR_HIP(hipMemcpyAsync(h, out, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    FR_HIP(hipStreamSynchronize(c->stream));
    bg = h[1] != 0.0 ? h[0] / h[1] : h[0];
    if (int r = launch_mask3d_apply(*c, nsam, d_a3dv, xc, ri2, ri3, rih, rii, halfw, bg, 1, part,
                                    out))
      return r;
    FR_HIP(hipMemcpyAsync(h, out, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    FR_HIP(hipStreamSynchronize(c->stream));
    sum3 = sum3 + h[0];
  }
  *fmask = (float)(sum3 / nsam / nsam / nsam);  // mask3d.f:119
  return 0;
}

int fr_fpart_from_mw(int nsam, float mw, float dalt, float psize, float* ssnr, float* fpart) {
  FR_CHECK(fpart && ssnr, "null argument");
  FR_CHECK(nsam >= 2, "bad NSAM");
  FR_CHECK(mw != 0.0f, "MW = 0 needs the iterative FPART search, which is not provided");
  FR_CHECK(dalt != 0.0f && psize != 0.0f, "DALT and PSIZE must be non-zero");
  // find_fpart.f:141-149
  const float box = psize * (float)nsam;
  float f = 1000.0f * mw / dalt / (box * box * box);
  f = fabsf(f);
  if (f != 0.0f)
    for (int i = 0; i < nsam / 2 + 1; i++) ssnr[i] = ssnr[i] / f;
  *fpart = f;
  return 0;
}

// ------------------------------------------------------------------ search
// The correlation annulus of ccp.f:110-131 and 194-216 on the half-plane:
// I >= 0, max(MAXR1^2, 1) <= I^2+J^2 < MAXR2^2 (DC is outside the NSAMR
// correlation loop), listed I-major.
static int search_annulus(int nsam, int maxr1, int maxr2, std::vector<int>& kidx) {
  kidx.clear();
  const int jc = nsam / 2 + 1;
  int lo = maxr1 * maxr1 < 1 ? 1 : maxr1 * maxr1, hi = maxr2 * maxr2;
  for (int i = 0; i <= jc - 2; i++)
    for (int j = -jc + 1; j <= jc - 1; j++) {
      int it = i * i + j * j;
      if (it >= lo && it < hi && abs(j) != jc - 1) {
        kidx.push_back(i);
        kidx.push_back(j);
      }
    }
  return (int)kidx.size() / 2;
}

static int search_tables(fr_ctx* c, int nsam, int maxr1, int maxr2) {
  if (c->skey[0] == nsam && c->skey[1] == maxr1 && c->skey[2] == maxr2) return 0;
  int nsamh = nsam / 2, nsamrh = nsamh / 2;
  if (2 * nsamrh != nsamh) nsamrh = (nsamh + 1) / 2;  // ccp.f:49-52
  const int nsamr = 2 * nsamrh;
  FR_CHECK(maxr2 <= nsamrh, "search band MAXR2=%d exceeds the NSAMR/2 grid (%d)", maxr2, nsamrh);
  std::vector<int> kidx;
  int kc = search_annulus(nsam, maxr1, maxr2, kidx);
  FR_CHECK(kc > 0, "empty search annulus");
  int half = (kc + 15) / 16 * 16;
  kidx.resize(2 * (size_t)half, 0);
  std::vector<float> shx(2 * (size_t)half * 15, 0.f), shy(2 * (size_t)half * 15, 0.f);
  for (int k = 0; k < kc; k++)
    for (int s = 0; s < 15; s++) {
      double ax = 2.0 * M_PI * kidx[2 * k] * (s - 7) / nsamr;
      double ay = 2.0 * M_PI * kidx[2 * k + 1] * (s - 7) / nsamr;
      shx[2 * (k * 15 + s)] = (float)cos(ax);
      shx[2 * (k * 15 + s) + 1] = (float)sin(ax);
      shy[2 * (k * 15 + s)] = (float)cos(ay);
      shy[2 * (k * 15 + s) + 1] = (float)sin(ay);
    }
  FR_HIP(hipStreamSynchronize(c->stream));
  if (c->d_kidx) (void)hipFree(c->d_kidx);
  if (c->d_shx) (void)hipFree(c->d_shx);
  if (c->d_shy) (void)hipFree(c->d_shy);
  c->d_kidx = nullptr;
  c->d_shx = c->d_shy = nullptr;
  FR_HIP(hipMalloc(&c->d_kidx, sizeof(int) * kidx.size()));
  FR_HIP(hipMalloc(&c->d_shx, sizeof(float) * shx.size()));
  FR_HIP(hipMalloc(&c->d_shy, sizeof(float) * shy.size()));
  FR_HIP(hipMemcpy(c->d_kidx, kidx.data(), sizeof(int) * kidx.size(), hipMemcpyHostToDevice));
  FR_HIP(hipMemcpy(c->d_shx, shx.data(), sizeof(float) * shx.size(), hipMemcpyHostToDevice));
  FR_HIP(hipMemcpy(c->d_shy, shy.data(), sizeof(float) * shy.size(), hipMemcpyHostToDevice));
  c->kc = kc;
  c->kpad = 2 * half;
  c->nsamr = nsamr;
  c->skey[0] = nsam;
  c->skey[1] = maxr1;
  c->skey[2] = maxr2;
  return 0;
}

int fr_search_layout(fr_ctx* c, int nsam, int maxr1, int maxr2, int* kc, int* kpad) {
  CTX_CHECK(c);
  SIZE_CHECK(nsam);
  FR_CHECK(kc && kpad, "null argument");
  if (int r = search_tables(c, nsam, maxr1, maxr2)) return r;
  *kc = c->kc;
  *kpad = c->kpad;
  return 0;
}

int fr_search_bank(fr_ctx* c, int nsam, int maxr1, int maxr2, const float* d_a3df, int norient,
                   const float* d_dm, int dtype, void* d_bank) {
  CTX_CHECK(c);
  SIZE_CHECK(nsam);
  FR_CHECK(d_a3df && d_dm && d_bank && norient >= 0, "null argument");
  FR_CHECK(dtype >= 0 && dtype <= 2, "dtype must be 0 (fp32), 1 (bf16) or 2 (bf16 separable)");
  if (int r = search_tables(c, nsam, maxr1, maxr2)) return r;
  if (dtype == 2) {
    FR_CHECK(maxr2 <= kSepMaxR2, "the separable bank holds MAXR2 <= %d (got %d)", kSepMaxR2,
             maxr2);
    return launch_bank_build_sep(*c, nsam, norient, maxr1 * maxr1 < 1 ? 1 : maxr1 * maxr1,
                                 maxr2 * maxr2, d_a3df, d_dm, d_bank);
  }
  return launch_bank_build(*c, nsam, norient, c->kc, c->kpad, d_a3df, d_dm, c->d_kidx, d_bank,
                           dtype == 1);
}

// Particle terms of the separable search (slots 23-25) and SUM2 (slot 5).
static int sep_terms(fr_ctx* c, int nsam, int maxr1, int maxr2, float rbfact, int np,
                     const float* d_outc, const float* d_ctff, void** xa, void** xb, float** c2s,
                     float** sum2) {
  FR_CHECK(maxr2 <= kSepMaxR2, "the separable search holds MAXR2 <= %d (got %d)", kSepMaxR2,
           maxr2);
  const size_t per = sizeof(unsigned) * 2048 * (size_t)np;
  int err = 0;
  if (!(*sum2 = (float*)c->slot(5, sizeof(float) * (size_t)np, &err))) return err;
  if (!(*xa = c->slot(23, per, &err))) return err;
  if (!(*xb = c->slot(24, per, &err))) return err;
  if (!(*c2s = (float*)c->slot(25, per, &err))) return err;
  if (int r = launch_search_feats(*c, nsam, np, c->kc, c->kpad, maxr2 * maxr2, rbfact, d_outc,
                                  d_ctff, c->d_kidx, nullptr, nullptr, *sum2))
    return r;
  return launch_sep_feats(*c, nsam, np, maxr1 * maxr1 < 1 ? 1 : maxr1 * maxr1, maxr2 * maxr2,
                          rbfact, d_outc, d_ctff, *xa, *xb, *c2s);
}

int fr_search_batch(fr_ctx* c, int nsam, int maxr1, int maxr2, float rbfact, int norient,
                    const void* d_bank, int dtype, int group, int np, const float* d_outc,
                    const float* d_ctff, int topk, int* d_top_o, int* d_top_row,
                    float* d_top_score) {
  CTX_CHECK(c);
  SIZE_CHECK(nsam);
  FR_CHECK(d_bank && d_outc && d_ctff && d_top_o && d_top_row && d_top_score, "null argument");
  FR_CHECK(dtype >= 0 && dtype <= 2, "dtype must be 0 (fp32), 1 (bf16) or 2 (bf16 separable)");
  FR_CHECK(np >= 0 && norient > 0 && topk >= 1 && topk <= 32, "bad sizes");
  FR_CHECK(group >= 1 && norient % group == 0, "orientations must come in whole view groups");
  if (np == 0) return 0;
  if (int r = search_tables(c, nsam, maxr1, maxr2)) return r;
  const int half = c->kpad / 2;
  int err = 0;
  if (dtype == 2) {
    void *xa, *xb;
    float *c2s, *sum2;
    if (int r = sep_terms(c, nsam, maxr1, maxr2, rbfact, np, d_outc, d_ctff, &xa, &xb, &c2s, &sum2))
      return r;
    float* score = (float*)c->slot(6, sizeof(float) * (size_t)np * norient, &err);
    if (!score) return err;
    int* brow = (int*)c->slot(7, sizeof(int) * (size_t)np * norient, &err);
    if (!brow) return err;
    int pslot = c->prof_begin(kProfSearchGemm);
    if (int r = launch_search_sep(*c, np, norient, c->nsamr, norient, nullptr, nullptr, 0, 0,
                                  nullptr, d_bank, xa, xb, c2s, sum2, score, brow))
      return r;
    c->prof_end(pslot);
    return launch_topk(*c, norient, np, norient, group, nullptr, nullptr, nullptr, 0, topk, score, brow,
                       d_top_o, d_top_row, d_top_score);
  }
  float* ap = (float*)c->slot(3, sizeof(float) * 2 * (size_t)np * half, &err);
  if (!ap) return err;
  float* c2 = (float*)c->slot(4, sizeof(float) * (size_t)np * half, &err);
  if (!c2) return err;
  float* sum2 = (float*)c->slot(5, sizeof(float) * (size_t)np, &err);
  if (!sum2) return err;
  float* score = (float*)c->slot(6, sizeof(float) * (size_t)np * norient, &err);
  if (!score) return err;
  int* brow = (int*)c->slot(7, sizeof(int) * (size_t)np * norient, &err);
  if (!brow) return err;
  float* nrm = (float*)c->slot(2, sizeof(float) * (size_t)np * norient, &err);
  if (!nrm) return err;
  if (int r = launch_search_feats(*c, nsam, np, c->kc, c->kpad, maxr2 * maxr2, rbfact, d_outc,
                                  d_ctff, c->d_kidx, ap, c2, sum2))
    return r;
  if (int r = launch_search_norm(*c, norient, c->kpad, np, d_bank, dtype == 1, c2, nrm)) return r;
  int pslot = c->prof_begin(kProfSearchGemm);
  if (int r = launch_search_gemm(*c, norient, norient, c->kpad, np, d_bank, dtype == 1, nullptr,
                                 ap, c2, nrm, sum2, c->d_shx, c->d_shy, score, brow))
    return r;
  c->prof_end(pslot);
  return launch_topk(*c, norient, np, norient, group, nullptr, nullptr, nullptr, 0, topk, score, brow,
                     d_top_o, d_top_row, d_top_score);
}

// ------------------------------------------------ learned orientation proposal
#define CNN_SHAPE(lo)                                    \
  fr::CnnLayout lo;                                      \
  if (int r_ = fr::cnn_layout(S, C, nconv, K, &lo)) return r_

int fr_cnn_param_count(int S, int C, int nconv, int K, long long* nparam) {
  FR_CHECK(nparam, "null argument");
  CNN_SHAPE(lo);
  *nparam = lo.total;
  return 0;
}

int fr_cnn_init(fr_ctx* c, int S, int C, int nconv, int K, unsigned long long seed,
                float* d_master, float* d_wq) {
  CTX_CHECK(c);
  FR_CHECK(d_master && d_wq, "null argument");
  CNN_SHAPE(lo);
  return cnn_init(*c, lo, seed, d_master, d_wq);
}

int fr_cnn_forward(fr_ctx* c, int S, int C, int nconv, int K, const float* d_wq, int nb,
                   const float* d_img, float* d_logits) {
  CTX_CHECK(c);
  FR_CHECK(d_wq && d_img && d_logits && nb >= 0, "null argument");
  CNN_SHAPE(lo);
  if (nb == 0) return 0;
  int err = 0;
  float* ws = (float*)c->slot(8, sizeof(float) * cnn_workspace(lo, nb, false), &err);
  if (!ws) return err;
  int pslot = c->prof_begin(kProfCnn);
  int r = cnn_forward(*c, lo, d_wq, nb, d_img, ws, d_logits);
  c->prof_end(pslot);
  return r;
}

int fr_cnn_forward_bf16(fr_ctx* c, int S, int C, int nconv, int K, const float* d_wq, int nb,
                        const float* d_img, float* d_logits) {
  CTX_CHECK(c);
  FR_CHECK(d_wq && d_img && d_logits && nb >= 0, "null argument");
  CNN_SHAPE(lo);
  FR_CHECK(cnn_bf16_ok(lo), "BF16 inference needs C = 64 channels and 2..8 convolutions");
  if (nb == 0) return 0;
  int err = 0;
  void* ws = c->slot(26, cnn_workspace_bf16(lo, nb), &err);
  if (!ws) return err;
  int pslot = c->prof_begin(kProfCnn);
  int r = cnn_forward_bf16(*c, lo, d_wq, nb, d_img, ws, d_logits);
  c->prof_end(pslot);
  return r;
}

int fr_cnn_gradients(fr_ctx* c, int S, int C, int nconv, int K, const float* d_wq, int nb,
                     const float* d_img, const int* d_label, float* d_loss, float* d_grad) {
  CTX_CHECK(c);
  FR_CHECK(d_wq && d_img && d_label && d_loss && d_grad && nb >= 1, "null argument");
  CNN_SHAPE(lo);
  int err = 0;
  float* ws = (float*)c->slot(9, sizeof(float) * cnn_workspace(lo, nb, true), &err);
  if (!ws) return err;
  return cnn_backward(*c, lo, d_wq, nb, d_img, d_label, ws, d_loss, d_grad);
}

int fr_cnn_gradients_bf16(fr_ctx* c, int S, int C, int nconv, int K, const float* d_wq, int nb,
                          const float* d_img, const int* d_label, float* d_loss, float* d_grad) {
  CTX_CHECK(c);
  FR_CHECK(d_wq && d_img && d_label && d_loss && d_grad && nb >= 1, "null argument");
  CNN_SHAPE(lo);
  int err = 0;
  float* ws = (float*)c->slot(9, sizeof(float) * cnn_workspace(lo, nb, true), &err);
  if (!ws) return err;
  return cnn_backward(*c, lo, d_wq, nb, d_img, d_label, ws, d_loss, d_grad, true);
}

int fr_cnn_adam(fr_ctx* c, int S, int C, int nconv, int K, float* d_master, float* d_wq,
                const float* d_grad, float* d_m, float* d_v, int step, float lr) {
  CTX_CHECK(c);
  FR_CHECK(d_master && d_wq && d_grad && d_m && d_v, "null argument");
  CNN_SHAPE(lo);
  return cnn_adam(*c, lo, d_master, d_wq, d_grad, d_m, d_v, step, lr);
}

int fr_cnn_propose(fr_ctx* c, int K, int nb, const float* d_logits, int topk, int* d_top,
                   float* d_prob) {
  CTX_CHECK(c);
  FR_CHECK(d_logits && d_top && d_prob && nb >= 0, "null argument");
  if (nb == 0) return 0;
  return cnn_propose(*c, K, nb, d_logits, topk, d_top, d_prob);
}

int fr_cnn_images(fr_ctx* c, int nsam, int S, int nb, const float* d_datc, const float* d_ctff,
                  float* d_img) {
  CTX_CHECK(c);
  SIZE_CHECK(nsam);
  FR_CHECK(d_datc && d_ctff && d_img && nb >= 0, "null argument");
  FR_CHECK(S >= 4 && S <= nsam && S % 2 == 0 && fft_supported(S), "bad network input size %d", S);
  if (nb == 0) return 0;
  int err = 0;
  float* buf = (float*)c->slot(10, sizeof(float) * (size_t)nb * S * (S + 2), &err);
  if (!buf) return err;
  return cnn_images(*c, nsam, S, nb, d_datc, d_ctff, buf, d_img);
}

int fr_cnn_synth(fr_ctx* c, int nsam, int S, int nb, const float* d_a3df, const float* d_dm,
                 const float* d_spar, const float* gpar, float disc, unsigned long long seed,
                 float* d_img) {
  CTX_CHECK(c);
  SIZE_CHECK(nsam);
  FR_CHECK(d_a3df && d_dm && d_spar && gpar && d_img && nb >= 0, "null argument");
  FR_CHECK(S >= 4 && S <= nsam && S % 2 == 0 && fft_supported(S), "bad network input size %d", S);
  FR_CHECK(disc > 0.0f, "disc must be positive");
  if (nb == 0) return 0;
  PrepGlobals g{gpar[0], gpar[1], gpar[2], gpar[3], gpar[4], gpar[5], gpar[6],
                gpar[7], gpar[8], gpar[9], gpar[10], 1};
  int err = 0;
  float* buf = (float*)c->slot(10, sizeof(float) * (size_t)nb * S * (S + 2), &err);
  if (!buf) return err;
  return cnn_synth(*c, nsam, S, nb, d_a3df, d_dm, d_spar, g, disc, seed, buf, d_img);
}

int fr_orient_nearest(fr_ctx* c, int n, const float* d_dm, int ncls, const float* d_dmc,
                      int nsym, const float* d_sym, int* d_label) {
  CTX_CHECK(c);
  FR_CHECK(d_dm && d_dmc && d_sym && d_label && n >= 0 && ncls >= 1, "null argument");
  if (n == 0) return 0;
  return orient_nearest(*c, n, d_dm, ncls, d_dmc, nsym, d_sym, d_label);
}

int fr_orient_within(fr_ctx* c, int na, const float* d_dma, int nb, const float* d_dmb,
                     int nsym, const float* d_sym, float max_deg, unsigned* d_mask) {
  CTX_CHECK(c);
  FR_CHECK(d_dma && d_dmb && d_sym && d_mask && na >= 0 && nb >= 0, "null argument");
  FR_CHECK(max_deg >= 0.0f && max_deg <= 180.0f, "max_deg must be in [0, 180]");
  if (na == 0 || nb == 0) return 0;
  const float tmin = (float)(1.0 + 2.0 * cos((double)max_deg * 3.14159265358979323846 / 180.0));
  return orient_within(*c, na, d_dma, nb, d_dmb, nsym, d_sym, tmin, d_mask);
}

int fr_search_list_batch(fr_ctx* c, int nsam, int maxr1, int maxr2, float rbfact, int norient,
                         const void* d_bank, int dtype, int np, const float* d_outc,
                         const float* d_ctff, int ncand, const int* d_cand, int group, int topk,
                         int* d_top_o, int* d_top_row, float* d_top_score) {
  CTX_CHECK(c);
  SIZE_CHECK(nsam);
  FR_CHECK(d_bank && d_outc && d_ctff && d_cand && d_top_o && d_top_row && d_top_score,
           "null argument");
  FR_CHECK(dtype >= 0 && dtype <= 2, "dtype must be 0 (fp32), 1 (bf16) or 2 (bf16 separable)");
  FR_CHECK(np >= 0 && norient > 0 && ncand > 0 && topk >= 1 && topk <= 32, "bad sizes");
  FR_CHECK(group >= 1 && ncand % group == 0, "candidate lists must come in whole groups");
  if (np == 0) return 0;
  if (int r = search_tables(c, nsam, maxr1, maxr2)) return r;
  const int half = c->kpad / 2;
  int err = 0;
  if (dtype == 2) {
    void *xa, *xb;
    float *c2s, *sum2;
    if (int r = sep_terms(c, nsam, maxr1, maxr2, rbfact, np, d_outc, d_ctff, &xa, &xb, &c2s, &sum2))
      return r;
    float* score = (float*)c->slot(6, sizeof(float) * (size_t)np * ncand, &err);
    if (!score) return err;
    int* brow = (int*)c->slot(7, sizeof(int) * (size_t)np * ncand, &err);
    if (!brow) return err;
    int pslot = c->prof_begin(kProfSearchList);
    if (int r = launch_search_sep(*c, np, norient, c->nsamr, ncand, nullptr, nullptr, 0, 0,
                                  d_cand, d_bank, xa, xb, c2s, sum2, score, brow))
      return r;
    c->prof_end(pslot);
    return launch_topk(*c, norient, np, ncand, group, nullptr, nullptr, d_cand, 0, topk, score, brow,
                       d_top_o, d_top_row, d_top_score);
  }
  float* ap = (float*)c->slot(3, sizeof(float) * 2 * (size_t)np * half, &err);
  if (!ap) return err;
  float* c2 = (float*)c->slot(4, sizeof(float) * (size_t)np * half, &err);
  if (!c2) return err;
  float* sum2 = (float*)c->slot(5, sizeof(float) * (size_t)np, &err);
  if (!sum2) return err;
  float* score = (float*)c->slot(6, sizeof(float) * (size_t)np * ncand, &err);
  if (!score) return err;
  int* brow = (int*)c->slot(7, sizeof(int) * (size_t)np * ncand, &err);
  if (!brow) return err;
  if (int r = launch_search_feats(*c, nsam, np, c->kc, c->kpad, maxr2 * maxr2, rbfact, d_outc,
                                  d_ctff, c->d_kidx, ap, c2, sum2))
    return r;
  int pslot = c->prof_begin(kProfSearchList);
  if (int r = launch_search_gemm(*c, ncand, norient, c->kpad, np, d_bank, dtype == 1, d_cand, ap,
                                 c2, nullptr, sum2, c->d_shx, c->d_shy, score, brow))
    return r;
  c->prof_end(pslot);
  return launch_topk(*c, norient, np, ncand, group, nullptr, nullptr, d_cand, 0, topk, score, brow,
                     d_top_o, d_top_row, d_top_score);
}

int fr_build_lists(fr_ctx* c, int np, int ktop, const int* d_top, const float* d_prob,
                   float threshold, int ncls, const int* d_cell_off, const int* d_cells, int group,
                   int nbank, int bank_group, int maxlen, int* d_loff, int* d_grp, int* d_cand,
                   int* d_info) {
  CTX_CHECK(c);
  FR_CHECK(d_top && d_prob && d_cell_off && d_cells && d_loff && d_grp && d_cand && d_info,
           "null argument");
  FR_CHECK(np >= 0 && ktop >= 1 && ncls >= 1 && nbank >= 1, "bad sizes");
  FR_CHECK(group >= 1 && bank_group >= 1 && nbank % bank_group == 0,
           "the bank (%d) must come in whole view groups of %d", nbank, bank_group);
  FR_CHECK(maxlen >= group && maxlen >= bank_group, "list capacity %d below one group", maxlen);
  FR_CHECK((long long)np * maxlen < (1LL << 31), "list capacity %d x %d exceeds 2^31 entries",
           np, maxlen);
  if (np == 0) return 0;
  int err = 0;
  int* len = (int*)c->slot(16, sizeof(int) * (size_t)np, &err);
  if (!len) return err;
  return launch_list_build(*c, np, ktop, d_top, d_prob, threshold, ncls, d_cell_off, d_cells, group,
                           nbank, bank_group, maxlen, len, d_loff, d_grp, d_info, d_cand);
}

int fr_search_ragged_batch(fr_ctx* c, int nsam, int maxr1, int maxr2, float rbfact, int norient,
                           const void* d_bank, int dtype, int np, const float* d_outc,
                           const float* d_ctff, const int* d_loff, const int* d_grp,
                           const int* d_cand, const int* d_info, int cap, int topk, int* d_top_o,
                           int* d_top_row, float* d_top_score) {
  CTX_CHECK(c);
  SIZE_CHECK(nsam);
  FR_CHECK(d_bank && d_outc && d_ctff && d_loff && d_grp && d_cand && d_top_o && d_top_row &&
               d_top_score,
           "null argument");
  FR_CHECK(dtype >= 0 && dtype <= 2, "dtype must be 0 (fp32), 1 (bf16) or 2 (bf16 separable)");
  FR_CHECK(np >= 0 && norient > 0 && cap > 0 && topk >= 1 && topk <= 32, "bad sizes");
  if (np == 0) return 0;
  if (int r = search_tables(c, nsam, maxr1, maxr2)) return r;
  const int half = c->kpad / 2;
  int err = 0;
  // tiles of lists within the capacity: at most cap/256 full tiles plus one
  // partial tile per particle
  const long long max_tiles = (long long)cap / 256 + np;
  if (dtype == 2) {
    void *xa, *xb;
    float *c2s, *sum2;
    if (int r = sep_terms(c, nsam, maxr1, maxr2, rbfact, np, d_outc, d_ctff, &xa, &xb, &c2s, &sum2))
      return r;
    float* score = (float*)c->slot(6, sizeof(float) * (size_t)cap, &err);
    if (!score) return err;
    int* brow = (int*)c->slot(7, sizeof(int) * (size_t)cap, &err);
    if (!brow) return err;
    int* toff = (int*)c->slot(17, sizeof(int) * ((size_t)np + 1), &err);
    if (!toff) return err;
    // with the confidence flags, the fallback particles (whole bank in order)
    // leave the ragged tiles for the shared-template form
    int* fb = nullptr;
    if (d_info) {
      fb = (int*)c->slot(34, sizeof(int) * ((size_t)np + 1), &err);
      if (!fb) return err;
    }
    // the tiles in the order of the particles' first list entries
    int* ord = (int*)c->slot(35, sizeof(int) * (size_t)np, &err);
    if (!ord) return err;
    if (int r = launch_tile_offsets(*c, np, d_loff, toff, d_info, fb, d_cand, ord)) return r;
    int pslot = c->prof_begin(kProfSearchList);
    if (int r = launch_search_sep(*c, np, norient, c->nsamr, 0, d_loff, toff, cap,
                                  (int)(max_tiles < (1LL << 30) ? max_tiles : (1LL << 30)), d_cand,
                                  d_bank, xa, xb, c2s, sum2, score, brow, fb, ord))
      return r;
    c->prof_end(pslot);
    return launch_topk(*c, norient, np, 0, 0, d_loff, d_grp, d_cand, cap, topk, score, brow, d_top_o,
                       d_top_row, d_top_score);
  }
  float* ap = (float*)c->slot(3, sizeof(float) * 2 * (size_t)np * half, &err);
  if (!ap) return err;
  float* c2 = (float*)c->slot(4, sizeof(float) * (size_t)np * half, &err);
  if (!c2) return err;
  float* sum2 = (float*)c->slot(5, sizeof(float) * (size_t)np, &err);
  if (!sum2) return err;
  float* score = (float*)c->slot(6, sizeof(float) * (size_t)cap, &err);
  if (!score) return err;
  int* brow = (int*)c->slot(7, sizeof(int) * (size_t)cap, &err);
  if (!brow) return err;
  int* toff = (int*)c->slot(17, sizeof(int) * ((size_t)np + 1), &err);
  if (!toff) return err;
  if (int r = launch_tile_offsets(*c, np, d_loff, toff)) return r;
  if (int r = launch_search_feats(*c, nsam, np, c->kc, c->kpad, maxr2 * maxr2, rbfact, d_outc,
                                  d_ctff, c->d_kidx, ap, c2, sum2))
    return r;
  int pslot = c->prof_begin(kProfSearchList);
  if (int r = launch_search_gemm_ragged(*c, np, norient, c->kpad, d_bank, dtype == 1, d_loff, toff,
                                        d_cand, cap, (int)(max_tiles < (1LL << 30) ? max_tiles
                                                                                     : (1LL << 30)),
                                        ap, c2, sum2, c->d_shx, c->d_shy, score, brow))
    return r;
  c->prof_end(pslot);
  return launch_topk(*c, norient, np, 0, 0, d_loff, d_grp, d_cand, cap, topk, score, brow, d_top_o,
                     d_top_row, d_top_score);
}

int fr_rescore_batch(fr_ctx* c, int nsam, int np, int k, const float* d_a3df, const float* d_outc,
                     const float* d_ctff, const int* d_top_o, const int* d_top_row, int nbank,
                     const float* d_dm_bank, int nsamr, int maxr1, int maxr2, float rbfact,
                     const float* d_fsct, const float* ring, float* d_score, int* d_best) {
  CTX_CHECK(c);
  SIZE_CHECK(nsam);
  FR_CHECK(d_a3df && d_outc && d_ctff && d_top_o && d_top_row && d_dm_bank && d_fsct && ring &&
               d_score && d_best,
           "null argument");
  FR_CHECK(np >= 0 && k >= 1 && nbank >= 1 && nsamr >= 2, "bad sizes");
  FR_CHECK((long long)np * k < (1LL << 31), "too many candidates");
  if (np == 0) return 0;
  const size_t n = (size_t)np * k;
  int err = 0;
  int* pidx = (int*)c->slot(18, sizeof(int) * n, &err);
  if (!pidx) return err;
  float* dm = (float*)c->slot(19, sizeof(float) * 9 * n, &err);
  if (!dm) return err;
  float* sh = (float*)c->slot(20, sizeof(float) * 2 * n, &err);
  if (!sh) return err;
  float* am = (float*)c->slot(21, sizeof(float) * n, &err);
  if (!am) return err;
  if (int r = launch_cand_items(*c, np, k, d_top_o, d_top_row, nbank, d_dm_bank, nsamr, pidx, dm,
                                sh, am))
    return r;
  if (int r = fr_cc3m_batch(c, nsam, (int)n, d_a3df, d_outc, d_ctff, pidx, dm, sh, am, maxr1,
                            maxr2, rbfact, d_fsct, ring, d_score))
    return r;
  return launch_best_cand(*c, np, k, d_top_o, d_score, d_best);
}

// ------------------------------------------------------------------ PREFINE
int fr_powell_minimize(int n, float* x, const float* e, float escale, int icon, int maxit,
                       float (*fn)(int n, const float* x, void* user), void* user, float* f,
                       int* nevals) {
  FR_CHECK(x && e && fn && f, "null argument");
  FR_CHECK(n >= 1 && n <= 6, "VA04A works on 1..6 parameters (got %d)", n);
  FR_CHECK(maxit >= 1 && escale > 0.0f, "bad MAXIT or ESCALE");
  fr::Powell p;
  p.start(x, e, n, escale, icon, maxit);
  while (p.run()) p.supply(fn(n, p.x(), user));
  for (int i = 0; i < n; i++) x[i] = p.x()[i];
  *f = p.value();
  if (nevals) *nevals = p.evaluations();
  return 0;
}

int fr_refine_batch(fr_ctx* c, int nsam, int n, const float* d_a3df, const float* d_spec,
                    const float* d_ctff, const int* pidx, const float* amag, float* par,
                    const int* mask, int maxr1, int maxr2, float rbfact, const float* d_fsct,
                    const float* ring, const float* prior, int maxit, int* nevals) {
  return fr::refine_items(c, nsam, n, d_a3df, d_spec, d_ctff, pidx, amag, par, mask, maxr1, maxr2,
                          rbfact, d_fsct, ring, prior, maxit, nullptr, nevals);
}
}  // extern "C"

namespace fr {
int refine_items(fr_ctx* c, int nsam, int n, const float* d_a3df, const float* d_spec,
                 const float* d_ctff, const int* pidx, const float* amag, float* par,
                 const int* mask, int maxr1, int maxr2, float rbfact, const float* d_fsct,
                 const float* ring, const float* prior, int maxit, const float* sm,
                 int* nevals) {
  CTX_CHECK(c);
  SIZE_CHECK(nsam);
  FR_CHECK(d_a3df && d_spec && d_ctff && pidx && amag && par && mask && d_fsct && ring && prior,
           "null argument");
  FR_CHECK(n >= 0 && maxit >= 1, "bad sizes");
  FR_CHECK(maxr1 >= 0 && maxr2 > maxr1, "bad resolution band %d..%d", maxr1, maxr2);
  for (int i = 0; i < n; i++) {
    FR_CHECK(pidx[i] >= 0, "negative particle index");
    FR_CHECK(amag[i] > 0.0f, "AMAG must be positive (magnification refinement is not supported)");
    FR_CHECK(prior[5 * i + 2] != 0.0f && prior[5 * i + 3] != 0.0f, "shift prior SX/SY must be non-zero");
  }
  if (n == 0) return 0;
  fr::RefineArgs a;
  a.n = n;
  a.pidx = pidx;
  a.amag = amag;
  a.mask = mask;
  a.prior = prior;
  a.maxit = maxit;
  a.sm = sm;
  // one round's items go over in one pinned copy (PIDX | DM | SHIFT | AMAG,
  // 13 words per item) and its scores come back in one
  auto cc3m = [&](int m, const int* hp, const float* hdm, const float* hsh, const float* ham,
                  float* hscore) -> int {
    int err = 0;
    const size_t words = 13 * (size_t)m;
    char* dbuf = (char*)c->slot(11, sizeof(float) * words, &err);
    if (!dbuf) return err;
    float* dsc = (float*)c->slot(15, sizeof(float) * (size_t)m, &err);
    if (!dsc) return err;
    char* h = (char*)c->host_pinned(sizeof(float) * (words + m), &err);
    if (!h) return err;
    const size_t o_dm = sizeof(int) * m, o_sh = o_dm + sizeof(float) * 9 * m,
                 o_am = o_sh + sizeof(float) * 2 * m, o_sc = o_am + sizeof(float) * m;
    memcpy(h, hp, sizeof(int) * m);
    memcpy(h + o_dm, hdm, sizeof(float) * 9 * m);
    memcpy(h + o_sh, hsh, sizeof(float) * 2 * m);
    memcpy(h + o_am, ham, sizeof(float) * m);
    FR_HIP(hipMemcpyAsync(dbuf, h, o_sc, hipMemcpyHostToDevice, c->stream));
    if (int r = fr_cc3m_batch(c, nsam, m, d_a3df, d_spec, d_ctff, (const int*)dbuf,
                              (const float*)(dbuf + o_dm), (const float*)(dbuf + o_sh),
                              (const float*)(dbuf + o_am), maxr1, maxr2, rbfact, d_fsct, ring,
                              dsc))
      return r;
    FR_HIP(hipMemcpyAsync(h + o_sc, dsc, sizeof(float) * m, hipMemcpyDeviceToHost, c->stream));
    FR_HIP(hipStreamSynchronize(c->stream));
    memcpy(hscore, h + o_sc, sizeof(float) * m);
    return 0;
  };
  return fr::refine_batch(*c, a, par, nevals, cc3m);
}
}  // namespace fr

extern "C" {

// ------------------------------------------------------------------- SIGMA
int fr_sigma_batch(fr_ctx* c, int nsam, int n, const float* d_a3df, const float* d_datc,
                   const float* d_ctff, const int* pidx, const float* par, const float* amag,
                   int maxr1, float rclas, const float* dmask, float sigp, const float* signset,
                   const float* sm, float* out) {
  CTX_CHECK(c);
  SIZE_CHECK(nsam);
  FR_CHECK(d_a3df && d_datc && d_ctff && pidx && par && amag && dmask && out && n >= 0,
           "null argument");
  FR_CHECK(sigp != 0.0f, "SIGP must be non-zero");
  FR_CHECK(maxr1 >= 0, "bad MAXR1");
  if (n == 0) return 0;
  const int jc = nsam / 2 + 1;
  std::vector<SigmaItem> it(n);
  for (int k = 0; k < n; k++) {
    FR_CHECK(pidx[k] >= 0, "negative particle index");
    FR_CHECK(amag[k] != 0.0f, "AMAG must be non-zero");
    const float* p = par + 5 * (size_t)k;
    SigmaItem& s = it[k];
    s.pidx = pidx[k];
    rotation(p[0], p[1], p[2], 0.0f, sm, s.dm);  // pextract.f:30-51
    s.shx = p[3];
    s.shy = p[4];
    float dm2[9];
    rotation(-p[2], -p[1], -p[0], amag[k], sm, dm2);  // sigma.f:110-111
    {
      const float d1 = dmask[0] - (float)jc, d2 = dmask[1] - (float)jc, d3 = dmask[2] - (float)jc;
      volatile float t;  // keep the reference's FP32 rounding of each product
      float xm = 0.f, ym = 0.f;
      t = dm2[0] * d1; xm = t;
      t = dm2[3] * d2; xm = xm + t;
      t = dm2[6] * d3; xm = xm + t;
      s.xm = xm + 1.0f;
      t = dm2[1] * d1; ym = t;
      t = dm2[4] * d2; ym = ym + t;
      t = dm2[7] * d3; ym = ym + t;
      s.ym = ym + 1.0f;
    }
    const float r = dmask[3] / amag[k];
    s.ri2f = r * r;
    s.signset = signset ? signset[k] : 0.0f;
    const int mrclas = (int)((float)nsam * rclas * fabsf(amag[k]));
    s.maxr12 = maxr1 * maxr1;
    s.maxr22 = mrclas * mrclas;
    s.pad_ = 0;
  }
  int err = 0;
  SigmaItem* d_it = (SigmaItem*)c->slot(11, sizeof(SigmaItem) * (size_t)n, &err);
  if (!d_it) return err;
  const size_t img = (size_t)nsam * (nsam + 2);
  float* buf = (float*)c->slot(0, sizeof(float) * img * 2 * (size_t)n, &err);
  if (!buf) return err;
  float* d_out = (float*)c->slot(15, sizeof(float) * 3 * (size_t)n, &err);
  if (!d_out) return err;
  FR_HIP(hipMemcpyAsync(d_it, it.data(), sizeof(SigmaItem) * n, hipMemcpyHostToDevice, c->stream));
  if (int r = launch_sigma(*c, nsam, n, d_a3df, d_it, d_datc, d_ctff, sigp, buf, d_out)) return r;
  FR_HIP(hipMemcpyAsync(out, d_out, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, c->stream));
  FR_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

}  // extern "C"

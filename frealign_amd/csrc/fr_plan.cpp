// fr_plan.cpp -- persistent, graph-captured batches and the launch autotuner.
//
// fr_autotune_cc3m: times the CC3M launch variants that apply to a box size
// (the fused band-pruned kernel at several persistent grid sizes, and the
// FFT path) on the caller's data, keeps the fastest for that NSAM in the
// context and returns the measurements as JSON.
//
// fr_plan: one box size, one batch size.  The score-and-insert pipeline of a
// batch -- CTFAPPLY (fr_prep_batch), CC3M at the particles' current
// parameters (fr_cc3m_batch), the A3D3 flush into the half-maps
// (fr_insert_batch) -- is captured once as a HIP graph with its host inputs
// in pinned staging buffers, and replayed per batch.  The host work left per
// batch is what must stay on the host for parity: the DM matrices and the
// symmetry copies (glibc cosf/sinf), written into the staging buffers.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/frealign_hip.h"
#include "fr_common.h"
#include "fr_ctx.h"
#include "fr_host.h"
#include "fr_kernels.h"

using namespace fr;

#define CTX_CHECK(c) FR_CHECK((c) != nullptr, "null fr_ctx")

struct fr_plan {
  fr_ctx* c = nullptr;
  int nsam = 0, nb = 0, ncopy = 0, fused = 0, blocks = 0;
  // bound arguments
  const float* a3df = nullptr;
  const float* pssnr = nullptr;
  const float* fsct = nullptr;
  PrepGlobals g{};
  Cc3mArgs ca{};
  Cc3mTables tabs;  // fused CC3M tables (owned by the context)
  int nsym = 0;
  int jsym[8] = {};
  float symop[72] = {};
  float pbc = 0, boff = 0, psize = 0;
  float *num0 = nullptr, *ctf20 = nullptr, *num1 = nullptr, *ctf21 = nullptr;
  long long* nksum = nullptr;
  double* ttd = nullptr;
  // device buffers
  float *d_img = nullptr, *d_ppar = nullptr, *d_ctff = nullptr, *d_outc = nullptr,
        *d_datc = nullptr, *d_dm = nullptr, *d_shift = nullptr, *d_amag = nullptr,
        *d_score = nullptr, *d_work = nullptr;
  int* d_pidx = nullptr;
  InsertTask* d_tasks = nullptr;
  unsigned char* d_ins = nullptr;  // insertion scratch
  float* d_part = nullptr;         // the preparation's ring sums (own: plans run concurrently)
  // pinned staging
  float *h_ppar = nullptr, *h_dm = nullptr, *h_shift = nullptr, *h_amag = nullptr,
        *h_score = nullptr;
  InsertTask* h_tasks = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // the plan's own stream (capture is not allowed on the legacy default
  // stream) and the event that orders it after the caller's stream
  hipStream_t stream = nullptr;
  hipEvent_t ready = nullptr;
  int launches = 0;
};

namespace {

int enqueue(fr_plan* p) {
  fr_ctx* c = p->c;
  const int nsam = p->nsam, nb = p->nb;
  FR_HIP(hipMemcpyAsync(p->d_ppar, p->h_ppar, sizeof(float) * 8 * nb, hipMemcpyHostToDevice,
                        c->stream));
  if (int r = launch_prep(*c, nsam, nb, p->d_img, p->d_ppar, p->pssnr, p->g, p->d_ctff, p->d_outc,
                          p->d_datc, reinterpret_cast<float2*>(p->d_part)))
    return r;
  FR_HIP(hipMemcpyAsync(p->d_dm, p->h_dm, sizeof(float) * 9 * nb, hipMemcpyHostToDevice,
                        c->stream));
  FR_HIP(hipMemcpyAsync(p->d_shift, p->h_shift, sizeof(float) * 2 * nb, hipMemcpyHostToDevice,
                        c->stream));
  FR_HIP(hipMemcpyAsync(p->d_amag, p->h_amag, sizeof(float) * nb, hipMemcpyHostToDevice,
                        c->stream));
  Cc3mArgs a = p->ca;
  a.n = nb;
  if (p->fused) {
    if (int r = launch_cc3m_fused(*c, a, p->a3df, p->d_outc, p->d_ctff, p->d_pidx, p->d_dm,
                                  p->d_shift, p->d_amag, p->fsct, p->tabs, p->d_score, p->blocks))
      return r;
  } else {
    if (int r = launch_cc3m(*c, a, p->a3df, p->d_outc, p->d_ctff, p->d_pidx, p->d_dm, p->d_shift,
                            p->d_amag, p->fsct, p->d_work, p->d_score))
      return r;
  }
  FR_HIP(hipMemcpyAsync(p->h_score, p->d_score, sizeof(float) * nb, hipMemcpyDeviceToHost,
                        c->stream));
  const int ntask = nb * p->ncopy;
  FR_HIP(hipMemcpyAsync(p->d_tasks, p->h_tasks, sizeof(InsertTask) * ntask, hipMemcpyHostToDevice,
                        c->stream));
  return launch_insert(*c, nsam, nb, ntask, p->d_tasks, p->d_datc, p->d_ctff, p->num0, p->ctf20,
                       p->num1, p->ctf21, reinterpret_cast<unsigned long long*>(p->nksum), p->ttd,
                       p->d_ins);
}

// Host staging for one batch: the CC3M DM (cc3m.f:48-65) and the insertion
// tasks of pinsert.f:42-93 (a skipped particle keeps its slots with an empty
// radius so the graph's launch shape never changes).
int stage(fr_plan* p, const float* ppar, const float* recb, const int* half) {
  const int nsam = p->nsam, nb = p->nb, nsamh = nsam / 2;
  memcpy(p->h_ppar, ppar, sizeof(float) * 8 * nb);
  std::vector<float> copies(9 * (size_t)p->ncopy);
  for (int b = 0; b < nb; b++) {
    const float* r = recb + 10 * (size_t)b;
    FR_CHECK(half[b] == 0 || half[b] == 1, "half-map flag must be 0 or 1");
    FR_CHECK(r[5] != 0.0f && r[7] != 0.0f, "AMAG and SIG2N of particle %d must be non-zero", b);
    rotation(r[0], r[1], r[2], r[5], nullptr, p->h_dm + 9 * (size_t)b);
    p->h_shift[2 * b] = r[3];
    p->h_shift[2 * b + 1] = r[4];
    p->h_amag[b] = fabsf(r[5]);
    const float on = r[6] / r[7];
    const float pref = p->pbc / 100.0f;
    const int irrec = (int)((float)nsam * r[9]);
    FR_CHECK(irrec >= 0 && irrec <= nsam, "RREC of particle %d out of range", b);
    float tmp = (r[8] / pref - p->boff / p->pbc) / (p->psize * p->psize);
    if (tmp > 2.0f) tmp = 2.0f;
    tmp = 2.0f * tmp / ((float)nsamh * (float)nsamh);
    float tm[9];
    rotation(r[0], r[1], r[2], r[5], nullptr, tm);
    const int nc = symmetry_copies(tm, p->nsym, p->jsym, p->symop, copies.data());
    FR_CHECK(nc == p->ncopy, "symmetry enumeration mismatch");
    for (int k = 0; k < nc; k++) {
      InsertTask& t = p->h_tasks[(size_t)b * p->ncopy + k];
      memcpy(t.dm, copies.data() + 9 * k, sizeof t.dm);
      t.shx = r[3];
      t.shy = r[4];
      t.tmp = tmp;
      t.occ = on;
      t.part = b;
      t.half = half[b];
      t.irrec2 = on == 0.0f ? -1 : irrec * irrec;  // pinsert.f:40: ON = 0 inserts nothing
      t.ilo = (short)-nsamh;
      t.ihi = (short)nsamh;
    }
  }
  return 0;
}

template <typename T>
int dalloc(T** p, size_t n) {
  FR_HIP(hipMalloc((void**)p, sizeof(T) * (n ? n : 1)));
  return 0;
}
template <typename T>
int halloc(T** p, size_t n) {
  FR_HIP(hipHostMalloc((void**)p, sizeof(T) * (n ? n : 1), hipHostMallocDefault));
  return 0;
}

}  // namespace

extern "C" {

int fr_plan_create(fr_ctx* c, int nsam, int nb, const float* d_a3df, const float* d_pssnr,
                   const float* gpar, int ictf, int maxr1, int maxr2, float rbfact,
                   const float* d_fsct, const float* ring, int nsym, const int* jsym,
                   const float* symop, float pbc, float boff, float psize, float* d_num0,
                   float* d_ctf2_0, float* d_num1, float* d_ctf2_1, long long* d_nksum,
                   double* d_ttd, fr_plan** out) {
  CTX_CHECK(c);
  FR_CHECK(out && d_a3df && d_pssnr && gpar && d_fsct && ring && jsym && symop && d_num0 &&
               d_ctf2_0 && d_num1 && d_ctf2_1 && d_nksum && d_ttd,
           "null argument");
  FR_CHECK(nsam >= 8 && nsam % 2 == 0 && fft_supported(nsam), "unsupported NSAM %d", nsam);
  FR_CHECK(nb >= 1 && nb <= 65535, "batch size must be 1..65535");
  FR_CHECK(maxr1 >= 0 && maxr2 > maxr1, "bad resolution band");
  FR_CHECK(nsym >= 1 && nsym <= 8, "nsym %d out of range", nsym);
  FR_CHECK(pbc != 0.0f && psize != 0.0f && gpar[6] != 0.0f, "PBC, PSIZE, HALFW must be non-zero");
  fr_plan* p = new fr_plan;
  *out = nullptr;
  p->c = c;
  p->nsam = nsam;
  p->nb = nb;
  p->a3df = d_a3df;
  p->pssnr = d_pssnr;
  p->fsct = d_fsct;
  p->g = PrepGlobals{gpar[0], gpar[1], gpar[2], gpar[3], gpar[4], gpar[5], gpar[6],
                     gpar[7], gpar[8], gpar[9], gpar[10], ictf};
  p->ca = Cc3mArgs{nsam, nb, maxr1, maxr2, rbfact, ring[0], ring[1], ring[2], ring[3]};
  p->nsym = nsym;
  int ncopy = 1;
  for (int k = 0; k < nsym; k++) {
    FR_CHECK(jsym[k] >= 1 && jsym[k] <= 60, "bad rotational order %d", jsym[k]);
    p->jsym[k] = jsym[k];
    ncopy *= jsym[k];
  }
  memcpy(p->symop, symop, sizeof(float) * 9 * nsym);
  p->ncopy = ncopy;
  p->pbc = pbc;
  p->boff = boff;
  p->psize = psize;
  p->num0 = d_num0;
  p->ctf20 = d_ctf2_0;
  p->num1 = d_num1;
  p->ctf21 = d_ctf2_1;
  p->nksum = d_nksum;
  p->ttd = d_ttd;
  const fr_ctx::Tune tn = c->tune_for(nsam);
  p->fused = cc3m_fused_ok(nsam, maxr2) && tn.cc3m_fused;
  p->blocks = tn.cc3m_blocks;
  const size_t ic = (size_t)nsam * (nsam / 2 + 1), img = (size_t)nsam * (nsam + 2);
  int r = 0;
  auto fail = [&](int rc) {
    fr_plan_destroy(p);
    return rc;
  };
  if ((r = dalloc(&p->d_img, img * nb)) || (r = dalloc(&p->d_ppar, 8 * (size_t)nb)) ||
      (r = dalloc(&p->d_ctff, 4 * ic * nb)) || (r = dalloc(&p->d_outc, 2 * ic * nb)) ||
      (r = dalloc(&p->d_datc, 2 * ic * nb)) || (r = dalloc(&p->d_dm, 9 * (size_t)nb)) ||
      (r = dalloc(&p->d_shift, 2 * (size_t)nb)) || (r = dalloc(&p->d_amag, (size_t)nb)) ||
      (r = dalloc(&p->d_score, (size_t)nb)) || (r = dalloc(&p->d_pidx, (size_t)nb)) ||
      (r = dalloc(&p->d_tasks, (size_t)nb * ncopy)) || (r = halloc(&p->h_ppar, 8 * (size_t)nb)) ||
      (r = halloc(&p->h_dm, 9 * (size_t)nb)) || (r = halloc(&p->h_shift, 2 * (size_t)nb)) ||
      (r = halloc(&p->h_amag, (size_t)nb)) || (r = halloc(&p->h_score, (size_t)nb)) ||
      (r = halloc(&p->h_tasks, (size_t)nb * ncopy)))
    return fail(r);
  if (!p->fused && (r = dalloc(&p->d_work, img * nb))) return fail(r);
  if ((r = dalloc(&p->d_ins, insert_scratch_bytes(nsam, nb, nb * ncopy)))) return fail(r);
  if ((r = dalloc(&p->d_part, 2 * prep_partials_count(nsam, nb)))) return fail(r);
  if (p->fused && (r = c->cc3m_tables(p->ca, &p->tabs))) return fail(r);
  std::vector<int> pid(nb);
  for (int b = 0; b < nb; b++) pid[b] = b;
  if (hipMemcpy(p->d_pidx, pid.data(), sizeof(int) * nb, hipMemcpyHostToDevice) != hipSuccess)
    return fail(fr::fail("pidx upload failed"));
  // zero the staging so that the capture reads defined values
  memset(p->h_ppar, 0, sizeof(float) * 8 * nb);
  memset(p->h_dm, 0, sizeof(float) * 9 * nb);
  memset(p->h_shift, 0, sizeof(float) * 2 * nb);
  for (int b = 0; b < nb; b++) p->h_amag[b] = 1.0f;
  memset(p->h_tasks, 0, sizeof(InsertTask) * nb * ncopy);
  for (int k = 0; k < nb * ncopy; k++) p->h_tasks[k].irrec2 = -1;
  // the FFT twiddles of this length are created lazily: make them before capture
  if (!c->twiddles(nsam)) return fail(fr::fail("twiddle table allocation failed"));
  if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&p->ready, hipEventDisableTiming) != hipSuccess)
    return fail(fr::fail("plan stream/event creation failed"));
  const bool prof = c->prof_on;
  hipStream_t caller = c->stream;
  if (hipStreamSynchronize(caller) != hipSuccess)
    return fail(fr::fail("synchronisation before capture failed"));
  c->prof_on = false;
  c->stream = p->stream;  // the launchers enqueue on the context's stream
  if (hipStreamBeginCapture(p->stream, hipStreamCaptureModeRelaxed) != hipSuccess) {
    c->stream = caller;
    c->prof_on = prof;
    return fail(fr::fail("stream capture could not start"));
  }
  r = enqueue(p);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(p->stream, &graph);
  c->stream = caller;
  c->prof_on = prof;
  if (r) {
    if (graph) (void)hipGraphDestroy(graph);
    return fail(r);
  }
  if (e != hipSuccess || !graph) return fail(fr::fail("stream capture failed: %s", hipGetErrorString(e)));
  p->graph = graph;
  if (hipGraphInstantiate(&p->exec, graph, nullptr, nullptr, 0) != hipSuccess)
    return fail(fr::fail("graph instantiation failed"));
  *out = p;
  return 0;
}

float* fr_plan_images(fr_plan* p) { return p ? p->d_img : nullptr; }

int fr_plan_load(fr_plan* p, const float* d_images) {
  FR_CHECK(p && d_images, "null argument");
  const size_t bytes = sizeof(float) * (size_t)p->nsam * (p->nsam + 2) * p->nb;
  FR_HIP(hipMemcpyAsync(p->d_img, d_images, bytes, hipMemcpyDeviceToDevice, p->c->stream));
  return 0;
}

int fr_plan_run(fr_plan* p, const float* ppar, const float* recb, const int* half, float* score) {
  FR_CHECK(p && ppar && recb && half, "null argument");
  if (int r = stage(p, ppar, recb, half)) return r;
  fr_ctx* c = p->c;
  // the images were loaded on the caller's stream: order the graph after it
  FR_HIP(hipEventRecord(p->ready, c->stream));
  FR_HIP(hipStreamWaitEvent(p->stream, p->ready, 0));
  hipStream_t caller = c->stream;
  c->stream = p->stream;
  int pslot = c->prof_begin(kProfPlan);
  hipError_t e = hipGraphLaunch(p->exec, p->stream);
  c->prof_end(pslot);
  c->stream = caller;
  FR_HIP(e);
  FR_HIP(hipStreamSynchronize(p->stream));  // staging and scores are reused by the next batch
  if (score) memcpy(score, p->h_score, sizeof(float) * p->nb);
  p->launches++;
  return 0;
}

int fr_plan_destroy(fr_plan* p) {
  if (!p) return 0;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  void* dev[] = {p->d_img, p->d_ppar, p->d_ctff, p->d_outc, p->d_datc, p->d_dm, p->d_shift,
                 p->d_amag, p->d_score, p->d_work, p->d_pidx, p->d_tasks, p->d_ins,
                 p->d_part};
  for (void* d : dev)
    if (d) (void)hipFree(d);
  void* host[] = {p->h_ppar, p->h_dm, p->h_shift, p->h_amag, p->h_score, p->h_tasks};
  for (void* h : host)
    if (h) (void)hipHostFree(h);
  if (p->ready) (void)hipEventDestroy(p->ready);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
  return 0;
}

int fr_set_cc3m_variant(fr_ctx* c, int nsam, int fused, int workgroups) {
  CTX_CHECK(c);
  FR_CHECK(workgroups >= 0, "workgroups must be >= 0");
  c->tune[nsam] = fr_ctx::Tune{fused ? 1 : 0, workgroups};
  return 0;
}

int fr_autotune_cc3m(fr_ctx* c, int nsam, int n, const float* d_a3df, const float* d_spec,
                     const float* d_ctff, const int* d_pidx, const float* d_dm,
                     const float* d_shift, const float* d_amag, int maxr1, int maxr2,
                     float rbfact, const float* d_fsct, const float* ring, float* d_score,
                     char* log, int logcap) {
  CTX_CHECK(c);
  FR_CHECK(n >= 1 && d_a3df && d_spec && d_ctff && d_pidx && d_dm && d_shift && d_amag &&
               d_fsct && ring && d_score,
           "null argument");
  struct Cand {
    int fused, blocks;
    float ms;
  };
  std::vector<Cand> cands;
  if (cc3m_fused_ok(nsam, maxr2)) {
    int cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) cus = prop.multiProcessorCount;
    cands.push_back({1, 0, 0.f});
    for (int k : {1, 2, 4, 8})
      if (k * cus < n) cands.push_back({1, k * cus, 0.f});
  }
  cands.push_back({0, 0, 0.f});
  hipEvent_t a, b;
  FR_HIP(hipEventCreate(&a));
  FR_HIP(hipEventCreate(&b));
  const fr_ctx::Tune saved = c->tune_for(nsam);
  const bool prof = c->prof_on;
  c->prof_on = false;
  int rc = 0;
  for (auto& cd : cands) {
    c->tune[nsam] = fr_ctx::Tune{cd.fused, cd.blocks};
    // one untimed call (tables, scratch), then the best of three
    rc = fr_cc3m_batch(c, nsam, n, d_a3df, d_spec, d_ctff, d_pidx, d_dm, d_shift, d_amag, maxr1,
                       maxr2, rbfact, d_fsct, ring, d_score);
    if (rc) break;
    float best = 1e30f;
    for (int rep = 0; rep < 3 && !rc; rep++) {
      FR_HIP(hipEventRecord(a, c->stream));
      rc = fr_cc3m_batch(c, nsam, n, d_a3df, d_spec, d_ctff, d_pidx, d_dm, d_shift, d_amag, maxr1,
                         maxr2, rbfact, d_fsct, ring, d_score);
      FR_HIP(hipEventRecord(b, c->stream));
      FR_HIP(hipEventSynchronize(b));
      float ms = 0.f;
      FR_HIP(hipEventElapsedTime(&ms, a, b));
      best = ms < best ? ms : best;
    }
    cd.ms = best;
  }
  c->prof_on = prof;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc) {
    c->tune[nsam] = saved;
    return rc;
  }
  size_t bi = 0;
  for (size_t i = 1; i < cands.size(); i++)
    if (cands[i].ms < cands[bi].ms) bi = i;
  c->tune[nsam] = fr_ctx::Tune{cands[bi].fused, cands[bi].blocks};
  if (log && logcap > 0) {
    std::string js = "{\"kernel\": \"cc3m\", \"nsam\": " + std::to_string(nsam) +
                     ", \"items\": " + std::to_string(n) + ", \"candidates\": [";
    for (size_t i = 0; i < cands.size(); i++) {
      char buf[128];
      snprintf(buf, sizeof buf, "%s{\"path\": \"%s\", \"workgroups\": %d, \"ms\": %.4f}",
               i ? ", " : "", cands[i].fused ? "fused" : "fft",
               cands[i].fused ? (cands[i].blocks ? cands[i].blocks : n) : 0, cands[i].ms);
      js += buf;
    }
    char buf[128];
    snprintf(buf, sizeof buf, "], \"choice\": %zu}", bi);
    js += buf;
    snprintf(log, (size_t)logcap, "%s", js.c_str());
  }
  return 0;
}

}  // extern "C"

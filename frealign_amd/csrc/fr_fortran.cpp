// fr_fortran.cpp -- per-call entry points with the reference's own symbols and
// argument lists (Fortran 77 implicit interfaces: every argument by
// reference, CHARACTER lengths appended), so that a flang- or gfortran-built
// LMAIN links against this library in place of the reference's routines:
//
//   SUBROUTINE CTFAPPLY(...)       ctfapply.f:8-161  -> ctfapply_
//   REAL FUNCTION CC3M(...)        cc3m.f:8-169      -> cc3m_
//   SUBROUTINE PEXTRACT(...)       pextract.f:8-101  -> pextract_
//   SUBROUTINE PRESB(...)          presb.f:8-160     -> presb_
//   SUBROUTINE SIGMA(...)          sigma.f:8-221     -> sigma_
//   SUBROUTINE PREFINE(...)        prefine.f:8-162   -> prefine_
//   SUBROUTINE PINSERT_TRILIN(...) pinsert_trilin.f:8-244 -> pinsert_trilin_
//   SUBROUTINE A3D3(...)           a3d3.f:8-146      -> a3d3_
//   SUBROUTINE PSEARCH(...)        psearch.f:8-121   -> psearch_
//
// Each call copies its per-particle host arrays to the device, runs the
// batched kernels on one item and copies the result back; the batched entry
// points of include/frealign_hip.h are the throughput path.  Two arrays stay
// resident instead of crossing PCIe on every call:
//   * the reference transform (C3DF / A3DF of CC3M, PEXTRACT, PRESB, SIGMA,
//     PREFINE and PSEARCH), uploaded when a call names a host array other than
//     the resident one, or the same array with a different fingerprint (1,024
//     words sampled across it), or after fr_percall_invalidate;
//   * A3D3's half-maps (A3DF/S3DF, D3DF/V3DF, NVOL1/NVOL2 private volumes),
//     accumulated on the device between the flushes and copied back to the
//     host arrays at the flush of the last particle (K = ILAST), which is when
//     LMAIN's loop ends, or by fr_percall_flush.
// The reference calls these routines serially per process (a library-wide
// context on device 0 is used; calls are not thread-safe).  The supported
// cases are the reference's defaults on the hot path: IPAD = 1, IRAD > 0,
// IEWALD = 0, non-helical ASYM, CC3M with IBUF = 0.  Anything else, or a
// device error, is fatal as the reference's own STOPs are: a message on
// stdout and exit(1).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/frealign_hip.h"
#include "fr_common.h"
#include "fr_ctx.h"
#include "fr_host.h"

namespace {

struct Dev {
  void* p = nullptr;
  size_t cap = 0;
  void* get(size_t bytes) {
    if (bytes <= cap) return p;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    cap = bytes;
    return p;
  }
};

// A host array mirrored on the device.  `dirty`: the device copy holds
// accumulations the host array does not have yet.
struct Resident {
  const void* host = nullptr;
  size_t bytes = 0;
  unsigned long long fp = 0;
  bool dirty = false;
  Dev dev;
};

struct State {
  fr_ctx* ctx = nullptr;
  Dev spec, ctff, fsct, dm, shift, amag, pidx, out;
  Dev vol, vol2, outc, datc, ppar, pssnr, task, sums, work;
  Dev stat[4];                    // KSUM, VSUM, PSUM, ASUM of PINSERT_TRILIN(_SE)
  Resident ref;                   // the reference transform
  Resident maps[4];               // A3DF, S3DF, D3DF, V3DF of A3D3
  Resident stats[4];              // KSUM, VSUM, PSUM, ASUM of A3D3 (NNSTAT = 1)
};

State& st() {
  static State s;
  return s;
}

[[noreturn]] void die(const char* routine, const char* what) {
  printf(" %s (frealign_hip): %s\n", routine, what);
  fflush(stdout);
  exit(1);
}

void need(bool ok, const char* routine, const char* what) {
  if (!ok) die(routine, what);
}

void check(int status, const char* routine) {
  if (status != 0) die(routine, fr_last_error());
}

fr_ctx* ctx(const char* routine) {
  State& s = st();
  if (!s.ctx) check(fr_create(0, &s.ctx), routine);
  return s.ctx;
}

void* upload(Dev& d, const void* host, size_t bytes, const char* routine) {
  void* p = d.get(bytes);
  need(p != nullptr, routine, "device allocation failed");
  need(hipMemcpy(p, host, bytes, hipMemcpyHostToDevice) == hipSuccess, routine,
       "host-to-device copy failed");
  return p;
}

void download(void* host, const void* dev, size_t bytes, const char* routine) {
  need(hipMemcpy(host, dev, bytes, hipMemcpyDeviceToHost) == hipSuccess, routine,
       "device-to-host copy failed");
}

// FNV-1a over 1,024 64-bit words spread evenly across the array, plus its last
// word: enough to notice a host array that was rewritten or reloaded
unsigned long long fingerprint(const void* host, size_t bytes) {
  const unsigned char* b = static_cast<const unsigned char*>(host);
  unsigned long long h = 1469598103934665603ull ^ bytes;
  const size_t nw = bytes / 8;
  auto mix = [&](size_t off) {
    unsigned long long w = 0;
    memcpy(&w, b + off, 8);
    h = (h ^ w) * 1099511628211ull;
  };
  if (nw > 0) {
    const size_t step = nw > 1024 ? nw / 1024 : 1;
    for (size_t k = 0; k < nw; k += step) mix(8 * k);
    mix(8 * (nw - 1));
  }
  return h;
}

// the device copy of a read-only host array, uploaded only when it changed
void* resident(Resident& r, const void* host, size_t bytes, const char* routine) {
  const unsigned long long fp = fingerprint(host, bytes);
  if (r.host == host && r.bytes == bytes && r.fp == fp && r.dev.p) return r.dev.p;
  void* p = upload(r.dev, host, bytes, routine);
  r.host = host;
  r.bytes = bytes;
  r.fp = fp;
  r.dirty = false;
  return p;
}

// copies an accumulated device map back into its host array
void write_back(Resident& r, const char* routine) {
  if (!r.dirty || !r.host) return;
  download(const_cast<void*>(r.host), r.dev.p, r.bytes, routine);
  r.fp = fingerprint(r.host, r.bytes);
  r.dirty = false;
}

void common_checks(const char* routine, int nsam, int irad, int ipad, int iewald) {
  need(nsam >= 8 && nsam % 2 == 0, routine, "NSAM must be even and >= 8");
  need(ipad == 1, routine, "only IPAD = 1 is supported");
  need(irad != 0, routine, "only trilinear interpolation (IRAD > 0) is supported");
  need(iewald == 0, routine, "Ewald-sphere extraction (IEWALD != 0) is not supported");
}

size_t half_plane(int nsam) { return (size_t)nsam * (nsam / 2 + 1); }
size_t half_volume(int nsam) { return (size_t)nsam * nsam * (nsam / 2 + 1); }

}  // namespace

extern "C" {

// cc3m.f:8-169
float cc3m_(const int* nsam, const int* irad, const float* amag, const float* spec,
            const float* a3df, const int* maxr1, const int* maxr2, const float* phi,
            const float* theta, const float* psi, const float* shx, const float* shy,
            const int* ibuf, float* /*pbuc*/, const float* rbfact, const float* /*sinclut*/,
            const int* ipad, float* /*rbuf*/, const float* fsct, const int* iewald,
            const float* /*thetatr*/, const float* ctff, const float* ri2, const float* ri3,
            const float* rih, const float* halfw, const float* /*pweights*/,
            const void* /*fftw_plans*/, const float* sm) {
  const char* R = "CC3M";
  const int n = *nsam;
  common_checks(R, n, *irad, *ipad, *iewald);
  need(*ibuf == 0, R, "only IBUF = 0 is supported");
  fr_ctx* c = ctx(R);
  State& s = st();
  const size_t ic = half_plane(n);
  void* dv = resident(s.ref, a3df, sizeof(float) * 2 * half_volume(n), R);
  void* ds = upload(s.spec, spec, sizeof(float) * 2 * ic, R);
  void* dc = upload(s.ctff, ctff, sizeof(float) * 4 * ic, R);
  void* df = upload(s.fsct, fsct, sizeof(float) * (n / 2 + 1), R);
  float dmh[9];
  fr::rotation(*phi, *theta, *psi, *amag, sm, dmh);  // cc3m.f:48-65
  void* dd = upload(s.dm, dmh, sizeof dmh, R);
  const float sh[2] = {*shx, *shy};
  void* dsh = upload(s.shift, sh, sizeof sh, R);
  const float am = fabsf(*amag);
  void* da = upload(s.amag, &am, sizeof am, R);
  const int p0 = 0;
  void* dp = upload(s.pidx, &p0, sizeof p0, R);
  void* dout = s.out.get(sizeof(float) * 8);
  need(dout != nullptr, R, "device allocation failed");
  const float ring[4] = {*ri2, *ri3, *rih, *halfw};
  check(fr_cc3m_batch(c, n, 1, (const float*)dv, (const float*)ds, (const float*)dc,
                      (const int*)dp, (const float*)dd, (const float*)dsh, (const float*)da,
                      *maxr1, *maxr2, *rbfact, (const float*)df, ring, (float*)dout),
        R);
  check(fr_synchronize(c), R);
  float score = 0.f;
  download(&score, dout, sizeof score, R);
  return score;
}

// pextract.f:8-101
void pextract_(const int* nsam, const int* irad, float* outc, const float* a3df, const float* phi,
               const float* theta, const float* psi, const float* /*sinclut*/, const int* ipad,
               const float* /*thet*/, const int* iewald, const float* ctff, const float* sm) {
  const char* R = "PEXTRACT";
  const int n = *nsam;
  common_checks(R, n, *irad, *ipad, *iewald);
  fr_ctx* c = ctx(R);
  State& s = st();
  const size_t ic = half_plane(n);
  void* dv = resident(s.ref, a3df, sizeof(float) * 2 * half_volume(n), R);
  void* dc = upload(s.ctff, ctff, sizeof(float) * 4 * ic, R);
  float dmh[9];
  fr::rotation(*phi, *theta, *psi, 0.0f, sm, dmh);  // pextract.f:30-51
  void* dd = upload(s.dm, dmh, sizeof dmh, R);
  void* dout = s.spec.get(sizeof(float) * 2 * ic);
  need(dout != nullptr, R, "device allocation failed");
  check(fr_pextract_batch(c, n, 1, (const float*)dv, (const float*)dd, nullptr, (const float*)dc,
                          (float*)dout),
        R);
  check(fr_synchronize(c), R);
  download(outc, dout, sizeof(float) * 2 * ic, R);
}

// presb.f:8-160
void presb_(const int* nsam, const int* irad, const float* amag, const float* spec,
            const float* a3df, const int* /*maxr1*/, const int* /*maxr2*/, const float* phi,
            const float* theta, const float* psi, const float* shx, const float* shy, float* rbin,
            const float* /*sinclut*/, const int* ipad, int* /*ns*/, float* /*bufa*/,
            float* /*bufb*/, float* /*bufab*/, const float* /*cs*/, const float* /*wl*/,
            const float* /*wgh1*/, const float* /*wgh2*/, const float* /*df1*/,
            const float* /*df2*/, const float* /*angast*/, const float* /*thet*/,
            const int* iewald, const float* ctff, const float* ri2, const float* ri3,
            const float* rih, const float* halfw, const char* asym, const void* /*fftw_plans*/,
            const float* sm, size_t /*asym_len*/) {
  const char* R = "PRESB";
  const int n = *nsam;
  common_checks(R, n, *irad, *ipad, *iewald);
  need(asym[0] != 'H', R, "helical symmetry is not supported");
  fr_ctx* c = ctx(R);
  State& s = st();
  const size_t ic = half_plane(n);
  void* dv = resident(s.ref, a3df, sizeof(float) * 2 * half_volume(n), R);
  void* ds = upload(s.spec, spec, sizeof(float) * 2 * ic, R);
  void* dc = upload(s.ctff, ctff, sizeof(float) * 4 * ic, R);
  float dmh[9];
  fr::rotation(*phi, *theta, *psi, *amag, sm, dmh);  // presb.f:47-62
  void* dd = upload(s.dm, dmh, sizeof dmh, R);
  const float sh[2] = {*shx, *shy};
  void* dsh = upload(s.shift, sh, sizeof sh, R);
  const float am = fabsf(*amag);
  void* da = upload(s.amag, &am, sizeof am, R);
  const int p0 = 0;
  void* dp = upload(s.pidx, &p0, sizeof p0, R);
  void* dr = upload(s.out, rbin, sizeof(float) * n, R);  // in/out: untouched shells keep values
  const float ring[4] = {*ri2, *ri3, *rih, *halfw};
  check(fr_presb_batch(c, n, 1, (const float*)dv, (const float*)ds, (const float*)dc,
                       (const int*)dp, (const float*)dd, (const float*)dsh, (const float*)da, ring,
                       (float*)dr),
        R);
  check(fr_synchronize(c), R);
  download(rbin, dr, sizeof(float) * n, R);
}

// sigma.f:8-221
void sigma_(const int* nsam, const int* irad, const float* amag, const float* a3df,
            const float* datc, const float* phi, const float* theta, const float* psi,
            const float* shx, const float* shy, const float* /*sinclut*/, const int* ipad,
            const int* iewald, const float* /*thetatr*/, const float* ctff, float* /*rbuf*/,
            const float* /*ri2*/, const int* maxr1, const int* /*maxr2*/, const float* signset,
            float* sigma2n, float* sigma2, float* logpc, const float* dmask, const char* asym,
            const int* /*k*/, const float* sigp, const float* /*pweights*/, const float* rclas,
            const void* /*fftw_plans*/, const float* sm, size_t /*asym_len*/) {
  const char* R = "SIGMA";
  const int n = *nsam;
  common_checks(R, n, *irad, *ipad, *iewald);
  need(asym[0] != 'H', R, "helical symmetry is not supported");
  fr_ctx* c = ctx(R);
  State& s = st();
  const size_t ic = half_plane(n);
  void* dv = resident(s.ref, a3df, sizeof(float) * 2 * half_volume(n), R);
  void* ds = upload(s.spec, datc, sizeof(float) * 2 * ic, R);
  void* dc = upload(s.ctff, ctff, sizeof(float) * 4 * ic, R);
  const int p0 = 0;
  const float par[5] = {*phi, *theta, *psi, *shx, *shy};
  float out[3];
  check(fr_sigma_batch(c, n, 1, (const float*)dv, (const float*)ds, (const float*)dc, &p0, par,
                       amag, *maxr1, *rclas, dmask, *sigp, signset, sm, out),
        R);
  *sigma2n = out[0];
  *sigma2 = out[1];
  *logpc = out[2];
}

// prefine.f:8-162 (IVFLAG = 0, AMAG > 0, non-helical): VA04A over the
// MASK-selected PHI, THETA, PSI, DSHX, DSHY minimising CALCFX (-CC3M - FSH and
// the range penalties, calcfx.f:152-191), the CC3M evaluations on the device.
// Returns the refined values, PRESA (the final score) and NPAR.
void prefine_(const int* nsam, const int* irad, float* amag, const int* maxr1, const int* maxr2,
              float* phi, float* theta, float* psi, float* dshx, float* dshy, float* presa,
              const int* mask, const float* rbfact, int* npar, const float* c3df,
              float* /*pbuf*/, const float* shx, const float* shy, const int* ipad,
              const float* /*cs*/, const float* /*wl*/, const float* /*wgh1*/,
              const float* /*wgh2*/, const float* /*thetatr*/, const float* ctff,
              const float* /*amagp*/, const float* rih, const float* halfw, const float* ri2,
              const float* ri3, const float* /*ri4*/, const float* /*rmax1*/,
              const float* /*rmax2*/, const float* /*xstd*/, float* /*mbuf*/,
              const int* /*ilst*/, const float* /*datc*/, const int* /*ibuf*/,
              const float* /*b3dv*/, float* /*datd*/, const float* /*sinclut*/,
              const int* ivflag, float* /*outd*/, const float* outc, float* /*qbuc*/,
              float* /*rbuf*/, const float* fsct, const int* iewald, const float* xm,
              const float* ym, const float* sx, const float* sy, const float* sig2,
              const char* asym, const float* /*thetam*/, const float* /*stheta*/,
              const float* /*psim*/, const float* /*spsi*/, const float* /*stif*/,
              const float* /*pweights*/, const void* /*fftw_plans*/, const float* sm,
              size_t /*asym_len*/) {
  const char* R = "PREFINE";
  const int n = *nsam;
  common_checks(R, n, *irad, *ipad, *iewald);
  need(n <= 1024, R, "NSAM above 1024 is not supported");
  need(asym[0] != 'H', R, "helical symmetry is not supported");
  need(*ivflag == 0, R, "only IVFLAG = 0 is supported");
  need(*amag > 0.0f, R, "magnification refinement (AMAG < 0) is not supported");
  need(*sx != 0.0f && *sy != 0.0f, R, "the shift prior SX, SY must be non-zero");
  fr_ctx* c = ctx(R);
  State& s = st();
  const size_t ic = half_plane(n);
  void* dv = resident(s.ref, c3df, sizeof(float) * 2 * half_volume(n), R);
  void* ds = upload(s.spec, outc, sizeof(float) * 2 * ic, R);
  void* dc = upload(s.ctff, ctff, sizeof(float) * 4 * ic, R);
  void* df = upload(s.fsct, fsct, sizeof(float) * (n / 2 + 1), R);
  const int p0 = 0;
  const float am = *amag;
  float par[6] = {*phi, *theta, *psi, *dshx, *dshy, 0.f};
  const float ring[4] = {*ri2, *ri3, *rih, *halfw};
  // FSH(SIG2, DSHX, DSHY, XM-SHX, YM-SHY, SX, SY) (prefine.f:74-75)
  const float prior[5] = {*xm - *shx, *ym - *shy, *sx, *sy, *sig2};
  check(fr::refine_items(c, n, 1, (const float*)dv, (const float*)ds, (const float*)dc, &p0, &am,
                         par, mask, *maxr1, *maxr2, *rbfact, (const float*)df, ring, prior, 50, sm,
                         nullptr),
        R);
  int np = 0;
  for (int j = 0; j < 5; j++) np += mask[j] == 1 ? 1 : 0;
  *npar = np;
  *phi = par[0];
  *theta = par[1];
  *psi = par[2];
  *dshx = par[3];
  *dshy = par[4];
  *presa = par[5];
}

// ctfapply.f:8-161 (MASKCOS, ASYM non-helical, XSTD >= 0).  OUTD and OUTC
// are one array (the driver passes OUTD for both, frealign_v9.f:679), the
// padded real image and its half-plane transform; the result is written
// through OUTD.  CTFF is recomputed on every call: the reference skips the
// CTF evaluation when DFMID1, DFMID2 and ANGAST equal the previous call's
// (ctfapply.f:54-58) and then reads the CTFF the caller passes back, which
// holds the same values.
void ctfapply_(const int* nsam, const float* spec, const float* shx, const float* shy,
               const float* cs, const float* wl, const float* wgh1, const float* wgh2,
               const float* dfmid1, const float* dfmid2, const float* angast,
               const float* thetatr, float* ctff, float* outd, float* /*outc*/,
               const float* amagp, const float* rih, const float* halfw, const float* ri2,
               const float* ri3, const float* /*ri4*/, float* /*datd*/, float* datc,
               const float* /*b3dv*/, const float* /*phi*/, const float* /*theta*/,
               const float* /*psi*/, float* /*w*/, const float* xstd, const int* /*ibuf*/,
               const float* tx, const float* ty, const char* asym, const float* pssnr,
               float* /*rbuf*/, float* /*pweights*/, const int* ictf,
               const void* /*fftw_plans*/, size_t /*asym_len*/) {
  const char* R = "CTFAPPLY";
  const int n = *nsam;
  need(n >= 8 && n % 2 == 0 && n <= 1024, R, "NSAM must be even, 8 .. 1024");
  need(asym[0] != 'H', R, "helical symmetry is not supported");
  need(*xstd >= 0.0f, R, "the MASKENV envelope (XSTD < 0) is not supported");
  need(*amagp != 0.0f, R, "AMAGP must be non-zero");
  need(*halfw != 0.0f, R, "HALFW must be non-zero");
  fr_ctx* c = ctx(R);
  State& s = st();
  const int jc = n / 2 + 1;
  const size_t ic = half_plane(n);
  void* ds = upload(s.spec, spec, sizeof(float) * 2 * ic, R);
  // PWEIGHT reads PSSNR(1..JC+1) (pweight.f)
  void* dps = upload(s.pssnr, pssnr, sizeof(float) * (jc + 1), R);
  const float pp[8] = {*shx, *shy, *dfmid1, *dfmid2, *angast, *amagp, 0.f, 0.f};
  void* dpp = upload(s.ppar, pp, sizeof pp, R);
  void* dc = s.ctff.get(sizeof(float) * 4 * ic);
  void* doc = s.outc.get(sizeof(float) * 2 * ic);
  void* ddc = s.datc.get(sizeof(float) * 2 * ic);
  need(dc && doc && ddc, R, "device allocation failed");
  const fr::PrepGlobals g{*cs, *wl, *wgh1, *wgh2, *thetatr, *rih, *halfw, *ri2, *ri3, *tx, *ty,
                          *ictf};
  check(fr::launch_ctfapply(*c, n, 1, (const float*)ds, (const float*)dpp, (const float*)dps, g,
                            (float*)dc, (float*)doc, (float*)ddc),
        R);
  check(fr_synchronize(c), R);
  download(ctff, dc, sizeof(float) * 4 * ic, R);
  download(outd, doc, sizeof(float) * 2 * ic, R);
  download(datc, ddc, sizeof(float) * 2 * ic, R);
}

}  // extern "C"

namespace {
// pinsert_trilin.f:8-244 and pinsert_trilin_se.f:8-359 (IRAD = 2 trilinear
// weights, IEWALD = 0, IPAD = 1): the stripe IP of width ST,
// I = -JC+1+NINT((IP-1)*ST) .. -JC+NINT(IP*ST) (pinsert_trilin.f:44-45),
// added into B3DF, T3DF, NKSUM and TTD, and with NNSTAT != 0 into the
// per-voxel KSUM, VSUM, PSUM, ASUM against C3DF (pinsert_trilin_se.f:186-194).
void insert_stripe(const char* R, const int* nsam, const int* irrec2, const float* spec,
                   const float* ctff, const float* occ, float* b3df, float* t3df, double* ttd,
                   const float* shx, const float* shy, float* asum, float* vsum, float* psum,
                   const float* c3df, int* ksum, const int* ipad, const int* iewald,
                   const int* nnstat, long long* nksum, const int* jc, const float* dm,
                   const float* tmp, const int* nsamh, const int* ic, const float* dx,
                   const float* dy, const float* stw, const int* ip) {
  const int n = *nsam;
  need(n >= 8 && n % 2 == 0 && n <= 1024, R, "NSAM must be even, 8 .. 1024");
  need(*ipad == 1, R, "only IPAD = 1 is supported");
  need(*iewald == 0, R, "Ewald-sphere insertion (IEWALD != 0) is not supported");
  const int j = n / 2 + 1;
  need(*jc == j && *nsamh == n / 2 && *ic == n * j, R, "JC, NSAMH or IC inconsistent with NSAM");
  // pinsert_trilin.f:44-45 (NINT: nearest, halves away from zero)
  int i1 = -j + 1 + (int)lroundf((float)(*ip - 1) * *stw);
  int i3 = -j + (int)lroundf((float)*ip * *stw);
  if (i1 < -(j - 1)) i1 = -(j - 1);
  if (i3 > j - 1) i3 = j - 1;
  if (i1 > i3 || *irrec2 < 0) return;  // an empty stripe adds nothing
  fr_ctx* c = ctx(R);
  State& s = st();
  const size_t hv = half_volume(n);
  void* ds = upload(s.spec, spec, sizeof(float) * 2 * half_plane(n), R);
  void* dc = upload(s.ctff, ctff, sizeof(float) * 4 * half_plane(n), R);
  void* db = upload(s.vol, b3df, sizeof(float) * 2 * hv, R);
  void* dt = upload(s.vol2, t3df, sizeof(float) * hv, R);
  fr::InsertTask t;
  memcpy(t.dm, dm, sizeof t.dm);  // DM(1..6)
  t.shx = *shx + *dx;             // pinsert_trilin.f:55,63: (SHX+DX)*I + (SHY+DY)*J
  t.shy = *shy + *dy;
  t.tmp = *tmp;
  t.occ = *occ;
  t.part = 0;
  t.half = 0;
  t.irrec2 = *irrec2;
  t.ilo = (short)i1;
  t.ihi = (short)i3;
  void* dtask = upload(s.task, &t, sizeof t, R);
  // NKSUM and TTD of this call (two halves' worth, only the first is used)
  void* dsum = s.sums.get(2 * sizeof(unsigned long long) + 2 * sizeof(double));
  need(dsum != nullptr, R, "device allocation failed");
  need(hipMemset(dsum, 0, 2 * sizeof(unsigned long long) + 2 * sizeof(double)) == hipSuccess, R,
       "device memset failed");
  unsigned long long* dk = (unsigned long long*)dsum;
  double* dd = (double*)(dk + 2);
  void* work = s.work.get(fr::insert_scratch_bytes(n, 1, 1));
  need(work != nullptr, R, "device allocation failed");
  fr::InsertStats stats{};
  void* sd[4] = {nullptr, nullptr, nullptr, nullptr};
  void* sh[4] = {ksum, vsum, psum, asum};
  if (*nnstat != 0) {
    need(c3df && ksum && vsum && psum && asum, R, "NNSTAT arrays missing");
    stats.ref = (const float*)resident(s.ref, c3df, sizeof(float) * 2 * hv, R);
    for (int q = 0; q < 4; q++) sd[q] = upload(s.stat[q], sh[q], sizeof(float) * hv, R);
    stats.ksum = (int*)sd[0];
    stats.vsum = (float*)sd[1];
    stats.psum = (float*)sd[2];
    stats.asum = (float*)sd[3];
  }
  // the second half-map is never addressed (t.half = 0): B3DF/T3DF stand in
  check(fr::launch_insert(*c, n, 1, 1, (const fr::InsertTask*)dtask, (const float*)ds,
                          (const float*)dc, (float*)db, (float*)dt, (float*)db, (float*)dt, dk,
                          dd, work, *nnstat != 0 ? &stats : nullptr),
        R);
  check(fr_synchronize(c), R);
  download(b3df, db, sizeof(float) * 2 * hv, R);
  download(t3df, dt, sizeof(float) * hv, R);
  if (*nnstat != 0)
    for (int q = 0; q < 4; q++) download(sh[q], sd[q], sizeof(float) * hv, R);
  unsigned long long k = 0;
  double td = 0.0;
  download(&k, dk, sizeof k, R);
  download(&td, dd, sizeof td, R);
  *nksum += (long long)k;
  *ttd += td;
}
}  // namespace

extern "C" {

// pinsert_trilin.f:8-244 (IEWALD = 0, IPAD = 1; the ASUM..KSUM statistics
// with NNSTAT != 0, as PINSERT only asks of the _SE variant)
void pinsert_trilin_(const int* nsam, const float* /*rii*/, const int* irrec2, const float* spec,
                     const float* ctff, const float* occ, float* b3df, float* t3df, double* ttd,
                     const float* shx, const float* shy, float* asum, float* vsum, float* psum,
                     float* c3df, int* ksum, const float* /*sinclut*/, const int* /*irada*/,
                     const int* ipad, const float* /*thet*/, const int* iewald,
                     const int* nnstat, long long* nksum, const int* jc, const float* dm,
                     const float* tmp, const int* nsamh, const int* ic, const float* dx,
                     const float* dy, const float* stw /* ST */, const int* ip, int* /*ns1*/) {
  insert_stripe("PINSERT_TRILIN", nsam, irrec2, spec, ctff, occ, b3df, t3df, ttd, shx, shy, asum,
                vsum, psum, c3df, ksum, ipad, iewald, nnstat, nksum, jc, dm, tmp, nsamh, ic, dx,
                dy, stw, ip);
}

// pinsert_trilin_se.f:8-359 (IEWALD = 0, IPAD = 1): the statistics variant,
// same arguments
void pinsert_trilin_se_(const int* nsam, const float* /*rii*/, const int* irrec2,
                        const float* spec, const float* ctff, const float* occ, float* b3df,
                        float* t3df, double* ttd, const float* shx, const float* shy, float* asum,
                        float* vsum, float* psum, float* c3df, int* ksum,
                        const float* /*sinclut*/, const int* /*irada*/, const int* ipad,
                        const float* /*thet*/, const int* iewald, const int* nnstat,
                        long long* nksum, const int* jc, const float* dm, const float* tmp,
                        const int* nsamh, const int* ic, const float* dx, const float* dy,
                        const float* stw, const int* ip, int* /*ns1*/) {
  insert_stripe("PINSERT_TRILIN_SE", nsam, irrec2, spec, ctff, occ, b3df, t3df, ttd, shx, shy,
                asum, vsum, psum, c3df, ksum, ipad, iewald, nnstat, nksum, jc, dm, tmp, nsamh, ic,
                dx, dy, stw, ip);
}

// a3d3.f:8-146 (INTERP = 1, IEWALD = 0, NNSTAT = 0, NONSYM false): the
// particle is recorded in RECBF/CTFBF/RECB/ISW as a3d3.f:39-85 does (ISWITCH
// alternating the half-maps); when the buffer holds NVOL particles, or at the
// last particle, slot I is inserted (a3d3.f:107-137 -> A3D3_S -> PINSERT ->
// PINSERT_TRILIN) into private volume IS of A3DF/S3DF (ISW 0) or D3DF/V3DF
// (ISW 1), IS = I, or (I-1)/2+1 when IFSC = 0 (a3d3_s.f:39-40), adding its
// voxel count to NKSUM1(I) and its CTF^2 sum to TTD1(IS) or TTD2(IS).  The
// slots go in order, one after another, where the reference runs them in an
// OpenMP loop.  PBC is declared INTEGER in a3d3.f and a3d3_s.f, which only
// pass it on; PINSERT reads the same storage as REAL (pinsert.f:31,43), as
// the driver's PBC is, so it is taken as REAL here.
void a3d3_(const int* nset, int* iswitch, const int* nsam, const int* irad, const float* /*ri*/,
           const float* rrec, const int* k, const int* ilast, const float* amagp,
           const float* datc, const float* ctff, float* asum, float* vsum, float* psum,
           float* c3df, int* ksum, const float* phi, const float* theta,
           const float* psi, const float* dshx, const float* dshy, const float* presa,
           const float* pbc, const float* boff, float* a3df, float* s3df, double* /*std*/,
           float* d3df, float* v3df, double* /*vtd*/, const int* nn1, const int* /*maxset*/,
           const float* /*sinclut*/, const int* /*irada*/, const int* ipad,
           const float* /*thetatr*/, const int* iewald, const int* nsym, const int* /*isymax*/,
           const int* jsym, const float* symop, const int* nnstat, long long* /*nksum*/,
           const int* ifsc, const int* /*imp*/, const int* nvol, const int* nvol1,
           const int* nvol2, const float* occ, const float* sig2n, const float* psize,
           const char* asym, const float* /*alpha*/, const float* /*rise*/, const int* /*nu*/,
           const int* /*hstart*/, long long* nksum1, double* ttd1, double* ttd2, int* /*ns*/,
           int* /*ns1*/, const int* interp, float* recbf, float* recb, int* isw, int* ibuf,
           float* ctfbf, const int* /*nn2*/, const int* nn3, const float* /*sm*/,
           const int* nonsym, const int* ism, size_t /*asym_len*/) {
  const char* R = "A3D3";
  const int n = *nsam;
  need(n >= 8 && n % 2 == 0 && n <= 1024 && n <= *nn1, R, "NSAM must be even, 8 .. min(NN1, 1024)");
  need(*nvol >= 1 && *nvol1 >= 1 && *nvol2 >= 1, R, "NVOL, NVOL1 and NVOL2 must be >= 1");
  need(*irad != 0 && *interp == 1, R, "only trilinear insertion (INTERP = 1) is supported");
  need(*ipad == 1 && *iewald == 0, R, "only IPAD = 1 and IEWALD = 0 are supported");
  need(asym[0] != 'H', R, "helical symmetry is not supported");
  need(*nonsym == 0, R, "the NONSYM insertion is not supported");
  const size_t ic = half_plane(n);
  const size_t hv = half_volume(n);
  const size_t nvox = (size_t)*nn3;  // A3DF(NN3, NVOL1): one private volume per column
  need(nvox >= hv, R, "NN3 is smaller than the NSAM half-volume");
  const size_t slot_c = (size_t)(*nn1 / 2) * *nn1;  // RECBF(NN1/2*NN1, NVOL), complex
  const size_t slot_f = (size_t)*nn1 * *nn1;        // CTFBF(NN1*NN1, NVOL), complex
  auto record = [&](int which) {                    // a3d3.f:41-56, 63-78
    const int b = (*ibuf)++;
    isw[b] = which;
    memcpy(recbf + 2 * slot_c * b, datc, sizeof(float) * 2 * ic);
    memcpy(ctfbf + 2 * slot_f * b, ctff, sizeof(float) * 2 * (size_t)n * (n + 2));
    float* r = recb + 10 * (size_t)b;
    r[0] = *phi;
    r[1] = *theta;
    r[2] = *psi;
    r[3] = *dshx;
    r[4] = *dshy;
    r[5] = *amagp;
    r[6] = *occ;
    r[7] = *sig2n;
    r[8] = *presa;
    r[9] = rrec[*nset - 1];
  };
  if (*ibuf < *nvol) {
    if (*iswitch == 0) {
      if (*occ != 0.0f) record(0);
      if (*ifsc <= 2 && *ism == 1) *iswitch = 1;
    } else {
      if (*ifsc == 0 && *occ != 0.0f) record(1);
      if (*ism == 1) *iswitch = 0;
    }
  }
  if (!(*ibuf == *nvol || *k == *ilast)) return;
  if (*ibuf >= 1) {
    fr_ctx* c = ctx(R);
    State& s = st();
    // the private volumes stay on the device between flushes
    float* host[4] = {a3df, s3df, d3df, v3df};
    const size_t bytes[4] = {sizeof(float) * 2 * nvox * *nvol1, sizeof(float) * nvox * *nvol1,
                             sizeof(float) * 2 * nvox * *nvol2, sizeof(float) * nvox * *nvol2};
    float* dev[4];
    for (int q = 0; q < 4; q++) {
      Resident& r = s.maps[q];
      if (r.dirty && r.host == host[q] && r.bytes == bytes[q])
        dev[q] = (float*)r.dev.p;
      else
        dev[q] = (float*)resident(r, host[q], bytes[q], R);
    }
    // NNSTAT != 0: KSUM, VSUM, PSUM, ASUM(NN3*NNSTAT+1, NVOL), column I per
    // slot (a3d3_s.f:70-71), resident like the half-maps; C3DF the reference
    void* sh[4] = {ksum, vsum, psum, asum};
    float* sdev[4] = {nullptr, nullptr, nullptr, nullptr};
    const float* dref = nullptr;
    const size_t scol = nvox + 1;
    if (*nnstat != 0) {
      need(c3df && ksum && vsum && psum && asum, R, "NNSTAT arrays missing");
      dref = (const float*)resident(s.ref, c3df, sizeof(float) * 2 * hv, R);
      const size_t sb = sizeof(float) * scol * *nvol;
      for (int q = 0; q < 4; q++) {
        Resident& r = s.stats[q];
        if (r.dirty && r.host == sh[q] && r.bytes == sb)
          sdev[q] = (float*)r.dev.p;
        else
          sdev[q] = (float*)resident(r, sh[q], sb, R);
      }
    }
    void* dsum = s.sums.get(2 * sizeof(long long) + 2 * sizeof(double));
    need(dsum != nullptr, R, "device allocation failed");
    long long* dk = (long long*)dsum;
    double* dd = (double*)(dk + 2);
    // a3d3.f:107-137 -> A3D3_S: slot I into private volume IS of its half
    // (IS = I, or (I-1)/2+1 when IFSC = 0), NKSUM1(I) and TTD1/TTD2(IS)
    for (int i = 0; i < *ibuf; i++) {
      const int h = isw[i];
      const int is = *ifsc == 0 ? i / 2 : i;
      need(is < (h == 0 ? *nvol1 : *nvol2), R, "buffer slot beyond NVOL1/NVOL2 private volumes");
      need(recb[10 * (size_t)i + 7] != 0.0f, R, "SIG2N is zero");
      void* ds = upload(s.spec, recbf + 2 * slot_c * i, sizeof(float) * 2 * ic, R);
      void* dc = upload(s.ctff, ctfbf + 2 * slot_f * i, sizeof(float) * 4 * ic, R);
      float* num = dev[2 * h] + 2 * nvox * is;
      float* ctf2 = dev[2 * h + 1] + nvox * is;
      need(hipMemset(dsum, 0, 2 * sizeof(long long) + 2 * sizeof(double)) == hipSuccess, R,
           "device memset failed");
      // one half is addressed (ISW): its arrays stand in for the other's
      if (*nnstat != 0) {
        const size_t o = scol * i;
        check(fr_insert_batch_se(c, n, 1, (const float*)ds, (const float*)dc,
                                 recb + 10 * (size_t)i, &h, *nsym, jsym, symop, *pbc, *boff,
                                 *psize, num, ctf2, num, ctf2, dk, dd, dref,
                                 (int*)sdev[0] + o, sdev[1] + o, sdev[2] + o, sdev[3] + o),
              R);
        for (int q = 0; q < 4; q++) s.stats[q].dirty = true;
      } else {
        check(fr_insert_batch(c, n, 1, (const float*)ds, (const float*)dc, recb + 10 * (size_t)i,
                              &h, *nsym, jsym, symop, *pbc, *boff, *psize, num, ctf2, num, ctf2,
                              dk, dd),
              R);
      }
      check(fr_synchronize(c), R);
      long long kc[2];
      double td[2];
      download(kc, dk, sizeof kc, R);
      download(td, dd, sizeof td, R);
      nksum1[i] += kc[h];  // PINSERT_TRILIN's NKSUM is NKSUM1(I)
      if (h == 0) ttd1[is] += td[0];
      else ttd2[is] += td[1];
      s.maps[2 * h].dirty = s.maps[2 * h + 1].dirty = true;
    }
  }
  *ibuf = 0;
  // the particle loop ends here: the host's half-maps become current
  if (*k == *ilast) {
    State& s = st();
    for (int q = 0; q < 4; q++) {
      write_back(s.maps[q], R);
      write_back(s.stats[q], R);
    }
  }
}

// psearch.f:8-121 (non-helical ASYM, IPAD = 1, IEWALD = 0): CCP with PAD and
// ROTATE2D for every SANG view on the device, the RANDOM jitter, TESTPAR's
// strict '>' insertion with CC > 0 admission and the running best on the
// host.  Returns PHI, THETA, PSI, DSHX, DSHY of the best view, TESTPAR and
// IRAN; PRESA is not touched (as psearch.f leaves it).
void psearch_(const int* nsam, const int* irad, const float* amag, const float* spec,
              const float* a3df, const float* dangin, const int* maxr1, const int* maxr2,
              float* phi, float* theta, float* psi, float* dshx, float* dshy,
              float* /*presa*/, const float* sang, const int* nsang, float* /*ccd*/,
              float* /*ccc*/, float* /*cbuf*/, const int* iquadmax, const int* mask,
              const float* rbfact, const float* /*sinclut*/, const int* ipad, float* /*rbuf*/,
              const int* iewald, const float* /*thetatr*/, const float* ctff,
              const float* /*ri2*/, const float* /*ri3*/, const float* /*rih*/,
              const float* /*halfw*/, float* testpar, const int* ipmax, const char* asym,
              int* iran, const void* /*fftw_plans*/, const float* sm, size_t /*asym_len*/) {
  const char* R = "PSEARCH";
  const int n = *nsam;
  common_checks(R, n, *irad, *ipad, *iewald);
  need(asym[0] != 'H', R, "helical symmetry is not supported");
  need(*nsang >= 1 && *ipmax >= 1, R, "NSANG and IPMAX must be >= 1");
  fr_ctx* c = ctx(R);
  State& s = st();
  const size_t ic = half_plane(n);
  void* dv = resident(s.ref, a3df, sizeof(float) * 2 * half_volume(n), R);
  void* ds = upload(s.spec, spec, sizeof(float) * 2 * ic, R);
  void* dc = upload(s.ctff, ctff, sizeof(float) * 4 * ic, R);
  float par[5] = {*phi, *theta, *psi, *dshx, *dshy};
  check(fr_psearch(c, n, 1, amag, (const float*)dv, (const float*)ds, (const float*)dc, *dangin,
                   *maxr1, *maxr2, par, sang, *nsang, *iquadmax, mask, *rbfact, testpar, *ipmax,
                   iran, sm, asym[0]),
        R);
  *phi = par[0];
  *theta = par[1];
  *psi = par[2];
  *dshx = par[3];
  *dshy = par[4];
}

// Not in the reference.  The host changed the reference transform in place:
// the next per-call routine uploads it again.  (Fortran: CALL FR_PERCALL_INVALIDATE)
void fr_percall_invalidate(void) {
  State& s = st();
  s.ref.host = nullptr;
  s.ref.fp = 0;
}
void fr_percall_invalidate_(void) { fr_percall_invalidate(); }

// Not in the reference.  Copies A3D3's device-resident half-maps back into the
// host arrays now (A3D3 itself does so at K = ILAST).  (Fortran: CALL FR_PERCALL_FLUSH)
void fr_percall_flush(void) {
  State& s = st();
  for (int q = 0; q < 4; q++) {
    write_back(s.maps[q], "FR_PERCALL_FLUSH");
    write_back(s.stats[q], "FR_PERCALL_FLUSH");
  }
}
void fr_percall_flush_(void) { fr_percall_flush(); }

}  // extern "C"

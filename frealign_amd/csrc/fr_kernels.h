// fr_kernels.h -- launchers shared between the kernel files and the C-ABI glue.
#pragma once
#include <vector>
#include "fr_common.h"

namespace fr {

// One insertion task: a particle under one symmetry copy (pinsert.f:74-93).
struct InsertTask {
  float dm[6];   // DM(1..6): the in-plane axes of the rotated plane
  float shx, shy;  // phase-ramp shift (DSHX, DSHY of a3d3.f:84-85)
  float tmp;     // B-factor exponent of pinsert.f:57-62
  float occ;     // ON = OCC/SIG2N (a3d3_s.f:44)
  int part;      // particle slot in the batch
  int half;      // 0: first half-map (A3DF/S3DF), 1: second (D3DF/V3DF)
  int irrec2;    // INT(NSAM*RREC)**2
  short ilo, ihi;  // I stripe of pinsert_trilin.f:44-45 (whole plane: -NSAM/2, NSAM/2)
};

// The per-voxel statistics of the _SE insertion (NNSTAT = 1,
// pinsert_trilin_se.f:186-194), over both half-maps' terms: ksum += 1,
// vsum += |BSAMP|^2, psum += PDIFF(BSAMP, ref), asum += |BSAMP|; ref is the
// reference transform in the half-volume layout (IPAD = 1).
struct InsertStats {
  const float* ref;
  int* ksum;
  float* vsum;
  float* psum;
  float* asum;
};

// scratch of launch_insert for nb particles and ntask tasks (grouped by particle)
size_t insert_scratch_bytes(int nsam, int nb, int ntask);
int launch_insert(Ctx& c, int nsam, int nb, int ntask, const InsertTask* d_tasks,
                  const float* d_datc, const float* d_ctff, float* num0, float* ctf2_0,
                  float* num1, float* ctf2_1, unsigned long long* d_nksum, double* d_ttd,
                  void* scratch, const InsertStats* stats = nullptr);

// FP64 shell sums of sheltest.f:48-88 (istp-wide shells); acc is nstp x 6 doubles:
// DFSC, MEAN12, MEAN22, PSSNR-sum, NPIX, NPIXT.  acc is cleared first, by a fill
// kernel when fill_kernel is set (stream-capturable), else by hipMemsetAsync.
int launch_shell_sums(Ctx& c, int nsam, int istp, int nstp, const float* a, const float* d,
                      const float* s, const float* v, double* acc, bool fill_kernel = false);
// Shells held by the file-scope accumulator array of the no-allocation shell
// sums (fr_stats.hip), and its per-device address: noalloc_scratch_init(dev)
// with dev current takes the address once; noalloc_scratch(dev) returns it
// (nullptr before the init).
constexpr int kNoAllocShells = 520;
int noalloc_scratch_init(int device);
double* noalloc_scratch(int device);
int launch_storeshift(Ctx& c, int nsam, double std, double vtd, int iredun, const float* a,
                      const float* s, const float* d, const float* v, float* b, float* cc);

// Volume finalisation (fr_final.hip).  COMBINE_ARRAYS: copy 1 += copies
// 2..ncopy (stride elements apart), complex a and real s.
int launch_combine(Ctx& c, int nsam, size_t stride, int ncopy, float* a, float* s);
// SHIFTVOL (NNSTAT = 0); stdi = STD*IREDUN after STD += VTD; pssnr/fsc on the device
int launch_shiftvol(Ctx& c, int nsam, double stdi, int irrec2, int ifsc, float fpart, float fmask,
                    float* a, const float* d, float* s, const float* v, const float* pssnr,
                    const float* fsc);
// CORRECT3D on the padded real volume; sinclut: SINCLUT(1..kSincLut) on the host
constexpr int kSincLut = 90;
int launch_correct3d(Ctx& c, int nsam, float scal, int interp, const float* sinclut,
                     float* a3dv);

// MASK3D passes (mask3d.f) over the padded real volume: block partial sums
// (kMask3dBlocks x 5 doubles in part) reduced in a fixed order into out[0..4]:
// edge-shell sum and count, then (stats) the RIC..RIO sums of A, A^2, count.
constexpr int kMask3dBlocks = 1024;
int launch_mask3d_sums(Ctx& c, int nsam, const float* a, float xc, float ri2, float ri3, float ri4,
                       float ri5, int stats, double* part, double* out);
// the cosine edge with background bg; out[0] = this pass's SUM3 terms
int launch_mask3d_apply(Ctx& c, int nsam, float* a, float xc, float ri2, float ri3, float rih,
                        float rii, float halfw, double bg, int inner, double* part, double* out);

// Particle preparation (lmain.f:459 r2c + ctfapply.f:8-161)
struct PrepGlobals {
  float cs, wl, wgh1, wgh2, thetatr, rih, halfw, ri2, ri3, tx, ty;
  int ictf;
};
// part: the ring partial sums of MASKCOS's mean (prep_partials_count(nsam,
// nb) float2; null: the context's buffer, for calls on the context's stream)
int launch_prep(Ctx& c, int nsam, int nb, float* data, const float* ppar, const float* pssnr,
                const PrepGlobals& g, float* ctff, float* outc, float* datc,
                float2* part = nullptr);
// ctfapply.f:8-161 alone, from nb spectra (NSAM x JC complex each, as the r2c leaves them)
int launch_ctfapply(Ctx& c, int nsam, int nb, const float* spec, const float* ppar,
                    const float* pssnr, const PrepGlobals& g, float* ctff, float* outc,
                    float* datc, float2* part = nullptr);
size_t prep_partials_count(int nsam, int nb);

// Central-section extraction (pextract.f:8-101), one output plane per item.
int launch_pextract(Ctx& c, int nsam, int n, const float* a3df, const float* dm, const int* pidx,
                    const float* ctff, float* outc);

// CC3M scores (cc3m.f:8-169) for n (particle, orientation, shift) items.
struct Cc3mArgs {
  int nsam, n, maxr1, maxr2;
  float rbfact, ri2, ri3, rih, halfw;
  int cc3 = 0;  // 1: CC3 (cc3.f): the section phase-flipped by the CTF sign, no |R| branch
};
// Scale by `scal`, then MASKCOS (origin=0) / MASKCOS2 (origin=1); one workgroup per image.
constexpr int kMaskThreads = 1024;
__global__ void k_scale_mask(int nsam, float* data, size_t img_stride, float scal,
                             const float* amag_arr, int amag_stride, float ri2, float ri3,
                             float rih, float halfw, int origin);
int launch_cc3m(Ctx& c, const Cc3mArgs& a, const float* a3df, const float* spec, const float* ctff,
                const int* pidx, const float* dm, const float* shift, const float* amag,
                const float* fsct, float* work, float* score);
int launch_presb(Ctx& c, int nsam, int n, const float* a3df, const float* spec, const float* ctff,
                 const int* pidx, const float* dm, const float* shift, const float* amag,
                 float ri2, float ri3, float rih, float halfw, float* work, float* rbin);
// SIGMA (sigma.f:8-221) item: PEXTRACT DM, shift, mask centre/radius^2,
// SIGNSET, particle and band.
struct SigmaItem {
  float dm[9];
  float shx, shy, xm, ym, ri2f, signset;
  int pidx, maxr12, maxr22, pad_;
};
int launch_sigma(Ctx& c, int nsam, int n, const float* a3df, const SigmaItem* items,
                 const float* datc, const float* ctff, float sigp, float* buf, float* out);
// CC3M with band-pruned DFT products on the matrix cores; one workgroup per item.
bool cc3m_fused_ok(int nsam, int maxr2);
size_t cc3m_fused_lds(int nsam);
// Constant tables of the fused kernel (device pointers, owned by the context):
//   dft  the folded DFT table TF[r][i] = cos(2 pi i r/N), TF[r][32+i] =
//        -sin(2 pi i r/N), r = 0..N/2, i < 32 (T[n][2i] = cos(2 pi i n/N),
//        T[n][2i+1] = -sin(2 pi i n/N) for every n is read from it);
//   pts  kCc3mShells+1 shell offsets, then the band points (I | (J+64) << 8)
//        sorted by shell INT(SQRT(I*I+J*J)+0.5), (I, J) order within a shell;
//   ring Q[c][k] = w_c sum_m T[m][k] R[m][c], R[m][k] the sum over the
//        MASKCOS2 ring pixels n of row m of T[n][k] (w_c the c2r weight): the
//        ring sum is sum Q B^T, for AMAG = 1 (null: the ring mean is taken by a
//        pass on the matrix cores);
//   wmap the MASKCOS2 weight of every pixel [m][n] at AMAG = 1 (null:
//        evaluated per pixel).
constexpr int kCc3mShells = 34;
constexpr int kCc3mPtsHdr = 36;
struct Cc3mTables {
  const float* dft = nullptr;
  const int* pts = nullptr;
  const float* ring = nullptr;
  int ring_cnt = 0;
  const float* wmap = nullptr;
};
void cc3m_fused_table(int nsam, std::vector<float>& tab);
void cc3m_fused_table_folded(int nsam, std::vector<float>& tab);
void cc3m_fused_points(int maxr12, int maxr22, std::vector<int>& pts);
int cc3m_fused_ring(int nsam, float ri2, float ri3, const std::vector<float>& dft,
                    std::vector<float>& ring);
int cc3m_fused_ring_q(int nsam, float ri2, float ri3, std::vector<float>& q);
void cc3m_fused_wmap(int nsam, float ri2, float ri3, float rih, float halfw,
                     std::vector<float>& w);
// A window of nsx x nsy shifts (odd sizes up to kCc3mWinMax) centred on each
// item's shift, dsx/dsy apart (phase units): the fused kernel then writes the
// best score of the window and its index (ix * nsy + iy, the first maximum).
constexpr int kCc3mWinMax = 11;
struct Cc3mWindow {
  int nsx, nsy;
  float dsx, dsy;
};
int launch_cc3m_fused(Ctx& c, const Cc3mArgs& a, const float* a3df, const float* spec,
                      const float* ctff, const int* pidx, const float* dm, const float* shift,
                      const float* amag, const float* fsct, const Cc3mTables& t, float* score,
                      int blocks = 0, const Cc3mWindow* win = nullptr, int* wcode = nullptr);

}  // namespace fr

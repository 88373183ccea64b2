// fr_finalize.cpp -- the end of a refinement cycle on the device in one call:
// frealign_v9.f:790-950 from the two accumulated half-maps to the FSC table's
// numbers and the final real-space map, by the library's own entry points in
// the reference's order.
#include <hip/hip_runtime.h>

#include <math.h>

#include <vector>

#include "../../include/frealign_hip.h"
#include "fr_ctx.h"

namespace {

// CALCSINC (calcsinc.f:17-19): SINCLUT(J) = SIN(X)/X, X = J*PI/180, J = 1..180
void calcsinc(std::vector<float>& lut) {
  const float pi = 3.1415926535897f;
  lut.resize(180);
  for (int j = 1; j <= 180; j++) {
    const float x = (float)j * pi / 180.0f;
    lut[j - 1] = sinf(x) / x;
  }
}

}  // namespace

extern "C" int fr_finalize(fr_ctx* c, int nsam, float* d_a3df, float* d_s3df,
                           const float* d_d3df, const float* d_v3df, const long long* nksum,
                           const double* ttd, int iredun, float psize, float ri, float ric,
                           float mw, float dalt, int ffilt, float rlim, float* d_b, float* d_c,
                           int nshell, float* fsc, float* pssnr, int* npix, int* npixt,
                           int* nstp, float* fmask, float* fpart) {
  if (!c) return fr::fail("null context");
  FR_CHECK(d_a3df && d_s3df && d_d3df && d_v3df && nksum && ttd && d_b && d_c && fsc && pssnr &&
               npix && npixt && nstp && fmask && fpart,
           "null argument");
  FR_CHECK(iredun >= 1 && psize > 0.0f && ri != 0.0f, "bad IREDUN, PSIZE or RI");
  const int jc = nsam / 2 + 1;
  FR_CHECK(nshell >= jc + 2, "nshell %d below NSAM/2 + 3", nshell);
  const long long nk = nksum[0] + nksum[1];
  FR_CHECK(nk > 0, "no voxel was inserted (NKSUM 0)");
  // frealign_v9.f:792-795: STD and VTD from the half-maps' CTF^2 sums (WC2 0.1)
  double std_ = ttd[0] / (double)nk * 0.1, vtd = ttd[1] / (double)nk * 0.1;
  if (int r = fr_storeshift(c, nsam, std_, vtd, iredun, d_a3df, d_s3df, d_d3df, d_v3df, d_b, d_c))
    return r;
  std::vector<float> lut;
  calcsinc(lut);
  // the half-maps in real space: CORRECT3D (INTERP 1), MASK3D at |RI| with
  // RIC and HALFW 6 (frealign_v9.f:825-849); FMASK from the first
  const float rio = fabsf(ri);
  float fm = 1.0f;
  float* halves[2] = {d_b, d_c};
  for (int k = 0; k < 2; k++) {
    float ave, sd, f;
    if (int r = fr_fft3d_c2r(c, nsam, halves[k])) return r;
    if (int r = fr_correct3d(c, nsam, halves[k], 1, 1, lut.data(), (int)lut.size())) return r;
    if (int r = fr_mask3d(c, nsam, halves[k], rio, ric, 6.0f, &f, &ave, &sd)) return r;
    if (k == 0) fm = f;
    if (int r = fr_fft3d_r2c(c, nsam, halves[k])) return r;
  }
  // SHELTEST's FSC and PSSNR (frealign_v9.f:867-869)
  std::vector<int> np_(nshell, 0), npt(nshell, 0);
  std::vector<double> dfsc(nshell, 0.0);
  std::vector<float> ps(nshell, 0.0f), fs(nshell, 0.0f);
  int ns = 0;
  if (int r = fr_sheltest(c, nsam, 1, d_b, d_c, d_s3df, d_v3df, fm, np_.data(), npt.data(),
                          dfsc.data(), ps.data(), &ns))
    return r;
  for (int i = 0; i < ns; i++) fs[i] = (float)dfsc[i];
  // FIND_FPART (frealign_v9.f:871-874): FPART, and PSSNR divided by it
  float fp = 0.0f;
  if (int r = fr_find_fpart(c, nsam, std_, vtd, d_a3df, d_d3df, d_s3df, d_v3df, iredun, rio, ric,
                            psize, mw, dalt, fs.data(), ps.data(), nshell, &fp))
    return r;
  // the final map (frealign_v9.f:888-950): SHIFTVOL into A3DF/S3DF, with the
  // Wiener filter only when FFILT; then c2r, CORRECT3D and, for RI > 0, MASK3D
  // (SHIFTVOL reads D3DF and writes A3DF, S3DF for IFSC 0)
  if (int r = fr_shiftvol(c, nsam, &std_, vtd, rlim, iredun, d_a3df, const_cast<float*>(d_d3df),
                          d_s3df, d_v3df, 0, ps.data(), ffilt ? fp : 0.0f, fs.data(), fm,
                          nshell))
    return r;
  if (int r = fr_fft3d_c2r(c, nsam, d_a3df)) return r;
  if (int r = fr_correct3d(c, nsam, d_a3df, 1, 1, lut.data(), (int)lut.size())) return r;
  if (ri > 0.0f) {
    float f, ave, sd;
    if (int r = fr_mask3d(c, nsam, d_a3df, ri, ric, 6.0f, &f, &ave, &sd)) return r;
  }
  for (int i = 0; i < nshell; i++) {
    fsc[i] = fs[i];
    pssnr[i] = ps[i];
    npix[i] = np_[i];
    npixt[i] = npt[i];
  }
  *nstp = ns;
  *fmask = fm;
  *fpart = fp;
  return 0;
}

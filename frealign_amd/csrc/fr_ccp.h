// fr_ccp.h -- PSEARCH -> CCP with PAD and ROTATE2D, reference-exact (fr_ccp.hip).
#pragma once
#include <vector>

#include "fr_common.h"

namespace fr {

constexpr int kCcpMaxR = 32;  // MAXR2 of the exact search (the band is 63 x 63 at 32)

struct CcpArgs {
  int nsam, nview, maxr1, maxr2;
  int np;            // in-plane steps scanned (NPSI, or 1 when MASK(3) = 0)
  int iq;            // quadrants scanned (IQUADMAX, or 1 when MASK(3) = 0)
  int mask4, mask5;  // shift search in x / y (MASK(4), MASK(5))
  float scal;        // PAD's SCAL = 1/NSAM/NSAM (FP32)
};

// Constant tables per (NSAM, MAXR2), built on the host in FP64 and stored FP32:
//   d       D(m), m = -8R..8R (complex)
//   k       K(L,M) = D(-L) D(-M) + B(L,M), rows M = -4R..4R, columns L = 0..4R
//   w       W(kx,ky) over the central cross, kx, ky = -R+1..R-1
//   e, e1   e^{2 pi i J n/NSAMR} (J = -R+1..R-1) and e^{2 pi i I n/NSAMR}
//           (I = 0..R-1), n = -7..7
//   region  per I, bit J+R-1: the correlation region of ccp.f:160-174
struct CcpTablesHost {
  std::vector<float> d, k, w, e, e1;
  std::vector<unsigned long long> region;
};
struct CcpTablesDev {
  const float *d = nullptr, *k = nullptr, *w = nullptr, *e = nullptr, *e1 = nullptr;
  const unsigned long long* region = nullptr;
};
void ccp_tables(int nsam, int maxr2, CcpTablesHost* t);
// bytes of the padded-transform band of one item
size_t ccp_band_bytes(int maxr2);
// items item0 .. item0+nitems-1 (item = particle * nview + view); fband holds
// nitems bands; score/code are indexed by item.  code = ((IPSI-1)*4 +
// IQUAD-1)*225 + (IXM+7)*15 + (IYM+7).  bw: EXP(-RBFACT r^2) for r^2 < MAXR2^2.
int launch_ccp(Ctx& c, const CcpArgs& a, int item0, int nitems, const float* a3df,
               const float* spec, const float* ctff, const float* dm, const float* bw,
               const CcpTablesDev& t, const float* psitab, float* fband, float* score,
               int* code);

}  // namespace fr

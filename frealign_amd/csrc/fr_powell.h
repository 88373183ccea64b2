// fr_powell.h -- VA04A (va04.f:8-369), Powell's conjugate-direction
// minimiser, as a resumable state machine so that many particles' PREFINE
// minimisations can run in lock-step and share one batched objective launch.
//
//   fr::Powell p;
//   p.start(x, e, n, escale, icon, maxit);
//   while (p.run()) p.supply(objective(p.x()));
//   // p.x(), p.value()
//
// The arithmetic is FP32 in the reference's order and the branch structure is
// the reference's (its three-way arithmetic IFs on FP32 differences, its
// computed GOTOs, its MAXX = 100*MAXIT evaluation safety exit), so the
// iterates are the reference's for the same objective values.
#pragma once

#include <functional>

#include "fr_common.h"

namespace fr {

class Powell {
 public:
  // x, e: n <= 6 values; icon and maxit as VA04A's ICON and MAXIT
  void start(const float* x, const float* e, int n, float escale, int icon, int maxit);
  // true: the objective is wanted at x(); pass it to supply() and run again
  bool run();
  void supply(float f) { F = f; }
  const float* x() const { return X + 1; }
  float value() const { return F; }
  int evaluations() const { return calls; }
  bool done() const { return at == kDone; }
  bool safety_exit() const { return safety; }

 private:
  enum Resume { kStart, kFirst, kLine, kRestart, kSafety, kDone };
  Resume at = kDone;
  int n = 0, icon = 1, maxit = 0, calls = 0;
  bool safety = false;
  float X[7] = {}, E[7] = {}, XS[7] = {}, W[61] = {};
  float escale = 0.f, F = 0.f, DDMAG = 0.f, SCER = 0.f, FKEEP = 0.f, FP = 0.f, SUM = 0.f;
  float DMAX = 0.f, DACC = 0.f, DMAG = 0.f, DDMAX = 0.f, DL = 0.f, D = 0.f, FPREV = 0.f;
  float FA = 0.f, DA = 0.f, DD = 0.f, FB = 0.f, DB = 0.f, FC = 0.f, DC = 0.f, A = 0.f, B = 0.f;
  float DI = 0.f, FI = 0.f, FHOLD = 0.f, AAA = 0.f;
  int JJ = 0, JJJ = 0, IND = 1, INN = 1, ITERC = 0, ISGRAD = 2, ITONE = 1, IDIRN = 0;
  int ILINE = 0, IS = 5, JIL = 1, ICNT = 0, MAXX = 0;
};

// PREFINE for a batch (prefine.f:8-162, IVFLAG 0, non-helical, AMAG > 0).
struct RefineArgs {
  int n = 0;
  const int* pidx = nullptr;    // host: particle of each item
  const float* amag = nullptr;  // host: |AMAG| of each item
  const int* mask = nullptr;    // 5 flags: refine PHI, THETA, PSI, DSHX, DSHY
  const float* prior = nullptr; // host n x 5: XM-SHX, YM-SHY, SX, SY, SIG2 (FSH of fsh.f)
  int maxit = 50;               // NCYCLS (prefine.f:29)
  const float* sm = nullptr;    // SM of the DMs (cc3m.f:66, MATMUL_T); null: identity
};

// Objective values in one launch: score[q] = CC3M of particle pidx[q] at
// (dm[9q..], shift[2q..], amag[q]) for q < m (host arrays).
using Cc3mFn = std::function<int(int m, const int* pidx, const float* dm, const float* shift,
                                 const float* amag, float* score)>;

// par: host n x 6, PHI THETA PSI DSHX DSHY in, refined values and PRESA out.
int refine_batch(Ctx& c, const RefineArgs& a, float* par, int* nevals, const Cc3mFn& cc3m);

}  // namespace fr

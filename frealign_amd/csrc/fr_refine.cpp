// fr_refine.cpp -- PREFINE (prefine.f:8-162) for a batch of particles.
//
// Every particle's VA04A minimisation (fr::Powell) runs on the host as a
// small state machine; all of them advance in lock-step, and the objective
// values they ask for in one round -- CALCFX's -CC3M - FSH + range penalties
// (calcfx.f:152-191) -- are computed by one batched CC3M launch on the GPU.
#include <hip/hip_runtime.h>
#include <math.h>

#include <vector>

#include "fr_common.h"
#include "fr_host.h"
#include "fr_kernels.h"
#include "fr_powell.h"

namespace fr {

void Powell::start(const float* x0, const float* e0, int n_, float escale_, int icon_,
                   int maxit_) {
  n = n_;
  icon = icon_;
  maxit = maxit_;
  escale = escale_;
  for (int i = 1; i <= n; i++) {
    X[i] = x0[i - 1];
    E[i] = e0[i - 1];
  }
  calls = 0;
  safety = false;
  at = kStart;
}

// va04.f:8-369.  Labels Lnn are the reference's statement labels.
bool Powell::run() {
  int i, j, k, ixp;
  switch (at) {
    case kStart: break;
    case kFirst: goto R_FIRST;
    case kLine: goto R_LINE;
    case kRestart: goto R_RESTART;
    case kSafety: goto R_SAFETY;
    case kDone: return false;
  }
  DDMAG = 0.1f * escale;
  SCER = 0.05f / escale;
  ICNT = 0;
  MAXX = 100 * maxit;
  for (i = 1; i <= n; i++) XS[i] = X[i];
  JJ = n * n + n;
  JJJ = JJ + n;
  k = n + 1;
  IND = 1;
  INN = 1;
  for (i = 1; i <= n; i++)
    for (j = 1; j <= n; j++) {  // unit directions scaled by E, step lengths ESCALE
      W[k] = (i == j) ? fabsf(E[i]) : 0.f;
      if (i == j) W[i] = escale;
      k++;
    }
  ITERC = 1;
  ISGRAD = 2;
  if (++ICNT > MAXX) goto L998;
  at = kFirst;
  calls++;
  return true;
R_FIRST:
  FKEEP = fabsf(F) + fabsf(F);
L5:  // a new iteration: remember the starting point
  ITONE = 1;
  FP = F;
  SUM = 0.f;
  ixp = JJ;
  for (i = 1; i <= n; i++) W[++ixp] = X[i];
  IDIRN = n + 1;
  ILINE = 1;
L7:  // line search along direction ILINE
  DMAX = W[ILINE];
  DACC = DMAX * SCER;
  DMAG = fminf(DDMAG, 0.1f * DMAX);
  DMAG = fmaxf(DMAG, 20.f * DACC);
  DDMAX = 10.f * DMAG;
  if (ITONE == 3) goto L71;
  DL = 0.f;
  D = DMAG;
  FPREV = F;
  IS = 5;
  FA = F;
  DA = DL;
L8:
  DD = D - DL;
  DL = D;
L58:  // step to D along the current direction and evaluate
  k = IDIRN;
  for (i = 1; i <= n; i++) X[i] += DD * W[k++];
  if (++ICNT > MAXX) goto L998;
  at = kLine;
  calls++;
  return true;
R_LINE:
  switch (IS) {
    case 1: goto L10;
    case 2: goto L11;
    case 3: goto L12;
    case 4: goto L13;
    case 6: goto L96;
    default: break;
  }
  // L14: bracket the minimum
  if (F - FA < 0.f) {
    FB = F;
    DB = D;
    goto L21;
  }
  if (F - FA == 0.f) {
    if (fabsf(D) - DMAX <= 0.f) {
      D = D + D;
      goto L8;
    }
    goto L20;  // the maximum change does not alter the function
  }
  FB = FA;
  DB = DA;
  FA = F;
  DA = D;
L21:
  if (ISGRAD == 1) goto L83;
L23:
  D = DB + DB - DA;
  IS = 1;
  goto L8;
L83:  // parabolic step from the gradient estimate
  D = 0.5f * (DA + DB - (FA - FB) / (DA - DB));
  IS = 4;
  if ((DA - D) * (D - DB) >= 0.f) goto L8;
L25:
  IS = 1;
  if (fabsf(D - DB) - DDMAX <= 0.f) goto L8;
L26:
  D = DB + copysignf(fabsf(DDMAX), DB - DA);
  IS = 1;
  DDMAX = DDMAX + DDMAX;
  DDMAG = DDMAG + DDMAG;
  if (DDMAX - DMAX > 0.f) DDMAX = DMAX;
  goto L8;
L13:
  if (F - FA >= 0.f) goto L23;
L28:
  FC = FB;
  DC = DB;
L29:
  FB = F;
  DB = D;
  goto L30;
L12:
  if (F - FB <= 0.f) goto L28;
  FA = F;
  DA = D;
  goto L30;
L11:
  if (F - FB < 0.f) {
    FA = FB;
    DA = DB;
    goto L29;
  }
  goto L10;
L71:
  DL = 1.f;
  DDMAX = 5.f;
  FA = FP;
  DA = -1.f;
  FB = FHOLD;
  DB = 0.f;
  D = 1.f;
L10:
  FC = F;
  DC = D;
L30:  // parabola through (DA,FA), (DB,FB), (DC,FC)
  A = (DB - DC) * (FA - FC);
  B = (DC - DA) * (FB - FC);
  if ((A + B) * (DA - DC) <= 0.f) {
    FA = FB;
    DA = DB;
    FB = FC;
    DB = DC;
    goto L26;
  }
  D = 0.5f * (A * (DB + DC) + B * (DA + DC)) / (A + B);
  DI = DB;
  FI = FB;
  if (FB - FC > 0.f) {
    DI = DC;
    FI = FC;
  }
  if (ITONE == 3) {
    ITONE = 2;
    goto L45;
  }
  if (fabsf(D - DI) - DACC <= 0.f) goto L41;
  if (fabsf(D - DI) - 0.03f * fabsf(D) <= 0.f) goto L41;
L45:
  if ((DA - DC) * (DC - D) >= 0.f) {
    FA = FB;
    DA = DB;
    FB = FC;
    DB = DC;
    goto L25;
  }
  IS = ((DB - D) * (D - DC) < 0.f) ? 3 : 2;
  goto L8;
L41:  // accept the line minimum; rescale the direction
  F = FI;
  D = DI - DL;
  DD = (DC - DB) * (DC - DA) * (DA - DB) / (A + B);
  if (DD < 0.f) DD = 1e-10f;
  DD = sqrtf(DD);
  for (i = 1; i <= n; i++) {
    X[i] += D * W[IDIRN];
    W[IDIRN] = DD * W[IDIRN];
    IDIRN++;
  }
  if (DD == 0.f) DD = 1e-10f;
  W[ILINE] = W[ILINE] / DD;
  ILINE++;
  if (ITONE == 2) goto L38;
  if (FPREV - F - SUM >= 0.f) {  // the largest decrease so far
    SUM = FPREV - F;
    JIL = ILINE;
  }
  if (IDIRN - JJ <= 0) goto L7;
  if (IND == 2) goto L72;
L92:  // extrapolate along the total displacement of this iteration
  FHOLD = F;
  IS = 6;
  ixp = JJ;
  for (i = 1; i <= n; i++) {
    ixp++;
    W[ixp] = X[i] - W[ixp];
  }
  DD = 1.f;
  goto L58;
L96:
  if (IND == 2) goto L87;
  if (FP - F <= 0.f) goto L37;
  D = 2.f * (FP + F - 2.f * FHOLD) / ((FP - F) * (FP - F));
  if (D * ((FP - FHOLD - SUM) * (FP - FHOLD - SUM)) - SUM >= 0.f) goto L37;
L87:  // replace direction JIL by the displacement
  j = JIL * n + 1;
  if (j - JJ <= 0) {
    for (i = j; i <= JJ; i++) W[i - n] = W[i];
    for (i = JIL; i <= n; i++) W[i - 1] = W[i];
  }
  IDIRN = IDIRN - n;
  ITONE = 3;
  k = IDIRN;
  ixp = JJ;
  AAA = 0.f;
  for (i = 1; i <= n; i++) {
    ixp++;
    W[k] = W[ixp];
    if (AAA - fabsf(W[k] / E[i]) < 0.f) AAA = fabsf(W[k] / E[i]);
    k++;
  }
  DDMAG = 1.f;
  if (AAA == 0.f) AAA = 1e-10f;
  W[n] = escale / AAA;
  ILINE = n;
  goto L7;
L37:
  ixp = JJ;
  AAA = 0.f;
  F = FHOLD;
  for (i = 1; i <= n; i++) {
    ixp++;
    X[i] -= W[ixp];
    if (AAA * fabsf(E[i]) - fabsf(W[ixp]) < 0.f) AAA = fabsf(W[ixp] / E[i]);
  }
  goto L72;
L38:
  AAA = AAA * (1.f + DI);
  if (IND == 2) goto L106;
L72:  // convergence tests
  if (IND == 2) goto L88;
  if (AAA - 0.1f > 0.f) goto L76;
  if (icon == 1) goto L20;
  IND = 2;
  if (INN == 2) goto L101;
  INN = 2;  // ICON = 2: perturb by 10 E and minimise again
  k = JJJ;
  for (i = 1; i <= n; i++) {
    W[++k] = X[i];
    X[i] += 10.f * E[i];
  }
  FKEEP = F;
  if (++ICNT > MAXX) goto L998;
  at = kRestart;
  calls++;
  return true;
R_RESTART:
  DDMAG = 0.f;
  goto L108;
L76:
  if (F - FP < 0.f) goto L35;
  goto L20;  // accuracy limited by errors in F
L88:
  IND = 1;
L35:
  {
    float tmp = FP - F;
    DDMAG = tmp > 0.f ? 0.4f * sqrtf(tmp) : 0.f;
  }
  ISGRAD = 1;
L108:
  ITERC++;
  if (ITERC - maxit <= 0) goto L5;
  if (F - FKEEP <= 0.f) goto L20;
  F = FKEEP;
  for (i = 1; i <= n; i++) X[i] = W[++JJJ];
  goto L20;
L101:
  JIL = 1;
  FP = FKEEP;
  if (F - FKEEP == 0.f) goto L20;
  if (F - FKEEP > 0.f) {
    JIL = 2;
    FP = F;
    F = FKEEP;
  }
  ixp = JJ;
  for (i = 1; i <= n; i++) {
    ixp++;
    k = ixp + n;
    if (JIL == 1) {
      W[ixp] = W[k];
    } else {
      W[ixp] = X[i];
      X[i] = W[k];
    }
  }
  JIL = 2;
  goto L92;
L106:
  if (AAA - 0.1f <= 0.f) goto L20;
  INN = 1;
  goto L35;
L20:
  at = kDone;
  return false;
L998:  // evaluation budget exhausted: back to the start point
  for (i = 1; i <= n; i++) X[i] = XS[i];
  safety = true;
  at = kSafety;
  calls++;
  return true;
R_SAFETY:
  at = kDone;
  return false;
}

namespace {

// fsh.f: the Gaussian shift prior
inline float fsh(float sig2, float x, float y, float xm, float ym, float sx, float sy) {
  return sig2 * (-(x - xm) * (x - xm) / 2.0f / (sx * sx) - (y - ym) * (y - ym) / 2.0f / (sy * sy));
}

}  // namespace

int refine_batch(Ctx& c, const RefineArgs& a, float* par, int* nevals, const Cc3mFn& cc3m) {
  (void)c;
  constexpr float kPi = 3.1415926535897f;
  const float epar[6] = {0.1f, 0.1f, 0.1f, 0.01f, 0.01f, 0.001f};  // prefine.f:46
  const int n = a.n;
  std::vector<Powell> pw(n);
  std::vector<int> npar(n, 0);
  std::vector<int> active;
  // prefine.f:56-66: the refined subset of (PHI, THETA, PSI, DSHX, DSHY)
  for (int i = 0; i < n; i++) {
    float xm[6], em[6];
    int np = 0;
    for (int j = 0; j < 5; j++)
      if (a.mask[j] == 1) {
        xm[np] = par[6 * i + j];
        em[np] = epar[j];
        np++;
      }
    npar[i] = np;
    if (np > 0) {
      pw[i].start(xm, em, np, 100.0f, 1, a.maxit);
      active.push_back(i);
    }
  }
  std::vector<int> want;
  std::vector<int> pidx;
  std::vector<float> dm, shift, amag, score;
  // objective at each item's current parameters: -CC3M - FSH + penalties
  auto evaluate = [&](const std::vector<int>& items, std::vector<float>& fval,
                      std::vector<float>& cc) -> int {
    const int m = (int)items.size();
    pidx.resize(m);
    dm.resize(9 * (size_t)m);
    shift.resize(2 * (size_t)m);
    amag.resize(m);
    score.resize(m);
    std::vector<float> xp(6 * (size_t)m);
    for (int q = 0; q < m; q++) {
      const int i = items[q];
      float* x = &xp[6 * (size_t)q];
      for (int j = 0; j < 5; j++) x[j] = par[6 * i + j];
      if (npar[i] > 0) {
        const float* v = pw[i].x();
        int ip = 0;
        for (int j = 0; j < 5; j++)
          if (a.mask[j] == 1) x[j] = v[ip++];
      }
      x[5] = a.amag[i];
      pidx[q] = a.pidx[i];
      rotation(x[0], x[1], x[2], x[5], a.sm, &dm[9 * (size_t)q]);
      shift[2 * q] = x[3];
      shift[2 * q + 1] = x[4];
      amag[q] = x[5];
    }
    if (int r = cc3m(m, pidx.data(), dm.data(), shift.data(), amag.data(), score.data())) return r;
    fval.resize(m);
    cc.resize(m);
    for (int q = 0; q < m; q++) {
      const int i = items[q];
      const float* x = &xp[6 * (size_t)q];
      const float* pr = a.prior + 5 * (size_t)i;
      float rft = -score[q] - fsh(pr[4], x[3], x[4], pr[0], pr[1], pr[2], pr[3]);
      if (fabsf(x[0]) > kPi) rft = rft + (kPi - fabsf(x[0])) * (kPi - fabsf(x[0]));
      if (fabsf(x[1]) > kPi) rft = rft + (kPi - fabsf(x[1])) * (kPi - fabsf(x[1]));
      if (fabsf(x[2]) > kPi) rft = rft + (kPi - fabsf(x[2])) * (kPi - fabsf(x[2]));
      if (fabsf(x[3]) > 1.0f) rft = rft + (1.0f - fabsf(x[3])) * (1.0f - fabsf(x[3]));
      if (fabsf(x[4]) > 1.0f) rft = rft + (1.0f - fabsf(x[4])) * (1.0f - fabsf(x[4]));
      fval[q] = rft;
      cc[q] = score[q];
    }
    return 0;
  };
  std::vector<float> fval, cc;
  // items with nothing to refine: PRESA = CC3M + FSH at the start (prefine.f:68-84, 155)
  {
    std::vector<int> idle;
    for (int i = 0; i < n; i++)
      if (npar[i] == 0) idle.push_back(i);
    if (!idle.empty()) {
      if (int r = evaluate(idle, fval, cc)) return r;
      for (size_t q = 0; q < idle.size(); q++) {
        const int i = idle[q];
        const float* pr = a.prior + 5 * (size_t)i;
        par[6 * i + 5] = cc[q] + fsh(pr[4], par[6 * i + 3], par[6 * i + 4], pr[0], pr[1], pr[2], pr[3]);
        if (nevals) nevals[i] = 1;
      }
    }
  }
  // lock-stepped minimisations
  for (int i : active) pw[i].run();
  while (!active.empty()) {
    if (int r = evaluate(active, fval, cc)) return r;
    std::vector<int> next;
    for (size_t q = 0; q < active.size(); q++) {
      const int i = active[q];
      pw[i].supply(fval[q]);
      if (pw[i].run()) next.push_back(i);
    }
    active.swap(next);
  }
  for (int i = 0; i < n; i++) {
    if (npar[i] == 0) continue;
    const float* v = pw[i].x();
    int ip = 0;
    for (int j = 0; j < 5; j++)
      if (a.mask[j] == 1) par[6 * i + j] = v[ip++];
    par[6 * i + 5] = -pw[i].value();  // PRESA = -F (prefine.f:140)
    if (nevals) nevals[i] = pw[i].evaluations();
  }
  return 0;
}

}  // namespace fr

// fr_ctf.h -- the CTF of ctf.f:8-64 at one half-plane point, shared by the
// particle preparation and the training-image synthesis of the proposal network.
#pragma once
#include <hip/hip_runtime.h>

#include "fr_common.h"
#include "fr_kernels.h"

namespace fr {

constexpr float kCtfTwoPi = 6.2831853071796f;

// ctf.f:8-64
__device__ inline cf ctf_dev(const PrepGlobals& g, float df1, float df2, float angast,
                             float thetatr, int ix, int iy) {
  float wgh2 = fabsf(g.wgh2);
  cf out;
  if (g.wgh2 != -1.0f) {
    float radf = (float)(ix * ix + iy * iy);
    if (radf != 0.0f) {
      float rad = sqrtf(radf);
      float angle = rad * thetatr;
      float angspt = atan2f((float)iy, (float)ix);
      float c1 = kCtfTwoPi * angle * angle / (2.0f * g.wl);
      float c2 = c1 * g.cs * angle * angle / 2.0f;
      float ccos = cosf(2.0f * (angspt - angast));
      float df = -0.5f * (df1 + df2 + ccos * (df1 - df2));
      float chi = c1 * df + c2;
      // beam tilt (ctf.f:44)
      chi = chi + kCtfTwoPi * g.cs * angle * angle * ((float)ix * g.tx + (float)iy * g.ty) * thetatr /
                      g.wl / 1000.0f;
      float s, c;
      sincosf(chi, &s, &c);
      out = cf{g.wgh1 * s - wgh2 * c, -g.wgh1 * c - wgh2 * s};
    } else {
      out = cf{-wgh2, 0.0f};
    }
    if (g.wgh2 < 0.0f) out = cf{-out.re, -out.im};
  } else {
    out = cf{1.0f, 0.0f};
  }
  return out;
}

// ctf.f:8-64 at (ix, iy) and at (-ix, -iy) together: the radius, the
// spherical-aberration terms and the defocus along the azimuth (from the
// first point's ATAN2) are shared, and the beam-tilt term of the second
// point is the first's negated (each factor of it is sign-symmetric), so
// the two phases differ only by that term.  Agrees with two ctf_dev calls
// except in the defocus of the second point, where the azimuth ATAN2(-IY,
// -IX) = ATAN2(IY, IX) -+ PI enters COS(2(ANGSPT-ANGAST)) after rounding.
__device__ inline void ctf_pair_dev(const PrepGlobals& g, float df1, float df2, float angast,
                                    float thetatr, int ix, int iy, cf& cr, cf& cl) {
  float wgh2 = fabsf(g.wgh2);
  if (g.wgh2 == -1.0f) {
    cr = cl = cf{1.0f, 0.0f};
    return;
  }
  float radf = (float)(ix * ix + iy * iy);
  if (radf != 0.0f) {
    float rad = sqrtf(radf);
    float angle = rad * thetatr;
    float angspt = atan2f((float)iy, (float)ix);
    float c1 = kCtfTwoPi * angle * angle / (2.0f * g.wl);
    float c2 = c1 * g.cs * angle * angle / 2.0f;
    float ccos = cosf(2.0f * (angspt - angast));
    float df = -0.5f * (df1 + df2 + ccos * (df1 - df2));
    float chi0 = c1 * df + c2;
    // beam tilt (ctf.f:44)
    float tilt = kCtfTwoPi * g.cs * angle * angle * ((float)ix * g.tx + (float)iy * g.ty) *
                 thetatr / g.wl / 1000.0f;
    float s, c;
    sincosf(chi0 + tilt, &s, &c);
    cr = cf{g.wgh1 * s - wgh2 * c, -g.wgh1 * c - wgh2 * s};
    sincosf(chi0 - tilt, &s, &c);
    cl = cf{g.wgh1 * s - wgh2 * c, -g.wgh1 * c - wgh2 * s};
  } else {
    cr = cl = cf{-wgh2, 0.0f};
  }
  if (g.wgh2 < 0.0f) {
    cr = cf{-cr.re, -cr.im};
    cl = cf{-cl.re, -cl.im};
  }
}

}  // namespace fr

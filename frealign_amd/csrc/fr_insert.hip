// fr_insert.hip -- CTF-weighted trilinear Fourier insertion, gathered per
// voxel tile.
//
// Replaces PINSERT_TRILIN (pinsert_trilin.f:8-244) as driven by PINSERT
// (pinsert.f:8-195) for a buffer of particles, the A3D3/A3D3_S flush
// (a3d3.f:107-143).  A task is (particle, symmetry copy); its plane points
// are the (I,J) of the full central plane with I^2+J^2 <= IRREC^2, at
// X3 = DM*(I,J), each adding to the voxels L in [max(0,FLOOR(X3)),
// min(NSAM/2,FLOOR(X3+1))] (and likewise M, N within +-NSAM/2) with the
// trilinear weight (1-|X3-L|)(1-|Y3-M|)(1-|Z3-N|).
//
// Instead of scattering every point with float atomics, each workgroup owns
// one 8x8x8 tile of the half-volume and each thread one voxel: the tile first
// lists (in task order) the tasks whose plane passes within reach of it, and
// each voxel then visits, per listed task, the few plane points (I,J) that can
// reach it -- those near its projection onto the plane, found through the
// inverse Gram matrix of the plane's axes -- and re-derives for each the
// reference's FP32 coordinates and voxel range (same operation order, no FMA
// contraction), so exactly the scatter's (point, voxel) pairs contribute,
// with bit-identical terms (the voxel-range test is done as two float
// comparisons per axis, equivalent to the reference's INT() rules; the J of a
// candidate row are bounded per axis by a form linear in I).  Tiles near the
// origin, which every plane crosses, are dispatched first so that they do not
// trail the launch.  Sums run in a fixed order: the volumes, NKSUM
// and TTD are deterministic, and no atomics are needed.  The per-point terms
// (sample x phase ramp x CTF x weight, and CTF^2 x weight) depend only on the
// particle, so they are formed once per particle.
//
// The _SE variant (PINSERT_TRILIN_SE with NNSTAT = 1,
// pinsert_trilin_se.f:186-194) also sums, per voxel, the count, |term|^2,
// |term| and the phase difference to the reference transform of the terms it
// receives; being per voxel, these ride on the same gather for free.
#include <hip/hip_runtime.h>

#include <math.h>

#include "fr_common.h"
#include "fr_kernels.h"

namespace fr {

namespace {

constexpr int kThreads = 256;
constexpr int kTile = 8;            // voxels per tile edge
constexpr int kGatherThreads = 512; // one per voxel of a tile
static_assert(kTile == 8 && kGatherThreads == 512, "the gather maps 8 waves to 4x4x4 sub-cubes");
constexpr int kChunk = 4096;        // tasks listed per pass
constexpr int kReduceThreads = 1024;

// per task, for the candidate search only (never in the sums)
struct InsertGeo {
  float nx, ny, nz, nl1;    // plane normal a x b and its L1 norm
  float h1x, h1y, h1z, r1;  // I = h1 . P for P on the plane; r1 = |h1|_1
  float h2x, h2y, h2z, r2;  // J = h2 . P; r2 = |h2|_1
  float rmax2;              // squared reach of the plane's points (+ margin)
  int valid;
  int zaxes;                // bit k: |b_k| too small for 1/b_k (axis k bounds no J)
  float pad_;
  float ibx, iby, ibz, pad2_;  // 1/b per axis (0 where |b| is tiny)
  // per axis k, the J of a point I a + J b within 1.02 of coordinate v_k:
  // J in v_k / b_k - I a_k / b_k -+ 1.02 / |b_k|
  float sx, sy, sz;         // a_k / b_k
  float ox, oy, oz;         // 1.02 / |b_k| (3e38 where b_k ~ 0)
  float pad3_[2];
};

__host__ __device__ inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

struct Scratch {
  int* first;        // [nb] first task of each particle (or -1)
  InsertGeo* geo;    // [ntask]
  float4* pts;       // [nb][w*w] (samp.re, samp.im, ctfv2, 0)
  unsigned long long* pcount;  // [ntiles][2]
  double* pttd;      // [ntiles][2]
};

Scratch scratch_layout(void* base, int nsam, int nb, int ntask, int ntiles) {
  const int jc = nsam / 2 + 1, w = 2 * jc - 1;
  char* p = static_cast<char*>(base);
  Scratch s;
  size_t o = 0;
  s.first = reinterpret_cast<int*>(p + o);
  o += align256(sizeof(int) * (size_t)nb);
  s.geo = reinterpret_cast<InsertGeo*>(p + o);
  o += align256(sizeof(InsertGeo) * (size_t)ntask);
  s.pts = reinterpret_cast<float4*>(p + o);
  o += align256(sizeof(float4) * (size_t)nb * w * w);
  s.pcount = reinterpret_cast<unsigned long long*>(p + o);
  o += align256(sizeof(unsigned long long) * 2 * (size_t)ntiles);
  s.pttd = reinterpret_cast<double*>(p + o);
  return s;
}

void tile_counts(int nsam, int* tx, int* ty, int* tz) {
  const int jc = nsam / 2 + 1;
  *tx = (jc + kTile - 1) / kTile;
  *ty = (nsam + kTile - 1) / kTile;
  *tz = *ty;
}

__global__ void __launch_bounds__(kThreads) k_first_clear(int nb, int* first) {
  const int p = blockIdx.x * kThreads + threadIdx.x;
  if (p < nb) first[p] = -1;
}

// the first task of each run of tasks of one particle (tasks come grouped by
// particle; a particle's copies share its shift, weight and radius)
__global__ void __launch_bounds__(kThreads)
    k_first_mark(int ntask, int nb, const InsertTask* __restrict__ tasks, int* first) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= ntask) return;
  const int part = tasks[k].part;
  if (part < 0 || part >= nb) return;
  if (k == 0 || tasks[k - 1].part != part) first[part] = k;
}

// pinsert_trilin.f:60-87: the weighted sample and CTF^2 of every plane point
// of a particle, exactly as the scatter formed them
__global__ void __launch_bounds__(kThreads)
    k_insert_points(int nsam, const int* __restrict__ first,
                    const InsertTask* __restrict__ tasks, const float2* __restrict__ datc,
                    const float2* __restrict__ ctff, float4* pts) {
  const int b = blockIdx.y;
  const int k = first[b];
  if (k < 0) return;
  const InsertTask t = tasks[k];
  const int jc = nsam / 2 + 1, w = 2 * jc - 1;
  const size_t ic = (size_t)nsam * jc;
  const int e = blockIdx.x * kThreads + threadIdx.x;
  if (e >= w * w) return;
  const int i = e % w - (jc - 1);
  const int j = e / w - (jc - 1);
  float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
  float g2;
  bool inside;
  {
#pragma clang fp contract(off)
    g2 = (float)(i * i + j * j);
    inside = g2 <= (float)t.irrec2;
  }
  if (inside) {
    const float2* sp = datc + (size_t)t.part * ic;
    const float2* cp = ctff + (size_t)t.part * 2 * ic;
    size_t id;
    bool mate = i < 0;
    if (!mate) id = (size_t)i + (size_t)jc * wrapn(j, nsam);
    else id = (size_t)(-i) + (size_t)jc * wrapn(-j, nsam);
    float2 s = sp[id];
    if (mate) s.y = -s.y;
    float2 cr = cp[id], cl = cp[id + ic];
    cf ctfs = cf{cr.x + cl.x, cr.y - cl.y};  // CTFR + CONJG(CTFL)
    float phase = t.shx * (float)i + t.shy * (float)j;
    float sn, cs;
    sincosf(phase, &sn, &cs);
    float wgt = expf(t.tmp * g2) * t.occ;
    cf samp = cmul(cmul(cf{s.x, s.y}, cf{cs, sn}), ctfs);
    float ctfv2 = 0.25f * (ctfs.re * ctfs.re + ctfs.im * ctfs.im) * wgt;
    samp = cscale(samp, wgt);
    out = make_float4(samp.re, samp.im, ctfv2, 0.f);
  }
  pts[(size_t)b * w * w + e] = out;
}

// per-task plane geometry for the candidate search (double, then rounded;
// the margins below cover the rounding)
__global__ void __launch_bounds__(kThreads)
    k_insert_geo(int ntask, int nsam, const InsertTask* __restrict__ tasks, InsertGeo* geo) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= ntask) return;
  const InsertTask t = tasks[k];
  const double ax = t.dm[0], ay = t.dm[1], az = t.dm[2];
  const double bx = t.dm[3], by = t.dm[4], bz = t.dm[5];
  InsertGeo g;
  const double nx = ay * bz - az * by, ny = az * bx - ax * bz, nz = ax * by - ay * bx;
  const double aa = ax * ax + ay * ay + az * az, bb = bx * bx + by * by + bz * bz;
  const double ab = ax * bx + ay * by + az * bz;
  const double det = aa * bb - ab * ab;
  g.valid = (t.irrec2 >= 0 && det > 1e-12) ? 1 : 0;
  const double id = g.valid ? 1.0 / det : 0.0;
  const double h1x = (bb * ax - ab * bx) * id, h1y = (bb * ay - ab * by) * id,
               h1z = (bb * az - ab * bz) * id;
  const double h2x = (aa * bx - ab * ax) * id, h2y = (aa * by - ab * ay) * id,
               h2z = (aa * bz - ab * az) * id;
  g.nx = (float)nx;
  g.ny = (float)ny;
  g.nz = (float)nz;
  g.nl1 = (float)(fabs(nx) + fabs(ny) + fabs(nz));
  g.h1x = (float)h1x;
  g.h1y = (float)h1y;
  g.h1z = (float)h1z;
  g.r1 = (float)(fabs(h1x) + fabs(h1y) + fabs(h1z));
  g.h2x = (float)h2x;
  g.h2y = (float)h2y;
  g.h2z = (float)h2z;
  g.r2 = (float)(fabs(h2x) + fabs(h2y) + fabs(h2z));
  // points reach |P| <= sqrt(IRREC^2 * lambda_max(G)) within the square
  // |I|,|J| <= NSAM/2; a voxel they touch lies within sqrt(3) more
  const double lmax = 0.5 * (aa + bb + sqrt((aa - bb) * (aa - bb) + 4.0 * ab * ab));
  const double jcm = (double)(nsam / 2);
  const double r2pts = t.irrec2 < 2.0 * jcm * jcm ? (double)t.irrec2 : 2.0 * jcm * jcm;
  const double reach = sqrt(r2pts * lmax) + 1.7320508 + 0.01;
  g.rmax2 = (float)(reach * reach);
  g.pad_ = 0.f;
  const bool zx = !(fabs(bx) > 1e-6), zy = !(fabs(by) > 1e-6), zz = !(fabs(bz) > 1e-6);
  g.zaxes = (zx ? 1 : 0) | (zy ? 2 : 0) | (zz ? 4 : 0);
  g.ibx = zx ? 0.f : (float)(1.0 / bx);
  g.iby = zy ? 0.f : (float)(1.0 / by);
  g.ibz = zz ? 0.f : (float)(1.0 / bz);
  g.pad2_ = 0.f;
  g.sx = zx ? 0.f : (float)(ax / bx);
  g.sy = zy ? 0.f : (float)(ay / by);
  g.sz = zz ? 0.f : (float)(az / bz);
  // an axis with b_k ~ 0 leaves J free: c_k = s_k = 0 and an unbounded o_k
  g.ox = zx ? 3.0e38f : (float)(1.02 / fabs(bx));
  g.oy = zy ? 3.0e38f : (float)(1.02 / fabs(by));
  g.oz = zz ? 3.0e38f : (float)(1.02 / fabs(bz));
  g.pad3_[0] = g.pad3_[1] = 0.f;
  geo[k] = g;
}

// the logical coordinates stored at index s of a wrapped axis: s itself for
// s < NSAM/2, s - NSAM for s > NSAM/2, and both +NSAM/2 and -NSAM/2 at s = NSAM/2
__device__ inline int logical_axis(int s, int nsam, int* v) {
  const int h = nsam / 2;
  if (s < h) { v[0] = s; return 1; }
  if (s > h) { v[0] = s - nsam; return 1; }
  v[0] = h;
  v[1] = -h;
  return 2;
}

// logical intervals covered by the storage range [s0, s1] of a wrapped axis
__device__ inline int logical_ranges(int s0, int s1, int nsam, int* lo, int* hi) {
  const int h = nsam / 2;
  int n = 0;
  if (s0 <= h) {  // the non-negative part s0 .. min(s1, h)
    lo[n] = s0;
    hi[n] = s1 < h ? s1 : h;
    n++;
  }
  if (s1 >= h) {  // the negative part max(s0, h) - nsam .. s1 - nsam
    lo[n] = (s0 > h ? s0 : h) - nsam;
    hi[n] = s1 - nsam;
    n++;
  }
  return n;
}

// does the plane through the origin with normal g reach the box [lo, hi]
// grown by one voxel (the trilinear reach), within the plane's disc?
__device__ inline bool plane_reaches(const InsertGeo& g, float xl, float xh, float yl, float yh,
                                     float zl, float zh) {
  xl -= 1.01f; yl -= 1.01f; zl -= 1.01f;
  xh += 1.01f; yh += 1.01f; zh += 1.01f;
  const float mn = (g.nx > 0.f ? g.nx * xl : g.nx * xh) + (g.ny > 0.f ? g.ny * yl : g.ny * yh) +
                   (g.nz > 0.f ? g.nz * zl : g.nz * zh);
  const float mx = (g.nx > 0.f ? g.nx * xh : g.nx * xl) + (g.ny > 0.f ? g.ny * yh : g.ny * yl) +
                   (g.nz > 0.f ? g.nz * zh : g.nz * zl);
  const float slack = 1e-3f * g.nl1 + 1e-3f;
  if (mn > slack || mx < -slack) return false;
  // nearest point of the box to the origin
  const float cx = xl > 0.f ? xl : (xh < 0.f ? xh : 0.f);
  const float cy = yl > 0.f ? yl : (yh < 0.f ? yh : 0.f);
  const float cz = zl > 0.f ? zl : (zh < 0.f ? zh : 0.f);
  return cx * cx + cy * cy + cz * cz <= g.rmax2;
}

// pinsert_trilin.f:88-111 sets, per axis, S = INT(A1) (minus 1 when A1 < 0)
// with A1 = (X3 + 0.5) - 0.5 and T = INT(A2) (minus 1 when A2 < 0) with
// A2 = X3 + 1, clamped to the axis's range, and fills the voxels S..T.  For
// an integer coordinate v inside the range, v >= S exactly when A1 < v + 1
// (A1 <= v + 1 when v + 1 < 0), and v <= T exactly when A2 >= v (A2 > v when
// v < 0); the clamps never exclude such a v.  The strict forms are the
// non-strict ones against the next float toward +infinity, so each bound is
// one comparison: A1 < range_lo(v) and A2 >= range_hi(v).
__device__ inline float next_up_neg(float x) {  // x < 0
  return __int_as_float(__float_as_int(x) - 1);
}
__device__ inline float range_lo(int v) {
  const float f = (float)(v + 1);
  return v + 1 >= 0 ? f : next_up_neg(f);
}
__device__ inline float range_hi(int v) {
  const float f = (float)v;
  return v >= 0 ? f : next_up_neg(f);
}

// pdiff.f:8-24: |phase of b conj(r)|, 0 where the product is 0
__device__ inline float pdiff_dev(float br, float bi, float2 r) {
  const float cr = br * r.x + bi * r.y, ci = bi * r.x - br * r.y;
  if (cr == 0.0f && ci == 0.0f) return 0.0f;
  return fabsf(atan2f(ci, cr));
}

template <bool kStats>
__global__ void __launch_bounds__(kGatherThreads)
    k_insert_gather(int nsam, int nb, int ntask, int tx, int ty,
                    const InsertTask* __restrict__ tasks, const InsertGeo* __restrict__ geo,
                    const float4* __restrict__ pts, float2* num0, float* ctf2_0, float2* num1,
                    float* ctf2_1, unsigned long long* pcount, double* pttd, InsertStats st) {
  __shared__ int list[kChunk];
  __shared__ int wave_tot[kGatherThreads / 64];
  const int jc = nsam / 2 + 1, w = 2 * jc - 1;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int b = blockIdx.x;
  // tiles go out slab by slab in order of |logical n|, rows in order of
  // |logical m| (storage rows 0, last, 1, last - 1, ...): the tiles near the
  // origin, which every plane crosses, start first instead of last
  const int tmf = (b / tx) % ty, tnf = b / (tx * ty);
  const int l0 = (b % tx) * kTile;
  const int m0 = ((tmf & 1) ? ty - 1 - (tmf >> 1) : (tmf >> 1)) * kTile;
  const int n0 = ((tnf & 1) ? ty - 1 - (tnf >> 1) : (tnf >> 1)) * kTile;
  // the tile's logical boxes (a wrapped axis may split in two)
  const int l1 = min(l0 + kTile - 1, jc - 1);
  int mlo[2], mhi[2], nlo[2], nhi[2];
  const int nmr = logical_ranges(m0, min(m0 + kTile - 1, nsam - 1), nsam, mlo, mhi);
  const int nnr = logical_ranges(n0, min(n0 + kTile - 1, nsam - 1), nsam, nlo, nhi);
  // this thread's voxel: wave w takes the 4x4x4 sub-cube w of the tile, so
  // that a plane leaves whole waves idle (their exec mask empty) rather than
  // most lanes of every wave
  const int l = l0 + 4 * (wave & 1) + (lane & 3);
  const int ms_ = m0 + 4 * ((wave >> 1) & 1) + ((lane >> 2) & 3);
  const int ns_ = n0 + 4 * (wave >> 2) + (lane >> 4);
  const bool mine = l < jc && ms_ < nsam && ns_ < nsam;
  int mv[2], nv[2];
  const int nm = mine ? logical_axis(ms_, nsam, mv) : 0;
  const int nn = mine ? logical_axis(ns_, nsam, nv) : 0;
  // the voxel's range tests as thresholds on X3 (see range_lo, range_hi)
  const float ltl = range_lo(l), utl = range_hi(l);
  float a0r = 0.f, a0i = 0.f, a0c = 0.f, a1r = 0.f, a1i = 0.f, a1c = 0.f;
  unsigned c0 = 0u, c1 = 0u;
  double t0 = 0.0, t1 = 0.0;
  // _SE statistics of this voxel (both halves), against its reference value
  float sv = 0.f, sp = 0.f, sa = 0.f;
  float2 refv = make_float2(0.f, 0.f);
  if (kStats && mine)
    refv = reinterpret_cast<const float2*>(
        st.ref)[(size_t)l + (size_t)jc * ((size_t)ms_ + (size_t)nsam * ns_)];
  for (int base = 0; base < ntask; base += kChunk) {
    __syncthreads();  // the previous chunk's list is no longer read
    const int cnt = min(kChunk, ntask - base);
    // list the tasks reaching this tile, in task order (ballot compaction)
    int nlist = 0;
    for (int k0 = 0; k0 < cnt; k0 += kGatherThreads) {
      const int k = k0 + tid;
      bool hit = false;
      if (k < cnt) {
        const InsertGeo g = geo[base + k];
        if (g.valid)
          for (int a = 0; a < nmr && !hit; a++)
            for (int c = 0; c < nnr && !hit; c++)
              hit = plane_reaches(g, (float)l0, (float)l1, (float)mlo[a], (float)mhi[a],
                                  (float)nlo[c], (float)nhi[c]);
      }
      const unsigned long long bal = __ballot(hit);
      const int before = __popcll(bal & ((1ull << lane) - 1ull));
      if (lane == 0) wave_tot[wave] = __popcll(bal);
      __syncthreads();
      int off = nlist;
      for (int q = 0; q < wave; q++) off += wave_tot[q];
      if (hit) list[off + before] = base + k;
      int tot = 0;
      for (int q = 0; q < kGatherThreads / 64; q++) tot += wave_tot[q];
      __syncthreads();
      nlist += tot;
    }
    // gather this voxel's contributions from the listed tasks
    if (mine) {
      for (int q = 0; q < nlist; q++) {
        // the same task for the whole workgroup: scalar loads of its records
        const int k = __builtin_amdgcn_readfirstlane(list[q]);
        const InsertGeo g = geo[k];
        const InsertTask t = tasks[k];
        if (t.part < 0 || t.part >= nb) continue;
        const float4* tp = pts + (size_t)t.part * w * w;
        for (int am = 0; am < nm; am++)
          for (int an = 0; an < nn; an++) {
            const int m = mv[am], n = nv[an];
            const float fl = (float)l, fm = (float)m, fn = (float)n;
            const float d = g.nx * fl + g.ny * fm + g.nz * fn;
            if (fabsf(d) > 1.01f * g.nl1 + 1e-3f) continue;
            const float ci = g.h1x * fl + g.h1y * fm + g.h1z * fn;
            const float cj = g.h2x * fl + g.h2y * fm + g.h2z * fn;
            int ilo = (int)ceilf(ci - g.r1 - 0.02f), ihi = (int)floorf(ci + g.r1 + 0.02f);
            int jlo = (int)ceilf(cj - g.r2 - 0.02f), jhi = (int)floorf(cj + g.r2 + 0.02f);
            ilo = max(ilo, (int)t.ilo);  // the task's I stripe, within -(JC-1)..JC-1
            ihi = min(ihi, (int)t.ihi);
            jlo = max(jlo, -(jc - 1));
            jhi = min(jhi, jc - 1);
            // per axis, the J interval of the points within 1.02 of the
            // voxel, linear in I: [c_k - I s_k - o_k, c_k - I s_k + o_k]
            // (an axis with b_k ~ 0 bounds no J: tested per I below)
            const float cx = (g.zaxes & 1) ? 0.f : fl * g.ibx;
            const float cy = (g.zaxes & 2) ? 0.f : fm * g.iby;
            const float cz = (g.zaxes & 4) ? 0.f : fn * g.ibz;
            const float ltm_ = range_lo(m), utm_ = range_hi(m), ltn_ = range_lo(n), utn_ = range_hi(n);
            for (int i = ilo; i <= ihi; i++) {
              const float fi = (float)i;
              const float ex = cx - fi * g.sx, ey = cy - fi * g.sy, ez = cz - fi * g.sz;
              const float jl = fmaxf(fmaxf((float)jlo, ex - g.ox), fmaxf(ey - g.oy, ez - g.oz));
              const float jh = fminf(fminf((float)jhi, ex + g.ox), fminf(ey + g.oy, ez + g.oz));
              if (g.zaxes) {
                bool none = false;
                const float slack = 1e-3f * (float)jc;
                if ((g.zaxes & 1) && fabsf(fl - fi * t.dm[0]) > 1.02f + slack * fabsf(t.dm[3]))
                  none = true;
                if ((g.zaxes & 2) && fabsf(fm - fi * t.dm[1]) > 1.02f + slack * fabsf(t.dm[4]))
                  none = true;
                if ((g.zaxes & 4) && fabsf(fn - fi * t.dm[2]) > 1.02f + slack * fabsf(t.dm[5]))
                  none = true;
                if (none) continue;
              }
              const int j0 = max((int)ceilf(jl - 0.01f), jlo);
              const int j1 = min((int)floorf(jh + 0.01f), jhi);
              float xi, yi, zi;
              {
#pragma clang fp contract(off)
                xi = t.dm[0] * fi;
                yi = t.dm[1] * fi;
                zi = t.dm[2] * fi;
              }
              const int ii = i * i;
              for (int j = j0; j <= j1; j++) {
                // pinsert_trilin.f:88-111 in the reference's FP32 order: the
                // point's X3 and whether this voxel is in its range
                float x3, y3, z3;
                bool in;
                {
#pragma clang fp contract(off)
                  const float fj = (float)j;
                  x3 = xi + t.dm[3] * fj;
                  y3 = yi + t.dm[4] * fj;
                  z3 = zi + t.dm[5] * fj;
                  in = ((x3 + 0.5f) - 0.5f < ltl) & (x3 + 1.0f >= utl) &
                       ((y3 + 0.5f) - 0.5f < ltm_) & (y3 + 1.0f >= utm_) &
                       ((z3 + 0.5f) - 0.5f < ltn_) & (z3 + 1.0f >= utn_);
                }
                // I^2 + J^2 <= IRREC^2 (exact in FP32 at these sizes)
                if (!in || ii + j * j > t.irrec2) continue;
                const float b3 = 1.0f - fabsf(z3 - fn);
                const float b2 = (1.0f - fabsf(y3 - fm));
                const float box = (1.0f - fabsf(x3 - fl)) * b2 * b3;
                const float4 pv = tp[(size_t)(j + jc - 1) * w + (i + jc - 1)];
                const float cv = pv.z * box;
                if (kStats) {
                  const float br = pv.x * box, bi = pv.y * box;
                  const float ab = hypotf(br, bi);
                  sv += ab * ab;
                  sa += ab;
                  sp += pdiff_dev(br, bi, refv);
                }
                if (t.half) {
                  a1r += pv.x * box;
                  a1i += pv.y * box;
                  a1c += cv;
                  c1++;
                  t1 += (double)cv;
                } else {
                  a0r += pv.x * box;
                  a0i += pv.y * box;
                  a0c += cv;
                  c0++;
                  t0 += (double)cv;
                }
              }
            }
          }
      }
    }
  }
  if (mine) {
    const size_t v = (size_t)l + (size_t)jc * ((size_t)ms_ + (size_t)nsam * ns_);
    if (c0) {
      float2 o = num0[v];
      num0[v] = make_float2(o.x + a0r, o.y + a0i);
      ctf2_0[v] += a0c;
    }
    if (c1) {
      float2 o = num1[v];
      num1[v] = make_float2(o.x + a1r, o.y + a1i);
      ctf2_1[v] += a1c;
    }
    if (kStats && (c0 | c1)) {
      st.ksum[v] += (int)(c0 + c1);
      st.vsum[v] += sv;
      st.psum[v] += sp;
      st.asum[v] += sa;
    }
  }
  // the tile's NKSUM and TTD: a fixed reduction order
  __shared__ unsigned long long rc[kGatherThreads / 64][2];
  __shared__ double rt[kGatherThreads / 64][2];
  unsigned long long u0 = c0, u1 = c1;
  for (int off = 32; off > 0; off >>= 1) {
    u0 += __shfl_xor(u0, off, 64);
    u1 += __shfl_xor(u1, off, 64);
    t0 += __shfl_xor(t0, off, 64);
    t1 += __shfl_xor(t1, off, 64);
  }
  if (lane == 0) {
    rc[wave][0] = u0;
    rc[wave][1] = u1;
    rt[wave][0] = t0;
    rt[wave][1] = t1;
  }
  __syncthreads();
  if (tid < 2) {
    unsigned long long cc = 0;
    double tt = 0.0;
    for (int q = 0; q < kGatherThreads / 64; q++) {
      cc += rc[q][tid];
      tt += rt[q][tid];
    }
    pcount[2 * (size_t)b + tid] = cc;
    pttd[2 * (size_t)b + tid] = tt;
  }
}

// NKSUM and TTD of the launch: tiles summed in a fixed order
__global__ void __launch_bounds__(kReduceThreads)
    k_insert_reduce(int ntiles, const unsigned long long* __restrict__ pcount,
                    const double* __restrict__ pttd, unsigned long long* nksum, double* ttd) {
  __shared__ unsigned long long rc[kReduceThreads / 64][2];
  __shared__ double rt[kReduceThreads / 64][2];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  unsigned long long u0 = 0, u1 = 0;
  double t0 = 0.0, t1 = 0.0;
  for (int b = tid; b < ntiles; b += kReduceThreads) {
    u0 += pcount[2 * (size_t)b];
    u1 += pcount[2 * (size_t)b + 1];
    t0 += pttd[2 * (size_t)b];
    t1 += pttd[2 * (size_t)b + 1];
  }
  for (int off = 32; off > 0; off >>= 1) {
    u0 += __shfl_xor(u0, off, 64);
    u1 += __shfl_xor(u1, off, 64);
    t0 += __shfl_xor(t0, off, 64);
    t1 += __shfl_xor(t1, off, 64);
  }
  if (lane == 0) {
    rc[wave][0] = u0;
    rc[wave][1] = u1;
    rt[wave][0] = t0;
    rt[wave][1] = t1;
  }
  __syncthreads();
  if (tid < 2) {
    unsigned long long cc = 0;
    double tt = 0.0;
    for (int q = 0; q < kReduceThreads / 64; q++) {
      cc += rc[q][tid];
      tt += rt[q][tid];
    }
    if (cc) {
      nksum[tid] += cc;
      ttd[tid] += tt;
    }
  }
}

}  // namespace

size_t insert_scratch_bytes(int nsam, int nb, int ntask) {
  const int jc = nsam / 2 + 1, w = 2 * jc - 1;
  int tx, ty, tz;
  tile_counts(nsam, &tx, &ty, &tz);
  const size_t ntiles = (size_t)tx * ty * tz;
  return align256(sizeof(int) * (size_t)nb) + align256(sizeof(InsertGeo) * (size_t)ntask) +
         align256(sizeof(float4) * (size_t)nb * w * w) +
         align256(sizeof(unsigned long long) * 2 * ntiles) + align256(sizeof(double) * 2 * ntiles);
}

int launch_insert(Ctx& c, int nsam, int nb, int ntask, const InsertTask* d_tasks,
                  const float* d_datc, const float* d_ctff, float* num0, float* ctf2_0,
                  float* num1, float* ctf2_1, unsigned long long* d_nksum, double* d_ttd,
                  void* scratch, const InsertStats* stats) {
  if (ntask <= 0 || nb <= 0) return 0;
  FR_CHECK(scratch != nullptr, "insertion scratch missing");
  const int jc = nsam / 2 + 1, w = 2 * jc - 1;
  int tx, ty, tz;
  tile_counts(nsam, &tx, &ty, &tz);
  const int ntiles = tx * ty * tz;
  Scratch s = scratch_layout(scratch, nsam, nb, ntask, ntiles);
  hipLaunchKernelGGL(k_first_clear, dim3((nb + kThreads - 1) / kThreads), dim3(kThreads), 0,
                     c.stream, nb, s.first);
  FR_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_first_mark, dim3((ntask + kThreads - 1) / kThreads), dim3(kThreads), 0,
                     c.stream, ntask, nb, d_tasks, s.first);
  FR_HIP(hipGetLastError());
  FR_CHECK(nb <= 65535, "too many particles in one insertion (%d)", nb);
  hipLaunchKernelGGL(k_insert_points, dim3((w * w + kThreads - 1) / kThreads, nb), dim3(kThreads),
                     0, c.stream, nsam, s.first, d_tasks, reinterpret_cast<const float2*>(d_datc),
                     reinterpret_cast<const float2*>(d_ctff), s.pts);
  FR_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_insert_geo, dim3((ntask + kThreads - 1) / kThreads), dim3(kThreads), 0,
                     c.stream, ntask, nsam, d_tasks, s.geo);
  FR_HIP(hipGetLastError());
  if (stats) {
    FR_CHECK(stats->ref && stats->ksum && stats->vsum && stats->psum && stats->asum,
             "insertion statistics arrays missing");
    hipLaunchKernelGGL(k_insert_gather<true>, dim3(ntiles), dim3(kGatherThreads), 0, c.stream,
                       nsam, nb, ntask, tx, ty, d_tasks, s.geo, s.pts,
                       reinterpret_cast<float2*>(num0), ctf2_0, reinterpret_cast<float2*>(num1),
                       ctf2_1, s.pcount, s.pttd, *stats);
  } else {
    hipLaunchKernelGGL(k_insert_gather<false>, dim3(ntiles), dim3(kGatherThreads), 0, c.stream,
                       nsam, nb, ntask, tx, ty, d_tasks, s.geo, s.pts,
                       reinterpret_cast<float2*>(num0), ctf2_0, reinterpret_cast<float2*>(num1),
                       ctf2_1, s.pcount, s.pttd, InsertStats{});
  }
  FR_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_insert_reduce, dim3(1), dim3(kReduceThreads), 0, c.stream, ntiles,
                     s.pcount, s.pttd, d_nksum, d_ttd);
  FR_HIP(hipGetLastError());
  return 0;
}

}  // namespace fr

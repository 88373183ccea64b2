// fr_lists.hip -- per-particle candidate lists of the learned search and the
// device side of the FP32 rescoring, so that a batch goes from the proposal
// network's classes to its winning orientations without a host round trip.
//
// A particle whose top class probability reaches the threshold searches the
// cells of its proposed classes (each cell: the bank orientations nearest to
// that class, in groups of `group` in-plane candidates of one view, -1
// padded); any other particle searches the whole bank in bank order, in the
// bank's own view groups, which is the exhaustive search of psearch.f:73-102.
// Lists are ragged: particle p's entries are cand[loff[p] .. loff[p+1]).
//
// The rescoring items are what LMAIN hands to PREFINE for each candidate
// (lmain.f:615-642): the candidate's angles (as the rotation matrix of the
// bank entry) and its shift on the NSAMR grid, SX = IXM/NSAMR*2pi
// (ccp.f:254-256); the winner is the first maximum of the rescored CC3M
// (strict '>', lmain.f:634).
#include <hip/hip_runtime.h>

#include "fr_common.h"
#include "fr_lists.h"

namespace fr {

namespace {

constexpr int kScan = 1024;  // threads of the single-workgroup scans
constexpr int kFill = 256;
constexpr int kTile = 256;   // template columns per GEMM tile
constexpr float kPi = 3.1415926535897f;

// Exclusive prefix sum of n values produced by val(i), written to out[0..n]
// (out[n] = total).  One workgroup of kScan threads; each thread owns a
// contiguous run of indices.
template <typename F>
__device__ void block_scan(int n, F val, int* out) {
  __shared__ long long part[kScan];
  const int tid = threadIdx.x;
  const int per = (n + kScan - 1) / kScan;
  const int i0 = tid * per, i1 = min(n, i0 + per);
  long long s = 0;
  for (int i = i0; i < i1; i++) s += val(i);
  part[tid] = s;
  __syncthreads();
  for (int off = 1; off < kScan; off <<= 1) {
    long long v = tid >= off ? part[tid - off] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  long long run = part[tid] - s;  // exclusive start of this thread's run
  for (int i = i0; i < i1; i++) {
    out[i] = (int)run;
    run += val(i);
  }
  if (tid == kScan - 1) out[n] = (int)part[kScan - 1];
}

// Stream compaction by the same contiguous runs: out[1 + k] = the k-th index
// i < n (ascending) with keep(i), out[0] = their count.
template <typename F>
__device__ void block_compact(int n, F keep, int* out) {
  __shared__ int cpart[kScan];
  const int tid = threadIdx.x;
  const int per = (n + kScan - 1) / kScan;
  const int i0 = tid * per, i1 = min(n, i0 + per);
  int s = 0;
  for (int i = i0; i < i1; i++) s += keep(i) ? 1 : 0;
  cpart[tid] = s;
  __syncthreads();
  for (int off = 1; off < kScan; off <<= 1) {
    int v = tid >= off ? cpart[tid - off] : 0;
    __syncthreads();
    cpart[tid] += v;
    __syncthreads();
  }
  int run = cpart[tid] - s;
  for (int i = i0; i < i1; i++)
    if (keep(i)) out[1 + run++] = i;
  if (tid == kScan - 1) out[0] = cpart[kScan - 1];
}

// List length, group and confidence of every particle, and the list offsets.
__global__ void __launch_bounds__(kScan)
    k_list_sizes(int np, int ktop, const int* __restrict__ top, const float* __restrict__ prob,
                 float threshold, int ncls, const int* __restrict__ cell_off, int group, int nbank,
                 int bank_group, int maxlen, int* len, int* loff, int* grp, int* info) {
  for (int p = threadIdx.x; p < np; p += kScan) {
    int l = 0;
    bool conf = prob[(size_t)p * ktop] >= threshold;
    if (conf) {
      for (int q = 0; q < ktop; q++) {
        int c = top[(size_t)p * ktop + q];
        if (c >= 0 && c < ncls) l += max(cell_off[c + 1] - cell_off[c], 0);
      }
      if (l == 0) conf = false;  // no cell to search: fall back
    }
    int g = conf ? group : bank_group;
    if (!conf) l = nbank;
    if (l > maxlen) l = maxlen / g * g;  // whole groups within the list capacity
    len[p] = l;
    grp[p] = g;
    info[2 * p] = conf ? 1 : 0;
  }
  __syncthreads();
  block_scan(np, [&](int i) { return len[i]; }, loff);
}

// Entries of each list: the proposed classes' cells in proposal order, or
// the bank in order; info[2p+1] counts the non-padding entries.
__global__ void __launch_bounds__(kFill)
    k_list_fill(int ktop, const int* __restrict__ top, int ncls, const int* __restrict__ cell_off,
                const int* __restrict__ cells, const int* __restrict__ loff, int* info,
                int* cand) {
  __shared__ int red[kFill / 64];
  const int p = blockIdx.x, tid = threadIdx.x;
  const int base = loff[p], len = loff[p + 1] - base;
  int valid = 0;
  if (info[2 * p]) {
    int pos = 0;
    for (int q = 0; q < ktop && pos < len; q++) {
      int c = top[(size_t)p * ktop + q];
      if (c < 0 || c >= ncls) continue;
      const int c0 = cell_off[c];
      int n = max(cell_off[c + 1] - c0, 0);
      if (n > len - pos) n = len - pos;
      for (int e = tid; e < n; e += kFill) {
        int o = cells[c0 + e];
        cand[base + pos + e] = o;
        valid += o >= 0;
      }
      pos += n;
    }
  } else {
    for (int e = tid; e < len; e += kFill) cand[base + e] = e;
    valid = tid < len ? (len - tid + kFill - 1) / kFill : 0;
  }
  for (int off = 32; off > 0; off >>= 1) valid += __shfl_down(valid, off, 64);
  if ((tid & 63) == 0) red[tid >> 6] = valid;
  __syncthreads();
  if (tid == 0) info[2 * p + 1] = red[0] + red[1] + red[2] + red[3];
}

// GEMM tile offsets of ragged lists: ceil(len / 256) tiles per particle (a
// decreasing offset pair counts as an empty list).  With `info` (the
// confidence flags of k_list_sizes) the particles that fall back to the
// whole bank get no tiles here; they are listed in fb instead (fb[0] = their
// count, fb[1..] the particles in order) for the shared-template search.
// With `ord` the tiles are laid out in the order of the particles' first
// list entries (their most probable class's first cell; ties by index):
// ord[q] is the particle whose tiles come q-th, so that particles proposing
// the same classes are searched at the same time and share their template
// reads in the caches.  Batches above kOrdMax particles keep index order.
constexpr int kOrdMax = 4096;

__global__ void __launch_bounds__(kScan)
    k_tile_offsets(int np, const int* __restrict__ loff, const int* __restrict__ info, int* toff,
                   int* fb, const int* __restrict__ cand, int* ord) {
  __shared__ long long skey[kOrdMax];
  const int tid = threadIdx.x;
  auto tiles = [&](int i) {
    if (info && !info[2 * i]) return 0;
    int n = loff[i + 1] - loff[i];
    return n > 0 ? (n + kTile - 1) / kTile : 0;
  };
  const bool sorted = ord && cand && np <= kOrdMax;
  if (sorted) {
    int n2 = 1;
    while (n2 < np) n2 <<= 1;
    for (int i = tid; i < n2; i += kScan) {
      long long key = 0x7fffffffffffffffLL;
      if (i < np) {
        const int k0 = tiles(i) > 0 ? cand[loff[i]] : 0x7fffffff;
        key = (long long)k0 * 4294967296LL + i;
      }
      skey[i] = key;
    }
    __syncthreads();
    // bitonic sort, ascending
    for (int k = 2; k <= n2; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = tid; i < n2; i += kScan) {
          const int l = i ^ j;
          if (l > i) {
            const long long x = skey[i], y = skey[l];
            if ((x > y) == ((i & k) == 0)) {
              skey[i] = y;
              skey[l] = x;
            }
          }
        }
        __syncthreads();
      }
    for (int q = tid; q < np; q += kScan) ord[q] = (int)(skey[q] & 0xffffffffLL);
  } else if (ord) {
    for (int q = tid; q < np; q += kScan) ord[q] = q;
  }
  block_scan(np, [&](int q) { return tiles(sorted ? (int)(skey[q] & 0xffffffffLL) : q); }, toff);
  // the fallback particles (not confident, non-empty list) in index order
  if (info && fb)
    block_compact(np, [&](int i) { return !info[2 * i] && loff[i + 1] > loff[i]; }, fb);
}

// CC3M items of the top-k candidates: pidx, rotation matrix, shift, AMAG.
// An empty candidate (-1) is scored at bank entry 0 with zero shift and
// ignored by the winner pick.
__global__ void __launch_bounds__(kFill)
    k_cand_items(int np, int k, const int* __restrict__ top_o, const int* __restrict__ top_row,
                 int nbank, const float* __restrict__ dm_bank, int nsamr, int* pidx, float* dm,
                 float* shift, float* amag) {
  const int i = blockIdx.x * kFill + threadIdx.x;
  if (i >= np * k) return;
  const int o = top_o[i];
  const bool ok = o >= 0 && o < nbank;
  const int oc = ok ? o : 0;
  const int row = ok ? top_row[i] : 7 * 15 + 7;
  for (int e = 0; e < 9; e++) dm[9 * (size_t)i + e] = dm_bank[9 * (size_t)oc + e];
  // ccp.f:254-256 on the 15x15 grid of -7..7 steps
  const float f = kPi * 2.0f;
  const int sx = row / 15 - 7, sy = row % 15 - 7;
  shift[2 * (size_t)i] = ((float)sx / (float)nsamr) * f;
  shift[2 * (size_t)i + 1] = ((float)sy / (float)nsamr) * f;
  pidx[i] = i / k;
  amag[i] = 1.0f;
}

// Winner per particle: the first maximum of the rescored CC3M over its valid
// candidates (scores of empty candidates become -inf).  One wave per
// particle, a lane per candidate (k <= 64): the wave maximum, then the lowest
// lane holding it (lane 0 when every score is -inf).
__global__ void __launch_bounds__(256)
    k_best_cand(int np, int k, const int* __restrict__ top_o, float* score, int* best) {
  const int p = blockIdx.x * 4 + (threadIdx.x >> 6), q = threadIdx.x & 63;
  if (p >= np) return;  // uniform across the wave
  const float ninf = -__builtin_inff();
  float v = ninf;
  if (q < k) {
    const size_t i = (size_t)p * k + q;
    v = top_o[i] >= 0 ? score[i] : ninf;
    score[i] = v;
  }
  float m = v;
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  const unsigned long long hit = __ballot(q < k && v == m);
  if (q == 0) best[p] = hit ? __ffsll((unsigned long long)hit) - 1 : 0;
}

}  // namespace

int launch_list_build(Ctx& c, int np, int ktop, const int* top, const float* prob, float threshold,
                      int ncls, const int* cell_off, const int* cells, int group, int nbank,
                      int bank_group, int maxlen, int* len_scratch, int* loff, int* grp, int* info,
                      int* cand) {
  hipLaunchKernelGGL(k_list_sizes, dim3(1), dim3(kScan), 0, c.stream, np, ktop, top, prob,
                     threshold, ncls, cell_off, group, nbank, bank_group, maxlen, len_scratch,
                     loff, grp, info);
  FR_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_list_fill, dim3(np), dim3(kFill), 0, c.stream, ktop, top, ncls, cell_off,
                     cells, loff, info, cand);
  FR_HIP(hipGetLastError());
  return 0;
}

int launch_tile_offsets(Ctx& c, int np, const int* loff, int* toff, const int* info, int* fb,
                        const int* cand, int* ord) {
  hipLaunchKernelGGL(k_tile_offsets, dim3(1), dim3(kScan), 0, c.stream, np, loff, info, toff, fb,
                     cand, ord);
  FR_HIP(hipGetLastError());
  return 0;
}

int launch_cand_items(Ctx& c, int np, int k, const int* top_o, const int* top_row, int nbank,
                      const float* dm_bank, int nsamr, int* pidx, float* dm, float* shift,
                      float* amag) {
  const int n = np * k;
  hipLaunchKernelGGL(k_cand_items, dim3((n + kFill - 1) / kFill), dim3(kFill), 0, c.stream, np, k,
                     top_o, top_row, nbank, dm_bank, nsamr, pidx, dm, shift, amag);
  FR_HIP(hipGetLastError());
  return 0;
}

int launch_best_cand(Ctx& c, int np, int k, const int* top_o, float* score, int* best) {
  FR_CHECK(k >= 1 && k <= 64, "at most 64 candidates per particle (got %d)", k);
  hipLaunchKernelGGL(k_best_cand, dim3((np + 3) / 4), dim3(256), 0, c.stream, np, k, top_o, score,
                     best);
  FR_HIP(hipGetLastError());
  return 0;
}

}  // namespace fr

// fr_ctx.h -- the library context behind the opaque fr_ctx handle.
#pragma once
#include <map>
#include <tuple>
#include <utility>
#include <vector>

#include "fr_ccp.h"
#include "fr_common.h"
#include "fr_kernels.h"

struct fr_ctx : public fr::Ctx {
  std::vector<fr::InsertTask> host_tasks;
  static constexpr int kSlots = 36;
  void* slots[kSlots] = {};
  size_t slot_bytes[kSlots] = {};
  void* slot(int k, size_t bytes, int* err);
  // grow-only pinned host staging (one buffer), for the small per-round
  // transfers of the lock-stepped minimisations
  void* hpin = nullptr;
  size_t hpin_bytes = 0;
  void* host_pinned(size_t bytes, int* err);
  // search annulus and shift-phase tables, cached per (NSAM, MAXR1, MAXR2)
  int skey[3] = {0, 0, 0};
  int kc = 0, kpad = 0, nsamr = 0;
  int* d_kidx = nullptr;
  float* d_shx = nullptr;
  float* d_shy = nullptr;
  // constant tables of the fused CC3M, built on first use and kept until the
  // context is destroyed (captured graphs hold their addresses)
  std::map<int, float*> dft_tabs;                        // per NSAM
  std::map<std::pair<int, int>, int*> pts_tabs;          // per (MAXR1^2, MAXR2^2)
  struct RingTab {
    float* d = nullptr;
    int cnt = 0;
  };
  std::map<std::tuple<int, float, float>, RingTab> ring_tabs;  // per (NSAM, RI2, RI3)
  // per (NSAM, RI2, RI3, RIH, HALFW)
  std::map<std::tuple<int, float, float, float, float>, float*> wmap_tabs;
  int cc3m_tables(const fr::Cc3mArgs& a, fr::Cc3mTables* t);
  // launch choices per box size (set by fr_autotune_cc3m)
  struct Tune {
    int cc3m_fused = 1;   // band-pruned fused CC3M where it applies
    int cc3m_blocks = -1;  // persistent workgroups (0: one per item; -1: one full wave)
  };
  std::map<int, Tune> tune;
  Tune tune_for(int nsam) const;
  // constant tables of the exact CCP search, per (NSAM, MAXR2), kept until
  // the context is destroyed
  std::map<std::pair<int, int>, fr::CcpTablesDev> ccp_tabs;
  int ccp_tables(int nsam, int maxr2, fr::CcpTablesDev* t);
};

namespace fr {
// fr_cc3m_batch (cc3 = 0) and fr_cc3_batch (cc3 = 1)
int cc3m_batch(fr_ctx* c, int nsam, int n, const float* d_a3df, const float* d_spec,
               const float* d_ctff, const int* d_pidx, const float* d_dm, const float* d_shift,
               const float* d_amag, int maxr1, int maxr2, float rbfact, const float* d_fsct,
               const float* ring, float* d_score, int cc3);
// fr_refine_batch with the symmetry matrix SM applied to every DM (null: identity)
int refine_items(fr_ctx* c, int nsam, int n, const float* d_a3df, const float* d_spec,
                 const float* d_ctff, const int* pidx, const float* amag, float* par,
                 const int* mask, int maxr1, int maxr2, float rbfact, const float* d_fsct,
                 const float* ring, const float* prior, int maxit, const float* sm,
                 int* nevals);
}  // namespace fr

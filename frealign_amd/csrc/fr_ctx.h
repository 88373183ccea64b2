// fr_ctx.h -- the library context behind the opaque fr_ctx handle.
#pragma once
#include <map>
#include <vector>

#include "fr_common.h"
#include "fr_kernels.h"

struct fr_ctx : public fr::Ctx {
  std::vector<fr::InsertTask> host_tasks;
  static constexpr int kSlots = 24;
  void* slots[kSlots] = {};
  size_t slot_bytes[kSlots] = {};
  void* slot(int k, size_t bytes, int* err);
  // search annulus and shift-phase tables, cached per (NSAM, MAXR1, MAXR2)
  int skey[3] = {0, 0, 0};
  int kc = 0, kpad = 0;
  int* d_kidx = nullptr;
  float* d_shx = nullptr;
  float* d_shy = nullptr;
  // DFT table of the fused CC3M, cached per NSAM
  int dft_nsam = 0;
  float* d_dft = nullptr;
  int ensure_dft(int nsam);
  // launch choices per box size (set by fr_autotune_cc3m)
  struct Tune {
    int cc3m_fused = 1;   // band-pruned fused CC3M where it applies
    int cc3m_blocks = 0;  // persistent workgroups (0: one per item)
  };
  std::map<int, Tune> tune;
  Tune tune_for(int nsam) const;
};

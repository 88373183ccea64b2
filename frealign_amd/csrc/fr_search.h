// fr_search.h -- launchers of the MFMA orientation x translation search.
#pragma once
#include "fr_common.h"

namespace fr {
size_t search_gemm_lds(bool bf16);
int launch_bank_build(Ctx& c, int nsam, int norient, int kc, int kpad, const float* a3df,
                      const float* dm, const int* kidx, void* bank, bool bf16);
int launch_search_feats(Ctx& c, int nsam, int np, int kc, int kpad, int maxr22, float rbfact,
                        const float* outc, const float* ctff, const int* kidx, float* ap,
                        float* c2, float* sum2);
int launch_search_norm(Ctx& c, int norient, int kpad, int np, const void* bank, bool bf16,
                       const float* c2, float* nrm);
// nrm (np x nrow) null: norms accumulated inside the GEMM tiles
int launch_search_gemm(Ctx& c, int nrow, int nbank, int kpad, int np, const void* bank,
                       bool bf16, const int* cand, const float* ap, const float* c2,
                       const float* nrm, const float* sum2, const float* shx_tab,
                       const float* shy_tab, float* score, int* best_row);
int launch_search_gemm_ragged(Ctx& c, int np, int nbank, int kpad, const void* bank, bool bf16,
                              const int* loff, const int* toff, const int* cand, int cap,
                              int max_tiles, const float* ap, const float* c2, const float* sum2,
                              const float* shx_tab, const float* shy_tab, float* score,
                              int* best_row);
// the separable BF16 search (fr_search_sep.hip): bank, particle terms and
// the search over uniform lists (loff null: nrow entries per particle, cand
// null for the whole bank) or ragged lists (loff, toff, cap)
constexpr int kSepMaxR2 = 32;     // the separable grid holds I < 32, |J| < 32
constexpr int kSepRow = 4096;     // bf16 values per template of the separable bank
int launch_bank_build_sep(Ctx& c, int nsam, int norient, int lo2, int hi2, const float* a3df,
                          const float* dm, void* bank);
int launch_sep_feats(Ctx& c, int nsam, int np, int lo2, int hi2, float rbfact, const float* outc,
                     const float* ctff, void* xa, void* xb, float* c2s);
// fb (ragged lists only, may be null): the fallback particles (fb[0] = count,
// fb[1..] particles) whose lists are the whole bank in order; they are
// searched four to a workgroup against shared templates, as the uniform form,
// in the same persistent launch as the listed particles' tiles.  ord (may be
// null): the particle of each toff slot (launch_tile_offsets' order)
int launch_search_sep(Ctx& c, int np, int nbank, int nsamr, int nrow, const int* loff,
                      const int* toff, int cap, int max_tiles, const int* cand, const void* bank,
                      const void* xa, const void* xb, const float* c2s, const float* sum2,
                      float* score, int* best_row, const int* fb = nullptr,
                      const int* ord = nullptr);
int launch_topk(Ctx& c, int nbank, int np, int n, int group, const int* loff, const int* grp,
                const int* cand, int cap, int topk, const float* score, const int* best_row,
                int* top_o, int* top_row, float* top_s);
}  // namespace fr

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
int launch_topk(Ctx& c, int np, int n, int group, const int* loff, const int* grp,
                const int* cand, int cap, int topk, const float* score, const int* best_row,
                int* top_o, int* top_row, float* top_s);
}  // namespace fr

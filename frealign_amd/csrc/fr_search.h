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
int launch_search_gemm(Ctx& c, int nrow, int nbank, int kpad, int np, const void* bank,
                       bool bf16, const int* cand, const float* ap, const float* nrm, const float* sum2,
                       const float* shx_tab, const float* shy_tab, float* score, int* best_row,
                       float* raw_peak);
int launch_search_norm_list(Ctx& c, int ncand, int nbank, int kpad, int np, const void* bank,
                            bool bf16, const int* cand, const float* c2, float* nrm);
int launch_map_cand(Ctx& c, int np, int ncand, int topk, const int* cand, int* top_o);
int launch_topk(Ctx& c, int np, int norient, int group, int topk, const float* score, const int* best_row,
                int* top_o, int* top_row, float* top_s);
}  // namespace fr

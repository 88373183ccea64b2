// fr_lists.h -- launchers of the ragged candidate lists and rescoring items.
#pragma once
#include "fr_common.h"

namespace fr {
int launch_list_build(Ctx& c, int np, int ktop, const int* top, const float* prob, float threshold,
                      int ncls, const int* cell_off, const int* cells, int group, int nbank,
                      int bank_group, int maxlen, int* len_scratch, int* loff, int* grp, int* info,
                      int* cand);
// toff[0..np] tile offsets of the lists; with info (k_list_sizes' flags) the
// fallback particles get no tiles and are listed in fb (fb[0] = count); with
// cand and ord the tiles follow the particles ord[0..np) (sorted by their
// first list entry), toff[q] being the first tile of particle ord[q]
int launch_tile_offsets(Ctx& c, int np, const int* loff, int* toff, const int* info = nullptr,
                        int* fb = nullptr, const int* cand = nullptr, int* ord = nullptr);
int launch_cand_items(Ctx& c, int np, int k, const int* top_o, const int* top_row, int nbank,
                      const float* dm_bank, int nsamr, int* pidx, float* dm, float* shift,
                      float* amag);
int launch_best_cand(Ctx& c, int np, int k, const int* top_o, float* score, int* best);
}  // namespace fr

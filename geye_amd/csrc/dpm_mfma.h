// dpm_mfma.h -- MFMA convolution launcher (dpm_mfma.hip).
#pragma once

#include "dpm_internal.h"

namespace geye {

// LDS bytes of the per-class kernel with nbuf filter buffers and tw-column tiles (split: 8
// chunks per cell).
size_t mfma_lds_bytes(int fh, int fw, int nt, bool split, int nbuf, int tw);
// whether a class runs 8-column tiles (the split forms, filters at most 8 wide, where that
// tile fits the LDS)
bool mfma_wide_tiles(int fh, int fw, int nt, bool split);
// Work-queue form: one launch for every tile of every filter class (tiles carry their class
// shape), `blocks` persistent workgroups taking tiles from *counter (0 on entry) until the
// list is exhausted; tiles are 4 columns x 16*MT rows.  lds = the largest class's need.
size_t mfma_q_lds_bytes(int fh, int fw, int nt, int MT, bool split);
int launch_conv_mfma_queue(hipStream_t s, int kind, bool split, int MT, int blocks, size_t lds,
                           const MfmaTile *tiles, int ntiles, int *counter, const void *feat,
                           const void *filt, const long long *map_off, const int2 *map_dims,
                           float *out);
// kind: 1 = bf16, 2 = fp16; split: hi/lo operand pairs, three MFMAs per tap.  tiles4 are the
// class's 4-column tiles, tiles8 (optional) the same outputs as 8-column tiles, taken where
// mfma_wide_tiles says so.  The persistent grid takes eighths / 8 of the workgroups the
// device holds at once (8: all).  Returns 0 on success, -1 when the shape does not fit.
int launch_conv_mfma(hipStream_t s, int kind, bool split, int nt, const MfmaTile *tiles4,
                     int ntiles4, const MfmaTile *tiles8, int ntiles8, int fh, int fw,
                     const void *feat, const void *filt, const long long *map_off,
                     const int2 *map_dims, float *out, int eighths = 8);

}  // namespace geye

// dpm_ascore.h -- accumulated-score map launcher.
#pragma once

#include "dpm_internal.h"

namespace geye {

// ac: IHEI x IWID column-major, read (the caller's prior values) and overwritten with the
// normalised accumulated score; maxbits must hold 0 on entry.
void launch_ascore(hipStream_t s, const ScoreMap *maps, int nmaps, const int *tab,
                   const float *score, int IWID, int IHEI, float thresh, float *ac,
                   unsigned int *maxbits);

}  // namespace geye

// dpm_rects.h -- launchers of the post-NMS box refinement and the part-score subtraction.
#pragma once

#include "dpm_internal.h"

namespace geye {

void launch_new_rects(hipStream_t s, const float *rows, int num, int GL, int ML,
                      const float *lsq, const int *lsq_off, const int *lsq_len, int ncomp,
                      int width, int height, float ratio, int *point, int *or_point, int *type,
                      float *scale, float *score, int *bad);
void launch_add_part(hipStream_t s, float *S, const float *M, int R0, int R1, int P0, int ax,
                     int ay);

}  // namespace geye

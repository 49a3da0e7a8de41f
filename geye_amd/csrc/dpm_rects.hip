// dpm_rects.hip -- box refinement after NMS, and the standalone part-score subtraction.
#include <hip/hip_runtime.h>

#include "dpm_rects.h"

namespace geye {

// get_new_rects (CT/tracking.cpp:90-166), one thread per surviving row.  The regressor
// input is Avec = [rX2 - rX1, (X1, Y1) of the root and of each part box]; each output
// coordinate is the sequential sum of Avec[k] * coefficient[k] in k order (product rounded,
// then added: this file is compiled without contraction).
__global__ __launch_bounds__(64) void k_new_rects(const float *__restrict__ rows, int num, int GL,
                                                  int ML, const float *__restrict__ lsq,
                                                  const int *__restrict__ lsq_off,
                                                  const int *__restrict__ lsq_len, int ncomp,
                                                  int width, int height, float ratio,
                                                  int *__restrict__ point,
                                                  int *__restrict__ or_point,
                                                  int *__restrict__ type,
                                                  float *__restrict__ scale,
                                                  float *__restrict__ score,
                                                  int *__restrict__ bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= num) return;
  const float *P = rows + (long long)i * GL;
  const int CNUM = (int)P[GL - 3];
  if (CNUM < 0 || CNUM >= ncomp || lsq_len[CNUM] < ML) {
    atomicOr(bad, 1);
    return;
  }
  const int MLc = lsq_len[CNUM];
  const float *x1 = lsq + lsq_off[CNUM];
  const float *y1 = x1 + MLc, *x2 = y1 + MLc, *y2 = x2 + MLc;
  float XP1 = 0.f, XP2 = 0.f, YP1 = 0.f, YP2 = 0.f;
  for (int k = 0; k < ML; k++) {
    float a;
    if (k == 0) {
      a = P[3] - P[1];
    } else {
      const int j = (k - 1) >> 1;
      a = ((k - 1) & 1) ? P[4 * j] : P[4 * j + 1];
    }
    YP1 += a * y1[k];
    YP2 += a * y2[k];
    XP1 += a * x1[k];
    XP2 += a * x2[k];
  }
  int PP[4];
  PP[0] = XP1 > 0.f ? (int)XP1 : 0;
  PP[1] = YP1 > 0.f ? (int)YP1 : 0;
  PP[2] = XP2 < (float)width ? (int)XP2 : width;
  PP[3] = YP2 < (float)height ? (int)YP2 : height;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    point[i * 4 + q] = PP[q];
    or_point[i * 4 + q] = (int)((float)PP[q] / ratio);
  }
  scale[i] = P[GL - 1];
  score[i] = P[GL - 2];
  type[i] = CNUM;
}

// add_part_calculation (CT/dt.cpp:38-55), one thread per root score cell.
__global__ __launch_bounds__(256) void k_add_part(float *__restrict__ S,
                                                  const float *__restrict__ M, int R0, int R1,
                                                  int P0, int ax, int ay) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (long long)R0 * R1) return;
  const int x = (int)(p / R0), y = (int)(p - (long long)x * R0);
  S[p] -= M[(ay - 1 + 2 * y) + (long long)(ax - 1 + 2 * x) * P0];
}

void launch_new_rects(hipStream_t s, const float *rows, int num, int GL, int ML,
                      const float *lsq, const int *lsq_off, const int *lsq_len, int ncomp,
                      int width, int height, float ratio, int *point, int *or_point, int *type,
                      float *scale, float *score, int *bad) {
  if (num <= 0) return;
  hipLaunchKernelGGL(k_new_rects, dim3((num + 63) / 64), dim3(64), 0, s, rows, num, GL, ML, lsq,
                     lsq_off, lsq_len, ncomp, width, height, ratio, point, or_point, type, scale,
                     score, bad);
}

void launch_add_part(hipStream_t s, float *S, const float *M, int R0, int R1, int P0, int ax,
                     int ay) {
  const long long n = (long long)R0 * R1;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_add_part, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, S, M, R0, R1,
                     P0, ax, ay);
}

}  // namespace geye

// dpm_ascore.hip -- the accumulated-score map (A_SCORE side output of get_boxes).
//
// calc_a_score (CT/get_boxes.cpp:265-300) scatters, for every valid (level, component)
// score map, score[Yn + Xn*R0] into ac[(jj + RY) + (ii + RX)*IHEI] with a strict max, where
// Xn/Yn are the image column/row ii/jj mapped into the score map.  For one map the mapping
// (ii, jj) -> (ii + RX, jj + RY) is a shift, so every output pixel receives at most one value
// per map: this kernel gathers instead (one thread per output pixel, maps in a loop), which
// is deterministic and needs no atomics.  The max is order-independent, so the result equals
// the reference's sequential scatter.  The threshold / offset / normalisation that follows
// (CT/get_boxes.cpp:1765-1787) is fused: pass 1 writes the thresholded+offset value and folds
// the maximum, pass 2 scales by (float)(1.0 / max).
#include <hip/hip_runtime.h>

#include "dpm_ascore.h"

namespace geye {

__global__ __launch_bounds__(256) void k_ascore_gather(const ScoreMap *__restrict__ maps,
                                                       int nmaps, const int *__restrict__ tab,
                                                       const float *__restrict__ score,
                                                       int IWID, int IHEI, float thresh,
                                                       float *__restrict__ ac,
                                                       unsigned int *__restrict__ maxbits) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long npx = (long long)IWID * IHEI;
  float t = 0.0f;
  if (p < npx) {
    const int imx = (int)(p / IHEI), imy = (int)(p - (long long)imx * IHEI);
    float a = ac[p];
    for (int k = 0; k < nmaps; k++) {
      const ScoreMap sm = maps[k];
      const int ii = imx - sm.RX, jj = imy - sm.RY;
      if (ii < 0 || jj < 0 || ii >= IWID || jj >= IHEI) continue;
      const int Xn = tab[sm.tab_off + ii];
      const int Yn = tab[sm.tab_off + IWID + jj];
      if (Yn < sm.R0 && Xn < sm.R1) {
        const float sc = score[sm.score_off + Yn + (long long)Xn * sm.R0];
        if (sc > a) a = sc;
      }
    }
    // A_SCORE < thresh -> 0, else += |thresh|
    t = (a < thresh) ? 0.0f : a + fabsf(thresh);
    ac[p] = t;
  }
  // t >= 0 here, so the float order equals the unsigned order of its bits
  unsigned int v = __float_as_uint(t);
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned int w = __shfl_xor(v, o);
    v = w > v ? w : v;
  }
  __shared__ unsigned int wmax[4];
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned int m = wmax[0];
    for (int w = 1; w < 4; w++) m = wmax[w] > m ? wmax[w] : m;
    if (m) atomicMax(maxbits, m);
  }
}

__global__ __launch_bounds__(256) void k_ascore_norm(float *__restrict__ ac, long long npx,
                                                     const unsigned int *__restrict__ maxbits) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const float mx = __uint_as_float(*maxbits);
  if (p >= npx || !(mx > 0.0f)) return;
  const float ratio = (float)(1.0 / (double)mx);
  ac[p] = ac[p] * ratio;
}

void launch_ascore(hipStream_t s, const ScoreMap *maps, int nmaps, const int *tab,
                   const float *score, int IWID, int IHEI, float thresh, float *ac,
                   unsigned int *maxbits) {
  const long long npx = (long long)IWID * IHEI;
  const int blocks = (int)((npx + 255) / 256);
  hipLaunchKernelGGL(k_ascore_gather, dim3(blocks), dim3(256), 0, s, maps, nmaps, tab, score,
                     IWID, IHEI, thresh, ac, maxbits);
  hipLaunchKernelGGL(k_ascore_norm, dim3(blocks), dim3(256), 0, s, ac, npx, maxbits);
}

}  // namespace geye

// dpm_nms.hip -- exact greedy NMS of CT/nms.cpp:94-217 on the device.
//
// The reference sorts scores with an unstable middle-pivot quicksort, then greedily picks
// the highest unchecked box and suppresses later boxes whose intersection over the smaller
// area exceeds `overlap`; its `while (cnum > 0)` loop stops when one unchecked box is left,
// so the last greedy pick is emitted only if it suppressed something; survivors come out in
// original candidate order.
//
// Device form:
//  1. prep: root box, area ((x2-x1+1)*(y2-y1+1) in double, rounded to float) and score per
//     candidate, and a 64-bit key (descending score, ascending index);
//  2. radix sort of the keys (a stable order sigma that agrees with the quicksort on every
//     pair of distinct scores);
//  3. rounds: the first W unchecked boxes in sigma form a window whose greedy outcome is
//     exact on its own (everything earlier is decided); a pairwise bitmask plus a one-wave
//     sequential scan picks within the window, then every later unchecked box is tested
//     against the window's picks in order (first suppressor recorded), and the unchecked
//     list is compacted;
//  4. finalise: the last pick is dropped unless some box names it as first suppressor, and
//     the picks are emitted in original order.
// Tie order: sigma differs from the quicksort only among equal scores.  If no box is
// suppressed by a pick of equal score, and the last pick's score is unique among picks,
// every order of equal scores gives the same result (induction over score groups); when
// that check fails the caller re-runs the exact host quicksort (a reported fallback).
#include <rocprim/device/device_radix_sort.hpp>

#include "dpm_nms.h"

namespace geye {

__device__ inline unsigned int order_desc(float f) {
  unsigned int u = __float_as_uint(f);
  u = (u & 0x80000000u) ? ~u : (u | 0x80000000u);  // ascending float order
  return ~u;                                      // descending
}

__global__ __launch_bounds__(256) void k_nms_prep(const ScoreMap *__restrict__ maps, int nmaps,
                                                  const float *__restrict__ score,
                                                  const int *__restrict__ cand, int n, int padx,
                                                  int pady, const int *__restrict__ rsize,
                                                  NmsBox *__restrict__ box,
                                                  unsigned long long *__restrict__ keys) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long e = cand[i];
  int lo = 0, hi = nmaps - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (maps[mid].score_off <= e) lo = mid; else hi = mid - 1;
  }
  const ScoreMap m = maps[lo];
  const long long k = e - m.score_off;
  const int x = (int)(k / m.R0), y = (int)(k - (long long)x * m.R0);
  const float s = m.scale;
  NmsBox b;
  b.y1 = ((float)y - (float)pady + 1.0f) * s;
  b.x1 = ((float)x - (float)padx + 1.0f) * s;
  b.y2 = (float)((double)(b.y1 + (float)rsize[m.comp * 2] * s) - 1.0);
  b.x2 = (float)((double)(b.x1 + (float)rsize[m.comp * 2 + 1] * s) - 1.0);
  b.area = (float)(((double)(b.x2 - b.x1) + 1.0) * ((double)(b.y2 - b.y1) + 1.0));
  b.score = score[e];
  b.idx = i;
  b.pad = 0;
  box[i] = b;
  keys[i] = ((unsigned long long)order_desc(b.score) << 32) | (unsigned int)i;
}

// prep from caller rows [y1, x1, y2, x2, ..., score, scale] (GL floats per row)
__global__ __launch_bounds__(256) void k_nms_prep_rows(const float *__restrict__ rows, int n,
                                                       int GL, NmsBox *__restrict__ box,
                                                       unsigned long long *__restrict__ keys) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float *r = rows + (size_t)i * GL;
  NmsBox b;
  b.y1 = r[0];
  b.x1 = r[1];
  b.y2 = r[2];
  b.x2 = r[3];
  b.area = (float)(((double)(b.x2 - b.x1) + 1.0) * ((double)(b.y2 - b.y1) + 1.0));
  b.score = r[GL - 2];
  b.idx = i;
  b.pad = 0;
  box[i] = b;
  keys[i] = ((unsigned long long)order_desc(b.score) << 32) | (unsigned int)i;
}

__global__ __launch_bounds__(256) void k_nms_gather(const unsigned long long *__restrict__ keys,
                                                    const NmsBox *__restrict__ box, int n,
                                                    NmsBox *__restrict__ sbox,
                                                    int *__restrict__ alive,
                                                    int *__restrict__ status,
                                                    int *__restrict__ first_sup) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  sbox[r] = box[(unsigned int)(keys[r] & 0xffffffffu)];
  alive[r] = r;
  status[r] = 0;
  first_sup[r] = -1;
}

// overlap test of CT/nms.cpp:145-175 (symmetric in a and b)
__device__ inline bool suppresses(const NmsBox &a, const NmsBox &b, float overlap) {
  const float yy1 = a.y1 >= b.y1 ? a.y1 : b.y1;
  const float xx1 = a.x1 >= b.x1 ? a.x1 : b.x1;
  const float yy2 = a.y2 <= b.y2 ? a.y2 : b.y2;
  const float xx2 = a.x2 <= b.x2 ? a.x2 : b.x2;
  const float w = (float)((double)(xx2 - xx1) + 1.0);
  const float h = (float)((double)(yy2 - yy1) + 1.0);
  const float R = a.area <= b.area ? a.area : b.area;
  if (w > 0 && h > 0) return w * h / R > overlap;
  return (a.y1 < b.y1 && a.x1 < b.x1 && a.y2 > b.y2 && a.x2 > b.x2) ||
         (a.y1 > b.y1 && a.x1 > b.x1 && a.y2 < b.y2 && a.x2 < b.x2);
}

// row i of the window bitmask: bit j (j > i) set when window box i suppresses box j
__global__ __launch_bounds__(64) void k_nms_winmask(const NmsBox *__restrict__ sbox,
                                                    const int *__restrict__ alive,
                                                    const int *__restrict__ nalive,
                                                    float overlap,
                                                    unsigned long long *__restrict__ mask) {
  const int W = min(*nalive, kNmsWindow);
  const int i = blockIdx.x;
  if (i >= W) return;
  const NmsBox a = sbox[alive[i]];
  for (int w = 0; w < kNmsWindow / 64; w++) {
    const int j = w * 64 + threadIdx.x;
    bool s = false;
    if (j > i && j < W) s = suppresses(a, sbox[alive[j]], overlap);
    unsigned long long bits = __ballot(s);
    if (threadIdx.x == 0) mask[(size_t)i * (kNmsWindow / 64) + w] = bits;
  }
}

// sequential greedy over the window, one wave; lane w owns word w of the removed set
__global__ __launch_bounds__(64) void k_nms_wingreedy(const int *__restrict__ alive,
                                                      const int *__restrict__ nalive,
                                                      const unsigned long long *__restrict__ mask,
                                                      int *__restrict__ status,
                                                      int *__restrict__ first_sup,
                                                      int *__restrict__ picks,
                                                      int *__restrict__ counters) {
  const int W = min(*nalive, kNmsWindow);
  const int lane = threadIdx.x;
  constexpr int NW = kNmsWindow / 64;
  unsigned long long removed = 0;
  int npick = counters[0];
  const int round_begin = npick;
  for (int i = 0; i < W; i++) {
    const unsigned long long word = __shfl(removed, i >> 6);
    if ((word >> (i & 63)) & 1ull) continue;
    const int ri = alive[i];
    if (lane == 0) {
      status[ri] = 1;
      picks[npick] = ri;
    }
    npick++;
    if (lane < NW) {
      const unsigned long long m = mask[(size_t)i * NW + lane];
      unsigned long long fresh = m & ~removed;
      removed |= m;
      while (fresh) {
        const int b = __ffsll((long long)fresh) - 1;
        fresh &= fresh - 1;
        const int rj = alive[lane * 64 + b];
        status[rj] = -1;
        first_sup[rj] = ri;
      }
    }
  }
  if (lane == 0) {
    counters[0] = npick;        // picks so far
    counters[1] = round_begin;  // first pick of this round
  }
}

// every unchecked box after the window against this round's picks, in pick order
__global__ __launch_bounds__(256) void k_nms_suppress(const NmsBox *__restrict__ sbox,
                                                      const int *__restrict__ alive,
                                                      const int *__restrict__ nalive,
                                                      const int *__restrict__ picks,
                                                      const int *__restrict__ counters,
                                                      float overlap, int *__restrict__ status,
                                                      int *__restrict__ first_sup,
                                                      int *__restrict__ keep) {
  __shared__ NmsBox sp[kNmsWindow];
  __shared__ int sr[kNmsWindow];
  const int na = *nalive;
  const int W = min(na, kNmsWindow);
  const int p0 = counters[1], p1 = counters[0];
  const int np = p1 - p0;
  for (int k = threadIdx.x; k < np; k += blockDim.x) {
    sr[k] = picks[p0 + k];
    sp[k] = sbox[sr[k]];
  }
  __syncthreads();
  const int t = W + blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= na) return;
  const int r = alive[t];
  const NmsBox b = sbox[r];
  int st = 0;
  for (int k = 0; k < np; k++)
    if (suppresses(sp[k], b, overlap)) {
      status[r] = -1;
      first_sup[r] = sr[k];
      st = -1;
      break;
    }
  keep[t - W] = st == 0 ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_nms_regather(const int *__restrict__ alive,
                                                      const int *__restrict__ nalive,
                                                      const int *__restrict__ sel, int nsel,
                                                      int *__restrict__ alive_out) {
  const int W = min(*nalive, kNmsWindow);
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nsel) alive_out[k] = alive[W + sel[k]];
}

// finalisation: conflicts, last-pick emission, survivors in original order
__global__ __launch_bounds__(256) void k_nms_check(const NmsBox *__restrict__ sbox, int n,
                                                   const int *__restrict__ status,
                                                   const int *__restrict__ first_sup,
                                                   const int *__restrict__ picks, int npick,
                                                   int *__restrict__ flags) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n || npick == 0) return;
  const int last = picks[npick - 1];
  if (status[r] == -1) {
    const int f = first_sup[r];
    if (sbox[f].score == sbox[r].score) atomicOr(&flags[0], 1);  // tie decided the outcome
    if (f == last) atomicOr(&flags[1], 1);                      // last pick suppressed something
  } else if (status[r] == 1 && r != last && sbox[r].score == sbox[last].score) {
    atomicOr(&flags[0], 2);  // the identity of the last pick depends on the tie order
  }
}

__global__ __launch_bounds__(256) void k_nms_keep(const NmsBox *__restrict__ sbox,
                                                  const int *__restrict__ picks, int npick,
                                                  const int *__restrict__ flags,
                                                  int *__restrict__ keep_orig) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npick) return;
  if (k == npick - 1 && flags[1] == 0) return;
  keep_orig[sbox[picks[k]].idx] = 1;
}

__global__ __launch_bounds__(256) void k_fill(int *__restrict__ p, long long n, int v) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void k_take(const int *__restrict__ src,
                                              const int *__restrict__ idx, int n,
                                              int *__restrict__ dst) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[k] = src[idx[k]];
}

size_t nms_sort_temp_bytes(int n) {
  size_t bytes = 0;
  unsigned long long *k = nullptr;
  (void)rocprim::radix_sort_keys(nullptr, bytes, k, k, (unsigned int)n, 0, 64, 0);
  return bytes;
}

void launch_nms_prep(hipStream_t s, const ScoreMap *maps, int nmaps, const float *score,
                     const int *cand, int n, int padx, int pady, const int *rsize, NmsBox *box,
                     unsigned long long *keys) {
  hipLaunchKernelGGL(k_nms_prep, dim3((n + 255) / 256), dim3(256), 0, s, maps, nmaps, score, cand,
                     n, padx, pady, rsize, box, keys);
}

void launch_nms_prep_rows(hipStream_t s, const float *rows, int n, int GL, NmsBox *box,
                          unsigned long long *keys) {
  hipLaunchKernelGGL(k_nms_prep_rows, dim3((n + 255) / 256), dim3(256), 0, s, rows, n, GL, box,
                     keys);
}

int launch_nms_sort(hipStream_t s, void *temp, size_t temp_bytes, unsigned long long *keys_in,
                    unsigned long long *keys_out, int n) {
  size_t bytes = temp_bytes;
  return rocprim::radix_sort_keys(temp, bytes, keys_in, keys_out, (unsigned int)n, 0, 64, s) ==
                 hipSuccess
             ? 0
             : -1;
}

void launch_nms_gather(hipStream_t s, const unsigned long long *keys, const NmsBox *box, int n,
                       NmsBox *sbox, int *alive, int *status, int *first_sup) {
  hipLaunchKernelGGL(k_nms_gather, dim3((n + 255) / 256), dim3(256), 0, s, keys, box, n, sbox,
                     alive, status, first_sup);
}

void launch_nms_round(hipStream_t s, const NmsBox *sbox, const int *alive, const int *nalive,
                      int nalive_host, float overlap, unsigned long long *mask, int *status,
                      int *first_sup, int *picks, int *counters, int *keep) {
  const int W = nalive_host < kNmsWindow ? nalive_host : kNmsWindow;
  hipLaunchKernelGGL(k_nms_winmask, dim3(W), dim3(64), 0, s, sbox, alive, nalive, overlap, mask);
  hipLaunchKernelGGL(k_nms_wingreedy, dim3(1), dim3(64), 0, s, alive, nalive, mask, status,
                     first_sup, picks, counters);
  const int rest = nalive_host - W;
  if (rest > 0)
    hipLaunchKernelGGL(k_nms_suppress, dim3((rest + 255) / 256), dim3(256), 0, s, sbox, alive,
                       nalive, picks, counters, overlap, status, first_sup, keep);
}

void launch_nms_regather(hipStream_t s, const int *alive, const int *nalive, const int *sel,
                         int nsel, int *alive_out) {
  if (nsel > 0)
    hipLaunchKernelGGL(k_nms_regather, dim3((nsel + 255) / 256), dim3(256), 0, s, alive, nalive,
                       sel, nsel, alive_out);
}

void launch_nms_check(hipStream_t s, const NmsBox *sbox, int n, const int *status,
                      const int *first_sup, const int *picks, int npick, int *flags) {
  hipLaunchKernelGGL(k_nms_check, dim3((n + 255) / 256), dim3(256), 0, s, sbox, n, status,
                     first_sup, picks, npick, flags);
}

void launch_nms_keep(hipStream_t s, const NmsBox *sbox, const int *picks, int npick,
                     const int *flags, int *keep_orig) {
  if (npick > 0)
    hipLaunchKernelGGL(k_nms_keep, dim3((npick + 255) / 256), dim3(256), 0, s, sbox, picks, npick,
                       flags, keep_orig);
}

void launch_fill(hipStream_t s, int *p, long long n, int v) {
  if (n > 0)
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}

void launch_take(hipStream_t s, const int *src, const int *idx, int n, int *dst) {
  if (n > 0)
    hipLaunchKernelGGL(k_take, dim3((n + 255) / 256), dim3(256), 0, s, src, idx, n, dst);
}

}  // namespace geye

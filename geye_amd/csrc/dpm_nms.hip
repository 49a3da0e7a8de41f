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
//     sequential scan picks within the window, then every unchecked box is tested against
//     the window's picks ranked before it (first suppressor recorded) -- with a spatial index
//     of the picks, only the picks it names -- and the unchecked list is compacted;
//  4. finalise: the last pick is dropped unless some box names it as first suppressor, and
//     the picks are emitted in original order.
// Tie order: sigma differs from the quicksort only among equal scores.  If no box is
// suppressed by a pick of equal score, and the last pick's score is unique among picks,
// every order of equal scores gives the same result (induction over score groups); when
// that check fails the caller re-runs the exact host quicksort (a reported fallback).
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>

#include "dpm_kernels.h"
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
                                                  unsigned int *__restrict__ keys,
                                                  int *__restrict__ vals) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long long e = cand[i];
  // candidates ascend, so a wave's cells lie in the map of its first active lane's cell or a
  // few after it: one search per wave, then each lane steps forward
  const int e0 = __builtin_amdgcn_readfirstlane((int)e);
  int lo = 0, hi = nmaps - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (maps[mid].score_off <= e0) lo = mid; else hi = mid - 1;
  }
  while (lo + 1 < nmaps && maps[lo + 1].score_off <= e) lo++;
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
  keys[i] = order_desc(b.score);
  vals[i] = i;
}

// prep from caller rows [y1, x1, y2, x2, ..., score, scale] (GL floats per row)
__global__ __launch_bounds__(256) void k_nms_prep_rows(const float *__restrict__ rows, int n,
                                                       int GL, NmsBox *__restrict__ box,
                                                       unsigned int *__restrict__ keys,
                                                       int *__restrict__ vals) {
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
  keys[i] = order_desc(b.score);
  vals[i] = i;
}

// Tile of a coordinate along one axis: floor(v) / 2^shift, clamped to the grid (monotone in
// v, so a point inside an interval falls in a tile of the interval's tile range).
__device__ inline int nms_tile(float v, int shift, int g) {
  const int t = (int)floorf(v) >> shift;
  return t < 0 ? 0 : (t >= g ? g - 1 : t);
}

// fine tiles, coarse tiles, and the words of all three mask arrays per mask word
__device__ __host__ inline int nms_coarse_tiles(const NmsGrid &g) {
  return ((g.gx + 3) >> 2) * ((g.gy + 3) >> 2);
}
__device__ __host__ inline int nms_grid_tiles(const NmsGrid &g) {
  return 2 * g.gx * g.gy + nms_coarse_tiles(g);
}

// sorted order: sbox[r] = box[ord[r]]; with a grid, each candidate's rank and decided flag,
// and the first round's tile counts
__global__ __launch_bounds__(256) void k_nms_gather(const int *__restrict__ ord,
                                                    const NmsBox *__restrict__ box, int n,
                                                    NmsBox *__restrict__ sbox,
                                                    int *__restrict__ alive,
                                                    int *__restrict__ status,
                                                    int *__restrict__ sup,
                                                    int *__restrict__ counters,
                                                    int *__restrict__ nalive0,
                                                    int *__restrict__ flags,
                                                    int *__restrict__ keep, NmsGrid g,
                                                    float *__restrict__ tie) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  // the pass's other initial state: pick counters, the first round's alive count, the tie
  // flags and set, and the survivor flags (written by the finalisation, after the rounds)
  if (r < 4) {
    counters[r] = 0;
    flags[r] = 0;
  }
  if (tie && r < kNmsTieCap) ((unsigned int *)tie)[r] = kNmsTieEmpty;  // the grid has >= 256
  if (r == 0) *nalive0 = n;
  if (g.ext && r < kNmsMaskWords * nms_grid_tiles(g)) g.ext[r] = 0;  // ext, ctr, ctrc
  if (r >= n) return;
  const int i = ord[r];
  sbox[r] = box[i];
  alive[r] = r;
  status[r] = 0;
  sup[r] = kNoSup;
  keep[r] = 0;
  if (g.ext) {
    g.rank_of[i] = r;
    g.dec[r] = 0;  // every candidate's flag cleared: r runs over all of them, in order
  }
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

// row i of the window bitmask, stored transposed ([word][row]): bit j (j > i) of row i set
// when window box i suppresses box j.  A workgroup stages the window's boxes in LDS (the
// alive indices, then the boxes, 4 per thread, loads in flight together) and each of its
// waves computes one row from there.
constexpr int kWinmaskRows = 4;  // rows per workgroup, one per wave

__global__ __launch_bounds__(256) void k_nms_winmask(const NmsBox *__restrict__ sbox,
                                                     const int *__restrict__ alive,
                                                     const int *__restrict__ nalive,
                                                     float overlap,
                                                     unsigned long long *__restrict__ mask) {
  __shared__ float s_y1[kNmsWindow], s_x1[kNmsWindow], s_y2[kNmsWindow], s_x2[kNmsWindow],
      s_ar[kNmsWindow];
  const int W = min(*nalive, kNmsWindow);
  if ((int)blockIdx.x * kWinmaskRows >= W) return;
  const int tid = threadIdx.x;
  constexpr int U = kNmsWindow / 256;
  // loads unconditional (position 0 stands in past W), so they are in flight together
  int a[U];
#pragma unroll
  for (int u = 0; u < U; u++) {
    const int j = tid + u * 256;
    a[u] = alive[j < W ? j : 0];
  }
  NmsBox b[U];
#pragma unroll
  for (int u = 0; u < U; u++) b[u] = sbox[a[u]];
#pragma unroll
  for (int u = 0; u < U; u++) {
    const int j = tid + u * 256;
    if (j < W) {
      s_y1[j] = b[u].y1;
      s_x1[j] = b[u].x1;
      s_y2[j] = b[u].y2;
      s_x2[j] = b[u].x2;
      s_ar[j] = b[u].area;
    }
  }
  __syncthreads();
  const int i = blockIdx.x * kWinmaskRows + (tid >> 6), lane = tid & 63;
  if (i >= W) return;
  NmsBox A;
  A.y1 = s_y1[i];
  A.x1 = s_x1[i];
  A.y2 = s_y2[i];
  A.x2 = s_x2[i];
  A.area = s_ar[i];
  for (int w = 0; w < kNmsWindow / 64; w++) {
    const int j = w * 64 + lane;
    bool sp = false;
    if (j > i && j < W) {
      NmsBox B;
      B.y1 = s_y1[j];
      B.x1 = s_x1[j];
      B.y2 = s_y2[j];
      B.x2 = s_x2[j];
      B.area = s_ar[j];
      sp = suppresses(A, B, overlap);
    }
    const unsigned long long bits = __ballot(sp);
    if (lane == 0) mask[(size_t)w * kNmsWindow + i] = bits;  // [word][row]
  }
}

// Greedy over the window.  The workgroup stages the window's bitmask ([word][row], all in
// LDS); then one wave runs the sequential decision with scalar state: block b of 64 rows is
// decided with word b alone (row j's word b read by readlane, folded in only when row j is
// still alive), and the rows picked in block b are OR-ed into the later words by their
// owner lanes (whole words: the decision needs only the removed set).  The first suppressor
// of a removed window box (the earliest pick whose row covers it) comes from the round's
// suppression pass, which tests every box ranked after each pick; in a round with no such
// pass (window_sup != 0) every thread takes window boxes that were not picked and scans the
// picks in order here.
__global__ __launch_bounds__(256) void k_nms_wingreedy(const int *__restrict__ alive,
                                                        const int *__restrict__ nalive,
                                                        const unsigned long long *__restrict__ mask,
                                                        int *__restrict__ status,
                                                        int *__restrict__ sup,
                                                        int *__restrict__ picks,
                                                        int *__restrict__ counters,
                                                        int window_sup) {
  constexpr int NW = kNmsWindow / 64;
  extern __shared__ unsigned long long smask[];  // [NW][kNmsWindow]
  int *salive = (int *)(smask + NW * kNmsWindow);
  __shared__ short s_pick[kNmsWindow];              // window indices of the picks, in order
  __shared__ unsigned long long s_picked[NW];       // picked rows of each 64-row block
  __shared__ int s_np;
  const int W = min(*nalive, kNmsWindow);
  const int nblk = (W + 63) / 64;
  // staging: this is the only workgroup on the device, so its loads are issued in wide,
  // unrolled batches to pay the memory latency once rather than per element
  {
    typedef __attribute__((ext_vector_type(4))) unsigned int u4;
    const u4 *src = (const u4 *)mask;
    u4 *dst = (u4 *)smask;
    const int n16 = nblk * kNmsWindow / 2;  // 16-byte chunks
    constexpr int U = (kNmsWindow * (kNmsWindow / 64) / 2) / 256;
    u4 v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int k = threadIdx.x + u * 256;
      if (k < n16) v[u] = src[k];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int k = threadIdx.x + u * 256;
      if (k < n16) dst[k] = v[u];
    }
    int a[kNmsWindow / 256];
#pragma unroll
    for (int u = 0; u < kNmsWindow / 256; u++) {
      const int i = threadIdx.x + u * 256;
      if (i < W) a[u] = alive[i];
    }
#pragma unroll
    for (int u = 0; u < kNmsWindow / 256; u++) {
      const int i = threadIdx.x + u * 256;
      if (i < W) salive[i] = a[u];
    }
  }
  __syncthreads();
  const int round_begin = counters[0];
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    unsigned long long removed = 0;  // lane w < nblk: word w of the removed set
    int np = 0;
    for (int b = 0; b < nblk; b++) {
      const int rows = min(64, W - 64 * b);
      // lane j holds row (64b + j)'s word b: the block's diagonal
      const unsigned long long diag = lane < rows ? smask[(size_t)b * kNmsWindow + 64 * b + lane] : 0ull;
      const unsigned long long rem0 =
          ((unsigned long long)(unsigned int)__builtin_amdgcn_readlane((int)(unsigned int)(removed >> 32), b) << 32) |
          (unsigned long long)(unsigned int)__builtin_amdgcn_readlane((int)(unsigned int)removed, b);
      const unsigned long long rowsmask = rows == 64 ? ~0ull : ((1ull << rows) - 1ull);
      // alive rows of the block, decided in order: the lowest alive row is a pick and
      // removes the rows its mask covers
      unsigned long long cand = ~rem0 & rowsmask;
      unsigned long long rem = rem0, picked = 0;
      while (cand) {
        const int i = __builtin_ctzll(cand);
        const unsigned int hi = (unsigned int)__builtin_amdgcn_readlane((int)(unsigned int)(diag >> 32), i);
        const unsigned int lo = (unsigned int)__builtin_amdgcn_readlane((int)(unsigned int)diag, i);
        const unsigned long long m = ((unsigned long long)hi << 32) | (unsigned long long)lo;
        picked |= 1ull << i;
        cand &= ~(m | (1ull << i));
        rem |= m;
      }
      if (lane == b) removed = rem;
      // owners of the later words fold in the picked rows (8 row reads in flight at a time)
      if (lane > b && lane < nblk) {
        const unsigned long long *row = smask + (size_t)lane * kNmsWindow + 64 * b;
        unsigned long long pk = picked;
        while (pk) {
          int ix[8];
#pragma unroll
          for (int u = 0; u < 8; u++) {
            ix[u] = pk ? __builtin_ctzll(pk) : -1;
            pk &= pk - 1;
          }
          unsigned long long m[8];
#pragma unroll
          for (int u = 0; u < 8; u++) m[u] = ix[u] >= 0 ? row[ix[u]] : 0ull;
#pragma unroll
          for (int u = 0; u < 8; u++) removed |= m[u];
        }
      }
      if (lane < rows && ((picked >> lane) & 1ull)) {
        const unsigned long long below = lane ? (picked & ((1ull << lane) - 1ull)) : 0ull;
        const int slot = np + __popcll(below);
        const int ri = salive[64 * b + lane];
        status[ri] = 1;
        picks[round_begin + slot] = ri;
        s_pick[slot] = (short)(64 * b + lane);
      }
      if (lane == 0) s_picked[b] = picked;
      np += __popcll(picked);
    }
    if (lane == 0) {
      counters[0] = round_begin + np;  // picks so far
      counters[1] = round_begin;       // first pick of this round
      s_np = np;
    }
  }
  __syncthreads();
  // first suppressors of the window boxes removed in this round, when no suppression pass
  // follows: every window box is either a pick or covered by an earlier pick's row
  if (window_sup) {
    const int np = s_np;
    for (int j = threadIdx.x; j < W; j += blockDim.x) {
      const int w = j >> 6, bit = j & 63;
      if ((s_picked[w] >> bit) & 1ull) continue;
      const unsigned long long *col = smask + (size_t)w * kNmsWindow;
      int f = -1;
      for (int k0 = 0; k0 < np && f < 0; k0 += 8) {
        int pp[8];
        unsigned long long v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) pp[u] = k0 + u < np ? s_pick[k0 + u] : -1;
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = pp[u] >= 0 ? col[pp[u]] : 0ull;
#pragma unroll
        for (int u = 0; u < 8; u++)
          if (f < 0 && ((v[u] >> bit) & 1ull)) f = pp[u];
      }
      if (f >= 0) sup[salive[j]] = salive[f];
    }
  }
}

// every unchecked box against this round's picks ranked before it, in pick order: the first
// suppressor of a removed window box or a later box (generic path for caller-supplied boxes)
__global__ __launch_bounds__(256) void k_nms_suppress(const NmsBox *__restrict__ sbox,
                                                      const int *__restrict__ alive,
                                                      const int *__restrict__ nalive,
                                                      const int *__restrict__ picks,
                                                      const int *__restrict__ counters,
                                                      float overlap, int *__restrict__ sup) {
  __shared__ NmsBox sp[kNmsWindow];
  __shared__ int sr[kNmsWindow];
  const int na = *nalive;
  const int p0 = counters[1], p1 = counters[0];
  const int np = p1 - p0;
  for (int k = threadIdx.x; k < np; k += blockDim.x) {
    sr[k] = picks[p0 + k];
    sp[k] = sbox[sr[k]];
  }
  __syncthreads();
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= na) return;
  const int r = alive[t];
  const NmsBox b = sbox[r];
  for (int k = 0; k < np && sr[k] < r; k++)
    if (suppresses(sp[k], b, overlap)) {
      sup[r] = sr[k];
      break;
    }
}

// Index of the round's picks, one wave per pick: the pick's box (its rank in .pad) to pbox,
// its candidate marked decided, its bit in the centre masks of its centre's tile and coarse
// tile and in the extent mask of every tile its grown box touches.
__global__ __launch_bounds__(256) void k_nms_grid_build(const NmsBox *__restrict__ sbox,
                                                        const int *__restrict__ picks,
                                                        const int *__restrict__ counters,
                                                        NmsGrid g, int round) {
  const int p0 = counters[1], np = counters[0] - p0;
  const int pl = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (pl >= np) return;
  const int rp = picks[p0 + pl];
  NmsBox P = sbox[rp];
  const int nt = g.gx * g.gy, cgx = (g.gx + 3) >> 2, nct = nms_coarse_tiles(g);
  const int w = pl >> 6;
  const unsigned long long bit = 1ull << (pl & 63);
  if (lane == 0) {
    g.dec[P.idx] = 1;
    const int ty = nms_tile((P.y1 + P.y2) * 0.5f, g.shift, g.gy);
    const int tx = nms_tile((P.x1 + P.x2) * 0.5f, g.shift, g.gx);
    atomicOr(&g.ctr[(size_t)w * nt + ty * g.gx + tx], bit);
    atomicOr(&g.ctrc[(size_t)w * nct + (ty >> 2) * cgx + (tx >> 2)], bit);
    P.pad = rp;
    g.pbox[pl] = P;
  }
  const int ty0 = nms_tile(P.y1 - kNmsTileMargin, g.shift, g.gy);
  const int ty1 = nms_tile(P.y2 + kNmsTileMargin, g.shift, g.gy);
  const int tx0 = nms_tile(P.x1 - kNmsTileMargin, g.shift, g.gx);
  const int tx1 = nms_tile(P.x2 + kNmsTileMargin, g.shift, g.gx);
  const int tw = tx1 - tx0 + 1, cells = tw * (ty1 - ty0 + 1);
  for (int k = lane; k < cells; k += 64) {
    const int dy = k / tw;
    atomicOr(&g.ext[(size_t)w * nt + (ty0 + dy) * g.gx + tx0 + (k - dy * tw)], bit);
  }
}

constexpr int kNmsStageWords = 2048;  // mask words a suppression workgroup stages in LDS
constexpr int kSupThreads = 512;      // threads (candidates) per suppression workgroup
constexpr int kSupWaves = kSupThreads / 64;

// Suppression pass over the grid, one thread per undecided candidate, in candidate order
// (neighbouring score cells, so a workgroup's boxes share tiles): round 0 takes every
// candidate, a later round the list of the candidates the previous rounds left undecided.  An
// undecided box gets the least rank of the round's picks ranked before it that suppress it.
// The workgroup stages the round's pick boxes and, when they fit, the masks of the tiles its
// boxes read (the bounding rectangle of their grown boxes) in LDS; each box then ORs, one
// mask word at a time, the extent mask of its centre's tile with the centre masks of the tiles
// its grown box touches (from global memory when the rectangle does not fit, over coarse tiles
// for a box touching more than 16), and tests the picks named there in pick order, which is
// rank order: the first that suppresses it, or the first ranked after it, ends the scan.  The
// boxes left undecided go, in thread order, to the workgroup's kSupThreads places in slot
// (the round's end compacts them into the next list).
// Completeness for overlap >= 0.5: suppression needs w h / R > overlap with R the smaller
// area, and the intersection's width and height are at most the smaller box's, so each
// exceeds half of the smaller box's (up to the rounding of w, h and R, well under a pixel
// here); an interval inside the smaller box's span and longer than half of it reaches to
// within half a pixel of that span's centre.  So the smaller box's centre lies in the larger
// box grown by half a pixel: a pick that suppresses B is in the extent mask of B's centre
// tile (B the smaller) or in the centre mask of a tile B's grown box touches (the pick the
// smaller), and the exact test of CT/nms.cpp decides.
__global__ __launch_bounds__(kSupThreads) void k_nms_grid_sup(const int *__restrict__ counters,
                                                              NmsGrid g, float overlap,
                                                              int round,
                                                              int *__restrict__ sup) {
  __shared__ float4 s_b[kNmsWindow];  // y1, x1, y2, x2
  __shared__ float s_a[kNmsWindow];
  __shared__ int s_r[kNmsWindow];
  __shared__ unsigned long long s_m[kNmsStageWords];  // [word][ext, ctr][rectangle tile]
  __shared__ int s_rect[4], s_cnt[kSupWaves];
  const int cnt = round == 0 ? g.n : *g.lcount;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  int i = -1;
  if (t < cnt) i = round == 0 ? t : g.list[t];
  const bool act = i >= 0 && !g.dec[i];
  const int nact = __syncthreads_count(act);
  if (nact == 0) {
    if (threadIdx.x == 0) g.bsum[blockIdx.x] = 0;
    return;
  }
  // a workgroup with few undecided boxes (the list's last) reads masks and picks from global
  // memory: staging costs it more than it saves
  const bool dense = nact > kSupThreads / 4;
  if (threadIdx.x == 0) {
    s_rect[0] = s_rect[1] = 0x7fffffff;
    s_rect[2] = s_rect[3] = -1;
  }
  const int p0 = counters[1], np = counters[0] - p0;
  const int nt = g.gx * g.gy, nw = (np + 63) >> 6;
  if (dense)
    for (int k = threadIdx.x; k < np; k += blockDim.x) {
      const NmsBox P = g.pbox[k];
      s_b[k] = make_float4(P.y1, P.x1, P.y2, P.x2);
      s_a[k] = P.area;
      s_r[k] = P.pad;
    }
  int r = 0, ty0 = 0, ty1 = -1, tx0 = 0, tx1 = -1, c = 0;
  NmsBox B;
  if (act) {
    r = g.rank_of[i];
    B = g.box[i];
    ty0 = nms_tile(B.y1 - kNmsTileMargin, g.shift, g.gy);
    ty1 = nms_tile(B.y2 + kNmsTileMargin, g.shift, g.gy);
    tx0 = nms_tile(B.x1 - kNmsTileMargin, g.shift, g.gx);
    tx1 = nms_tile(B.x2 + kNmsTileMargin, g.shift, g.gx);
    c = nms_tile((B.y1 + B.y2) * 0.5f, g.shift, g.gy) * g.gx +
        nms_tile((B.x1 + B.x2) * 0.5f, g.shift, g.gx);  // inside the grown box's tiles
  }
  int RY0 = 0, RX0 = 0, rw = 0, nrect = 0;
  bool staged = false;
  if (dense) {
    // the workgroup's rectangle: wave reductions, then one LDS atomic per wave and bound
    int a0 = act ? ty0 : 0x7fffffff, a1 = act ? tx0 : 0x7fffffff, a2 = act ? ty1 : -1,
        a3 = act ? tx1 : -1;
    for (int o = 1; o < 64; o <<= 1) {
      a0 = min(a0, __shfl_xor(a0, o));
      a1 = min(a1, __shfl_xor(a1, o));
      a2 = max(a2, __shfl_xor(a2, o));
      a3 = max(a3, __shfl_xor(a3, o));
    }
    __syncthreads();  // s_rect initialised
    if ((threadIdx.x & 63) == 0 && a2 >= 0) {
      atomicMin(&s_rect[0], a0);
      atomicMin(&s_rect[1], a1);
      atomicMax(&s_rect[2], a2);
      atomicMax(&s_rect[3], a3);
    }
    __syncthreads();
    RY0 = s_rect[0];
    RX0 = s_rect[1];
    rw = s_rect[3] - RX0 + 1;
    nrect = rw * (s_rect[2] - RY0 + 1);
    staged = 2 * nw * nrect <= kNmsStageWords;
    if (staged) {
      // [word][ext, ctr][tile]: up to 4 loads per thread, all in flight before the stores
      constexpr int U = kNmsStageWords / kSupThreads;
      const int total = 2 * nw * nrect;
      unsigned long long v[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int k = threadIdx.x + u * kSupThreads;
        v[u] = 0ull;
        if (k < total) {
          const int wk = k / nrect, lt = k - wk * nrect, dy = lt / rw;
          const int tile = (RY0 + dy) * g.gx + RX0 + (lt - dy * rw);
          v[u] = ((wk & 1) ? g.ctr : g.ext)[(size_t)(wk >> 1) * nt + tile];
        }
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int k = threadIdx.x + u * kSupThreads;
        if (k < total) s_m[k] = v[u];
      }
    }
    __syncthreads();
  }
  int best = kNoSup;
  if (act) {
    // global path: a box touching more than 16 tiles reads the coarse centre masks
    const unsigned long long *cm = g.ctr;
    int rowlen = g.gx, ntl = nt, qy0 = ty0, qy1 = ty1, qx0 = tx0, qx1 = tx1;
    if (!staged && (ty1 - ty0 + 1) * (tx1 - tx0 + 1) > 16) {
      qy0 >>= 2; qy1 >>= 2; qx0 >>= 2; qx1 >>= 2;
      cm = g.ctrc;
      rowlen = (g.gx + 3) >> 2;
      ntl = nms_coarse_tiles(g);
    }
    const int cy = c / g.gx, lc = (cy - RY0) * rw + (c - cy * g.gx - RX0);
    bool done = false;
    for (int w = 0; w < nw && !done; w++) {
      unsigned long long bits;
      if (staged) {
        const unsigned long long *e = s_m + (size_t)(2 * w) * nrect, *m = e + nrect;
        bits = e[lc];
        for (int ty = qy0; ty <= qy1; ty++) {
          const unsigned long long *row = m + (ty - RY0) * rw - RX0;
          for (int tx = qx0; tx <= qx1; tx++) bits |= row[tx];
        }
      } else {
        bits = g.ext[(size_t)w * nt + c];
        const unsigned long long *m = cm + (size_t)w * ntl;
        for (int ty = qy0; ty <= qy1; ty++)
          for (int tx = qx0; tx <= qx1; tx++) bits |= m[ty * rowlen + tx];
      }
      while (bits && !done) {
        const int pl = w * 64 + __builtin_ctzll(bits);
        bits &= bits - 1;
        NmsBox P;
        if (dense) {
          const float4 q = s_b[pl];
          P.y1 = q.x;
          P.x1 = q.y;
          P.y2 = q.z;
          P.x2 = q.w;
          P.area = s_a[pl];
          P.pad = s_r[pl];
        } else {
          P = g.pbox[pl];
        }
        if (P.pad >= r) {
          done = true;  // this pick and every later one rank after the box
        } else if (suppresses(P, B, overlap)) {
          best = P.pad;
          done = true;
        }
      }
    }
    if (best != kNoSup) {
      sup[r] = best;
      g.dec[i] = 1;
    }
  }
  // the boxes still undecided, in thread order, to the workgroup's places in slot
  const bool keep = act && best == kNoSup;
  const unsigned long long bal = __ballot(keep);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) s_cnt[wave] = __popcll(bal);
  __syncthreads();
  if (keep) {
    int pos = blockIdx.x * kSupThreads + __popcll(bal & ((1ull << lane) - 1ull));
    for (int w = 0; w < wave; w++) pos += s_cnt[w];
    g.slot[pos] = i;
  }
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int w = 0; w < kSupWaves; w++) tot += s_cnt[w];
    g.bsum[blockIdx.x] = tot;
  }
}


constexpr int kNmsCompactBlock = 1024;  // positions per workgroup of the per-round counts

// End of a round: every alive position t < na is either decided (a window box, or a later
// box some pick suppressed) or still alive.  Each workgroup counts the alive later boxes of
// its kNmsCompactBlock positions (4 consecutive per thread) into blocksum; positions past the
// true count add nothing, so a grid over a host bound counts exactly.  With a grid, the
// workgroups also clear its masks for the next round's index (the suppression pass that read
// them has finished: the words the round's picks used), and workgroups past the first nb
// compact the boxes the pass left undecided into the next round's candidate list.
__global__ __launch_bounds__(256) void k_nms_post(const int *__restrict__ alive,
                                                  const int *__restrict__ nalive,
                                                  const int *__restrict__ sup,
                                                  const int *__restrict__ counters, NmsGrid g,
                                                  int nb, int *__restrict__ blocksum) {
  __shared__ int s_w[4];
  if (g.ext) {
    const int nt = g.gx * g.gy, nct = nms_coarse_tiles(g);
    const int nw = (counters[0] - counters[1] + 63) >> 6;
    const int step = gridDim.x * blockDim.x;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nw * nt; k += step) {
      g.ext[k] = 0ull;
      g.ctr[k] = 0ull;
    }
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nw * nct; k += step) g.ctrc[k] = 0ull;
    // workgroups nb and up: the suppression pass's workgroup b = blockIdx.x - nb moves its
    // undecided boxes to the next list, after those of the workgroups before it
    if ((int)blockIdx.x >= nb) {
      const int b = blockIdx.x - nb;
      const int before = preceding_count(g.bsum, b, s_w);
      const int kb = g.bsum[b];
      for (int k = threadIdx.x; k < kb; k += blockDim.x)
        g.list[before + k] = g.slot[b * kSupThreads + k];
      if (b == (int)gridDim.x - nb - 1 && threadIdx.x == 0) *g.lcount = before + kb;
      return;
    }
  }
  const int na = *nalive;
  const int W = min(na, kNmsWindow);
  const int t0 = blockIdx.x * kNmsCompactBlock + threadIdx.x * 4;
  // the four positions' loads go out together at each step of the chain
  int r[4], su[4];
#pragma unroll
  for (int k = 0; k < 4; k++) r[k] = t0 + k < na ? alive[t0 + k] : -1;
#pragma unroll
  for (int k = 0; k < 4; k++) su[k] = r[k] >= 0 ? sup[r[k]] : kNoSup;
  int c = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) c += (r[k] >= 0 && t0 + k >= W && su[k] == kNoSup) ? 1 : 0;
  for (int m = 1; m < 64; m <<= 1) c += __shfl_xor(c, m);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) blocksum[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// The alive later boxes, in order, to alive_out: blocksum holds each workgroup's count (from
// k_nms_post, positions as there); a workgroup's offset is the sum of the counts before it,
// and the last workgroup writes the total, the next round's alive count.
__global__ __launch_bounds__(256) void k_nms_compact(const int *__restrict__ alive,
                                                     const int *__restrict__ nalive,
                                                     const int *__restrict__ sup,
                                                     const int *__restrict__ blocksum,
                                                     int *__restrict__ nalive_out,
                                                     int *__restrict__ alive_out,
                                                     int *host_total) {
  __shared__ int s_w[4], s_red[4];
  const int before = preceding_count(blocksum, blockIdx.x, s_red);
  const int na = *nalive;
  const int W = min(na, kNmsWindow);
  const int t0 = blockIdx.x * kNmsCompactBlock + threadIdx.x * 4;
  int r[4], su[4];
#pragma unroll
  for (int k = 0; k < 4; k++) r[k] = t0 + k >= W && t0 + k < na ? alive[t0 + k] : -1;
#pragma unroll
  for (int k = 0; k < 4; k++) su[k] = r[k] >= 0 ? sup[r[k]] : 0;
  bool f[4];
  int c = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    f[k] = r[k] >= 0 && su[k] == kNoSup;
    c += f[k] ? 1 : 0;
  }
  // inclusive scan of c over the wave, then the preceding waves' totals
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int inc = c;
  for (int m = 1; m < 64; m <<= 1) {
    const int v = __shfl_up(inc, m);
    if (lane >= m) inc += v;
  }
  if (lane == 63) s_w[wave] = inc;
  __syncthreads();
  int pos = before + inc - c;
  for (int w = 0; w < wave; w++) pos += s_w[w];
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
    const int total = before + s_w[0] + s_w[1] + s_w[2] + s_w[3];
    *nalive_out = total;
    if (host_total) *(volatile int *)host_total = total;  // coherent pinned host memory
  }
#pragma unroll
  for (int k = 0; k < 4; k++)
    if (f[k]) alive_out[pos++] = r[k];
}



// finalisation: conflicts, last-pick emission, survivors in original order
// a tie on score v decided an outcome: flag it unless v's equal-score run is already in the
// reference's order, and enter v in the tie set for the host (open addressing on its bits)
__device__ inline void nms_tie(int *flags, int bit, float v, const float *resolved, int nresolved,
                               float *tie_vals) {
  for (int k = 0; k < nresolved; k++)
    if (resolved[k] == v) return;
  atomicOr(&flags[0], bit);
  const unsigned int b = __float_as_uint(v);
  unsigned int *tab = (unsigned int *)tie_vals;
  unsigned int h = (b * 2654435761u) >> 24;  // kNmsTieCap = 256 slots
  for (int k = 0; k < kNmsTieCap; k++, h = (h + 1) & (kNmsTieCap - 1)) {
    const unsigned int old = atomicCAS(&tab[h], kNmsTieEmpty, b);
    if (old == kNmsTieEmpty) {
      atomicAdd(&flags[2], 1);
      return;
    }
    if (old == b) return;
  }
  atomicOr(&flags[3], 1);  // the set is full
}

__global__ __launch_bounds__(256) void k_nms_check(const NmsBox *__restrict__ sbox, int n,
                                                   const int *__restrict__ status,
                                                   const int *__restrict__ sup,
                                                   const int *__restrict__ picks,
                                                   const int *__restrict__ counters,
                                                   int *__restrict__ flags,
                                                   const float *__restrict__ resolved,
                                                   int nresolved,
                                                   float *__restrict__ tie_vals) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int npick = counters[0];
  if (r >= n || npick == 0) return;
  const int last = picks[npick - 1];
  if (status[r] != 1) {
    const int f = sup[r];
    if (f == kNoSup) {
      atomicOr(&flags[0], 4);  // undecided box: cannot happen; forces the exact host path
      return;
    }
    if (sbox[f].score == sbox[r].score)  // a tie decided the outcome
      nms_tie(flags, 1, sbox[r].score, resolved, nresolved, tie_vals);
    if (f == last) atomicOr(&flags[1], 1);  // last pick suppressed something
  } else if (r != last && sbox[r].score == sbox[last].score) {
    // the identity of the last pick depends on the tie order
    nms_tie(flags, 2, sbox[r].score, resolved, nresolved, tie_vals);
  }
}

__global__ __launch_bounds__(256) void k_nms_keep(const NmsBox *__restrict__ sbox,
                                                  const int *__restrict__ picks,
                                                  const int *__restrict__ counters,
                                                  const int *__restrict__ flags,
                                                  int *__restrict__ keep_orig,
                                                  int *__restrict__ host_tie) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  // the check has finished: its tie flag goes to the mapped host slot
  if (k == 0 && host_tie) *(volatile int *)host_tie = flags[0];
  const int npick = counters[0];
  if (k >= npick) return;
  if (k == npick - 1 && flags[1] == 0) return;
  keep_orig[sbox[picks[k]].idx] = 1;
}

__global__ __launch_bounds__(256) void k_fill(int *__restrict__ p, long long n, int v) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// rocPRIM picks a merge sort below about 1 M items (about 20 dependent launches for a
// 1080p candidate set); onesweep is four digit passes over the same keys, so the merge
// path is switched off.  Onesweep tiles of 256 x 16 keys ranked with the match algorithm:
// 0.131 ms for 1.03 M keys against 0.148 with the default ranking (11-bit digits do not fit
// the LDS at 256 threads; 512-thread tiles neither).
using NmsSortConfig = rocprim::radix_sort_config<
    rocprim::default_config, rocprim::default_config,
    rocprim::radix_sort_onesweep_config<rocprim::kernel_config<256, 12>,
                                        rocprim::kernel_config<256, 16>, 8,
                                        rocprim::block_radix_rank_algorithm::match>,
    0>;

size_t nms_sort_temp_bytes(int n) {
  size_t bytes = 0;
  unsigned int *k = nullptr;
  int *v = nullptr;
  (void)rocprim::radix_sort_pairs<NmsSortConfig>(nullptr, bytes, k, k, v, v, (unsigned int)n, 0,
                                                 32, 0);
  return bytes;
}

void launch_nms_prep(hipStream_t s, const ScoreMap *maps, int nmaps, const float *score,
                     const int *cand, int n, int padx, int pady, const int *rsize, NmsBox *box,
                     unsigned int *keys, int *vals) {
  hipLaunchKernelGGL(k_nms_prep, dim3((n + 255) / 256), dim3(256), 0, s, maps, nmaps, score, cand,
                     n, padx, pady, rsize, box, keys, vals);
}

void launch_nms_prep_rows(hipStream_t s, const float *rows, int n, int GL, NmsBox *box,
                          unsigned int *keys, int *vals) {
  hipLaunchKernelGGL(k_nms_prep_rows, dim3((n + 255) / 256), dim3(256), 0, s, rows, n, GL, box,
                     keys, vals);
}

int launch_nms_sort(hipStream_t s, void *temp, size_t temp_bytes, unsigned int *keys_in,
                    unsigned int *keys_out, int *vals_in, int *vals_out, int n) {
  size_t bytes = temp_bytes;
  return rocprim::radix_sort_pairs<NmsSortConfig>(temp, bytes, keys_in, keys_out, vals_in,
                                                  vals_out, (unsigned int)n, 0, 32, s) == hipSuccess
             ? 0
             : -1;
}

void launch_nms_gather(hipStream_t s, const int *ord, const NmsBox *box, int n, NmsBox *sbox,
                       int *alive, int *status, int *sup, int *counters, int *nalive0,
                       int *flags, int *keep, const NmsGrid *grid, float *tie) {
  NmsGrid g;
  if (grid) g = *grid;
  const int threads = std::max(grid ? std::max(n, kNmsMaskWords * nms_grid_tiles(g)) : n,
                               kNmsTieCap);
  hipLaunchKernelGGL(k_nms_gather, dim3((threads + 255) / 256), dim3(256), 0, s, ord, box, n,
                     sbox, alive, status, sup, counters, nalive0, flags, keep, g, tie);
}

void nms_grid_shape(int W, int H, int *gx, int *gy, int *shift) {
  int sh = 6;  // 64-pixel tiles, coarser when the frame would need more than kNmsMaxTiles
  for (;;) {
    const long long x = ((long long)std::max(W, 1) + (1LL << sh) - 1) >> sh;
    const long long y = ((long long)std::max(H, 1) + (1LL << sh) - 1) >> sh;
    if (x * y <= kNmsMaxTiles) {
      *gx = (int)x;
      *gy = (int)y;
      *shift = sh;
      return;
    }
    sh++;
  }
}

size_t nms_grid_bytes(int gx, int gy) {
  NmsGrid g;
  g.gx = gx;
  g.gy = gy;
  return sizeof(NmsBox) * kNmsWindow +
         sizeof(unsigned long long) * kNmsMaskWords * (size_t)nms_grid_tiles(g);
}

// layout: pick boxes, then the extent, centre and coarse centre masks (contiguous, so the
// gather kernel clears them as one array)
void nms_grid_bind(NmsGrid &g, void *buf) {
  const size_t nt = (size_t)g.gx * g.gy;
  g.pbox = (NmsBox *)buf;
  g.ext = (unsigned long long *)(g.pbox + kNmsWindow);
  g.ctr = g.ext + kNmsMaskWords * nt;
  g.ctrc = g.ctr + kNmsMaskWords * nt;
}

int launch_nms_round(hipStream_t s, const NmsBox *sbox, const int *alive, const int *nalive,
                     int nalive_host, float overlap, unsigned long long *mask, int *status,
                      int *sup, int *picks, int *counters, int *blocksum,
                      const NmsGrid *grid, int round) {
  const int W = nalive_host < kNmsWindow ? nalive_host : kNmsWindow;
  const size_t glds = sizeof(unsigned long long) * kNmsWindow * (kNmsWindow / 64) + 4 * kNmsWindow;
  // set once per process (a function-local static is initialised once, also when serving
  // workers launch concurrently)
  static const bool attr = hipFuncSetAttribute((const void *)k_nms_wingreedy,
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)glds) == hipSuccess;
  if (!attr) return -1;
  hipLaunchKernelGGL(k_nms_winmask, dim3((W + kWinmaskRows - 1) / kWinmaskRows), dim3(256), 0, s,
                     sbox, alive, nalive, overlap, mask);
  // the suppression pass below (launched when the bound leaves boxes after the window) also
  // gives the window's removed boxes their first suppressor
  const int rest = nalive_host - W;
  hipLaunchKernelGGL(k_nms_wingreedy, dim3(1), dim3(256), glds, s, alive, nalive, mask, status,
                     sup, picks, counters, rest <= 0 ? 1 : 0);
  if (rest <= 0) return 0;
  const int nb = (nalive_host + kNmsCompactBlock - 1) / kNmsCompactBlock;
  int nbc = 0;  // suppression workgroups whose undecided boxes the round's end compacts
  if (grid) {
    hipLaunchKernelGGL(k_nms_grid_build, dim3((W + 3) / 4), dim3(256), 0, s, sbox, picks,
                       counters, *grid, round);
    // round 0 takes every candidate; a later round's list holds the alive boxes, at most the
    // host's bound
    nbc = ((round == 0 ? grid->n : std::min(grid->n, nalive_host)) + kSupThreads - 1) /
          kSupThreads;
    hipLaunchKernelGGL(k_nms_grid_sup, dim3(nbc), dim3(kSupThreads), 0, s, counters, *grid,
                       overlap, round, sup);
  } else {
    hipLaunchKernelGGL(k_nms_suppress, dim3((nalive_host + 255) / 256), dim3(256), 0, s, sbox,
                       alive, nalive, picks, counters, overlap, sup);
  }
  NmsGrid none;
  hipLaunchKernelGGL(k_nms_post, dim3(nb + nbc), dim3(256), 0, s, alive, nalive, sup, counters,
                     grid ? *grid : none, nb, blocksum);
  return 0;
}

void launch_nms_compact(hipStream_t s, const int *alive, const int *nalive, const int *sup,
                        int *blocksum, int bound, int *nalive_out, int *alive_out,
                        int *host_total) {
  const int nb = (bound + kNmsCompactBlock - 1) / kNmsCompactBlock;
  hipLaunchKernelGGL(k_nms_compact, dim3(nb), dim3(256), 0, s, alive, nalive, sup, blocksum,
                     nalive_out, alive_out, host_total);
}

void launch_nms_check(hipStream_t s, const NmsBox *sbox, int n, const int *status, const int *sup,
                      const int *picks, const int *counters, int *flags, const float *resolved,
                      int nresolved, float *tie_vals) {
  hipLaunchKernelGGL(k_nms_check, dim3((n + 255) / 256), dim3(256), 0, s, sbox, n, status, sup,
                     picks, counters, flags, resolved, nresolved, tie_vals);
}

void launch_nms_keep(hipStream_t s, const NmsBox *sbox, const int *picks, const int *counters,
                     int bound, const int *flags, int *keep_orig, int *host_tie) {
  if (bound > 0)
    hipLaunchKernelGGL(k_nms_keep, dim3((bound + 255) / 256), dim3(256), 0, s, sbox, picks,
                       counters, flags, keep_orig, host_tie);
}

void launch_fill(hipStream_t s, int *p, long long n, int v) {
  if (n > 0)
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}

}  // namespace geye

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

// sorted order: sbox[r] = box[ord[r]]; optional rank lookup by flat score index
__global__ __launch_bounds__(256) void k_nms_gather(const int *__restrict__ ord,
                                                    const NmsBox *__restrict__ box, int n,
                                                    NmsBox *__restrict__ sbox,
                                                    int *__restrict__ alive,
                                                    int *__restrict__ status,
                                                    int *__restrict__ sup,
                                                    const int *__restrict__ cand,
                                                    int *__restrict__ rank_of_flat,
                                                    int *__restrict__ counters,
                                                    int *__restrict__ nalive0,
                                                    int *__restrict__ flags,
                                                    int *__restrict__ keep) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  // the pass's other initial state: pick counters, the first round's alive count, the tie
  // flags and the survivor flags (written by the finalisation, after the rounds)
  if (r < 4) {
    counters[r] = 0;
    flags[r] = 0;
  }
  if (r == 0) *nalive0 = n;
  if (r >= n) return;
  const int i = ord[r];
  sbox[r] = box[i];
  alive[r] = r;
  status[r] = 0;
  sup[r] = kNoSup;
  keep[r] = 0;
  if (rank_of_flat) rank_of_flat[cand[i]] = r;
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

// Map path: each pick scatters its suppression over the score-map cells whose root boxes
// can intersect it (a conservative cell rectangle per map).  The workgroup lays the cells
// of all maps end to end and every thread walks its share with a running map index, so the
// rank loads are independent.  The candidate's box is rebuilt from its cell with the prep
// kernel's arithmetic and the exact overlap test decides; the first suppressor is the
// minimum pick rank.
__global__ __launch_bounds__(256) void k_nms_scatter(const NmsBox *__restrict__ sbox,
                                                     const int *__restrict__ picks,
                                                     const int *__restrict__ counters,
                                                     const ScoreMap *__restrict__ maps,
                                                     int nmaps, const int *__restrict__ rsize,
                                                     int padx, int pady,
                                                     const int *__restrict__ rank_of_flat,
                                                     float overlap, int *__restrict__ sup) {
  // per-map tables in dynamic LDS sized to nmaps (40 B per map): a workgroup takes only what
  // the plan's maps need, so more of them (and other kernels) share a CU
  extern __shared__ long long s_dyn[];
  long long *s_off = s_dyn;                                 // [nmaps]
  int *s_xa = (int *)(s_off + nmaps);                       // [nmaps] each
  int *s_ya = s_xa + nmaps, *s_h = s_ya + nmaps, *s_R0 = s_h + nmaps;
  float *s_s = (float *)(s_R0 + nmaps), *s_fh = s_s + nmaps, *s_fw = s_fh + nmaps;
  int *s_beg = (int *)(s_fw + nmaps);                       // [nmaps + 1]
  __shared__ int s_part[256];
  const int p0 = counters[1], p1 = counters[0];
  if ((int)blockIdx.x >= p1 - p0) return;
  const int rp = picks[p0 + blockIdx.x];
  const NmsBox P = sbox[rp];
  const int tid = threadIdx.x;
  // per-map rectangles; thread t owns maps [t*per, (t+1)*per).  A cell's box B (width
  // bw = rw*s, height bh = rh*s) can be suppressed only if its overlap with the pick, ix by iy,
  // exceeds overlap * R with R the smaller area; iy <= min(bh, ph) then needs ix > tx =
  // overlap * R / min(bh, ph) (and likewise iy > ty), which bounds B's corner on both sides.
  // The bounds are relaxed by 0.1% and one cell each way; the exact test decides.
  const int per = (nmaps + 255) / 256;
  const float pw = P.x2 - P.x1 + 1.0f, ph = P.y2 - P.y1 + 1.0f;
  int mycnt = 0;
  for (int k = tid * per; k < min(nmaps, (tid + 1) * per); k++) {
    const ScoreMap m = maps[k];
    const float s = m.scale;
    const int rh = rsize[m.comp * 2], rw = rsize[m.comp * 2 + 1];
    const float bw = (float)rw * s, bh = (float)rh * s;
    const float R = fminf(bw * bh, P.area) * 0.999f;
    const float tx = overlap > 0.0f ? overlap * R / fminf(bh, ph) : 0.0f;
    const float ty = overlap > 0.0f ? overlap * R / fminf(bw, pw) : 0.0f;
    const int xa = max((int)floorf((P.x1 - bw + tx) / s) + padx - 2, 0);
    const int xb = min((int)ceilf((P.x2 + 1.0f - tx) / s) + padx + 1, m.R1 - 1);
    const int ya = max((int)floorf((P.y1 - bh + ty) / s) + pady - 2, 0);
    const int yb = min((int)ceilf((P.y2 + 1.0f - ty) / s) + pady + 1, m.R0 - 1);
    const int h = yb - ya + 1, w = xb - xa + 1;
    s_xa[k] = xa;
    s_ya[k] = ya;
    s_h[k] = h > 0 ? h : 1;
    s_R0[k] = m.R0;
    s_off[k] = m.score_off;
    s_s[k] = s;
    s_fh[k] = (float)rh * s;
    s_fw[k] = (float)rw * s;
    s_beg[k] = (h > 0 && w > 0) ? w * h : 0;  // count for now
    mycnt += s_beg[k];
  }
  s_part[tid] = mycnt;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {  // inclusive scan of per-thread counts
    const int v = tid >= o ? s_part[tid - o] : 0;
    __syncthreads();
    s_part[tid] += v;
    __syncthreads();
  }
  int run = s_part[tid] - mycnt;
  for (int k = tid * per; k < min(nmaps, (tid + 1) * per); k++) {
    const int c = s_beg[k];
    s_beg[k] = run;
    run += c;
  }
  if (tid == 255) s_beg[nmaps] = s_part[255];
  __syncthreads();
  const int total = s_beg[nmaps];
  // each wave takes one contiguous chunk of the cells laid end to end (lanes consecutive,
  // so the rank loads stay coalesced) and crosses only the maps that chunk spans; the
  // chunk's first map comes from a binary search of the offsets
  constexpr int B = 4;  // cells per batch: their rank loads are in flight together
  const int lane = tid & 63;
  const int nw = 4 * gridDim.y, wv = blockIdx.y * 4 + (tid >> 6);
  const int chunk = (total + nw * 64 * B - 1) / (nw * 64 * B) * (64 * B);
  const int c0 = wv * chunk, c1 = min(total, c0 + chunk);
  int k = 0;
  {
    int lo = 0, hi = nmaps - 1;  // last map with s_beg[k] <= c0
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_beg[mid] <= c0) lo = mid; else hi = mid - 1;
    }
    k = lo;
  }
  const int stride = 64;
  for (int base = c0 + lane; base < c1; base += B * stride) {
    int kk[B], xs[B], ys[B], rr[B];
#pragma unroll
    for (int u = 0; u < B; u++) {
      const int idx = base + u * stride;
      rr[u] = -1;
      if (idx < c1) {
        while (s_beg[k + 1] <= idx) k++;
        const int off = idx - s_beg[k];
        const int h = s_h[k];
        const int xo = off / h;
        kk[u] = k;
        xs[u] = s_xa[k] + xo;
        ys[u] = s_ya[k] + (off - xo * h);
        rr[u] = rank_of_flat[s_off[k] + (long long)xs[u] * s_R0[k] + ys[u]];
      }
    }
#pragma unroll
    for (int u = 0; u < B; u++) {
      const int r = rr[u];
      if (r <= rp) continue;  // not a candidate (-1), or ranked before the pick
      const int km = kk[u];
      const float s = s_s[km];
      NmsBox b;
      b.y1 = ((float)ys[u] - (float)pady + 1.0f) * s;
      b.x1 = ((float)xs[u] - (float)padx + 1.0f) * s;
      b.y2 = (b.y1 + s_fh[km]) - 1.0f;  // == (float)((double)(y1 + rh*s) - 1.0)
      b.x2 = (b.x1 + s_fw[km]) - 1.0f;
      b.area = (float)(((double)(b.x2 - b.x1) + 1.0) * ((double)(b.y2 - b.y1) + 1.0));
      // sup only decreases: a box that already names an earlier suppressor needs no atomic
      if (suppresses(P, b, overlap) &&
          __hip_atomic_load(&sup[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > rp)
        atomicMin(&sup[r], rp);
    }
  }
}


constexpr int kNmsCompactBlock = 1024;  // positions per workgroup of the per-round counts

// End of a round, fused: every alive position t < na is either decided (a window box, or a
// later box some pick suppressed), whose score cell then leaves rank_of_flat (-1, when the
// map path is on), or still alive.  Each workgroup counts the alive later boxes of its
// kNmsCompactBlock positions (4 consecutive per thread) into blocksum; positions past the
// true count add nothing, so a grid over a host bound counts exactly.

__global__ __launch_bounds__(256) void k_nms_post(const int *__restrict__ alive,
                                                  const int *__restrict__ nalive,
                                                  const int *__restrict__ sup,
                                                  const NmsBox *__restrict__ sbox,
                                                  const int *__restrict__ cand,
                                                  int *__restrict__ rank_of_flat,
                                                  int *__restrict__ blocksum) {
  __shared__ int s_w[4];
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
  bool ret[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const bool stay = r[k] >= 0 && t0 + k >= W && su[k] == kNoSup;
    c += stay ? 1 : 0;
    ret[k] = r[k] >= 0 && !stay;
  }
  if (rank_of_flat) {
    int ci[4];
#pragma unroll
    for (int k = 0; k < 4; k++) ci[k] = ret[k] ? sbox[r[k]].idx : -1;
    int fl[4];
#pragma unroll
    for (int k = 0; k < 4; k++) fl[k] = ci[k] >= 0 ? cand[ci[k]] : -1;
#pragma unroll
    for (int k = 0; k < 4; k++)
      if (fl[k] >= 0) rank_of_flat[fl[k]] = -1;
  }
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
__global__ __launch_bounds__(256) void k_nms_check(const NmsBox *__restrict__ sbox, int n,
                                                   const int *__restrict__ status,
                                                   const int *__restrict__ sup,
                                                   const int *__restrict__ picks,
                                                   const int *__restrict__ counters,
                                                   int *__restrict__ flags) {
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
    if (sbox[f].score == sbox[r].score) atomicOr(&flags[0], 1);  // a tie decided the outcome
    if (f == last) atomicOr(&flags[1], 1);                      // last pick suppressed something
  } else if (r != last && sbox[r].score == sbox[last].score) {
    atomicOr(&flags[0], 2);  // the identity of the last pick depends on the tie order
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
                       int *alive, int *status, int *sup, const int *cand, int *rank_of_flat,
                       int *counters, int *nalive0, int *flags, int *keep) {
  hipLaunchKernelGGL(k_nms_gather, dim3((n + 255) / 256), dim3(256), 0, s, ord, box, n, sbox,
                     alive, status, sup, cand, rank_of_flat, counters, nalive0, flags, keep);
}

int launch_nms_round(hipStream_t s, const NmsBox *sbox, const int *alive, const int *nalive,
                     int nalive_host, float overlap, unsigned long long *mask, int *status,
                      int *sup, int *picks, int *counters, int *blocksum,
                      const NmsMaps *maps) {
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
  if (maps) {
    // s_off 8 B, seven 4-byte tables and s_beg (nmaps + 1 entries) per map
    const size_t slds = 40 * (size_t)maps->nmaps + 4;
    hipLaunchKernelGGL(k_nms_scatter, dim3(W, kNmsScatterSplit), dim3(256), slds, s, sbox, picks,
                       counters, maps->maps,
                       maps->nmaps, maps->rsize, maps->padx, maps->pady, maps->rank_of_flat,
                       overlap, sup);
  } else {
    hipLaunchKernelGGL(k_nms_suppress, dim3((nalive_host + 255) / 256), dim3(256), 0, s, sbox,
                       alive, nalive, picks, counters, overlap, sup);
  }
  const bool retire = maps && maps->cand;
  hipLaunchKernelGGL(k_nms_post, dim3(nb), dim3(256), 0, s, alive, nalive, sup, sbox,
                     retire ? maps->cand : nullptr, retire ? maps->rank_of_flat : nullptr,
                     blocksum);
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
                      const int *picks, const int *counters, int *flags) {
  hipLaunchKernelGGL(k_nms_check, dim3((n + 255) / 256), dim3(256), 0, s, sbox, n, status, sup,
                     picks, counters, flags);
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

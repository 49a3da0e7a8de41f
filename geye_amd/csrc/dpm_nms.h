// dpm_nms.h -- device NMS launchers (dpm_nms.hip).
#pragma once

#include "dpm_internal.h"

namespace geye {

constexpr int kNmsWindow = 1024;
constexpr int kNoSup = 0x7fffffff;
// tie records of the finalisation check (a hash set of the scores whose equal-score order
// decided an outcome, empty slots kNmsTieEmpty, a NaN no score takes) and the scores whose
// equal-score order the host has set to the reference's
constexpr int kNmsTieCap = 256;
constexpr int kNmsResolvedCap = 32;
constexpr unsigned int kNmsTieEmpty = 0x7fc00001u;

struct NmsBox {
  float y1, x1, y2, x2, area, score;
  int idx, pad;
};

// Spatial index of one round's picks for the suppression pass (overlap >= 0.5).  The frame
// is cut into square tiles of 2^shift pixels (edge tiles also take what lies outside the
// frame), and into coarse tiles of 4 x 4 of them.  Each tile holds two bitmasks over the
// round's picks (bit k = the round's k-th pick): the picks whose box, grown by kNmsTileMargin
// pixels, touches the tile (extent masks), and the picks whose centre lies in it (centre
// masks, also per coarse tile).  Masks are stored word-major: word w of every tile, then
// word w + 1.
constexpr int kNmsMaskWords = kNmsWindow / 64;
constexpr int kNmsMaxTiles = 4096;
constexpr float kNmsTileMargin = 2.0f;
struct NmsGrid {
  int gx = 0, gy = 0, shift = 0;   // fine tiles (the coarse grid: ceil(g / 4), shift + 2)
  unsigned long long *ext = nullptr;   // [kNmsMaskWords][gx * gy]
  unsigned long long *ctr = nullptr;   // [kNmsMaskWords][gx * gy]
  unsigned long long *ctrc = nullptr;  // [kNmsMaskWords][coarse tiles]
  NmsBox *pbox = nullptr;          // [kNmsWindow] the round's picks, their rank in .pad
  int *rank_of = nullptr;          // candidate -> position in the sorted order
  unsigned char *dec = nullptr;    // candidate decided (picked, or its first suppressor known)
  const NmsBox *box = nullptr;     // the candidates' boxes in candidate order
  int n = 0;                       // candidates
  // the candidates still undecided after a round, in candidate order (read by the next
  // round's suppression pass), and their count; the pass leaves each workgroup's undecided
  // boxes at its places of slot (n rounded up to 1024 entries) and their count in bsum, which
  // the round's end compacts
  int *list = nullptr, *lcount = nullptr, *slot = nullptr, *bsum = nullptr;
};
// tile edge and grid of a W x H frame (at most kNmsMaxTiles tiles)
void nms_grid_shape(int W, int H, int *gx, int *gy, int *shift);
// device bytes of one grid's masks and pick boxes
size_t nms_grid_bytes(int gx, int gy);
// points grid's masks and pbox into a buffer of nms_grid_bytes(gx, gy)
void nms_grid_bind(NmsGrid &g, void *buf);

size_t nms_sort_temp_bytes(int n);
void launch_nms_prep(hipStream_t s, const ScoreMap *maps, int nmaps, const float *score,
                     const int *cand, int n, int padx, int pady, const int *rsize, NmsBox *box,
                     unsigned int *keys, int *vals);
void launch_nms_prep_rows(hipStream_t s, const float *rows, int n, int GL, NmsBox *box,
                          unsigned int *keys, int *vals);
int launch_nms_sort(hipStream_t s, void *temp, size_t temp_bytes, unsigned int *keys_in,
                    unsigned int *keys_out, int *vals_in, int *vals_out, int n);
// also zeroes counters[0..3], flags[0..3] and keep[0..n), empties the tie set (tie, when
// non-null) and sets *nalive0 = n; with a grid, fills its rank_of, clears dec and the masks
void launch_nms_gather(hipStream_t s, const int *ord, const NmsBox *box, int n, NmsBox *sbox,
                       int *alive, int *status, int *sup, int *counters, int *nalive0,
                       int *flags, int *keep, const NmsGrid *grid, float *tie);
// One round: window bitmask, greedy, suppression of the later boxes, then the fused
// retire-and-count pass that leaves per-block counts of the surviving later boxes in
// blocksum.  nalive_host is a host upper bound of *nalive: the kernels read the true count
// and do nothing past it.
// grid (overlap >= 0.5 only) selects the indexed suppression pass, one thread per undecided
// candidate testing the picks the index names; without it every later box is tested against
// every pick.  Returns -1 (nothing launched) when the greedy kernel's LDS attribute could not
// be set.
int launch_nms_round(hipStream_t s, const NmsBox *sbox, const int *alive, const int *nalive,
                      int nalive_host, float overlap, unsigned long long *mask, int *status,
                      int *sup, int *picks, int *counters, int *blocksum, const NmsGrid *grid,
                      int round);
// The round's surviving later boxes, in order, to alive_out, their count to *nalive_out;
// bound >= *nalive (positions of the round).
// host_total (optional): coherent pinned host memory that also receives the new alive count
void launch_nms_compact(hipStream_t s, const int *alive, const int *nalive, const int *sup,
                        int *blocksum, int bound, int *nalive_out, int *alive_out,
                        int *host_total);
// the pick count is read from counters[0] on the device; bound >= that count.  A tie that
// decided an outcome sets flags[0] unless its score is one of the nresolved scores in
// resolved (their equal-score runs are in the reference's order); each such score is also
// entered in the tie set tie_vals (flags[2] counts the distinct scores entered, flags[3] is
// set when the set is full).
void launch_nms_check(hipStream_t s, const NmsBox *sbox, int n, const int *status, const int *sup,
                      const int *picks, const int *counters, int *flags, const float *resolved,
                      int nresolved, float *tie_vals);
void launch_nms_keep(hipStream_t s, const NmsBox *sbox, const int *picks, const int *counters,
                     int bound, const int *flags, int *keep_orig,
                     int *host_tie);
void launch_fill(hipStream_t s, int *p, long long n, int v);

}  // namespace geye

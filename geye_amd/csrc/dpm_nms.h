// dpm_nms.h -- device NMS launchers (dpm_nms.hip).
#pragma once

#include "dpm_internal.h"

namespace geye {

constexpr int kNmsWindow = 1024;
constexpr int kNmsMaxMaps = 1024;
constexpr int kNoSup = 0x7fffffff;
constexpr int kNmsScatterSplit = 8;  // workgroups per pick in the scatter pass

struct NmsBox {
  float y1, x1, y2, x2, area, score;
  int idx, pad;
};

// score-map geometry for the scatter form of the suppression pass
struct NmsMaps {
  const ScoreMap *maps;
  int nmaps;
  const int *rsize;
  int padx, pady;
  int *rank_of_flat;
  const int *cand;   // candidate -> flat score index (retires decided cells between rounds)
};

size_t nms_sort_temp_bytes(int n);
void launch_nms_prep(hipStream_t s, const ScoreMap *maps, int nmaps, const float *score,
                     const int *cand, int n, int padx, int pady, const int *rsize, NmsBox *box,
                     unsigned int *keys, int *vals);
void launch_nms_prep_rows(hipStream_t s, const float *rows, int n, int GL, NmsBox *box,
                          unsigned int *keys, int *vals);
int launch_nms_sort(hipStream_t s, void *temp, size_t temp_bytes, unsigned int *keys_in,
                    unsigned int *keys_out, int *vals_in, int *vals_out, int n);
// also zeroes counters[0..3], flags[0..3] and keep[0..n), and sets *nalive0 = n
void launch_nms_gather(hipStream_t s, const int *ord, const NmsBox *box, int n, NmsBox *sbox,
                       int *alive, int *status, int *sup, const int *cand, int *rank_of_flat,
                       int *counters, int *nalive0, int *flags, int *keep);
// One round: window bitmask, greedy, suppression of the later boxes, then the fused
// retire-and-count pass that leaves per-block counts of the surviving later boxes in
// blocksum.  nalive_host is a host upper bound of *nalive: the kernels read the true count
// and do nothing past it.
// maps selects the score-map scatter for the suppression of the later boxes; without it
// every later box is tested against every pick.  Returns -1 (nothing launched) when the
// greedy kernel's LDS attribute could not be set.
int launch_nms_round(hipStream_t s, const NmsBox *sbox, const int *alive, const int *nalive,
                      int nalive_host, float overlap, unsigned long long *mask, int *status,
                      int *sup, int *picks, int *counters, int *blocksum, const NmsMaps *maps);
// The round's surviving later boxes, in order, to alive_out, their count to *nalive_out;
// bound >= *nalive (positions of the round).
// host_total (optional): coherent pinned host memory that also receives the new alive count
void launch_nms_compact(hipStream_t s, const int *alive, const int *nalive, const int *sup,
                        int *blocksum, int bound, int *nalive_out, int *alive_out,
                        int *host_total);
// the pick count is read from counters[0] on the device; bound >= that count
void launch_nms_check(hipStream_t s, const NmsBox *sbox, int n, const int *status, const int *sup,
                      const int *picks, const int *counters, int *flags);
void launch_nms_keep(hipStream_t s, const NmsBox *sbox, const int *picks, const int *counters,
                     int bound, const int *flags, int *keep_orig,
                     int *host_tie);
void launch_fill(hipStream_t s, int *p, long long n, int v);

}  // namespace geye

// dpm_kernels.h -- launchers of the gfx950 DPM kernels (defined in the .hip files).
#pragma once

#include "dpm_internal.h"

namespace geye {

void launch_resize_rows(hipStream_t s, const ResizeOp *ops, int nops, int blocks,
                        const int2 *tab, const int *si, const float *alpha, const float *src,
                        float *dst);
void launch_resize_cols(hipStream_t s, const ResizeOp *ops, int nops, int blocks,
                        const int2 *tab, const int *si, const float *alpha, const float *src,
                        float *dst);
// HOG histogram in two launches: gradients per pyramid image, then votes per level
// (HogLevel tiles: kHogVoteCols cell columns of kHogVoteRows cells per workgroup).  grid holds
// the magnitude plane (4 * grid_total bytes) followed by the orientation plane (grid_total).
// bgr (optional): the round reads the frames' BGR bytes (rows ws bytes apart, frames
// frame_bytes apart) in place of the planar source image (the first round)
void launch_resize_fused(hipStream_t s, const FusedResize *ops, int nops, int blocks,
                         const int2 *tab, const int4 *si4, const float4 *al4, const float *src,
                         float *dst, const unsigned char *bgr = nullptr, int ws = 0,
                         long long frame_bytes = 0);
// bgr / ws / frame_bytes: the frames of the images whose GradImg::frame is >= 0
void launch_hog_grad(hipStream_t s, const GradImg *imgs, int nimg, int blocks, const float *img,
                     void *grid, long long grid_total, const unsigned char *bgr = nullptr,
                     int ws = 0, long long frame_bytes = 0);
void launch_hog_votes(hipStream_t s, const HogLevel *lv, int nlv, int blocks, const void *grid,
                      long long grid_total, float *hist, float *norm);
void launch_hog_feat(hipStream_t s, const HogLevel *lv, int nlv, int blocks, const float *hist,
                     const float *norm, float *feat, void *featb, int hkind);
void launch_conv_f32(hipStream_t s, int nf, const ConvTile *tiles, int ntiles, int fh, int fw,
                     const float *feat, const float *filt, const long long *map_off,
                     const int2 *map_dims, float *out);
// The DT launchers return -1 (nothing launched) when the dynamic LDS the jobs need could
// not be enabled for the kernel.
int launch_dt(hipStream_t s, const DtJob *jobs, int njobs, size_t lds, const int4 *nodes,
              const int *depths, const float *src, float *dst, int *ptr);
// the 16-bit forms' DT: windowed minimum, the same outputs as launch_dt
int launch_dt_window(hipStream_t s, const DtJob *jobs, int njobs, size_t lds, const int4 *nodes,
                     const int *depths, const float *src, float *dst, int *ptr);
void launch_score(hipStream_t s, const ScoreMap *maps, int nmaps, const ScorePart *parts,
                  const float *conv, const float *M, float thresh, int half, float *score,
                  int *flags, int blocks);
int compact_blocks(long long n);
// host_total (optional): mapped host memory that also receives the count
void launch_compact(hipStream_t s, const int *flags, long long n, int *blocksum, int *total,
                    int *out, int *host_total = nullptr);
// out[f] = first candidate of frame f of a batch, out[nf] = *count (out: mapped host memory)
void launch_frame_offsets(hipStream_t s, const int *cand, const int *count, long long score_one,
                          int nf, int *out);
void launch_rows(hipStream_t s, const ScoreMap *maps, int nmaps, const ScorePart *parts,
                 const float *score, const int *cand, const int *sel, int ncand,
                 const int *tmpIx, const int *tmpIy, int half, int GL, int padx, int pady,
                 const int *rsize, float *rows);

}  // namespace geye

// dpm_internal.h -- shared host/device definitions of the gfx950 DPM detection path.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/geye_dpm.h"

namespace geye {

// ---------------------------------------------------------------------------------------
// Device work descriptors (plain structs, copied to HBM once per plan).
// ---------------------------------------------------------------------------------------

// One separable-resize pass over one image: out[c][dy][x] (rows pass) or
// out[c][y][dx] (columns pass) as a weighted sum of input samples; (start, count) index
// the interpolation table per destination index.
struct ResizeOp {
  long long in_off, out_off;   // element offsets into the image arena
  int in_h, in_w;              // input image (row-major [3][in_h][in_w])
  int out_h, out_w;            // output image
  int tab_off;                 // offset of this op's (start, count) pairs
  int blocks_begin;            // first flat block of this op in the launch
};

// Gradient grid of one pyramid image: per pixel (x, y), 1 <= x <= w-2, 1 <= y <= h-2, the
// strongest channel's gradient magnitude v (float) and snapped orientation o (one byte), in
// two column-major planes of the grid arena -- v at mag[grid_off + x*gs + y], o at
// ori[grid_off + x*gs + y] -- with the column stride gs = grid_stride(h), h rounded up to
// kGridAlign so that every 4-row group starts 16-byte aligned.  Tiles of 64 (x) x kGradTY (y)
// pixels.
constexpr int kGradTY = 16, kGridAlign = 32;
__host__ __device__ inline int grid_stride(int h) { return (h + kGridAlign - 1) & ~(kGridAlign - 1); }
struct GradImg {
  long long img_off, grid_off;
  int h, w;
  int tiles_x, tiles_y;
  int blocks_begin;
  int frame;                   // >= 0: the image is that frame of the batch, read as BGR bytes
};

// One fused resize of a whole image: out = cols(rows(in)), 16 x 64 output tiles.  rtab /
// ctab index the (start, count) tap lists of the row and column passes.
struct FusedResize {
  long long in_off, out_off;
  int in_h, in_w, out_h, out_w;
  int rtab_off, ctab_off;
  int tiles_x, tiles_y, blocks_begin;
  int frame;                   // frame of a batch (the first round reads that frame's bytes)
};
// fused resize tile: 8 x 64 outputs (13 KB of LDS, so a CU holds as many workgroups as its
// wave slots allow; on 1080p 8 rows beat 16 and 32); span: source columns
constexpr int kResizeTY = 8, kResizeTX = 64, kResizeSpan = 136;

// HOG vote workgroup: kHogVoteCols adjacent cell columns (one wave each) of kHogVoteRows
// cells (one lane per cell); one column per workgroup measured fastest (2 and 4 share the
// staged pixel columns but leave waves idle between barriers)
constexpr int kHogVoteRows = 64, kHogVoteCols = 1;

struct HogLevel {
  long long img_off;           // image arena offset ([3][h][w] row-major)
  long long grid_off;          // gradient grid of the image (GradImg)
  long long hist_off;          // hist arena offset ([18][b1][b0])
  long long norm_off;          // norm arena offset ([b1][b0])
  long long feat_off;          // feature arena offset ([31][F1][F0])
  long long featb_off;         // bf16 channel-last arena offset ([F1][F0][32])
  int h, w, sbin, b0, b1, F0, F1;
  int tiles_x, tiles_y;        // vote tiles: ceil(b1 / kHogVoteCols) x ceil(b0 / kHogVoteRows)
  int blocks_begin;            // first flat block (hist or feat launch)
  int T;                       // cell rows per vote tile (kHogVoteRows)
};

// One convolution tile: NF filters of one shape against one feature level.
struct ConvTile {
  long long feat_off;          // [31][F1][F0]
  int F0, F1;
  int H, W;                    // output map size
  int x0, y0;
  long long filt_off;          // float offset of the group in the packed filter buffer
  int out_slot;                // first output map slot of the group
  int nvalid;                  // real filters in the group (the rest are zero padding)
};

// One MFMA convolution tile: 4 output columns x 64 rows of one level for up to 16*NT
// filters of one class.
struct MfmaTile {
  long long featb_off;         // 16-bit channel-last feature arena offset ([F1][F0][32])
  int F0, F1;
  int x0, y0;
  long long filt_off;          // 16-bit element offset of the group's packed filters
  int out_slot;                // slot of the group's first member
  int nvalid;                  // real filters in the group
  int fh, fw, nt;              // filter class shape and 16-filter tiles (work-queue kernel)
  int cost;                    // relative work, for ordering the queue
};

// Sum of the counts of the workgroups before workgroup b (a 256-thread workgroup reads them
// together; s_red holds 4 ints).  Stream compactions use it in place of a separate scan
// launch: each workgroup reads at most a few thousand counts.
__device__ inline int preceding_count(const int *__restrict__ blocksum, int b, int *s_red) {
  int part = 0;
  for (int j = threadIdx.x; j < b; j += 256) part += blocksum[j];
  for (int m = 32; m > 0; m >>= 1) part += __shfl_xor(part, m);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = part;
  __syncthreads();
  const int r = s_red[0] + s_red[1] + s_red[2] + s_red[3];
  __syncthreads();
  return r;
}

// Slots of a context's mapped host counts: 0 and 1 alternate between NMS rounds; the
// candidate count of the last scoring pass; the survivor count and the tie flag of the last
// NMS.
constexpr int kHostCandidates = 2, kHostSurvivors = 3, kHostTie = 4;

// Halo on each side of an LDS line in the windowed DT (at least its widest window
// half-width, 16).
constexpr int kDtHalo = 24;
// dynamic LDS limit of the windowed DT kernel (160 KB less its static reduction buffer)
constexpr int kDtWindowMaxLds = 160 * 1024 - 256;

// LDS words per line of the windowed DT: the line and one halo, which it shares with the
// next line (line l at l * stride + kDtHalo, kDtHalo more words after the last line; every
// halo word holds +inf, a cost that never wins).  Padded so that the row pass's lanes, lines
// fastest, fall on distinct banks of the 32 that serve each half wave: with NL <= 32 lines
// the stride is 32 / NL modulo 32, otherwise odd.
__host__ __device__ inline int dt_line_stride(int n, int NL) {
  int w = n + kDtHalo;
  if (NL <= 32) {
    const int r = 32 / NL;
    w += ((r - w) % 32 + 32) % 32;
  } else {
    w |= 1;
  }
  return w;
}

// One DT job: a block of lines (columns or rows) of one map.
struct DtJob {
  long long src_off, dst_off;  // map base offsets (part score arena / DT arenas)
  int n, nlines, line0;        // line length, lines in this job, first line
  int line_stride, elem_stride;
  int node_off, depth_off, depth_count;  // node table for length n
  int G;                       // lanes per line
  int neg;                     // negate the source on load
  float a, b;
  // outputs the windowed DT writes (the 16-bit forms): 0 every output; 1 (columns pass) M
  // everywhere, the argmin on rows y == ay mod 2; 2 (rows pass) on rows y == ay - 1 mod 2
  // M at columns x == ax - 1 mod 2, on rows y == ay mod 2 the argmin at x == ax mod 2 --
  // exactly the cells the scoring and the part boxes read.  With sel != 0 the selected
  // argmins and the rows pass's M go to half grids, cell (y, x) at (x >> 1) * ceil(P0 / 2)
  // + (y >> 1) (the columns pass's argmins at x * ceil(P0 / 2) + (y >> 1)), argmins as
  // 16-bit indices
  int sel, ax, ay;
  // windowed form: lines per batch of the divide and conquer it falls back to when the
  // window would be too wide (float lines, argmins and node table within the window lines'
  // LDS)
  int fb_lines;
};

struct ScoreMap {              // one (valid level, component) root score map
  long long root_off;          // root conv output offset
  long long score_off;         // flat score array offset (reference output order)
  int R0, R1;
  int level, comp;
  int npart;
  int part_first;              // index into ScorePart[]
  float off;                   // offw[ridx[comp]]
  float scale;                 // sbin / scales[level]
  int RX, RY;                  // accumulated-score shift of this component at this level
  int tab_off;                 // level's (Xn per image column, Yn per image row) table
  int blocks_begin;            // first 256-cell block of this map in the score launch
};

struct ScorePart {
  long long M_off;             // DT output map
  long long Ix_off;            // tmpIx arena offset (same map geometry)
  long long Iy_off;            // tmpIy arena offset
  int P0, P1, ax, ay, ps0, ps1;
};

// ---------------------------------------------------------------------------------------
// Host-side plan for one (frame size, model) pair.
// ---------------------------------------------------------------------------------------
struct Level {
  int img_h, img_w, sbin, F0, F1;
  float scale;
  int img_index;               // which arena image feeds this level
  bool valid;                  // root level with F0 >= max_Y && F1 >= max_X
};

struct DeviceBuffer {
  void *p = nullptr;
  size_t bytes = 0;
};

struct Plan {
  int W = 0, H = 0, ws = 0;
  int max_scale = 0, LEN = 0, interval = 0, sbin = 0;
  // A batch plan covers nb frames: every descriptor of a one-frame plan repeated per frame,
  // frame f's arenas at f times the one-frame sizes below (its features at f * feat_stride /
  // featb_stride, which may be a deeper model's one-frame size when the pyramid is shared).
  int nb = 1;
  long long img_one = 0, grid_one = 0, hist_one = 0, norm_one = 0, conv_one = 0, dt_one = 0,
            score_one = 0, feat_stride = 0, featb_stride = 0;
  int nmaps_one = 0, nparts_one = 0;
  std::vector<Level> lv;
  // images: index -> (h, w, arena offset)
  std::vector<int> img_h, img_w;
  std::vector<long long> img_off;
  long long img_total = 0;
  // resize rounds: each round is (rows ops, cols ops) over flat blocks
  struct Round {
    std::vector<ResizeOp> rows, cols;
    int rows_blocks = 0, cols_blocks = 0;
    long long tmp_total = 0;
  };
  std::vector<Round> rounds;
  std::vector<std::vector<FusedResize>> fused;  // per round
  std::vector<int> fused_blocks;
  std::vector<int2> resize_tab;      // (start, count) per destination index, all ops
  std::vector<int> resize_si;
  std::vector<float> resize_alpha;
  std::vector<int4> resize_si4;       // the same taps padded to 4 per destination index
  std::vector<float4> resize_al4;
  // HOG
  std::vector<HogLevel> hog;          // [T16 levels][T8 levels][feature levels]
  std::vector<GradImg> grad;          // one per pyramid image
  int grad_blocks = 0;
  long long grid_total = 0;           // pixels of the gradient planes (5 bytes each)
  int hog16_n = 0, hog8_n = 0, hogf_n = 0;
  int hog_hist_blocks = 0, hog8_blocks = 0, hog_feat_blocks = 0;
  long long hist_total = 0, norm_total = 0, feat_total = 0, featb_total = 0;
  std::vector<long long> feat_off;    // per level, -1 when the level has no features
  // convolution
  std::vector<ConvTile> conv_tiles;    // all tiles, grouped by shape class
  struct ConvClass {
    int h, w, nf, tile_begin, tile_count, filt_off;
  };
  std::vector<ConvClass> conv_classes;
  std::vector<MfmaTile> mfma_tiles, mfma_tiles8;  // 4-column tiles; the same as 8-column tiles
  struct MfmaClass {
    int h, w, nt, tile_begin, tile_count;
    int tile8_begin = 0, tile8_count = 0;
  };
  std::vector<MfmaClass> mfma_classes;
  // work-queue form: every class's tiles in one list, largest first; one list per tile
  // height (16 * MT rows, MT = 2 or 4)
  std::vector<MfmaTile> q_tiles[2];
  size_t q_lds[2] = {0, 0};
  // autotuned MFMA schedule per precision (index 1 = bf16, 2 = fp16): -1 = not tuned yet,
  // 0 = one launch per class, else 1 + (MT index) * 8 + blocks per CU - 1
  int mfma_choice[5] = {-1, -1, -1, -1, -1};
  std::vector<std::pair<std::string, float>> mfma_tuning[5];
  std::vector<long long> map_off;      // output map slot -> offset in conv arena
  std::vector<int2> map_dims;          // output map slot -> (H, W)
  long long conv_total = 0;
  std::vector<long long> root_slot;    // [level][NoR] -> slot (or -1)
  std::vector<long long> part_slot;    // [L][NoP] -> slot (or -1)
  // DT
  std::vector<DtJob> dt_pass1, dt_pass2;
  std::vector<DtJob> dt_pass1w, dt_pass2w;  // the windowed DT's jobs (lines per job fit LDS)
  std::vector<int4> dt_nodes;          // (d, lo, hi, 0) per node, per length
  std::vector<int> dt_depths;          // node index where each depth starts (+ end)
  std::unordered_map<int, std::pair<int, int>> dt_len_tab;  // n -> (node_off, depth_off)
  long long dt_total = 0;              // cells per DT arena
  size_t dt_lds1 = 0, dt_lds2 = 0;     // dynamic LDS of the two DT launches
  size_t dt_lds1w = 0, dt_lds2w = 0;   // the same for the windowed DT of the 16-bit forms
  // scoring
  std::vector<ScoreMap> smaps;
  std::vector<ScorePart> sparts;
  long long score_total = 0;
  int score_blocks = 0;        // 256-cell blocks of the score launch (each within one map)
  int GL = 0;
  std::vector<int> ascore_tab;         // per valid level: Xn[W] then Yn[H]
  // device copies of the tables
  DeviceBuffer d_resize_ops, d_resize_tab, d_resize_si, d_resize_alpha, d_hog, d_conv_tiles,
      d_mfma_tiles, d_mfma_tiles8, d_map_dims, d_map_off, d_dt1, d_dt2, d_dt1w, d_dt2w, d_resize_si4,
      d_resize_al4, d_nodes, d_depths, d_smaps, d_sparts,
      d_ascore_tab, d_grad, d_grid, d_qtiles0, d_qtiles1, d_qcounter, d_fused;
  // device arenas
  DeviceBuffer d_frame, d_img, d_tmp, d_hist, d_norm, d_feat, d_featb, d_conv, d_tmpM,
      d_tmpIy, d_M, d_tmpIx, d_score, d_flags, d_blocksum, d_cand, d_count, d_rows, d_ascore;
  long long cand_cap = 0;
  // device NMS
  DeviceBuffer n_box, n_keys, n_keys2, n_sbox, n_alive0, n_alive1, n_status, n_fsup, n_picks,
      n_counters, n_mask, n_sort, n_flags, n_keep_orig, n_kept,
      n_total, n_vals, n_ord, n_rank, n_dec, n_alist, n_acount, n_grid, n_tie;
  long long nms_fallbacks = 0;
  std::map<long long, hipGraphExec_t> graphs;  // front of the frame, per (precision, ws, thresh)
  // generation: new whenever the plan's graphs are dropped (a buffer they use moved) or the
  // plan is made; graphs over several plans (a batch's front) record it
  unsigned long long gen = 0;
  // batch plans: per frame's NMS job its small buffers (counters, flags, totals), window
  // bitmask, compaction counts, sort storage and pick index, two events, and mapped host slots
  // (4 per job, then the nb + 1 frame offsets of the candidate list)
  std::vector<DeviceBuffer> b_small, b_mask, b_blocksum, b_sort, b_grid, b_acount;
  std::vector<hipEvent_t> b_ev;
  int *b_host = nullptr, *b_dev = nullptr;
};

struct Model {
  geye_ctx *ctx = nullptr;
  int numcomponent = 0, sbin = 0, interval = 0, max_Y = 0, max_X = 0, padx = 0, pady = 0;
  int maxpart = 0, NoR = 0, NoP = 0, DefL = 0;
  std::vector<int> ridx, oidx, rsize, numpart, pidx, didx, psize, anchor, root_size, part_size;
  std::vector<float> offw, def;
  std::vector<std::vector<float>> rootf, partf;   // reference layout [31][w][h]
  // packed filters per shape class (device): [group][31][w][h][NF]
  // A class is a set of filters of one kind convolved together: (h, w) is the bounding
  // size; a member smaller than it is zero-padded at its far edges, which leaves every
  // product sum (and the FP32 path's exact accumulation) unchanged.
  struct FilterClass {
    int h, w;
    bool root;
    std::vector<int> members;        // filter ids
    std::vector<int2> msize;         // true (h, w) per member
    int nf = 1;                      // FP32 kernel: filters per group
    long long off = 0;               // FP32 packed offset (floats)
    int nt = 1;                      // MFMA kernel: 16-filter tiles per group
    long long hoff = 0;              // MFMA packed offset (16-bit elements)
    int nslots = 0;                  // output slots per level (members + padding)
  };
  std::vector<FilterClass> classes;
  // least-squares box coefficients (x1, y1, x2, y2 per component), may be empty
  std::vector<float> lsq;
  std::vector<int> lsq_idx;                       // [offsets C][lengths C]
  DeviceBuffer d_filters, d_rsize, d_lsq, d_lsq_idx;
  DeviceBuffer d_filt16[5];  // MFMA filter packs, indexed by precision (GEYE_BF16 .. FP16X3)
};

struct KernelStat {
  double ms = 0;
  long long count = 0;
};

}  // namespace geye

struct geye_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;
  bool profile = false;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::unordered_map<std::string, geye::KernelStat> stats;
  std::vector<hipEvent_t> event_pool;
  std::vector<float> last_scores;
  std::vector<geye::DeviceBuffer> scratch_bufs;
  geye::Plan *last_plan = nullptr;
  bool graphs = true;                 // replay the per-frame front as a HIP graph
  int dt_form = GEYE_DT_AUTO;         // geye_dt_form
  hipStream_t job_streams[4] = {};    // streams of a batch's concurrent NMS jobs (created on use)
  int num_cu = 256;
  // eighths of the device's resident workgroups the persistent MFMA convolution takes (8:
  // all).  A serving worker leaves a quarter to the other frames' kernels, which cannot share
  // a CU with a split-form convolution workgroup (its LDS fills the CU)
  int conv_eighths = 8;
  int *h_count = nullptr;             // pinned host slots for counts (coherent, mapped: the
  int *d_hcount = nullptr;            // compaction kernels write them directly; see kHost*)
  hipEvent_t ev_count[2] = {nullptr, nullptr};
  long long nms_fallbacks = 0;
  long long nms_tie_repairs = 0;  // NMS calls whose tie order the host set by a partial quicksort
  // batch fronts captured as one graph, per (form, frames, models' plans): the graph and the
  // generations of the plans it was captured against
  std::map<std::vector<long long>, std::pair<hipGraphExec_t, std::vector<unsigned long long>>>
      batch_graphs;
};

struct geye_model {
  geye::Model m;
};

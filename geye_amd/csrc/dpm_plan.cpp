// dpm_plan.cpp -- per-(model, frame size) execution plan of the gfx950 DPM path.
//
// The plan fixes every geometric decision on the host once per frame size, in the same
// integer/float expressions as the reference, and packs the work descriptors the kernels
// consume: the level plan (CT/featurepyramid.cpp.new2:174-239, CT/featurepyramid.cpp.org:
// 570-630), the resize interpolation tables (CT/resize.cc:68-108), HOG level geometry
// (CT/featurepyramid.cpp.new2:271-289), convolution output sizes and the level skip rule
// (CT/fconvsMT.cpp:409-495), DT jobs and the score-map order of get_boxes
// (CT/get_boxes.cpp:1410-1570).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "dpm_mfma.h"
#include "dpm_runtime.h"

namespace geye {

// Pyramid depth beyond the first octave (Model_info.max_scale as ini_scales sets it,
// CT/featurepyramid.cpp.new2:179-219).  The frame's cell grid (H / sbin by W / sbin) shrinks
// by 2^(-1/interval) per level; the depth ends one level before the first level (counted from
// interval) whose grid is smaller than half the smallest root filter in either direction,
// capped at 36 levels and never below interval.  The grid is tracked in float with the
// reference's operations, because the level where the comparison first holds depends on
// that rounding.
int compute_max_scale(const Model &m, int W, int H) {
  constexpr int kDepthCap = 36;
  const float shrink = 1 / (float)std::pow(2.0, 1.0 / (double)m.interval);
  float half_rows = (float)m.rsize[0], half_cols = (float)m.rsize[1];
  for (int c = 1; c < m.numcomponent; c++) {
    half_rows = std::min(half_rows, (float)m.rsize[2 * c]);
    half_cols = std::min(half_cols, (float)m.rsize[2 * c + 1]);
  }
  half_rows /= 2;
  half_cols /= 2;
  float rows = (float)H / (float)m.sbin, cols = (float)W / (float)m.sbin;
  for (int level = 0; level < m.interval + kDepthCap; level++) {
    rows *= shrink;
    cols *= shrink;
    if (rows < half_rows || cols < half_cols) return std::max(level - m.interval - 1, m.interval);
  }
  return std::max(kDepthCap, m.interval);
}

static int blocks_of(int n, int sbin) { return (int)std::floor(double(n) / double(sbin) + 0.5); }

void level_geometry(int max_scale, int sbin, int interval, int W, int H, std::vector<Level> &lv,
                    std::vector<int> &img_h, std::vector<int> &img_w,
                    std::vector<int> &img_src, std::vector<float> &img_scale) {
  const int LEN = max_scale + interval;
  lv.assign(LEN, Level{});
  img_h.clear(); img_w.clear(); img_src.clear(); img_scale.clear();
  const int sbin2 = (int)std::floor((double)sbin / 2.0);
  const float sc = (float)std::pow(2, (1.0 / (double)interval));
  std::vector<int> level_img(LEN, -1);
  for (int ii = 0; ii < interval && ii < LEN; ii++) {
    const float st = (float)(1.0 / std::pow((double)sc, (double)ii));
    int h = H, w = W;
    if (st != 1.0f) {
      h = (int)((double)((float)H * st) + 0.5);
      w = (int)((double)((float)W * st) + 0.5);
    }
    int im = (int)img_h.size();
    img_h.push_back(h); img_w.push_back(w);
    img_src.push_back(ii == 0 ? -1 : 0);
    img_scale.push_back(st);
    lv[ii] = Level{h, w, sbin2, 0, 0, st * 2, im, false};
    if (ii + interval < LEN) lv[ii + interval] = Level{h, w, sbin, 0, 0, st, im, false};
  }
  for (int ii = 0; ii < interval && ii < LEN; ii++) {
    int prev_img = lv[ii].img_index;
    int ph = lv[ii].img_h, pw = lv[ii].img_w;
    for (int jj = ii + interval; jj < max_scale; jj += interval) {
      const int l = jj + interval;
      const int nh = (int)((double)((float)ph * 0.5f) + 0.5);
      const int nw = (int)((double)((float)pw * 0.5f) + 0.5);
      int im = (int)img_h.size();
      img_h.push_back(nh); img_w.push_back(nw);
      img_src.push_back(prev_img);
      img_scale.push_back(0.5f);
      lv[l] = Level{nh, nw, sbin, 0, 0, (float)(0.5 * (double)lv[jj].scale), im, false};
      prev_img = im;
      ph = nh;
      pw = nw;
    }
  }
  for (int l = 0; l < LEN; l++) {
    Level &L = lv[l];
    L.F0 = std::max(blocks_of(L.img_h, L.sbin) - 2, 0);
    L.F1 = std::max(blocks_of(L.img_w, L.sbin) - 2, 0);
  }
}

// resize1dtran's interpolation table (CT/resize.cc:68-108), grouped per destination index.
void resize_tables(int sheight, int dheight, std::vector<int2> &tab, std::vector<int> &si,
                         std::vector<float> &alpha) {
  const float scale = (float)dheight / (float)sheight;
  const float invscale = (float)sheight / (float)dheight;
  const size_t first = si.size();
  for (int dy = 0; dy < dheight; dy++) {
    const int start = (int)si.size();
    const float fsy1 = (float)dy * invscale;
    const float fsy2 = fsy1 + invscale;
    const int sy1 = (int)((double)fsy1 + 0.99);
    const int sy2 = (int)fsy2;
    const float e1 = (float)sy1 - fsy1;
    if ((double)e1 > 1e-3) { si.push_back(sy1 - 1); alpha.push_back(e1 * scale); }
    for (int sy = sy1; sy < sy2; sy++) { si.push_back(sy); alpha.push_back(scale); }
    const float e2 = fsy2 - (float)sy2;
    if ((double)e2 > 1e-3) { si.push_back(sy2); alpha.push_back(e2 * scale); }
    tab.push_back(int2{start, (int)si.size() - start});
  }
  // the reference's table never indexes outside the source; keep that an invariant
  for (size_t k = first; k < si.size(); k++)
    if (si[k] < 0 || si[k] >= sheight) si[k] = si[k] < 0 ? 0 : sheight - 1;
}

static void dt_nodes_for(Plan &P, int n) {
  if (P.dt_len_tab.count(n)) return;
  const int node_off = (int)P.dt_nodes.size();
  const int depth_off = (int)P.dt_depths.size();
  std::vector<std::pair<int, int>> cur{{0, n - 1}}, next;
  int idx = 0;
  while (!cur.empty()) {
    P.dt_depths.push_back(idx);
    next.clear();
    for (auto [d1, d2] : cur) {
      const int d = (d1 + d2) >> 1;
      P.dt_nodes.push_back(int4{d, d1 - 1, d2 + 1 >= n ? n : d2 + 1, 0});
      idx++;
      if (d - 1 >= d1) next.push_back({d1, d - 1});
      if (d2 >= d + 1) next.push_back({d + 1, d2});
    }
    std::swap(cur, next);
  }
  P.dt_depths.push_back(idx);
  P.dt_len_tab[n] = {node_off, depth_off};
}

static int dt_depth_count(const Plan &P, int n) {
  // depths between consecutive entries; recomputed from the table
  auto it = P.dt_len_tab.find(n);
  int depth_off = it->second.second;
  int k = 0;
  int total = n;
  while (P.dt_depths[depth_off + k] < total) k++;
  return k;
}

// per line: source (n + 1 floats, padded) and argmins (n ints); per block: the node table
// (n int2, 8-byte aligned)
static size_t dt_lds_bytes(int n, int G) {
  const size_t NL = (size_t)(256 / G);
  const size_t words = NL * (size_t)(n + 1) + NL * (size_t)n;
  return (words + (words & 1)) * 4 + (size_t)n * 8;
}

// the windowed DT: float lines (n plus halos, padded), the source staged into them
static size_t dt_lds_bytes_window_lines(int n, int G) {
  const int NL = 256 / G;
  return ((size_t)NL * (size_t)dt_line_stride(n, NL) + kDtHalo) * 4;
}

// its divide-and-conquer fallback, fb lines at a time: float lines, argmins, node table
static size_t dt_lds_bytes_fallback(int n, int fb) {
  const size_t words = 2 * (size_t)fb * (size_t)n;
  return (words + (words & 1)) * 4 + (size_t)n * 8;
}

// the most fallback lines per batch within the window lines' LDS (one always fits: 16 n bytes
// against at least 4 (n + 24) + 24 words)
static int dt_fallback_lines(int n, int G) {
  int fb = 256 / G;
  while (fb > 1 && dt_lds_bytes_fallback(n, fb) > dt_lds_bytes_window_lines(n, G)) fb--;
  return fb;
}

static size_t dt_lds_bytes_window(int n, int G) {
  return std::max(dt_lds_bytes_window_lines(n, G),
                  dt_lds_bytes_fallback(n, dt_fallback_lines(n, G)));
}

// lanes per line: the fewest whose block fits 48 KB of LDS (several blocks per CU), else a
// whole wave per line with up to 160 KB
static int dt_pick_G(int n) {
  for (int G = 1; G <= 64; G *= 2)
    if (dt_lds_bytes(n, G) <= 48 * 1024) return G;
  return dt_lds_bytes(n, 64) <= (size_t)kDtWindowMaxLds ? 64 : -1;
}

// the same for the windowed kernel, whose lines carry halos, within kDtWindowJobLds: each
// job is a load - compute - store sequence with barriers between, so several jobs per CU
// overlap their phases (on 1080p part maps 24 KB of lines beat 12, 16, 20, 28, 32 and 40 KB)
constexpr size_t kDtWindowJobLds = 24 * 1024;
static int dt_pick_G_window(int n) {
  for (int G = 1; G <= 64; G *= 2)
    if (dt_lds_bytes_window(n, G) <= kDtWindowJobLds) return G;
  return dt_lds_bytes_window(n, 64) <= (size_t)kDtWindowMaxLds ? 64 : -1;
}

static void add_dt_jobs(Plan &P, std::vector<DtJob> &jobs, long long src_off, long long dst_off,
                        int n, int nlines, int line_stride, int elem_stride, int neg, float a,
                        float b, bool window = false, int sel = 0, int ax = 0, int ay = 0) {
  dt_nodes_for(P, n);
  const auto tabs = P.dt_len_tab[n];
  const int G = window ? dt_pick_G_window(n) : dt_pick_G(n);
  const int NL = 256 / G;
  const int dc = dt_depth_count(P, n);
  for (int l0 = 0; l0 < nlines; l0 += NL) {
    DtJob J;
    J.src_off = src_off;
    J.dst_off = dst_off;
    J.n = n;
    J.nlines = std::min(NL, nlines - l0);
    J.line0 = l0;
    J.line_stride = line_stride;
    J.elem_stride = elem_stride;
    J.node_off = tabs.first;
    J.depth_off = tabs.second;
    J.depth_count = dc;
    J.G = G;
    J.neg = neg;
    J.a = a;
    J.b = b;
    J.sel = sel;
    J.ax = ax;
    J.ay = ay;
    J.fb_lines = window ? dt_fallback_lines(n, G) : 0;
    jobs.push_back(J);
  }
}

int build_plan(Model &m, int W, int H, Plan &P, std::string &err, int max_scale_override) {
  if (m.sbin > 8 || m.sbin < 2) {
    err = "sbin outside [2, 8] is not supported";
    return GEYE_ERR_UNSUPPORTED;
  }
  if (W < 8 || H < 8) {
    err = "frame too small";
    return GEYE_ERR_ARG;
  }
  P.W = W;
  P.H = H;
  P.interval = m.interval;
  P.sbin = m.sbin;
  P.max_scale = max_scale_override >= 0 ? max_scale_override : compute_max_scale(m, W, H);
  P.LEN = P.max_scale + m.interval;
  std::vector<int> img_src;
  std::vector<float> img_scale;
  level_geometry(P.max_scale, m.sbin, m.interval, W, H, P.lv, P.img_h, P.img_w, img_src,
                 img_scale);
  const int LEN = P.LEN, interval = m.interval;
  for (int l = 0; l < LEN; l++) {
    Level &L = P.lv[l];
    if (L.F0 > 0 && L.F1 > 0 && (L.img_h < 3 || L.img_w < 3)) {
      err = "level image smaller than 3 pixels";
      return GEYE_ERR_ARG;
    }
    L.valid = (l >= interval) && !(L.F0 + 2 * m.pady < m.max_Y || L.F1 + 2 * m.padx < m.max_X);
  }
  // ---- images and resize rounds
  const int nimg = (int)P.img_h.size();
  // image 0 is the frame itself: its readers (the first resize round and its gradients) read
  // the frame's BGR bytes, so it takes no room in the arena
  P.img_off.assign(nimg, 0);
  long long off = 0;
  for (int i = 0; i < nimg; i++) {
    P.img_off[i] = off;
    if (img_src[i] >= 0) off += 3LL * P.img_h[i] * P.img_w[i];
  }
  P.img_total = off;
  // round of an image = depth of its resize chain
  std::vector<int> round(nimg, 0);
  int maxr = 0;
  for (int i = 0; i < nimg; i++) {
    if (img_src[i] < 0) round[i] = -1;
    else if (img_src[i] == 0) round[i] = 0;
    else round[i] = round[img_src[i]] + 1;
    maxr = std::max(maxr, round[i]);
  }
  P.rounds.assign(maxr + 1, Plan::Round{});
  P.resize_tab.clear();
  P.resize_si.clear();
  P.resize_alpha.clear();
  for (int r = 0; r <= maxr; r++) {
    Plan::Round &R = P.rounds[r];
    long long toff = 0;
    for (int i = 0; i < nimg; i++) {
      if (round[i] != r) continue;
      const int s = img_src[i];
      const int ih = P.img_h[s], iw = P.img_w[s], oh = P.img_h[i], ow = P.img_w[i];
      ResizeOp rows{P.img_off[s], toff, ih, iw, oh, iw, (int)P.resize_tab.size(), R.rows_blocks};
      resize_tables(ih, oh, P.resize_tab, P.resize_si, P.resize_alpha);
      R.rows_blocks += (int)((3LL * oh * iw + 255) / 256);
      R.rows.push_back(rows);
      ResizeOp cols{toff, P.img_off[i], oh, iw, oh, ow, (int)P.resize_tab.size(), R.cols_blocks};
      resize_tables(iw, ow, P.resize_tab, P.resize_si, P.resize_alpha);
      R.cols_blocks += (int)((3LL * oh * ow + 255) / 256);
      R.cols.push_back(cols);
      toff += 3LL * oh * iw;
    }
    R.tmp_total = toff;
  }
  // fused form: the same tables, one launch per round
  P.fused.assign(P.rounds.size(), {});
  P.fused_blocks.assign(P.rounds.size(), 0);
  for (size_t r = 0; r < P.rounds.size(); r++) {
    const Plan::Round &R = P.rounds[r];
    for (size_t k = 0; k < R.rows.size(); k++) {
      const ResizeOp &ro = R.rows[k], &co = R.cols[k];
      FusedResize f{};
      f.in_off = ro.in_off;
      f.out_off = co.out_off;
      f.in_h = ro.in_h;
      f.in_w = ro.in_w;
      f.out_h = co.out_h;
      f.out_w = co.out_w;
      f.rtab_off = ro.tab_off;
      f.ctab_off = co.tab_off;
      f.tiles_x = (f.out_w + kResizeTX - 1) / kResizeTX;
      f.tiles_y = (f.out_h + kResizeTY - 1) / kResizeTY;
      f.blocks_begin = P.fused_blocks[r];
      f.frame = 0;
      P.fused_blocks[r] += f.tiles_x * f.tiles_y;
      // every tile's source-column span must fit the kernel's LDS rows
      for (int tx = 0; tx < f.tiles_x; tx++) {
        const int d0 = tx * kResizeTX, d1 = std::min(f.out_w, d0 + kResizeTX) - 1;
        const int2 a = P.resize_tab[f.ctab_off + d0], b = P.resize_tab[f.ctab_off + d1];
        const int c0 = P.resize_si[a.x], c1 = P.resize_si[b.x + b.y - 1];
        if (c1 - c0 + 1 > kResizeSpan) {
          err = "resize scale below the fused kernel's range";
          return GEYE_ERR_UNSUPPORTED;
        }
      }
      // and every output must use at most 4 source samples per pass
      for (int d = 0; d < f.out_w; d++)
        if (P.resize_tab[f.ctab_off + d].y > 4) {
          err = "resize scale below the fused kernel's range";
          return GEYE_ERR_UNSUPPORTED;
        }
      for (int d = 0; d < f.out_h; d++)
        if (P.resize_tab[f.rtab_off + d].y > 4) {
          err = "resize scale below the fused kernel's range";
          return GEYE_ERR_UNSUPPORTED;
        }
      P.fused[r].push_back(f);
    }
  }
  // ---- gradient grids (one per pyramid image, shared by its sbin 4 and sbin 8 levels)
  P.grad.clear();
  std::vector<long long> grid_of(P.img_h.size(), 0);
  {
    long long go = 0;
    int gb = 0;
    for (size_t i = 0; i < P.img_h.size(); i++) {
      GradImg g{};
      g.img_off = P.img_off[i];
      g.grid_off = go;
      g.h = P.img_h[i];
      g.w = P.img_w[i];
      g.tiles_x = (g.w + 63) / 64;
      g.tiles_y = (g.h + kGradTY - 1) / kGradTY;
      g.blocks_begin = gb;
      g.frame = img_src[i] < 0 ? 0 : -1;
      gb += g.tiles_x * g.tiles_y;
      grid_of[i] = go;
      go += (long long)grid_stride(g.h) * g.w;
      P.grad.push_back(g);
    }
    P.grad_blocks = gb;
    P.grid_total = go;
  }
  // ---- HOG levels (one vote launch over cell-column tiles, then one feature launch)
  P.hog.clear();
  P.feat_off.assign(LEN, -1);
  long long ho = 0, no = 0, fo = 0, fbo = 0;
  std::vector<HogLevel> h16, h8, hf;
  int b16 = 0, b8 = 0, bf = 0;
  for (int l = 0; l < LEN; l++) {
    const Level &L = P.lv[l];
    if (L.F0 <= 0 || L.F1 <= 0) continue;
    HogLevel h{};
    h.img_off = P.img_off[L.img_index];
    h.grid_off = grid_of[L.img_index];
    h.h = L.img_h;
    h.w = L.img_w;
    h.sbin = L.sbin;
    h.b0 = blocks_of(L.img_h, L.sbin);
    h.b1 = blocks_of(L.img_w, L.sbin);
    h.F0 = L.F0;
    h.F1 = L.F1;
    h.hist_off = ho;
    h.norm_off = no;
    h.feat_off = fo;
    h.featb_off = fbo;
    P.feat_off[l] = fo;
    ho += 18LL * h.b0 * h.b1;
    no += (long long)h.b0 * h.b1;
    fo += 31LL * L.F0 * L.F1;
    fbo += 32LL * L.F0 * L.F1;
    h.T = kHogVoteRows;
    h.tiles_x = (h.b1 + kHogVoteCols - 1) / kHogVoteCols;
    h.tiles_y = (h.b0 + h.T - 1) / h.T;
    HogLevel hh = h;
    hh.blocks_begin = b16;
    b16 += h.tiles_x * h.tiles_y;
    h16.push_back(hh);
    HogLevel hft = h;
    hft.blocks_begin = bf;
    bf += (int)((1LL * L.F0 * L.F1 + 255) / 256);
    hf.push_back(hft);
  }
  P.hist_total = ho;
  P.norm_total = no;
  P.feat_total = fo;
  P.featb_total = fbo;
  P.hog = h16;
  P.hog.insert(P.hog.end(), h8.begin(), h8.end());
  P.hog.insert(P.hog.end(), hf.begin(), hf.end());
  P.hog_hist_blocks = b16;  // T16 blocks; T8 count stored below
  P.hog_feat_blocks = bf;
  P.hog16_n = (int)h16.size();
  P.hog8_n = (int)h8.size();
  P.hog8_blocks = b8;
  P.hogf_n = (int)hf.size();
  // ---- convolution: output slots per (level, class), FP32 and MFMA tiles
  P.conv_tiles.clear();
  P.conv_classes.clear();
  P.mfma_tiles.clear();
  P.mfma_tiles8.clear();
  P.mfma_classes.clear();
  P.q_tiles[0].clear();
  P.q_tiles[1].clear();
  P.q_lds[0] = P.q_lds[1] = 0;
  P.map_off.clear();
  P.map_dims.clear();
  P.root_slot.assign((size_t)LEN * m.NoR, -1);
  P.part_slot.assign((size_t)LEN * m.NoP, -1);
  std::vector<long long> featb_off(LEN, -1);
  for (const HogLevel &h : hf)
    for (int l = 0; l < LEN; l++)
      if (P.feat_off[l] == h.feat_off) featb_off[l] = h.featb_off;
  long long co = 0;
  for (size_t ci = 0; ci < m.classes.size(); ci++) {
    const Model::FilterClass &fc = m.classes[ci];
    const int nmem = (int)fc.members.size();
    Plan::ConvClass cc{fc.h, fc.w, fc.nf, (int)P.conv_tiles.size(), 0, (int)ci};
    Plan::MfmaClass mc{fc.h, fc.w, fc.nt, (int)P.mfma_tiles.size(), 0};
    mc.tile8_begin = (int)P.mfma_tiles8.size();
    for (int l = interval; l < LEN; l++) {
      if (!P.lv[l].valid) continue;
      const int fl = fc.root ? l : l - interval;
      const Level &FL = P.lv[fl];
      const int base = (int)P.map_off.size();
      for (int k = 0; k < fc.nslots; k++) {
        if (k < nmem) {
          const int Hh = FL.F0 - fc.msize[k].x + 1, Ww = FL.F1 - fc.msize[k].y + 1;
          if (Hh < 1 || Ww < 1) {
            err = "filter larger than a valid feature level";
            return GEYE_ERR_ARG;
          }
          P.map_off.push_back(co);
          P.map_dims.push_back(int2{Hh, Ww});
          co += (long long)Hh * Ww;
          const int fid = fc.members[k];
          if (fc.root) P.root_slot[(size_t)fl * m.NoR + fid] = base + k;
          else P.part_slot[(size_t)fl * m.NoP + fid] = base + k;
        } else {
          P.map_off.push_back(0);
          P.map_dims.push_back(int2{0, 0});
        }
      }
      auto group_extent = [&](int first, int count, int &Hm, int &Wm) {
        Hm = Wm = 0;
        for (int k = first; k < std::min(nmem, first + count); k++) {
          Hm = std::max(Hm, P.map_dims[base + k].x);
          Wm = std::max(Wm, P.map_dims[base + k].y);
        }
      };
      const int g32 = (nmem + fc.nf - 1) / fc.nf;
      for (int g = 0; g < g32; g++) {
        int Hm, Wm;
        group_extent(g * fc.nf, fc.nf, Hm, Wm);
        for (int x0 = 0; x0 < Wm; x0 += 16)
          for (int y0 = 0; y0 < Hm; y0 += 64) {
            ConvTile t;
            t.feat_off = P.feat_off[fl];
            t.F0 = FL.F0;
            t.F1 = FL.F1;
            t.H = Hm;
            t.W = Wm;
            t.x0 = x0;
            t.y0 = y0;
            t.filt_off = fc.off + (long long)g * 31 * fc.w * fc.h * fc.nf;
            t.out_slot = base + g * fc.nf;
            t.nvalid = std::min(fc.nf, nmem - g * fc.nf);
            P.conv_tiles.push_back(t);
          }
      }
      const int gsz = 16 * fc.nt;
      const int g16 = (nmem + gsz - 1) / gsz;
      for (int g = 0; g < g16; g++) {
        int Hm, Wm;
        group_extent(g * gsz, gsz, Hm, Wm);
        // rows of tiles along x: the MFMA kernel carries the shared feature columns from
        // one tile to the next; the same outputs as 4- and as 8-column tiles
        for (int tw = 4; tw <= 8; tw += 4)
          for (int y0 = 0; y0 < Hm; y0 += 64)
            for (int x0 = 0; x0 < Wm; x0 += tw) {
              MfmaTile t;
              t.featb_off = featb_off[fl];
              t.F0 = FL.F0;
              t.F1 = FL.F1;
              t.x0 = x0;
              t.y0 = y0;
              t.filt_off = fc.hoff + (long long)g * fc.w * fc.h * fc.nt * 512;
              t.out_slot = base + g * gsz;
              t.nvalid = std::min(gsz, nmem - g * gsz);
              t.fh = fc.h;
              t.fw = fc.w;
              t.nt = fc.nt;
              t.cost = 0;
              (tw == 4 ? P.mfma_tiles : P.mfma_tiles8).push_back(t);
            }
        // work-queue tiles: 4 columns x 32 or 64 rows
        for (int mi = 0; mi < 2; mi++) {
          const int rows = mi == 0 ? 32 : 64;
          for (int x0 = 0; x0 < Wm; x0 += 4)
            for (int y0 = 0; y0 < Hm; y0 += rows) {
              MfmaTile t;
              t.featb_off = featb_off[fl];
              t.F0 = FL.F0;
              t.F1 = FL.F1;
              t.x0 = x0;
              t.y0 = y0;
              t.filt_off = fc.hoff + (long long)g * fc.w * fc.h * fc.nt * 512;
              t.out_slot = base + g * gsz;
              t.nvalid = std::min(gsz, nmem - g * gsz);
              t.fh = fc.h;
              t.fw = fc.w;
              t.nt = fc.nt;
              t.cost = fc.h * fc.w * fc.nt * std::min(rows, Hm - y0);
              P.q_tiles[mi].push_back(t);
            }
          P.q_lds[mi] = std::max(P.q_lds[mi], mfma_q_lds_bytes(fc.h, fc.w, fc.nt, mi == 0 ? 2 : 4, false));
        }
      }
    }
    cc.tile_count = (int)P.conv_tiles.size() - cc.tile_begin;
    mc.tile_count = (int)P.mfma_tiles.size() - mc.tile_begin;
    mc.tile8_count = (int)P.mfma_tiles8.size() - mc.tile8_begin;
    P.conv_classes.push_back(cc);
    P.mfma_classes.push_back(mc);
  }
  P.conv_total = co;
  for (int mi = 0; mi < 2; mi++)
    std::stable_sort(P.q_tiles[mi].begin(), P.q_tiles[mi].end(),
                     [](const MfmaTile &a, const MfmaTile &b) { return a.cost > b.cost; });
  // ---- DT jobs and score maps
  P.dt_pass1.clear();
  P.dt_pass2.clear();
  P.dt_pass1w.clear();
  P.dt_pass2w.clear();
  P.dt_nodes.clear();
  P.dt_depths.clear();
  P.dt_len_tab.clear();
  P.smaps.clear();
  P.sparts.clear();
  P.ascore_tab.clear();
  long long dto = 0, so = 0;
  int sblocks = 0;
  const int MP = m.maxpart;
  for (int l = interval; l < LEN; l++) {
    if (!P.lv[l].valid) continue;
    const int L = l - interval;
    // accumulated-score geometry (calc_a_score, CT/get_boxes.cpp:265-300): image column ii
    // reads score column Xn = (int)((float)ii / scale + padx), row jj reads Yn likewise;
    // the same float operations as the reference, evaluated once per level here.
    const float lscale = (float)m.sbin / P.lv[l].scale;
    const int tab_off = (int)P.ascore_tab.size();
    for (int ii = 0; ii < W; ii++) P.ascore_tab.push_back((int)((float)ii / lscale + (float)m.padx));
    for (int jj = 0; jj < H; jj++) P.ascore_tab.push_back((int)((float)jj / lscale + (float)m.pady));
    const int block_pad = (int)((double)lscale / 2.0);
    for (int jj = 0; jj < m.numcomponent; jj++) {
      if (jj >= m.NoR) {
        err = "component without a root filter";
        return GEYE_ERR_ARG;
      }
      const long long rslot = P.root_slot[(size_t)l * m.NoR + jj];
      ScoreMap sm{};
      sm.root_off = P.map_off[rslot];
      sm.R0 = P.map_dims[rslot].x;
      sm.R1 = P.map_dims[rslot].y;
      sm.score_off = so;
      sm.level = l;
      sm.comp = jj;
      sm.npart = m.numpart[jj];
      sm.part_first = (int)P.sparts.size();
      sm.off = m.offw[m.ridx[jj]];
      sm.scale = lscale;
      sm.RY = (int)((double)((float)m.rsize[jj * 2] * lscale) / 2.0 - 1.0 + (double)block_pad);
      sm.RX = (int)((double)((float)m.rsize[jj * 2 + 1] * lscale) / 2.0 - 1.0 + (double)block_pad);
      sm.tab_off = tab_off;
      sm.blocks_begin = sblocks;
      sblocks += (int)(((long long)sm.R0 * sm.R1 + 255) / 256);
      so += (long long)sm.R0 * sm.R1;
      for (int kk = 0; kk < m.numpart[jj]; kk++) {
        const int pid = m.pidx[jj * MP + kk], did = m.didx[jj * MP + kk];
        const long long pslot = P.part_slot[(size_t)L * m.NoP + pid];
        const int P0 = P.map_dims[pslot].x, P1 = P.map_dims[pslot].y;
        ScorePart q{};
        q.M_off = dto;
        q.Ix_off = dto;
        q.Iy_off = dto;
        q.P0 = P0;
        q.P1 = P1;
        q.ax = m.anchor[did * 2] + 1;
        q.ay = m.anchor[did * 2 + 1] + 1;
        q.ps0 = m.psize[(jj * MP + kk) * 2];
        q.ps1 = m.psize[(jj * MP + kk) * 2 + 1];
        if (q.ay < 2 || q.ax < 2 || q.ay - 1 + 2 * (sm.R0 - 1) >= P0 ||
            q.ax - 1 + 2 * (sm.R1 - 1) >= P1) {
          err = "part anchor outside the part score map";
          return GEYE_ERR_ARG;
        }
        const float *df = &m.def[did * 4];
        if (dt_pick_G(P0) < 0 || dt_pick_G(P1) < 0) {
          err = "part score map too large for the DT kernel";
          return GEYE_ERR_UNSUPPORTED;
        }
        // pass 1: columns (along y, n = P0) with (ay, by); source negated (inverse_Q)
        add_dt_jobs(P, P.dt_pass1, P.map_off[pslot], dto, P0, P1, P0, 1, 1, df[2], df[3]);
        add_dt_jobs(P, P.dt_pass1w, P.map_off[pslot], dto, P0, P1, P0, 1, 1, df[2], df[3],
                    true, 1, q.ax, q.ay);
        // pass 2: rows (along x, n = P1) with (ax, bx)
        add_dt_jobs(P, P.dt_pass2, dto, dto, P1, P0, 1, P0, 0, df[0], df[1]);
        add_dt_jobs(P, P.dt_pass2w, dto, dto, P1, P0, 1, P0, 0, df[0], df[1], true, 2, q.ax,
                    q.ay);
        P.sparts.push_back(q);
        dto += (long long)P0 * P1;
      }
      P.smaps.push_back(sm);
    }
  }
  P.dt_total = dto;
  P.score_total = so;
  P.score_blocks = sblocks;
  P.dt_lds1 = 0;
  P.dt_lds2 = 0;
  P.dt_lds1w = 0;
  P.dt_lds2w = 0;
  for (auto &j : P.dt_pass1) P.dt_lds1 = std::max(P.dt_lds1, dt_lds_bytes(j.n, j.G));
  for (auto &j : P.dt_pass2) P.dt_lds2 = std::max(P.dt_lds2, dt_lds_bytes(j.n, j.G));
  for (auto &j : P.dt_pass1w) P.dt_lds1w = std::max(P.dt_lds1w, dt_lds_bytes_window(j.n, j.G));
  for (auto &j : P.dt_pass2w) P.dt_lds2w = std::max(P.dt_lds2w, dt_lds_bytes_window(j.n, j.G));
  P.GL = m.numcomponent > 0 ? (m.numpart[0] + 1) * 4 + 3 : 0;
  for (int jj = 1; jj < m.numcomponent; jj++)
    if (m.numpart[jj] != m.numpart[0]) {
      err = "components with different part counts are not supported (row width)";
      return GEYE_ERR_UNSUPPORTED;
    }
  return GEYE_OK;
}

void replicate_plan(const Plan &one, int nb, long long feat_stride, long long featb_stride,
                    Plan &P) {
  P = Plan();
  P.W = one.W;
  P.H = one.H;
  P.max_scale = one.max_scale;
  P.LEN = one.LEN;
  P.interval = one.interval;
  P.sbin = one.sbin;
  P.nb = nb;
  P.img_one = one.img_total;
  P.grid_one = one.grid_total;
  P.hist_one = one.hist_total;
  P.norm_one = one.norm_total;
  P.conv_one = one.conv_total;
  P.dt_one = one.dt_total;
  P.score_one = one.score_total;
  P.feat_stride = feat_stride > 0 ? feat_stride : one.feat_total;
  P.featb_stride = featb_stride > 0 ? featb_stride : one.featb_total;
  P.nmaps_one = (int)one.smaps.size();
  P.nparts_one = (int)one.sparts.size();
  P.lv = one.lv;
  P.img_h = one.img_h;
  P.img_w = one.img_w;
  P.img_off = one.img_off;
  P.feat_off = one.feat_off;
  P.img_total = one.img_total * nb;
  P.grid_total = one.grid_total * nb;
  P.hist_total = one.hist_total * nb;
  P.norm_total = one.norm_total * nb;
  P.feat_total = P.feat_stride * nb;
  P.featb_total = P.featb_stride * nb;
  P.conv_total = one.conv_total * nb;
  P.dt_total = one.dt_total * nb;
  P.score_total = one.score_total * nb;
  P.GL = one.GL;
  // the shared tables (resize taps, DT node tables, accumulated-score columns)
  P.resize_tab = one.resize_tab;
  P.resize_si = one.resize_si;
  P.resize_alpha = one.resize_alpha;
  P.dt_nodes = one.dt_nodes;
  P.dt_depths = one.dt_depths;
  P.dt_len_tab = one.dt_len_tab;
  P.ascore_tab = one.ascore_tab;
  P.dt_lds1 = one.dt_lds1;
  P.dt_lds2 = one.dt_lds2;
  P.dt_lds1w = one.dt_lds1w;
  P.dt_lds2w = one.dt_lds2w;
  // resize rounds, gradients, HOG votes and features
  P.fused.assign(one.fused.size(), {});
  P.fused_blocks.assign(one.fused.size(), 0);
  for (size_t r = 0; r < one.fused.size(); r++) {
    for (int f = 0; f < nb; f++)
      for (FusedResize op : one.fused[r]) {
        op.in_off += f * one.img_total;
        op.out_off += f * one.img_total;
        op.blocks_begin += f * one.fused_blocks[r];
        op.frame = f;
        P.fused[r].push_back(op);
      }
    P.fused_blocks[r] = one.fused_blocks[r] * nb;
  }
  for (int f = 0; f < nb; f++)
    for (GradImg g : one.grad) {
      g.img_off += f * one.img_total;
      g.grid_off += f * one.grid_total;
      g.blocks_begin += f * one.grad_blocks;
      if (g.frame >= 0) g.frame = f;
      P.grad.push_back(g);
    }
  P.grad_blocks = one.grad_blocks * nb;
  const int nvote = one.hog16_n + one.hog8_n;
  auto shift_hog = [&](HogLevel h, int f, int blocks) {
    h.img_off += f * one.img_total;
    h.grid_off += f * one.grid_total;
    h.hist_off += f * one.hist_total;
    h.norm_off += f * one.norm_total;
    h.feat_off += f * P.feat_stride;
    h.featb_off += f * P.featb_stride;
    h.blocks_begin += f * blocks;
    return h;
  };
  for (int f = 0; f < nb; f++)
    for (int k = 0; k < nvote; k++)
      P.hog.push_back(shift_hog(one.hog[k], f, one.hog_hist_blocks + one.hog8_blocks));
  for (int f = 0; f < nb; f++)
    for (int k = nvote; k < nvote + one.hogf_n; k++)
      P.hog.push_back(shift_hog(one.hog[k], f, one.hog_feat_blocks));
  P.hog16_n = nvote * nb;
  P.hog8_n = 0;
  P.hogf_n = one.hogf_n * nb;
  P.hog_hist_blocks = (one.hog_hist_blocks + one.hog8_blocks) * nb;
  P.hog8_blocks = 0;
  P.hog_feat_blocks = one.hog_feat_blocks * nb;
  // convolution: each class's tiles for every frame, contiguous; output slots per frame
  const int nslots = (int)one.map_off.size();
  for (int f = 0; f < nb; f++)
    for (int k = 0; k < nslots; k++) {
      P.map_off.push_back(one.map_off[k] + f * one.conv_total);
      P.map_dims.push_back(one.map_dims[k]);
    }
  for (const Plan::ConvClass &cc : one.conv_classes) {
    Plan::ConvClass c2 = cc;
    c2.tile_begin = (int)P.conv_tiles.size();
    for (int f = 0; f < nb; f++)
      for (int k = cc.tile_begin; k < cc.tile_begin + cc.tile_count; k++) {
        ConvTile t = one.conv_tiles[k];
        t.feat_off += f * P.feat_stride;
        t.out_slot += f * nslots;
        P.conv_tiles.push_back(t);
      }
    c2.tile_count = cc.tile_count * nb;
    P.conv_classes.push_back(c2);
  }
  for (const Plan::MfmaClass &mc : one.mfma_classes) {
    Plan::MfmaClass c2 = mc;
    c2.tile_begin = (int)P.mfma_tiles.size();
    for (int f = 0; f < nb; f++)
      for (int k = mc.tile_begin; k < mc.tile_begin + mc.tile_count; k++) {
        MfmaTile t = one.mfma_tiles[k];
        t.featb_off += f * P.featb_stride;
        t.out_slot += f * nslots;
        P.mfma_tiles.push_back(t);
      }
    c2.tile_count = mc.tile_count * nb;
    c2.tile8_begin = (int)P.mfma_tiles8.size();
    for (int f = 0; f < nb; f++)
      for (int k = mc.tile8_begin; k < mc.tile8_begin + mc.tile8_count; k++) {
        MfmaTile t = one.mfma_tiles8[k];
        t.featb_off += f * P.featb_stride;
        t.out_slot += f * nslots;
        P.mfma_tiles8.push_back(t);
      }
    c2.tile8_count = mc.tile8_count * nb;
    P.mfma_classes.push_back(c2);
  }
  // a batch runs one launch per filter class (the schedule the autotuner picks for 1080p)
  for (int k = 0; k < 5; k++) P.mfma_choice[k] = 0;
  // DT jobs
  // (pass 1 reads the convolution arena, pass 2 the first pass's output in the DT arena)
  auto rep_jobs = [&](const std::vector<DtJob> &src, std::vector<DtJob> &dst, long long src_frame) {
    for (int f = 0; f < nb; f++)
      for (DtJob j : src) {
        j.src_off += f * src_frame;
        j.dst_off += f * one.dt_total;
        dst.push_back(j);
      }
  };
  rep_jobs(one.dt_pass1, P.dt_pass1, one.conv_total);
  rep_jobs(one.dt_pass2, P.dt_pass2, one.dt_total);
  rep_jobs(one.dt_pass1w, P.dt_pass1w, one.conv_total);
  rep_jobs(one.dt_pass2w, P.dt_pass2w, one.dt_total);
  // score maps and parts
  for (int f = 0; f < nb; f++) {
    for (ScoreMap m : one.smaps) {
      m.root_off += f * one.conv_total;
      m.score_off += f * one.score_total;
      m.part_first += f * (int)one.sparts.size();
      m.blocks_begin += f * one.score_blocks;
      P.smaps.push_back(m);
    }
    for (ScorePart q : one.sparts) {
      q.M_off += f * one.dt_total;
      q.Ix_off += f * one.dt_total;
      q.Iy_off += f * one.dt_total;
      P.sparts.push_back(q);
    }
  }
  P.score_blocks = one.score_blocks * nb;
}

int dt_line_geometry(int n, bool window, int *G, size_t *lds) {
  const int g = window ? dt_pick_G_window(n) : dt_pick_G(n);
  if (g < 0) return GEYE_ERR_UNSUPPORTED;
  *G = g;
  *lds = window ? dt_lds_bytes_window(n, g) : dt_lds_bytes(n, g);
  return GEYE_OK;
}

int dt_single_jobs(int n0, int n1, float ax, float bx, float ay, float by,
                   std::vector<DtJob> &p1, std::vector<DtJob> &p2, std::vector<int4> &nodes,
                   std::vector<int> &depths, size_t &lds1, size_t &lds2, bool window) {
  if (dt_pick_G(n0) < 0 || dt_pick_G(n1) < 0) return GEYE_ERR_UNSUPPORTED;
  Plan P;
  add_dt_jobs(P, P.dt_pass1, 0, 0, n0, n1, n0, 1, 0, ay, by, window);
  add_dt_jobs(P, P.dt_pass2, 0, 0, n1, n0, 1, n0, 0, ax, bx, window);
  p1 = P.dt_pass1;
  p2 = P.dt_pass2;
  nodes = P.dt_nodes;
  depths = P.dt_depths;
  lds1 = window ? dt_lds_bytes_window(n0, dt_pick_G_window(n0)) : dt_lds_bytes(n0, dt_pick_G(n0));
  lds2 = window ? dt_lds_bytes_window(n1, dt_pick_G_window(n1)) : dt_lds_bytes(n1, dt_pick_G(n1));
  return GEYE_OK;
}

}  // namespace geye

// dpm_kernels.hip -- gfx950 kernels of the FP32 DPM detection path.
//
// Every kernel reproduces the reference's per-element arithmetic order (the library is
// built with -ffp-contract=off, so `a += b * c` stays a multiply and an add), which is what
// makes the FP32 path bit-identical to the CPU restatement of the reference:
//   resize      CT/resize.cc:52-124        gather form of alphacopy, table order per output
//   HOG hist    CT/featurepyramid.cpp.new2:323-378  gather per cell, pixels in x-major order
//   HOG feat    CT/featurepyramid.cpp.new2:395-492  double-promoted normaliser and texture
//   conv        CT/GPU_function.cu:302-358 per output: f outer, (xp, yp) inner, fused MAC
//   DT          CT/dt.cpp:61-136           divide-and-conquer, level-synchronous
//   score       CT/get_boxes.cpp:1022-1041 (+bias) and CT/dt.cpp:38-55 (part add)
//   boxes       CT/get_boxes.cpp:185-258, 1497-1546
#include "dpm_kernels.h"

#include <hip/hip_fp16.h>

namespace geye {

// Workgroups are dealt to the 8 XCDs round-robin (block b on XCD b % 8), each XCD with its
// own L2.  This hands every XCD runs of C consecutive jobs -- neighbouring jobs share cache
// lines at their edges, so those lines meet in one L2 -- while successive runs still rotate
// over the XCDs, which keeps the heavy and light parts of the job list spread across them.
// Blocks past the last whole group of 8 * C keep their own index.
__device__ inline int xcd_local_block(int b, int nb) {
  constexpr int X = 8, C = 16, GRP = X * C;
  if (b >= (nb / GRP) * GRP) return b;
  const int g = b / GRP, l = b - g * GRP;
  return g * GRP + (l % X) * C + l / X;
}


// Index of the descriptor whose block range holds block bx: the last of n descriptors (in
// ascending blocks_begin order) with blocks_begin <= bx.  A binary search, uniform across the
// workgroup: a batch of frames lays many frames' descriptors end to end.
template <typename T>
__device__ inline int find_blocks(const T *__restrict__ d, int n, int bx) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (d[mid].blocks_begin <= bx) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// ---------------------------------------------------------------------------------------
// Resize, one separable pass.  Each output sums alpha*src over its table entries in table
// order starting from 0, exactly the per-destination sequence alphacopy produces.
// rows pass: out[c][dy][x] = sum_k alpha_k * in[c][si_k][x]   (in: [3][in_h][in_w])
// cols pass: out[c][y][dx] = sum_k alpha_k * in[c][y][si_k]
// ---------------------------------------------------------------------------------------
__device__ inline int find_op(const ResizeOp *ops, int nops, int b) {
  return find_blocks(ops, nops, b);
}

__global__ __launch_bounds__(256) void k_resize_rows(const ResizeOp *__restrict__ ops, int nops,
                                                     const int2 *__restrict__ tab,
                                                     const int *__restrict__ si,
                                                     const float *__restrict__ alpha,
                                                     const float *__restrict__ src,
                                                     float *__restrict__ dst) {
  const ResizeOp op = ops[find_op(ops, nops, blockIdx.x)];
  // 32-bit index arithmetic: an op covers at most 3 x 2^24 elements
  const unsigned e = (blockIdx.x - op.blocks_begin) * blockDim.x + threadIdx.x;
  const unsigned per_c = (unsigned)op.out_h * (unsigned)op.in_w;
  if (e >= 3u * per_c) return;
  const unsigned c = e / per_c;
  const unsigned r = e - c * per_c;
  const unsigned dy = r / (unsigned)op.in_w, x = r - dy * (unsigned)op.in_w;
  const float *in = src + op.in_off + (c * (unsigned)op.in_h) * (unsigned)op.in_w + x;
  const int2 t = tab[op.tab_off + dy];
  float acc = 0.0f;
  for (int k = 0; k < t.y; k++) acc += alpha[t.x + k] * in[(unsigned)si[t.x + k] * (unsigned)op.in_w];
  dst[op.out_off + e] = acc;
}

__global__ __launch_bounds__(256) void k_resize_cols(const ResizeOp *__restrict__ ops, int nops,
                                                     const int2 *__restrict__ tab,
                                                     const int *__restrict__ si,
                                                     const float *__restrict__ alpha,
                                                     const float *__restrict__ src,
                                                     float *__restrict__ dst) {
  const ResizeOp op = ops[find_op(ops, nops, blockIdx.x)];
  const unsigned e = (blockIdx.x - op.blocks_begin) * blockDim.x + threadIdx.x;
  const unsigned per_c = (unsigned)op.out_h * (unsigned)op.out_w;
  if (e >= 3u * per_c) return;
  const unsigned c = e / per_c;
  const unsigned r = e - c * per_c;
  const unsigned y = r / (unsigned)op.out_w, dx = r - y * (unsigned)op.out_w;
  const float *in = src + op.in_off + (c * (unsigned)op.in_h + y) * (unsigned)op.in_w;
  const int2 t = tab[op.tab_off + dx];
  float acc = 0.0f;
  for (int k = 0; k < t.y; k++) acc += alpha[t.x + k] * in[si[t.x + k]];
  dst[op.out_off + e] = acc;
}

// Both passes of one resize, per kResizeTY x 64 output tile (CT/resize.cc:135-160): the row
// pass for the tile's output rows over the source columns its 64 outputs tap goes to LDS, the
// column pass reads it there.  Each element is the same ordered tap sum (from 0, table
// order) as the two separate passes, so the image is bit-identical.
//
// BGR: the first round, whose source is the frame itself -- the row pass reads the frame's
// BGR bytes (frame op.frame of a batch, frame_bytes apart, rows ws bytes apart) instead of
// the planar float copy: a quarter of the bytes, the same values (the conversion is exact),
// so the same bits.
template <bool BGR>
__global__ __launch_bounds__(256) void k_resize_fused(const FusedResize *__restrict__ ops,
                                                      int nops, const int2 *__restrict__ tab,
                                                      const int4 *__restrict__ si4,
                                                      const float4 *__restrict__ al4,
                                                      const float *__restrict__ src,
                                                      const unsigned char *__restrict__ bgr,
                                                      int ws, long long frame_bytes,
                                                      float *__restrict__ dst) {
  constexpr int MAXTAP = 4;  // scales >= 0.5: at most 4 source samples per output
  __shared__ float s_t[3][kResizeTY][kResizeSpan];
  __shared__ int s_rn[kResizeTY], s_cn[kResizeTX];
  // column taps stored tap-major so that a wave's lanes (consecutive x) hit distinct banks
  __shared__ int s_rsi[kResizeTY][MAXTAP], s_csi[MAXTAP][kResizeTX];
  __shared__ float s_ral[kResizeTY][MAXTAP], s_cal[MAXTAP][kResizeTX];
  const int bx = xcd_local_block(blockIdx.x, gridDim.x);
  const FusedResize op = ops[find_blocks(ops, nops, bx)];
  const int tile = bx - op.blocks_begin;
  const int tyi = tile / op.tiles_x, txi = tile - tyi * op.tiles_x;
  const int dy0 = tyi * kResizeTY, dx0 = txi * kResizeTX;
  const int ny = min(kResizeTY, op.out_h - dy0), nx = min(kResizeTX, op.out_w - dx0);
  const int tid = threadIdx.x;
  // the tile's tap lists (row pass for its rows, column pass for its columns)
  if (tid < kResizeTY + kResizeTX) {
    const bool row = tid < kResizeTY;
    const int j = row ? tid : tid - kResizeTY;
    const bool live = row ? j < ny : j < nx;
    const int k = (row ? op.rtab_off + dy0 : op.ctab_off + dx0) + j;
    const int n = live ? min(tab[k].y, MAXTAP) : 0;
    const int4 sv = live ? si4[k] : make_int4(0, 0, 0, 0);
    const float4 av = live ? al4[k] : make_float4(0.f, 0.f, 0.f, 0.f);
    const int svq[MAXTAP] = {sv.x, sv.y, sv.z, sv.w};
    const float avq[MAXTAP] = {av.x, av.y, av.z, av.w};
#pragma unroll
    for (int q = 0; q < MAXTAP; q++) {
      if (row) {
        s_rsi[j][q] = svq[q];
        s_ral[j][q] = avq[q];
      } else {
        s_csi[q][j] = svq[q];
        s_cal[q][j] = avq[q];
      }
    }
    (row ? s_rn : s_cn)[j] = n;
  }
  __syncthreads();
  const int c0 = s_csi[0][0], span = s_csi[s_cn[nx - 1] - 1][nx - 1] - c0 + 1;
  // row pass, the three planes of an element together: up to 12 loads in flight
  const long long plane = (long long)op.in_h * op.in_w;
  const float *in = src + op.in_off + c0;
  const unsigned char *inb = bgr + (BGR ? op.frame * frame_bytes + 3LL * c0 : 0);
  constexpr int U = 3;  // elements per thread per step
  for (int i0 = tid; i0 < ny * span; i0 += U * 256) {
    float v[U][3][MAXTAP];
    int rr[U], xx[U], nn[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = i0 + u * 256;
      rr[u] = i / span;
      xx[u] = i - rr[u] * span;
      nn[u] = i < ny * span ? s_rn[rr[u]] : 0;
#pragma unroll
      for (int q = 0; q < MAXTAP; q++)
        if (q < nn[u]) {
          if (BGR) {
            const unsigned char *p = inb + (long long)s_rsi[rr[u]][q] * ws + 3 * xx[u];
#pragma unroll
            for (int c = 0; c < 3; c++) v[u][c][q] = (float)p[c];
          } else {
            const float *p = in + (long long)s_rsi[rr[u]][q] * op.in_w + xx[u];
#pragma unroll
            for (int c = 0; c < 3; c++) v[u][c][q] = p[c * plane];
          }
        }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (nn[u] == 0) continue;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        float acc = 0.0f;
#pragma unroll
        for (int q = 0; q < MAXTAP; q++)
          if (q < nn[u]) acc += s_ral[rr[u]][q] * v[u][c][q];
        s_t[c][rr[u]][xx[u]] = acc;
      }
    }
  }
  __syncthreads();
  const long long oplane = (long long)op.out_h * op.out_w;
  float *out = dst + op.out_off + (long long)dy0 * op.out_w + dx0;
  for (int i = tid; i < ny * kResizeTX; i += 256) {
    const int r = i / kResizeTX, x = i - r * kResizeTX;
    if (x >= nx) continue;
    const int n = s_cn[x];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      float acc = 0.0f;
#pragma unroll
      for (int q = 0; q < MAXTAP; q++)
        if (q < n) acc += s_cal[q][x] * s_t[c][r][s_csi[q][x] - c0];
      out[c * oplane + (long long)r * op.out_w + x] = acc;
    }
  }
}

void launch_resize_fused(hipStream_t s, const FusedResize *ops, int nops, int blocks,
                         const int2 *tab, const int4 *si4, const float4 *al4, const float *src,
                         float *dst, const unsigned char *bgr, int ws, long long frame_bytes) {
  if (blocks <= 0) return;
  if (bgr)
    hipLaunchKernelGGL(k_resize_fused<true>, dim3(blocks), dim3(256), 0, s, ops, nops, tab, si4,
                       al4, src, bgr, ws, frame_bytes, dst);
  else
    hipLaunchKernelGGL(k_resize_fused<false>, dim3(blocks), dim3(256), 0, s, ops, nops, tab, si4,
                       al4, src, nullptr, 0, 0LL, dst);
}

// ---------------------------------------------------------------------------------------
// HOG orientation histogram + block energy.
// A workgroup owns a TxT tile of cells.  Stage 1 computes, once per pixel of the tile's
// support, the strongest-channel gradient, its snapped orientation and magnitude
// (CT/featurepyramid.cpp.new2:325-364) into LDS.  Stage 2 gives each thread one cell and
// accumulates the bilinear votes of its pixels in the reference's x-major, y-minor order
// (CT/featurepyramid.cpp.new2:323-378), so each histogram bin sees the same sequence of
// float additions as the reference's scatter loop.  Stage 3 writes hist[o][x][y] and
// norm[x][y] = sum_{o<9} (h_o + h_{o+9})^2 (CT/featurepyramid.cpp.new2:391-404).
// ---------------------------------------------------------------------------------------
__constant__ float c_hcos[9] = {1.0000f, 0.9397f, 0.7660f, 0.5000f, 0.1736f,
                                -0.1736f, -0.5000f, -0.7660f, -0.9397f};
__constant__ float c_hsin[9] = {0.0000f, 0.3420f, 0.6428f, 0.8660f, 0.9848f,
                                0.9848f, 0.8660f, 0.6428f, 0.3420f};

__device__ inline int floor_div(int a, int b) {  // b > 0
  int q = a / b;
  return (a % b != 0 && a < 0) ? q - 1 : q;
}

// first pixel coordinate whose left interpolation cell is >= c:
// ixp(x) = floor((2x + 1 - sbin) / (2 sbin)) >= c  <=>  x >= ceil((2 sbin c + sbin - 1) / 2)
__device__ inline int first_pix(int c, int sbin) { return floor_div(2 * sbin * c + sbin, 2); }

__device__ inline void interp_coord(int x, int sbin, int *ix, float *v0, float *v1) {
  float xp = (float)(((double)(float)x + 0.5) / (double)(float)sbin - 0.5);
  int i = (int)floorf(xp);
  float a = xp - (float)i;
  *ix = i;
  *v0 = a;
  *v1 = (float)(1.0 - (double)a);
}

// Gradient of the strongest channel at pixel (x, y) of a [3][h][w] tile staged in LDS
// (CT/featurepyramid.cpp.new2:325-364): central differences per channel, the first channel
// with the largest squared magnitude (strict >), the orientation of the largest |dot| with
// the 4-decimal direction table (first maximum; the sign picks o or o + 9), v = sqrt.
__device__ inline int2 hog_grad_at(const float *t0, const float *t1, const float *t2, int i,
                                   int row) {
  float dy = t0[i + row] - t0[i - row];
  float dx = t0[i + 1] - t0[i - 1];
  float v = dx * dx + dy * dy;
  const float dy2 = t1[i + row] - t1[i - row];
  const float dx2 = t1[i + 1] - t1[i - 1];
  const float v2 = dx2 * dx2 + dy2 * dy2;
  const float dy3 = t2[i + row] - t2[i - row];
  const float dx3 = t2[i + 1] - t2[i - 1];
  const float v3 = dx3 * dx3 + dy3 * dy3;
  if (v2 > v) { v = v2; dx = dx2; dy = dy2; }
  if (v3 > v) { v = v3; dx = dx3; dy = dy3; }
  float best_dot = 0.0f;
  int best_o = 0;
#pragma unroll
  for (int o = 0; o < 9; o++) {
    const float dot = c_hcos[o] * dx + c_hsin[o] * dy;
    if (dot > best_dot) {
      best_dot = dot;
      best_o = o;
    } else if (-dot > best_dot) {
      best_dot = -dot;
      best_o = o + 9;
    }
  }
  return make_int2(__float_as_int(sqrtf(v)), best_o);
}

// One block = 64 (x) x kGradTY (y) pixels of one image: the 66 x (kGradTY + 2) x 3
// neighbourhood is read row-major (coalesced) into LDS -- from the float image, or for the
// frame itself (G.frame >= 0) from its BGR bytes, plane c being byte c of a pixel as in
// Ipl_to_FLOAT (CT/featurepyramid.cpp.new2:816-850; the conversion is exact, so the values are
// the planar copy's).  Each thread then takes 4 consecutive
// y of one column and writes their magnitudes as one 16-byte store and their orientations as
// one 4-byte store into the column-major planes.  Columns outside [1, w-2] are not written,
// and rows outside [1, h-2] hold zeros: readers clamp to that range, exactly as the reference
// clamps the pixel it differentiates.
__global__ __launch_bounds__(256) void k_hog_grad(const GradImg *__restrict__ imgs, int nimg,
                                                  const float *__restrict__ img_arena,
                                                  const unsigned char *__restrict__ bgr, int ws,
                                                  long long frame_bytes,
                                                  float *__restrict__ mag_arena,
                                                  unsigned char *__restrict__ ori_arena) {
  constexpr int TXP = 64, TYP = kGradTY, SX = TXP + 2, SY = TYP + 2, QY = TYP / 4;
  static_assert(kGridAlign % TYP == 0 && 256 % QY == 0 && TXP * QY % 256 == 0, "grad tile");
  __shared__ float s_t[3][SY * SX];
  const int bx = xcd_local_block(blockIdx.x, gridDim.x);
  const GradImg G = imgs[find_blocks(imgs, nimg, bx)];
  const int tile = bx - G.blocks_begin;
  const int tx = tile / G.tiles_y, ty = tile - tx * G.tiles_y;
  const int x0 = tx * TXP, y0 = ty * TYP;
  const int h = G.h, w = G.w;
  const float *img = img_arena + G.img_off;
  const unsigned char *fr = bgr + (G.frame >= 0 ? G.frame * frame_bytes : 0);
  // every element's three loads are issued before the first store (a thread stages up to
  // U elements), so the tile arrives in one round trip
  constexpr int U = (SY * SX + 255) / 256;
  {
    float v[U][3];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = min((int)threadIdx.x + 256 * u, SY * SX - 1);
      const int r = i / SX, q = i - r * SX;
      const int yy = min(max(y0 - 1 + r, 0), h - 1), xx = min(max(x0 - 1 + q, 0), w - 1);
      if (G.frame >= 0) {
        const unsigned char *p = fr + (long long)yy * ws + 3 * xx;
        v[u][0] = (float)p[0];
        v[u][1] = (float)p[1];
        v[u][2] = (float)p[2];
      } else {
        const long long o = (long long)yy * w + xx;
        v[u][0] = img[o];
        v[u][1] = img[o + (long long)h * w];
        v[u][2] = img[o + 2LL * h * w];
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = (int)threadIdx.x + 256 * u;
      if (i < SY * SX) {
        s_t[0][i] = v[u][0];
        s_t[1][i] = v[u][1];
        s_t[2][i] = v[u][2];
      }
    }
  }
  __syncthreads();
  const int q = threadIdx.x % QY;
  const long long gs = grid_stride(h);
#pragma unroll
  for (int k = 0; k < TXP * QY / 256; k++) {
    const int lx = (int)threadIdx.x / QY + (256 / QY) * k;
    const int x = x0 + lx;
    if (x < 1 || x > w - 2) continue;
    float m[4];
    unsigned int o = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int ly = 4 * q + j, y = y0 + ly;
      int2 g = make_int2(0, 0);
      if (y >= 1 && y <= h - 2) g = hog_grad_at(s_t[0], s_t[1], s_t[2], (ly + 1) * SX + lx + 1, SX);
      m[j] = __int_as_float(g.x);
      o |= (unsigned int)g.y << (8 * j);
    }
    // y0 + 4q + 3 < gs: tiles_y * kGradTY <= grid_stride(h)
    const long long e = G.grid_off + (long long)x * gs + y0 + 4 * q;
    *reinterpret_cast<float4 *>(mag_arena + e) = make_float4(m[0], m[1], m[2], m[3]);
    *reinterpret_cast<unsigned int *>(ori_arena + e) = o;
  }
}

// Orientation histogram + block energy.  A workgroup owns kHogVoteCols adjacent cell columns
// of kHogVoteRows cells, one wave per cell column and one lane per cell.  Each cell adds its
// bilinear votes in the reference's pixel order, x outer and y inner
// (CT/featurepyramid.cpp.new2:323-378), so every bin sees the same sequence of float additions
// as the reference's scatter loop.  Then hist[o][x][y] and norm[x][y] = sum_{o<9} (h_o +
// h_{o+9})^2 (CT/featurepyramid.cpp.new2:391-404).
//
// The workgroup walks the union of its cell columns' pixel columns ((kHogVoteCols + 1) sbin
// columns): all its threads stage a pixel column of its rows in LDS with contiguous loads from
// the two planes, and the waves whose cell column the pixel column votes into read their rows
// from there; the next column's loads are in flight while the current one votes.  A lane's
// bins live at s_acc[o][lane] of its wave.  Measured on 1080p pyramids, the kernel is bound
// neither by HBM nor by the read-add-write chains nor by the loads' latency: the 5-byte grid
// (8 bytes before) left its time unchanged; reading a half column's bins together and
// forwarding repeats in registers ran 20% slower (its compares), two columns in flight 5%
// slower, LDS float atomics 5x slower.  Sharing the staging across 2 or 4 cell columns (each
// pixel column read once instead of twice) ran 10% and 23% slower: only the waves whose
// columns a pixel column votes into work between the barriers.
constexpr int kHogVoteJM = 16;                                         // 2 * sbin <= 16
constexpr int kHogVoteThreads = 64 * kHogVoteCols;
constexpr int kHogVoteStage =                                          // staged rows (>= (R + 1) sbin + 1)
    (kHogVoteRows * kHogVoteJM / 2 + kHogVoteJM / 2 + 1 + 63) / 64 * 64;
constexpr int kHogVoteLoads = (kHogVoteStage + kHogVoteThreads - 1) / kHogVoteThreads;
constexpr int kHogVotePad = kHogVoteStage + kHogVoteStage / 8;         // row r at r + r / 8

__device__ inline int hog_stage_index(int r) { return r + (r >> 3); }

template <int NJ>
__device__ inline void hog_vote_column(const int2 *s_col, float *acc, float wx, int ry0, int ry1,
                                       const float *wy, int ny) {
  // two halves of NJ / 2 rows: their LDS reads go out together
#pragma unroll
  for (int h0 = 0; h0 < NJ; h0 += NJ / 2) {
    int2 g[NJ / 2];
#pragma unroll
    for (int j = 0; j < NJ / 2; j++) g[j] = s_col[hog_stage_index(min(ry0 + h0 + j, ry1))];
#pragma unroll
    for (int j = 0; j < NJ / 2; j++)
      if (h0 + j < ny) acc[g[j].y * kHogVoteRows] += wx * wy[h0 + j] * __int_as_float(g[j].x);
  }
}

template <int NJ>
__device__ inline void hog_votes_cells(const HogLevel &L, const float *__restrict__ mag,
                                       const unsigned char *__restrict__ ori, int cx0, int cy0,
                                       int2 *s_col, float *s_acc) {
  constexpr int R = kHogVoteRows, NL = kHogVoteLoads;
  const int lane = threadIdx.x & 63, cx = cx0 + (int)(threadIdx.x >> 6);
  const int sbin = L.sbin, h = L.h, w = L.w;
  const long long gs = grid_stride(h);
  const int vis0 = L.b0 * sbin, vis1 = L.b1 * sbin;
  const int cy = cy0 + lane;
  // pixels voting into a cell: interpolation cell in {c - 1, c}, within [1, vis - 2]; the
  // workgroup's columns run from its first cell column's first to its last one's last
  const int cxl = min(cx0 + kHogVoteCols - 1, L.b1 - 1);
  const int Xa = max(1, first_pix(cx0 - 1, sbin)), Xb = min(vis1 - 2, first_pix(cxl + 1, sbin) - 1);
  const bool col_live = cx <= cxl;
  const int xa = max(1, first_pix(cx - 1, sbin)), xb = min(vis1 - 2, first_pix(cx + 1, sbin) - 1);
  const int cyl = min(cy0 + R - 1, L.b0 - 1);
  const int Ya = max(1, first_pix(cy0 - 1, sbin));
  const int Yb = min(vis0 - 2, first_pix(cyl + 1, sbin) - 1);
  const int ns = min(Yb - Ya + 1, kHogVoteStage);
  int y0 = Ya, y1 = Ya, ny = 0;
  if (cy < L.b0) {
    y0 = max(1, first_pix(cy - 1, sbin));
    y1 = min(vis0 - 2, first_pix(cy + 1, sbin) - 1);
    ny = y1 - y0 + 1;
  }
  float wy[NJ];
#pragma unroll
  for (int j = 0; j < NJ; j++) {
    const int y = min(y0 + j, y1);
    int iy;
    float v0, v1;
    interp_coord(y, sbin, &iy, &v0, &v1);
    wy[j] = (iy == cy) ? v1 : v0;
  }
  // staged rows of this lane's cell (clamped into the stage)
  const int ry0 = min(max(y0 - Ya, 0), kHogVoteStage - 1);
  const int ry1 = min(max(y1 - Ya, 0), kHogVoteStage - 1);
  // staged row i of a column is grid row min(Ya + min(i, ns - 1), h - 2), as the reference
  // clamps the pixel it differentiates; every thread issues the same NL loads per column
  // (per plane), and the next column's loads are in flight while one votes
  if (Xa > Xb) return;  // uniform across the workgroup
  const int ylast = max(ns - 1, 0);
  float vm[NL];
  unsigned int vo[NL];
  auto fetch = [&](int x) {
    const long long col = (long long)min(x, w - 2) * gs;
#pragma unroll
    for (int k = 0; k < NL; k++) {
      const long long e =
          col + max(0, min(Ya + min((int)threadIdx.x + kHogVoteThreads * k, ylast), h - 2));
      vm[k] = mag[e];
      vo[k] = ori[e];
    }
  };
  fetch(Xa);
  for (int x = Xa; x <= Xb; x++) {
#pragma unroll
    for (int k = 0; k < NL; k++) {
      const int i = (int)threadIdx.x + kHogVoteThreads * k;
      if (i < kHogVoteStage) s_col[hog_stage_index(i)] = make_int2(__float_as_int(vm[k]), (int)vo[k]);
    }
    __syncthreads();
    if (x < Xb) fetch(x + 1);
    if (col_live && x >= xa && x <= xb) {
      int ix;
      float v0, v1;
      interp_coord(x, sbin, &ix, &v0, &v1);
      const float wx = (ix == cx) ? v1 : v0;
      hog_vote_column<NJ>(s_col, s_acc + lane, wx, ry0, ry1, wy, ny);
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kHogVoteThreads) void k_hog_votes(
    const HogLevel *__restrict__ lv, int nlv, const float *__restrict__ mag_arena,
    const unsigned char *__restrict__ ori_arena, float *__restrict__ hist_arena,
    float *__restrict__ norm_arena) {
  constexpr int R = kHogVoteRows;
  __shared__ int2 s_col[kHogVotePad];
  __shared__ float s_acc[kHogVoteCols][18 * R];
  const int bx = xcd_local_block(blockIdx.x, gridDim.x);
  const HogLevel L = lv[find_blocks(lv, nlv, bx)];
  const int tile = bx - L.blocks_begin;
  const int cx0 = tile / L.tiles_y * kHogVoteCols, cy0 = (tile % L.tiles_y) * R;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  float *acc = s_acc[wave];
#pragma unroll
  for (int o = 0; o < 18; o++) acc[o * R + lane] = 0.0f;
  const float *mag = mag_arena + L.grid_off;
  const unsigned char *ori = ori_arena + L.grid_off;
  if (2 * L.sbin <= kHogVoteJM / 2)
    hog_votes_cells<kHogVoteJM / 2>(L, mag, ori, cx0, cy0, s_col, acc);
  else
    hog_votes_cells<kHogVoteJM>(L, mag, ori, cx0, cy0, s_col, acc);
  __syncthreads();
  const int cx = cx0 + wave, cy = cy0 + lane;
  if (cx >= L.b1 || cy >= L.b0) return;
  const long long BSQ = (long long)L.b0 * L.b1;
  float *hist = hist_arena + L.hist_off + (long long)cx * L.b0 + cy;
  acc += lane;
  float nrm = 0.0f;
#pragma unroll
  for (int o = 0; o < 18; o++) hist[o * BSQ] = acc[o * R];
#pragma unroll
  for (int o = 0; o < 9; o++) {
    const float t = acc[o * R] + acc[(o + 9) * R];
    nrm += t * t;
  }
  norm_arena[L.norm_off + (long long)cx * L.b0 + cy] = nrm;
}

// ---------------------------------------------------------------------------------------
// HOG features: 18 contrast-sensitive, 9 insensitive, 4 texture (CT/featurepyramid.cpp.new2:
// 407-492).  The four normalisers are 1/sqrt((double)sum + 1e-4) rounded to float and the
// texture features 0.2357*t in double, as the reference's promotions compute them.
// Writes [31][F1][F0] float when feat_arena is non-null and, when featb_arena is non-null,
// the MFMA forms' channel-last copy (they read only that): hkind 1 / 2 = [F1][F0][32] BF16 /
// FP16 (the rounding of each feature, channel 31 = 0); hkind 3 / 4 = [F1][F0][64], the split
// operand pair of each feature -- channels 0..31 hold hi = round16(f), channels 32..63 hold
// lo = round16(f - hi) (f - hi is exact in FP32) -- at twice the cell offsets.
// ---------------------------------------------------------------------------------------
__device__ inline float min_2(float x) { return (x < 0.2f) ? x : 0.2f; }

__device__ inline float norm_factor(const float *p, int b0) {
  float s = p[0] + p[1] + p[b0] + p[b0 + 1];
  return (float)(1.0 / sqrt((double)s + 0.0001));
}

__device__ inline unsigned short half_bits(float f, int kind) {
  if (kind == 1 || kind == 3) {
    unsigned int u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite values)
    return (unsigned short)(u >> 16);
  }
  return __half_as_ushort(__float2half_rn(f));
}

__global__ __launch_bounds__(256) void k_hog_feat(const HogLevel *__restrict__ lv, int nlv,
                                                  const float *__restrict__ hist_arena,
                                                  const float *__restrict__ norm_arena,
                                                  float *__restrict__ feat_arena,
                                                  unsigned short *__restrict__ featb_arena,
                                                  int hkind) {
  // the block's channel-last cells, staged so that the global stores of the block's
  // contiguous region are contiguous across lanes: cell c's chunk q at c * CPC + (q ^ swz(c))
  __shared__ uint4 s_cells[256 * 8];
  const HogLevel L = lv[find_blocks(lv, nlv, (int)blockIdx.x)];
  const long long e0 = (long long)(blockIdx.x - L.blocks_begin) * blockDim.x;
  const long long e = e0 + threadIdx.x;
  const long long OSQ = (long long)L.F0 * L.F1;
  const int cpc = hkind <= 2 ? 4 : 8;               // 16-byte chunks per cell
  auto swz = [&](long long c) { return cpc == 4 ? (int)((c >> 1) & 3) : (int)(c & 7); };
  if (e < OSQ) {
    const int x = (int)(e / L.F0), y = (int)(e - (long long)x * L.F0);
    const int b0 = L.b0;
    const long long BSQ = (long long)b0 * L.b1;
    // the 3x3 norm neighbourhood (x..x+2, y..y+2) once; each normaliser sums its 2x2 window
    // in the reference's order p[0] + p[1] + p[b0] + p[b0 + 1]
    const float *nb = norm_arena + L.norm_off + (long long)x * b0 + y;
    float nv[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) nv[i][j] = nb[(long long)i * b0 + j];
    auto nf = [&](int i, int j) {
      const float sum = nv[i][j] + nv[i][j + 1] + nv[i + 1][j] + nv[i + 1][j + 1];
      return (float)(1.0 / sqrt((double)sum + 0.0001));
    };
    const float n1 = nf(1, 1), n2 = nf(1, 0), n3 = nf(0, 1), n4 = nf(0, 0);
    const float *src = hist_arena + L.hist_off + (long long)(x + 1) * b0 + (y + 1);
    float fv[31];
    float t1 = 0.0f, t2 = 0.0f, t3 = 0.0f, t4 = 0.0f;
    float hs[18];
#pragma unroll
    for (int o = 0; o < 18; o++) hs[o] = src[o * BSQ];
#pragma unroll
    for (int o = 0; o < 18; o++) {
      const float hv = hs[o];
      float h1 = min_2(hv * n1), h2 = min_2(hv * n2), h3 = min_2(hv * n3), h4 = min_2(hv * n4);
      fv[o] = 0.5f * (h1 + h2 + h3 + h4);
      t1 += h1; t2 += h2; t3 += h3; t4 += h4;
    }
#pragma unroll
    for (int o = 0; o < 9; o++) {
      float sum = hs[o] + hs[o + 9];
      float h1 = min_2(sum * n1), h2 = min_2(sum * n2), h3 = min_2(sum * n3), h4 = min_2(sum * n4);
      fv[18 + o] = 0.5f * (h1 + h2 + h3 + h4);
    }
    fv[27] = (float)(0.2357 * (double)t1);
    fv[28] = (float)(0.2357 * (double)t2);
    fv[29] = (float)(0.2357 * (double)t3);
    fv[30] = (float)(0.2357 * (double)t4);
    if (feat_arena) {
      float *dst = feat_arena + L.feat_off + e;
#pragma unroll
      for (int o = 0; o < 31; o++) dst[o * OSQ] = fv[o];
    }
    if (featb_arena) {
      // 16-bit channels 0..31 (the last one 0), and for the split forms the lo halves as
      // channels 32..63; featb_off is a multiple of 32 elements, so every cell starts on a
      // 64-byte (128-byte) boundary
      unsigned short hb[32], lb[32];
#pragma unroll
      for (int c = 0; c < 32; c++) {
        hb[c] = c < 31 ? half_bits(fv[c], hkind) : (unsigned short)0;
        lb[c] = 0;
      }
      if (hkind > 2) {
#pragma unroll
        for (int c = 0; c < 31; c++) {
          const float hv = hkind == 3 ? __uint_as_float((unsigned int)hb[c] << 16)
                                      : __half2float(__ushort_as_half(hb[c]));
          lb[c] = half_bits(fv[c] - hv, hkind);
        }
      }
      const int t = threadIdx.x;
#pragma unroll
      for (int q = 0; q < 8; q++) {
        if (q >= cpc) break;
        const unsigned short *h = q < 4 ? hb : lb;
        const int c0 = (q & 3) * 8;
        s_cells[t * cpc + (q ^ swz(t))] =
            make_uint4(h[c0] | ((unsigned)h[c0 + 1] << 16), h[c0 + 2] | ((unsigned)h[c0 + 3] << 16),
                       h[c0 + 4] | ((unsigned)h[c0 + 5] << 16), h[c0 + 6] | ((unsigned)h[c0 + 7] << 16));
      }
    }
  }
  if (!featb_arena) return;  // uniform over the launch
  __syncthreads();
  // the block's cells [e0, e0 + 256) are contiguous in the arena: write them chunk by chunk
  uint4 *db = (uint4 *)(featb_arena + (cpc == 4 ? L.featb_off : 2 * L.featb_off)) + e0 * cpc;
  const long long ncells = OSQ - e0 < 256 ? OSQ - e0 : 256;
  for (int k = threadIdx.x; k < (int)ncells * cpc; k += 256) {
    const int c = k / cpc, q = k - c * cpc;
    db[k] = s_cells[c * cpc + (q ^ swz(c))];
  }
}

// ---------------------------------------------------------------------------------------
// FP32 filter convolution, LDS-tiled.  A workgroup computes a 16 (x) x 64 (y) tile of
// output positions for NF filters of one shape; each thread owns 4 consecutive y x NF
// filters.  Per output the arithmetic is exactly CT/GPU_function.cu:302-358:
//   add = 0; for f: { val = 0; for xp: for yp: val = fma(A, B, val); add += val; }
// Channels stream through double-buffered LDS tiles.
// ---------------------------------------------------------------------------------------
template <int NF>
__global__ __launch_bounds__(256) void k_conv_f32(const ConvTile *__restrict__ tiles, int fh,
                                                  int fw, const float *__restrict__ feat_arena,
                                                  const float *__restrict__ filt,
                                                  const long long *__restrict__ map_off,
                                                  const int2 *__restrict__ map_dims,
                                                  float *__restrict__ out_arena) {
  constexpr int TX = 16, TY = 64, PY = 4;
  constexpr int AXM = TX + 15, AYP = TY + 16;  // fw, fh <= 16
  __shared__ float As[2][AXM * AYP];
  __shared__ float Bs[2][16 * 16 * NF];
  const ConvTile t = tiles[blockIdx.x];
  const int tid = threadIdx.x;
  const int tx = tid >> 4, ty = tid & 15;
  const int AX = TX + fw - 1, AY = TY + fh - 1;
  const long long FSQ = (long long)t.F0 * t.F1;
  const float *A0 = feat_arena + t.feat_off;
  const float *B0 = filt + t.filt_off;
  const int nb = fw * fh * NF;

  auto stage = [&](int f, int buf) {
    const float *A = A0 + f * FSQ;
    for (int i = tid; i < AX * AY; i += 256) {
      int ax = i / AY, ay = i - ax * AY;
      int gx = t.x0 + ax, gy = t.y0 + ay;
      As[buf][ax * AYP + ay] = (gx < t.F1 && gy < t.F0) ? A[(long long)gx * t.F0 + gy] : 0.0f;
    }
    const float *B = B0 + (long long)f * nb;
    for (int i = tid; i < nb; i += 256) Bs[buf][i] = B[i];
  };

  float acc[PY][NF];
#pragma unroll
  for (int i = 0; i < PY; i++)
#pragma unroll
    for (int n = 0; n < NF; n++) acc[i][n] = 0.0f;

  stage(0, 0);
  __syncthreads();
  for (int f = 0; f < 31; f++) {
    const int cur = f & 1;
    if (f + 1 < 31) stage(f + 1, cur ^ 1);
    float val[PY][NF];
#pragma unroll
    for (int i = 0; i < PY; i++)
#pragma unroll
      for (int n = 0; n < NF; n++) val[i][n] = 0.0f;
    const float *Ab = &As[cur][tx * AYP + ty * PY];
    const float *Bb = &Bs[cur][0];
    for (int xp = 0; xp < fw; xp++) {
      for (int yp = 0; yp < fh; yp++) {
        float b[NF];
#pragma unroll
        for (int n = 0; n < NF; n++) b[n] = Bb[(xp * fh + yp) * NF + n];
#pragma unroll
        for (int i = 0; i < PY; i++) {
          float a = Ab[xp * AYP + i + yp];
#pragma unroll
          for (int n = 0; n < NF; n++) val[i][n] = __builtin_fmaf(a, b[n], val[i][n]);
        }
      }
    }
#pragma unroll
    for (int i = 0; i < PY; i++)
#pragma unroll
      for (int n = 0; n < NF; n++) acc[i][n] += val[i][n];
    __syncthreads();
  }
  const int x = t.x0 + tx;
#pragma unroll
  for (int n = 0; n < NF; n++) {
    if (n >= t.nvalid) continue;
    const int2 hw = map_dims[t.out_slot + n];  // the member's own output size
    if (x >= hw.y) continue;
    float *o = out_arena + map_off[t.out_slot + n] + (long long)x * hw.x;
#pragma unroll
    for (int i = 0; i < PY; i++) {
      int y = t.y0 + ty * PY + i;
      if (y < hw.x) o[y] = 0.0f + acc[i][n];
    }
  }
}

// ---------------------------------------------------------------------------------------
// Distance transform, one 1-D pass over a block of lines (columns or rows) of one map.
// The reference's recursion (CT/dt.cpp:61-85) visits node d of interval [d1, d2] with the
// source range [s1, s2] = [ptr[d1-1] or 0, ptr[d2+1] or n-1] and keeps the first index of
// minimal cost ((src + (a*t)*t) + b*t), t = d - s.  Processing the recursion tree depth by
// depth gives the same (value, index) at every node; a group of G lanes owns one line, and
// at depths with fewer nodes than lanes a node's range is split across lanes and the
// partial results are merged lexicographically on (cost, index), which is exactly "first
// minimum" of the sequential scan.
// ---------------------------------------------------------------------------------------
__device__ inline float dt_cost(float src, float a, float b, int t) {
  float ft = (float)t;
  return src + a * ft * ft + b * ft;
}

// Scans p in [p0, p1] in ascending order after (best, bi) was set from p0 - 1, keeping the
// first minimum (the reference's strict "best > cost" switch).  Loads go out U at a time
// so the LDS latency overlaps the compare chain.
template <int U>
__device__ inline void dt_scan(const float *ls, float a, float b, int d, int p0, int p1,
                               float &best, int &bi) {
  int p = p0;
  for (; p + U - 1 <= p1; p += U) {
    float v[U];
#pragma unroll
    for (int u = 0; u < U; u++) v[u] = ls[p + u];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const float c = dt_cost(v[u], a, b, d - (p + u));
      if (best > c) {
        best = c;
        bi = p + u;
      }
    }
  }
  for (; p <= p1; p++) {
    const float c = dt_cost(ls[p], a, b, d - p);
    if (best > c) {
      best = c;
      bi = p;
    }
  }
}


// Stages the job's lines into LDS, line l at s_src[l * LS + H0 + p], negated when the job
// says so.  MINMAX also returns this thread's minimum and maximum of the staged values.
// The dimension contiguous in memory (positions of a line, or the lines at one position) is
// read four elements per load (16 bytes at 4-byte alignment, which gfx950 global loads
// accept), U such loads in flight per thread, so a job's lines arrive in few round trips.
struct __attribute__((packed, aligned(4))) DtQuad {
  float v0, v1, v2, v3;
};

template <bool MINMAX>
__device__ inline void dt_stage(const DtJob &J, const float *__restrict__ src_arena, float *s_src,
                                int LS, int H0, float &vmin, float &vmax) {
  const int n = J.n, tid = threadIdx.x;
  const float *src = src_arena + J.src_off + (long long)J.line0 * J.line_stride;
  constexpr int U = 4;
  const bool cols = J.elem_stride == 1;
  // inner: the contiguous dimension (positions for columns, lines for rows); outer: the other
  const int inner = cols ? n : J.nlines, outer = cols ? J.nlines : n;
  const long long ostride = cols ? J.line_stride : J.elem_stride;
  const int qn = (inner + 3) >> 2, total = outer * qn;
  const float sgn = J.neg ? -1.0f : 1.0f;
  for (int q0 = tid; q0 < total; q0 += 256 * U) {
    float v[U][4];
    int oo[U], ii[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int q = q0 + u * 256;
      const int o = q < total ? q / qn : 0;
      const int i = q < total ? 4 * (q - o * qn) : inner;  // i == inner: nothing to load
      oo[u] = o;
      ii[u] = i;
      const float *p = src + (long long)o * ostride + i;
      if (i + 3 < inner) {
        const DtQuad t = *(const DtQuad *)p;
        v[u][0] = t.v0;
        v[u][1] = t.v1;
        v[u][2] = t.v2;
        v[u][3] = t.v3;
      } else {
#pragma unroll
        for (int k = 0; k < 4; k++) v[u][k] = i + k < inner ? p[k] : 0.0f;
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int i = ii[u] + k;
        if (i >= inner) continue;
        const int l = cols ? oo[u] : i, p = cols ? i : oo[u];
        const float x = v[u][k] * sgn;
        s_src[l * LS + H0 + p] = x;
        if (MINMAX) {
          vmin = fminf(vmin, x);
          vmax = fmaxf(vmax, x);
        }
      }
    }
  }
}

// divide and conquer, one tree depth at a time: a node (d1, d2) takes d = (d1 + d2) / 2
// and searches the sources between the argmins of its bounding ancestors d1 - 1 and
// d2 + 1 (CT/dt.cpp:61-85).  The nodes of each depth come from the LDS copy of the
// breadth-first node table of length n.
__device__ inline void dt_divide_conquer(const DtJob &J, const int *__restrict__ depths,
                                         const float *s_src, int LS, int H0, int *s_ptr,
                                         const int2 *s_node) {
  const int n = J.n, G = J.G;
  const int tid = threadIdx.x;
  const int line = tid / G, g = tid - line * G;
  const bool active_line = line < J.nlines;
  const float *ls = s_src + line * LS + H0;
  int *lp = s_ptr + line * n;
  const float a = J.a, b = J.b;
  for (int k = 0; k < J.depth_count; k++) {
    const int nb = depths[J.depth_off + k], ne = depths[J.depth_off + k + 1];
    const int nk = ne - nb;
    if (nk >= G) {
      for (int i = nb + g; i < ne; i += G) {
        if (!active_line) break;
        const int2 nd = s_node[i];
        const int d = nd.x & 0xffff, d1 = nd.x >> 16, d2 = nd.y;
        const int s1 = d1 == 0 ? 0 : lp[d1 - 1];
        const int s2 = d2 == n - 1 ? n - 1 : lp[d2 + 1];
        int s = s1;
        float best = dt_cost(ls[s1], a, b, d - s1);
        dt_scan<4>(ls, a, b, d, s1 + 1, s2, best, s);
        lp[d] = s;
      }
    } else {
      int S = 1;
      while (S * 2 * nk <= G) S *= 2;
      const int ni = g / S, part = g - ni * S;
      const bool act = active_line && ni < nk;
      float best = __builtin_inff();
      int bi = 0x7fffffff;
      int d = 0;
      if (act) {
        const int2 nd = s_node[nb + ni];
        d = nd.x & 0xffff;
        const int d1 = nd.x >> 16, d2 = nd.y;
        const int s1 = d1 == 0 ? 0 : lp[d1 - 1];
        const int s2 = d2 == n - 1 ? n - 1 : lp[d2 + 1];
        const int len = s2 - s1 + 1;
        const int chunk = (len + S - 1) / S;
        const int p0 = s1 + part * chunk;
        const int p1 = min(s2, p0 + chunk - 1);
        if (p0 <= p1) {
          best = dt_cost(ls[p0], a, b, d - p0);
          bi = p0;
          dt_scan<4>(ls, a, b, d, p0 + 1, p1, best, bi);
        }
      }
      for (int m = 1; m < S; m <<= 1) {
        float ob = __shfl_xor(best, m);
        int oi = __shfl_xor(bi, m);
        if (ob < best || (ob == best && oi < bi)) {
          best = ob;
          bi = oi;
        }
      }
      if (act && part == 0) lp[d] = bi;
    }
    __syncthreads();
  }
}

__device__ inline void dt_load_nodes(const DtJob &J, const int4 *__restrict__ nodes,
                                     int2 *s_node) {
  for (int i = threadIdx.x; i < J.n; i += 256) {
    const int4 nd = nodes[J.node_off + i];
    s_node[i] = make_int2(nd.x | ((nd.y + 1) << 16), nd.z - 1);
  }
}

// writes M and the argmin back with the staging pattern (coalesced along the memory
// order); M is the reference's cost expression at the chosen argmin, hence the same bits
// as the sequential scan whenever the argmin is the same
__device__ inline void dt_write(const DtJob &J, const float *s_src, int LS, int H0,
                                const int *s_ptr, float *__restrict__ dst_arena,
                                int *__restrict__ ptr_arena) {
  const int n = J.n, tid = threadIdx.x;
  const float a = J.a, b = J.b;
  float *dst = dst_arena + J.dst_off + (long long)J.line0 * J.line_stride;
  int *pdst = ptr_arena + J.dst_off + (long long)J.line0 * J.line_stride;
  if (J.elem_stride == 1) {
    for (int i = tid; i < J.nlines * n; i += 256) {
      const int l = i / n, p = i - l * n;
      const int sp = s_ptr[l * n + p];
      const long long o = (long long)l * J.line_stride + p;
      dst[o] = dt_cost(s_src[l * LS + H0 + sp], a, b, p - sp);
      pdst[o] = sp;
    }
  } else {
    for (int i = tid; i < J.nlines * n; i += 256) {
      const int p = i / J.nlines, l = i - p * J.nlines;
      const int sp = s_ptr[l * n + p];
      const long long o = (long long)l * J.line_stride + (long long)p * J.elem_stride;
      dst[o] = dt_cost(s_src[l * LS + H0 + sp], a, b, p - sp);
      pdst[o] = sp;
    }
  }
}

// The reference's DT: a group of G lanes per line, LDS lines of n + 1 floats.
__global__ __launch_bounds__(256) void k_dt(const DtJob *__restrict__ jobs,
                                            const int4 *__restrict__ nodes,
                                            const int *__restrict__ depths,
                                            const float *__restrict__ src_arena,
                                            float *__restrict__ dst_arena,
                                            int *__restrict__ ptr_arena) {
  extern __shared__ float smem[];
  const DtJob J = jobs[xcd_local_block(blockIdx.x, gridDim.x)];
  const int n = J.n, NLn = 256 / J.G;
  float *s_src = smem;                                  // [NLn][n+1]
  int *s_ptr = (int *)(smem + NLn * (n + 1));           // [NLn][n]
  int2 *s_node = (int2 *)(s_ptr + NLn * n + (NLn * n & 1));  // [n] (d | d1 << 16, d2)
  dt_load_nodes(J, nodes, s_node);
  float unused0 = 0.0f, unused1 = 0.0f;
  dt_stage<false>(J, src_arena, s_src, n + 1, 0, unused0, unused1);
  __syncthreads();
  dt_divide_conquer(J, depths, s_src, n + 1, 0, s_ptr, s_node);
  dt_write(J, s_src, n + 1, 0, s_ptr, dst_arena, ptr_arena);
}

// ---------------------------------------------------------------------------------------
// Distance transform of the 16-bit forms: the reference's DT, with each output's minimum
// searched in a bounded window instead of the divide and conquer's source range.
//
// For output d let g(d) be the first source of least cost c_d(s) = ((f(s) + (a*t)*t) + b*t),
// t = d - s, over the whole line, with every cost rounded as the reference rounds it.
//  - Window.  c_d(d) = f(d) exactly, and a source at |t| >= T loses to it whenever
//    a t^2 - |b| |t| exceeds the line block's value range plus the costs' rounding error, so
//    g(d) lies within T of d for the T the job takes from that bound.
//  - Equality with the divide and conquer (CT/dt.cpp:61-85).  In exact arithmetic the
//    cost difference of two sources s' < s changes by -2a (d' - d)(s - s') <= -2a from d to
//    d' > d.  With eps a bound of the rounding error of every cost in the windows, the
//    rounded first argmins g therefore cannot decrease along the line when a >= 2 eps
//    (each decrease would need a rounded difference to move against an exact change of at
//    least 2a, by more than 4 eps).  The divide and conquer scans node d over
//    [s1, s2] = [g(d1 - 1), g(d2 + 1)], which holds g(d) when g is monotone, and keeps the
//    first least cost there, which is g(d) (the first global minimum; earlier sources cost
//    strictly more).  So its argmin at every node is g(d), and M is the same expression at
//    the same source: the outputs are bit-identical to dt()'s.
// The kernel checks both conditions per job from the staged values; a job that fails one
// (a <= 0, a window wider than 16, or a too small against eps) runs the divide and conquer.
// Each output then takes the first least cost over the 2T + 1 sources of its window in
// ascending order, with the reference's cost expression and strict comparison.
// ---------------------------------------------------------------------------------------
// B0: b == 0 (either sign), where b*t is a zero and (x + Q) + b*t == x + Q for every x, Q
// with Q >= 0 (a sum that is exactly zero rounds to +0, and +0 + -0 = +0), so each window
// position costs one add less.
template <int T, bool ROWS, bool B0>
__device__ inline void dt_window(const DtJob &J, const float *s_f, int LSf,
                                 float *__restrict__ dst_arena, int *__restrict__ ptr_arena) {
  constexpr int NW = 2 * T + 1;
  const int n = J.n, nl = J.nlines;
  const float a = J.a, b = J.b;
  // the cost terms (a*t)*t and b*t of window position j (t = T - j) are the same for every
  // lane, so they live in scalar registers
  float Q[NW], B[NW];
#pragma unroll
  for (int j = 0; j < NW; j++) {
    const float t = (float)(T - j);
    Q[j] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int((a * t) * t)));
    B[j] = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(b * t)));
  }
  float *dst = dst_arena + J.dst_off + (long long)J.line0 * J.line_stride;
  int *pdst = ptr_arena + J.dst_off + (long long)J.line0 * J.line_stride;
  // the selected outputs go to half grids of the map (H = ceil(P0 / 2) cells per column):
  // the argmins as 16-bit indices in the first half of the map's int area, the rows pass's M
  // at the start of the map's float area
  float *dstc = dst_arena + J.dst_off;
  unsigned short *pc = (unsigned short *)ptr_arena + 2 * J.dst_off;
  const int H = ((ROWS ? J.elem_stride : n) + 1) >> 1;
  const int tid = threadIdx.x;
  // ROWS (lines interleaved in memory): lanes take consecutive lines of one position, so the
  // stores are contiguous; otherwise consecutive positions of one line
  const int NL = 256 / J.G;
  // output selection (DtJob::sel): rows pass, each line keeps every other position, and
  // writes M (on rows y == ay - 1 mod 2) or the argmin (on rows y == ay mod 2)
  const bool sel_rows = ROWS && J.sel == 2, sel_cols = !ROWS && J.sel == 1;
  int l, d, dstep, lstep = 0;
  int par = 0, kind = 0;  // rows pass: first position and output kind of this lane's line
  if (ROWS) {
    l = tid & (NL - 1);
    d = tid / NL;
    dstep = 256 / NL;
    if (sel_rows) {
      kind = (J.line0 + l - J.ay + 1) & 1;  // 0: M, 1: argmin
      par = (J.ax - 1 + kind) & 1;
    }
  } else {
    l = tid / n;
    d = tid - l * n;
    lstep = 256 / n;
    dstep = 256 - lstep * n;
  }
  const int dlim = sel_rows ? (n + 1) / 2 : n;
  while (ROWS ? d < dlim : l < nl) {
    const int dd = sel_rows ? 2 * d + par : d;
    if (!ROWS || (l < nl && dd < n)) {
      // sources dd - T .. dd + T; outside the line the halo holds +inf, which never wins
      const float *w = s_f + l * LSf + kDtHalo - T + dd;
      float best = __builtin_inff();
      int bj = T;
#pragma unroll
      for (int j = 0; j < NW; j++) {
        const float c = B0 ? w[j] + Q[j] : (w[j] + Q[j]) + B[j];
        if (best > c) {
          best = c;
          bj = j;
        }
      }
      const int s = dd - T + bj;
      if (sel_rows) {
        // half grid: (y, x) at (x >> 1) * H + (y >> 1), M or the 16-bit argmin
        const long long oc = (long long)(dd >> 1) * H + ((J.line0 + l) >> 1);
        if (kind == 0) dstc[oc] = best;
        else pc[oc] = (unsigned short)s;
      } else {
        const long long o = (long long)l * J.line_stride + (long long)dd * J.elem_stride;
        dst[o] = best;
        if (!sel_cols) pdst[o] = s;
        else if (((dd - J.ay) & 1) == 0)
          pc[(long long)(J.line0 + l) * H + (dd >> 1)] = (unsigned short)s;  // column-major half
      }
    }
    d += dstep;
    if (!ROWS) {
      l += lstep;
      if (d >= n) {
        d -= n;
        l++;
      }
    }
  }
}

// The selected outputs (DtJob::sel 1 or 2) of a job decided by the divide and conquer, in the
// windowed form's half-grid layout (the same cells and places as dt_window writes).
__device__ inline void dt_write_selected(const DtJob &J, const float *s_src, const int *s_ptr,
                                         float *__restrict__ dst_arena,
                                         int *__restrict__ ptr_arena) {
  const int n = J.n, tid = threadIdx.x;
  const float a = J.a, b = J.b;
  float *dstc = dst_arena + J.dst_off;
  unsigned short *pc = (unsigned short *)ptr_arena + 2 * J.dst_off;
  if (J.sel == 1) {  // columns pass: M everywhere, the argmin on rows y == ay mod 2
    const int H = (n + 1) >> 1;
    float *dst = dstc + (long long)J.line0 * J.line_stride;
    for (int i = tid; i < J.nlines * n; i += 256) {
      const int l = i / n, p = i - l * n;
      const int sp = s_ptr[l * n + p];
      dst[(long long)l * J.line_stride + p] = dt_cost(s_src[l * n + sp], a, b, p - sp);
      if (((p - J.ay) & 1) == 0)
        pc[(long long)(J.line0 + l) * H + (p >> 1)] = (unsigned short)sp;
    }
  } else {  // rows pass: M on rows y == ay - 1 at x == ax - 1, the argmin on y == ay at x == ax
    const int H = (J.elem_stride + 1) >> 1;
    for (int i = tid; i < J.nlines * n; i += 256) {
      const int p = i / J.nlines, l = i - p * J.nlines;
      const int y = J.line0 + l;
      const int kind = (y - J.ay + 1) & 1;  // 0: M, 1: argmin
      if (((p - (J.ax - 1 + kind)) & 1) != 0) continue;
      const int sp = s_ptr[l * n + p];
      const long long oc = (long long)(p >> 1) * H + (y >> 1);
      if (kind == 0) dstc[oc] = dt_cost(s_src[l * n + sp], a, b, p - sp);
      else pc[oc] = (unsigned short)sp;
    }
  }
}

__global__ __launch_bounds__(256) void k_dt_window(const DtJob *__restrict__ jobs,
                                                   const int4 *__restrict__ nodes,
                                                   const int *__restrict__ depths,
                                                   const float *__restrict__ src_arena,
                                                   float *__restrict__ dst_arena,
                                                   int *__restrict__ ptr_arena) {
  extern __shared__ float smem[];
  __shared__ float s_red[2][4];
  const DtJob J = jobs[xcd_local_block(blockIdx.x, gridDim.x)];
  const int n = J.n, NLn = 256 / J.G, LSf = dt_line_stride(n, NLn);
  const int tid = threadIdx.x;
  // [NLn][LSf] + kDtHalo float lines that share their halos (line l at l * LSf + kDtHalo):
  // the halo words hold +inf, the staged source the rest
  {
    const int gap = LSf - n;
    for (int i = tid; i < kDtHalo + J.nlines * gap; i += 256) {
      int at = i;
      if (i >= kDtHalo) {
        const int k = i - kDtHalo, ln = k / gap;
        at = ln * LSf + kDtHalo + n + (k - ln * gap);
      }
      smem[at] = __builtin_inff();
    }
  }
  float vmin = __builtin_inff(), vmax = -__builtin_inff();
  dt_stage<true>(J, src_arena, smem, LSf, kDtHalo, vmin, vmax);
  for (int m = 1; m < 64; m <<= 1) {
    vmin = fminf(vmin, __shfl_xor(vmin, m));
    vmax = fmaxf(vmax, __shfl_xor(vmax, m));
  }
  if ((tid & 63) == 0) {
    s_red[0][tid >> 6] = vmin;
    s_red[1][tid >> 6] = vmax;
  }
  __syncthreads();
  vmin = fminf(fminf(s_red[0][0], s_red[0][1]), fminf(s_red[0][2], s_red[0][3]));
  vmax = fmaxf(fmaxf(s_red[1][0], s_red[1][1]), fmaxf(s_red[1][2], s_red[1][3]));
  // Window half-width.  A cost (f + (a*t)*t) + b*t is within u (2|f| + 4 a t^2 + 2|b t|)
  // (1 + 1e-3) of the exact f + a t^2 + b t (u = 2^-24), so a source at offset t loses to
  // the centre (cost f(d), exact) once a (1 - 2^-20) t^2 - |b| (1 + 2^-20) |t| exceeds the
  // value range plus 2^-22 max|f|: beyond the positive root x of that quadratic.
  const float a = J.a, b = J.b;
  const float fmag = fmaxf(fabsf(vmin), fabsf(vmax));
  int T = 0;
  if (a > 0.0f && fmag < 1e30f) {
    const float a1 = a * (1.0f - 9.5367431640625e-7f), b1 = fabsf(b) * (1.0f + 9.5367431640625e-7f);
    const float R = (vmax - vmin) + fmag * 2.384185791015625e-7f + 1e-30f;
    const float x = (b1 + sqrtf(b1 * b1 + 4.0f * a1 * R)) / (2.0f * a1);
    if (x < 16.0f) {
      const int need = (int)floorf(x * (1.0f + 1.52587890625e-5f) + 9.765625e-4f) + 1;
      const int Tw = need < 8 ? 8 : need;  // 8 .. 16
      // monotone argmins: a >= 2 eps with eps = 2^-23 (2 max|f| + 4 a T^2 + 2 |b| T) (twice
      // the rounding bound above)
      const float ft = (float)Tw;
      const float eps = 1.1920928955078125e-7f * (2.0f * fmag + 4.0f * a * ft * ft + 2.0f * fabsf(b) * ft);
      if (need <= 16 && a >= 2.0f * eps) T = Tw;
    }
  }
  if (T == 0) {
    // the reference's divide and conquer, J.fb_lines lines at a time (float lines, argmins
    // and the node table fit the window lines' area); the lines are staged again from memory
    const int fb = J.fb_lines;
    float *s_f = smem;                                                // [fb][n]
    int *s_ptr = (int *)(smem + fb * n);                              // [fb][n]
    int2 *s_node = (int2 *)(s_ptr + fb * n + ((2 * fb * n) & 1));     // [n]
    dt_load_nodes(J, nodes, s_node);
    for (int l0 = 0; l0 < J.nlines; l0 += fb) {
      DtJob Js = J;
      Js.line0 = J.line0 + l0;
      Js.nlines = min(fb, J.nlines - l0);
      float u0 = 0.0f, u1 = 0.0f;
      dt_stage<false>(Js, src_arena, s_f, n, 0, u0, u1);
      __syncthreads();
      dt_divide_conquer(Js, depths, s_f, n, 0, s_ptr, s_node);
      if (J.sel == 0) dt_write(Js, s_f, n, 0, s_ptr, dst_arena, ptr_arena);
      else dt_write_selected(Js, s_f, s_ptr, dst_arena, ptr_arena);
      __syncthreads();
    }
    return;
  }
  const bool rows = J.elem_stride != 1;
  const bool b0 = b == 0.0f;
#define GEYE_DT_WINDOW(TT)                                                                \
  (rows ? (b0 ? dt_window<TT, true, true>(J, smem, LSf, dst_arena, ptr_arena)              \
              : dt_window<TT, true, false>(J, smem, LSf, dst_arena, ptr_arena))            \
        : (b0 ? dt_window<TT, false, true>(J, smem, LSf, dst_arena, ptr_arena)             \
              : dt_window<TT, false, false>(J, smem, LSf, dst_arena, ptr_arena)))
  switch (T) {
    case 8: GEYE_DT_WINDOW(8); break;
    case 9: GEYE_DT_WINDOW(9); break;
    case 10: GEYE_DT_WINDOW(10); break;
    case 11: GEYE_DT_WINDOW(11); break;
    case 12: GEYE_DT_WINDOW(12); break;
    case 13: GEYE_DT_WINDOW(13); break;
    case 14: GEYE_DT_WINDOW(14); break;
    case 15: GEYE_DT_WINDOW(15); break;
    default: GEYE_DT_WINDOW(16); break;
  }
#undef GEYE_DT_WINDOW
}

// ---------------------------------------------------------------------------------------
// Scores: root + bias, minus each part's DT value at (ay-1+2y, ax-1+2x) in part order
// (CT/get_boxes.cpp:1022-1041, CT/dt.cpp:38-55), written in the reference's output order;
// flags mark score > thresh (CT/get_boxes.cpp:185-220).
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_score(const ScoreMap *__restrict__ maps, int nmaps,
                                               const ScorePart *__restrict__ parts,
                                               const float *__restrict__ conv_arena,
                                               const float *__restrict__ M_arena, float thresh,
                                               int half, float *__restrict__ score,
                                               int *__restrict__ flags) {
  // each workgroup covers 256 cells of one map (ScoreMap::blocks_begin), so the map and its
  // parts are the same for every lane
  const int bx = blockIdx.x;
  int lo = 0, hi = nmaps - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (maps[mid].blocks_begin <= bx) lo = mid; else hi = mid - 1;
  }
  const ScoreMap m = maps[lo];
  const long long i = (long long)(bx - m.blocks_begin) * 256 + threadIdx.x;
  if (i >= (long long)m.R0 * m.R1) return;
  const long long e = m.score_off + i;
  const int x = (int)(i / m.R0), y = (int)(i - (long long)x * m.R0);
  float s = conv_arena[m.root_off + i] + m.off;
  // up to 8 parts' values are loaded together, then subtracted in part order
  for (int k0 = 0; k0 < m.npart; k0 += 8) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      v[u] = 0.0f;
      if (k0 + u < m.npart) {
        const ScorePart q = parts[m.part_first + k0 + u];
        // half grid (the windowed DT's selected outputs): cell (ay - 1 + 2y, ax - 1 + 2x)
        // sits at (x + (ax - 1) / 2) * ceil(P0 / 2) + y + (ay - 1) / 2
        v[u] = half ? M_arena[q.M_off + (long long)(x + ((q.ax - 1) >> 1)) * ((q.P0 + 1) >> 1) +
                              y + ((q.ay - 1) >> 1)]
                    : M_arena[q.M_off + (q.ay - 1 + 2 * y) + (long long)(q.ax - 1 + 2 * x) * q.P0];
      }
    }
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (k0 + u < m.npart) s -= v[u];
  }
  score[e] = s;
  flags[e] = s > thresh ? 1 : 0;
}

// ---------------------------------------------------------------------------------------
// Order-preserving stream compaction of flagged indices: block counts, then a scatter pass
// whose workgroups add up the counts before them.
// ---------------------------------------------------------------------------------------
constexpr int kCompactBlock = 1024;

__global__ __launch_bounds__(256) void k_flag_count(const int *__restrict__ flags, long long n,
                                                    int *__restrict__ blocksum) {
  __shared__ int s[256];
  long long base = (long long)blockIdx.x * kCompactBlock;
  int c = 0;
  for (int i = threadIdx.x; i < kCompactBlock; i += 256)
    if (base + i < n) c += flags[base + i];
  s[threadIdx.x] = c;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) blocksum[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(256) void k_flag_scatter(const int *__restrict__ flags,
                                                      long long n,
                                                      const int *__restrict__ blocksum,
                                                      int *__restrict__ total,
                                                      int *__restrict__ host_total,
                                                      int *__restrict__ out) {
  __shared__ int s[256];
  __shared__ int s_red[4];
  const int before = preceding_count(blocksum, blockIdx.x, s_red);
  long long base = (long long)blockIdx.x * kCompactBlock;
  // each thread owns 4 consecutive elements
  int f[4], c = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    long long i = base + threadIdx.x * 4 + k;
    f[k] = (i < n) ? flags[i] : 0;
    c += f[k];
  }
  s[threadIdx.x] = c;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {
    int t = (int)threadIdx.x >= o ? s[threadIdx.x - o] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  int pos = before + s[threadIdx.x] - c;
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 255) {
    *total = before + s[255];
    if (host_total) *(volatile int *)host_total = before + s[255];  // mapped host memory
  }
#pragma unroll
  for (int k = 0; k < 4; k++)
    if (f[k]) out[pos++] = (int)(base + threadIdx.x * 4 + k);
}

// ---------------------------------------------------------------------------------------
// Candidate rows (CT/get_boxes.cpp:1497-1546): root box (rootbox, :227-235), part boxes
// (partbox, :243-258, including its probe at (2(y-1)+ay, 2(x-1)+ax)), component, score,
// scale.  Ix/Iy are composed at the probe from the two DT passes' argmins
// (CT/dt.cpp:120-131).
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_rows(const ScoreMap *__restrict__ maps, int nmaps,
                                              const ScorePart *__restrict__ parts,
                                              const float *__restrict__ score,
                                              const int *__restrict__ cand,
                                              const int *__restrict__ sel, int ncand,
                                              const int *__restrict__ tmpIx,
                                              const int *__restrict__ tmpIy, int half, int GL,
                                              int padx, int pady, const int *__restrict__ rsize,
                                              float *__restrict__ rows) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncand) return;
  // row c is candidate sel[c] (the NMS survivors, in their original order) or candidate c
  const long long e = cand[sel ? sel[c] : c];
  int lo = 0, hi = nmaps - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (maps[mid].score_off <= e) lo = mid; else hi = mid - 1;
  }
  const ScoreMap m = maps[lo];
  const long long i = e - m.score_off;
  const int x = (int)(i / m.R0), y = (int)(i - (long long)x * m.R0);
  const float scale = m.scale;
  float *r = rows + (long long)c * GL;
  float r0 = ((float)y - (float)pady + 1.0f) * scale;
  float r1 = ((float)x - (float)padx + 1.0f) * scale;
  r[0] = r0;
  r[1] = r1;
  r[2] = (float)((double)(r0 + (float)rsize[m.comp * 2] * scale) - 1.0);
  r[3] = (float)((double)(r1 + (float)rsize[m.comp * 2 + 1] * scale) - 1.0);
  // up to 8 parts at a time: their Ix loads go out together, then their Iy loads
  for (int k0 = 0; k0 < m.npart; k0 += 8) {
    ScorePart q[8];
    int ix[8], iy[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      ix[u] = 0;
      if (k0 + u >= m.npart) continue;
      q[u] = parts[m.part_first + k0 + u];
      const int probex = (x - 1) * 2 + q[u].ax, probey = (y - 1) * 2 + q[u].ay;
      // 16-bit half grids: (y, x) at (x >> 1) * ceil(P0 / 2) + (y >> 1)
      ix[u] = half ? ((const unsigned short *)tmpIx)[2 * q[u].Ix_off +
                                                     (long long)(probex >> 1) * ((q[u].P0 + 1) >> 1) +
                                                     (probey >> 1)]
                   : tmpIx[q[u].Ix_off + probey + (long long)probex * q[u].P0];
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      iy[u] = 0;
      if (k0 + u >= m.npart) continue;
      const int probey = (y - 1) * 2 + q[u].ay;
      // the columns pass's argmins: every column, rows y == ay mod 2 (the probe's parity)
      iy[u] = half ? ((const unsigned short *)tmpIy)[2 * q[u].Iy_off +
                                                     (long long)ix[u] * ((q[u].P0 + 1) >> 1) +
                                                     (probey >> 1)]
                   : tmpIy[q[u].Iy_off + (long long)ix[u] * q[u].P0 + probey];
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      if (k0 + u >= m.npart) continue;
      const float px = (float)ix[u] + 1.0f, py = (float)iy[u] + 1.0f;
      float b0 = (float)((((double)py - 2.0) / 2.0 + 1.0 - (double)(float)pady) * (double)scale);
      float b1 = (float)((((double)px - 2.0) / 2.0 + 1.0 - (double)(float)padx) * (double)scale);
      float *o = r + 4 + 4 * (k0 + u);
      o[0] = b0;
      o[1] = b1;
      o[2] = (float)(((double)b0 + (double)((float)q[u].ps0 * scale) / 2.0) - 1.0);
      o[3] = (float)(((double)b1 + (double)((float)q[u].ps1 * scale) / 2.0) - 1.0);
    }
  }
  r[4 + 4 * m.npart] = (float)m.comp;
  r[5 + 4 * m.npart] = score[e];
  r[6 + 4 * m.npart] = scale;
}

// ---------------------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------------------
void launch_resize_rows(hipStream_t s, const ResizeOp *ops, int nops, int blocks,
                        const int2 *tab, const int *si, const float *alpha, const float *src,
                        float *dst) {
  hipLaunchKernelGGL(k_resize_rows, dim3(blocks), dim3(256), 0, s, ops, nops, tab, si, alpha,
                     src, dst);
}

void launch_resize_cols(hipStream_t s, const ResizeOp *ops, int nops, int blocks,
                        const int2 *tab, const int *si, const float *alpha, const float *src,
                        float *dst) {
  hipLaunchKernelGGL(k_resize_cols, dim3(blocks), dim3(256), 0, s, ops, nops, tab, si, alpha,
                     src, dst);
}

void launch_hog_grad(hipStream_t s, const GradImg *imgs, int nimg, int blocks, const float *img,
                     void *grid, long long grid_total, const unsigned char *bgr, int ws,
                     long long frame_bytes) {
  if (blocks > 0)
    hipLaunchKernelGGL(k_hog_grad, dim3(blocks), dim3(256), 0, s, imgs, nimg, img, bgr, ws,
                       frame_bytes, (float *)grid, (unsigned char *)grid + 4 * grid_total);
}

void launch_hog_votes(hipStream_t s, const HogLevel *lv, int nlv, int blocks, const void *grid,
                      long long grid_total, float *hist, float *norm) {
  if (blocks > 0)
    hipLaunchKernelGGL(k_hog_votes, dim3(blocks), dim3(kHogVoteThreads), 0, s, lv, nlv,
                       (const float *)grid, (const unsigned char *)grid + 4 * grid_total, hist,
                       norm);
}

void launch_hog_feat(hipStream_t s, const HogLevel *lv, int nlv, int blocks, const float *hist,
                     const float *norm, float *feat, void *featb, int hkind) {
  hipLaunchKernelGGL(k_hog_feat, dim3(blocks), dim3(256), 0, s, lv, nlv, hist, norm, feat,
                     (unsigned short *)featb, hkind);
}

void launch_conv_f32(hipStream_t s, int nf, const ConvTile *tiles, int ntiles, int fh, int fw,
                     const float *feat, const float *filt, const long long *map_off,
                     const int2 *map_dims, float *out) {
  switch (nf) {
    case 8:
      hipLaunchKernelGGL(k_conv_f32<8>, dim3(ntiles), dim3(256), 0, s, tiles, fh, fw, feat, filt,
                         map_off, map_dims, out);
      break;
    case 4:
      hipLaunchKernelGGL(k_conv_f32<4>, dim3(ntiles), dim3(256), 0, s, tiles, fh, fw, feat, filt,
                         map_off, map_dims, out);
      break;
    case 2:
      hipLaunchKernelGGL(k_conv_f32<2>, dim3(ntiles), dim3(256), 0, s, tiles, fh, fw, feat, filt,
                         map_off, map_dims, out);
      break;
    default:
      hipLaunchKernelGGL(k_conv_f32<1>, dim3(ntiles), dim3(256), 0, s, tiles, fh, fw, feat, filt,
                         map_off, map_dims, out);
      break;
  }
}

int launch_dt(hipStream_t s, const DtJob *jobs, int njobs, size_t lds, const int4 *nodes,
              const int *depths, const float *src, float *dst, int *ptr) {
  static const bool attr = hipFuncSetAttribute((const void *)k_dt,
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               160 * 1024) == hipSuccess;
  if (!attr && lds > 64 * 1024) return -1;
  hipLaunchKernelGGL(k_dt, dim3(njobs), dim3(256), lds, s, jobs, nodes, depths, src, dst, ptr);
  return 0;
}

int launch_dt_window(hipStream_t s, const DtJob *jobs, int njobs, size_t lds, const int4 *nodes,
                     const int *depths, const float *src, float *dst, int *ptr) {
  // the kernel's own static LDS counts against the same 160 KB
  static const bool attr = hipFuncSetAttribute((const void *)k_dt_window,
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               kDtWindowMaxLds) == hipSuccess;
  if (!attr && lds > 64 * 1024) return -1;
  hipLaunchKernelGGL(k_dt_window, dim3(njobs), dim3(256), lds, s, jobs, nodes, depths, src, dst,
                     ptr);
  return 0;
}

void launch_score(hipStream_t s, const ScoreMap *maps, int nmaps, const ScorePart *parts,
                  const float *conv, const float *M, float thresh, int half, float *score,
                  int *flags, int blocks) {
  if (blocks > 0)
    hipLaunchKernelGGL(k_score, dim3(blocks), dim3(256), 0, s, maps, nmaps, parts, conv, M,
                       thresh, half, score, flags);
}

int compact_blocks(long long n) { return (int)((n + kCompactBlock - 1) / kCompactBlock); }

// flagged indices of [0, n) in order to out, their count to *total (one workgroup at least,
// so that an empty input writes 0)
void launch_compact(hipStream_t s, const int *flags, long long n, int *blocksum, int *total,
                    int *out, int *host_total) {
  const int nb = compact_blocks(n) > 0 ? compact_blocks(n) : 1;
  hipLaunchKernelGGL(k_flag_count, dim3(nb), dim3(256), 0, s, flags, n, blocksum);
  hipLaunchKernelGGL(k_flag_scatter, dim3(nb), dim3(256), 0, s, flags, n, blocksum, total,
                     host_total, out);
}

// The first candidate of each frame of a batch (candidates ascend by flat score index, frame
// f's cells from f * score_one on) and the total at [nf]: one thread per frame.
__global__ void k_frame_offsets(const int *__restrict__ cand, const int *__restrict__ count,
                                long long score_one, int nf, int *__restrict__ out) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f > nf) return;
  const int n = *count;
  if (f == nf) {
    *(volatile int *)(out + f) = n;
    return;
  }
  const long long key = (long long)f * score_one;
  int lo = 0, hi = n;  // first position with cand >= key
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if ((long long)cand[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  *(volatile int *)(out + f) = lo;
}

void launch_frame_offsets(hipStream_t s, const int *cand, const int *count, long long score_one,
                          int nf, int *out) {
  hipLaunchKernelGGL(k_frame_offsets, dim3((nf + 1 + 255) / 256), dim3(256), 0, s, cand, count,
                     score_one, nf, out);
}

void launch_rows(hipStream_t s, const ScoreMap *maps, int nmaps, const ScorePart *parts,
                 const float *score, const int *cand, const int *sel, int ncand,
                 const int *tmpIx, const int *tmpIy, int half, int GL, int padx, int pady,
                 const int *rsize, float *rows) {
  if (ncand <= 0) return;
  hipLaunchKernelGGL(k_rows, dim3((ncand + 255) / 256), dim3(256), 0, s, maps, nmaps, parts,
                     score, cand, sel, ncand, tmpIx, tmpIy, half, GL, padx, pady, rsize, rows);
}

}  // namespace geye

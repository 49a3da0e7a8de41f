// dpm_mfma.hip -- BF16 / FP16 filter convolution as an implicit GEMM on gfx950 MFMA.
//
// C[n][x][y] = sum_{xp, yp, c} F[x+xp][y+yp][c] * W_n[c][xp][yp]   (CT/GPU_function.cu:302-358)
// with M = output positions, N = filters, K = taps x 32 channels (31 HOG + a zero pad).
// Features are channel-last 16-bit ([x][y][32], 64 bytes per cell), so one
// v_mfma_f32_16x16x32_{bf16,f16} consumes exactly one tap for 16 positions x 16 filters,
// with FP32 accumulation.
//
// Workgroup (4 waves) = 4 output columns x 64 rows x up to 48 filters.  The feature tile
// the workgroup needs ((4 + fw - 1) x (64 + fh - 1) cells) is staged into LDS once and read
// by every tap; the packed filters stream through LDS one tap column (xp) at a time,
// double-buffered.  Each wave owns one column: 4 M-tiles x NT N-tiles of accumulators.
// Operand layout (per lane l of the wave): A row l%16, channels 8*(l/16)..+8; B filter
// l%16, the same channels; accumulator rows 4*(l/16)..+4 of column l%16.
//
// Split forms (SPLIT = true): every operand x is the pair hi = round16(x), lo = round16(x -
// hi).  A cell then holds 64 channels ([hi 32][lo 32], 8 chunks of 16 bytes) and a filter
// tap 8 chunks per 16 filters; each tap runs three MFMAs into the same accumulator, hi*hi,
// hi*lo, lo*hi (lo*lo, below 2^-16 of the product, is dropped).  CPC below is the number of
// 16-byte chunks per cell or per filter tap: 4, or 8 when split.
#include "dpm_mfma.h"

#include <algorithm>

namespace geye {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(8))) _Float16 f16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;
typedef __attribute__((ext_vector_type(4))) unsigned int u32x4;

// 4 consecutive output rows, stored as one 16-byte write at 4-byte alignment
struct __attribute__((packed, aligned(4))) RowQuad {
  float v0, v1, v2, v3;
};

template <int KIND>
__device__ inline f32x4 mfma16(u32x4 a, u32x4 b, f32x4 c) {
  if constexpr (KIND == 1)
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a),
                                                  __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
}

// Feature tile in LDS: a ring of AX = 4 + fw - 1 column slots, feature column gx in slot
// gx mod AX, each slot AY = 64 + fh - 1 cells of 4 16-byte chunks.  The chunk index of cell c
// (c = slot * AY + row) is XOR-swizzled by (c >> 1) & 3, so the 8 consecutive cells that 8
// lanes read in one LDS cycle fall on 8 distinct 16-byte bank groups.
// With 8 chunks per cell (split forms) the chunk index is XOR-swizzled by cell & 7.
template <int CPC>
__device__ inline int conv_cell_chunk(int cell, int part) {
  if constexpr (CPC == 4) return cell * 4 + (part ^ ((cell >> 1) & 3));
  else return cell * 8 + (part ^ (cell & 7));
}

// Global chunk of feature column gx, chunk rem (= row * CPC + part) of the tile's rows; the
// load is unconditional from a clamped in-level address and `valid` says whether it is real.
template <int CPC>
__device__ inline u32x4 conv_feat_load(const u32x4 *__restrict__ F, const MfmaTile &t, int gx,
                                       int rem, bool &valid) {
  const int gy = t.y0 + rem / CPC;
  valid = gx < t.F1 && gy < t.F0;
  const int cx = min(gx, t.F1 - 1), cy = min(gy, t.F0 - 1);
  return F[(unsigned)((cx * t.F0 + cy) * CPC + (rem & (CPC - 1)))];  // a level holds < 2^31 chunks
}

// Feature arena and filter pack of a tile in 16-byte chunks (the tile's offsets count the
// 16-bit elements of the 32-channel forms; the split forms' cells and filter taps are twice
// as long, at twice the offsets).
template <int CPC>
__device__ inline const u32x4 *conv_feat_base(const u32x4 *feat, const MfmaTile &t) {
  return feat + t.featb_off * (CPC / 4) / 8;
}
template <int CPC>
__device__ inline const u32x4 *conv_filt_base(const u32x4 *filt, const MfmaTile &t) {
  return filt + t.filt_off * (CPC / 4) / 8;
}

// Registers per thread: PN chunks of the next tile's 4 new feature columns, PRE chunks of the
// next tap column's filters (NTH threads per workgroup).
template <int NT, bool SPLIT, int NTH, int TW> struct ConvRegs {
  // TW new columns of up to 79 cells (filter height <= 16)
  static constexpr int PN = (TW * 79 * (SPLIT ? 8 : 4) + NTH - 1) / NTH;
  static constexpr int PRE = SPLIT && NTH == 256 ? (NT == 3 ? 9 : 8) : (NT == 3 ? 5 : 4);
};

// Stages all AX feature columns of tile t into the ring (8 loads in flight per thread; a
// column holds at least NTH chunks -- AY >= 64 cells of CPC chunks, NTH <= 64 CPC -- so one
// step of NTH crosses at most one column boundary).
template <int CPC, int NTH>
__device__ inline void conv_stage_tile(u32x4 *As, const u32x4 *__restrict__ feat,
                                       const MfmaTile &t, int AX, int AY, int tid) {
  const int CS = AY * CPC, NA = AX * CS;
  const u32x4 *F = conv_feat_base<CPC>(feat, t);
  int c = 0, rem = tid, slot = t.x0 % AX;
  for (int i0 = 0; i0 < NA; i0 += 8 * NTH) {
    u32x4 v[8];
    bool ok[8];
    int cc[8], rr[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      cc[u] = slot;
      rr[u] = rem;
      v[u] = conv_feat_load<CPC>(F, t, t.x0 + c, rem, ok[u]);
      rem += NTH;
      if (rem >= CS) {
        rem -= CS;
        c++;
        slot = slot + 1 == AX ? 0 : slot + 1;
      }
    }
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (i0 + u * NTH + tid < NA)
        As[conv_cell_chunk<CPC>(cc[u] * AY + rr[u] / CPC, rr[u] & (CPC - 1))] =
            ok[u] ? v[u] : (u32x4){0u, 0u, 0u, 0u};
  }
}

// Loads the TW feature columns tile tn adds to its predecessor (tn's relative columns
// AX-TW..AX-1) into registers.  live == false still issues the loads, all from one cell, so
// that the code around them stays free of branches and the wait counts stay exact.
template <int CPC, int PN, int NTH, int TW>
__device__ inline void conv_fetch_new(u32x4 (&nf)[PN], unsigned &vmask,
                                      const u32x4 *__restrict__ feat, const MfmaTile &tn,
                                      bool live, int AX, int CS, int tid) {
  const u32x4 *F = conv_feat_base<CPC>(feat, tn);
  int c = 0, rem = tid;
  vmask = 0;
#pragma unroll
  for (int k = 0; k < PN; k++) {
    bool v;
    nf[k] = conv_feat_load<CPC>(F, tn, live ? tn.x0 + AX - TW + c : tn.x0, live ? rem : 0, v);
    vmask |= (unsigned)(v && live) << k;
    rem += NTH;
    if (rem >= CS) { rem -= CS; c++; }
  }
}

// Writes those columns into the slots of the current tile's relative columns 0..TW-1 (sx is
// the slot of its relative column 0): feature column x0 + AX + k lives in slot (x0 + k) mod AX.
template <int CPC, int PN, int NTH, int TW>
__device__ inline void conv_put_new(u32x4 *As, const u32x4 (&nf)[PN], unsigned vmask, int sx,
                                    int AX, int AY, int tid) {
  const int CS = AY * CPC;
  int c = 0, rem = tid;
#pragma unroll
  for (int k = 0; k < PN; k++) {
    if (k * NTH + tid < TW * CS) {
      int s2 = sx + c;
      if (s2 >= AX) s2 -= AX;
      As[conv_cell_chunk<CPC>(s2 * AY + rem / CPC, rem & (CPC - 1))] =
          ((vmask >> k) & 1) ? nf[k] : (u32x4){0u, 0u, 0u, 0u};
    }
    rem += NTH;
    if (rem >= CS) { rem -= CS; c++; }
  }
}

// One tap column for one wave: fh taps x MT M-tiles x NT N-tiles.  Bcol is the column's
// filters in LDS ([fh][NT][CPC chunks][16 filters]); cell0 the ring cell of the wave's first
// row.  Two register sets: tap yp + 1's fragments are read while tap yp's MFMAs run (the
// scheduling barriers keep the compiler from sinking the reads next to their use); the reads
// are unconditional (the last ones repeat tap fh - 1) so the wait counts are fixed.  Split:
// fragments [0, NT) / [0, MT) are the hi halves, [NT, 2NT) / [MT, 2MT) the lo halves; the
// three products of a tap go to every accumulator in the order hi*hi, hi*lo, lo*hi, each
// product over all MT x NT accumulators before the next (independent MFMAs between
// dependent ones).
template <int KIND, int NT, bool SPLIT, int MT>
__device__ inline void conv_column(f32x4 (&acc)[MT][NT], const u32x4 *Bcol, const u32x4 *As,
                                   int cell0, int fh, int r, int chunk) {
  constexpr int CPC = SPLIT ? 8 : 4, NH = SPLIT ? 2 : 1;
  u32x4 b0[NH * NT], a0[NH * MT], b1[NH * NT], a1[NH * MT];
  auto load = [&](u32x4 *b, u32x4 *a, int yp) {
#pragma unroll
    for (int h = 0; h < NH; h++)
#pragma unroll
      for (int nt = 0; nt < NT; nt++)
        b[h * NT + nt] = Bcol[((yp * NT + nt) * CPC + 4 * h + chunk) * 16 + r];
#pragma unroll
    for (int h = 0; h < NH; h++)
#pragma unroll
      for (int mt = 0; mt < MT; mt++)
        a[h * MT + mt] = As[conv_cell_chunk<CPC>(cell0 + mt * 16 + yp, 4 * h + chunk)];
  };
  auto mma = [&](const u32x4 *b, const u32x4 *a) {
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
      for (int nt = 0; nt < NT; nt++) acc[mt][nt] = mfma16<KIND>(a[mt], b[nt], acc[mt][nt]);
    if constexpr (SPLIT) {
#pragma unroll
      for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int nt = 0; nt < NT; nt++) acc[mt][nt] = mfma16<KIND>(a[mt], b[NT + nt], acc[mt][nt]);
#pragma unroll
      for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int nt = 0; nt < NT; nt++) acc[mt][nt] = mfma16<KIND>(a[MT + mt], b[nt], acc[mt][nt]);
    }
  };
  load(b0, a0, 0);
  for (int yp = 0; yp < fh; yp += 2) {
    load(b1, a1, min(yp + 1, fh - 1));
    __builtin_amdgcn_sched_barrier(0);
    mma(b0, a0);
    __builtin_amdgcn_sched_barrier(0);
    load(b0, a0, min(yp + 2, fh - 1));
    __builtin_amdgcn_sched_barrier(0);
    if (yp + 1 < fh) mma(b1, a1);
    __builtin_amdgcn_sched_barrier(0);
  }
}

// Writes one wave's outputs (column x, rows y0 .. y0 + 16 MT - 1, the group's filters).  A
// lane holds 4 consecutive rows of one filter's map: one 16-byte store (4-byte aligned, which
// gfx950 global stores accept) unless the map ends inside them.
template <int NT, int MT>
__device__ inline void conv_store(const f32x4 (&acc)[MT][NT], const MfmaTile &t, int x, int y0,
                                  int r, int chunk, const long long *__restrict__ map_off,
                                  const int2 *__restrict__ map_dims, float *__restrict__ out) {
#pragma unroll
  for (int nt = 0; nt < NT; nt++) {
    const int n = nt * 16 + r;
    if (n >= t.nvalid) continue;
    const int slot = t.out_slot + n;
    const int2 hw = map_dims[slot];
    if (x >= hw.y) continue;
    float *o = out + map_off[slot] + (long long)x * hw.x;
#pragma unroll
    for (int mt = 0; mt < MT; mt++) {
      const int y = y0 + mt * 16 + chunk * 4;
      if (y + 3 < hw.x) {
        *(RowQuad *)(o + y) = RowQuad{acc[mt][nt][0], acc[mt][nt][1], acc[mt][nt][2],
                                      acc[mt][nt][3]};
      } else {
#pragma unroll
        for (int i = 0; i < 4; i++)
          if (y + i < hw.x) o[y + i] = acc[mt][nt][i];
      }
    }
  }
}

// Per-class kernel, persistent: workgroup b of G takes the contiguous tile range
// [b*n/G, (b+1)*n/G) of a list of TW-column tiles ordered along x within (level, group, row of
// tiles).  When the next tile is the x + TW neighbour, its TW new feature columns are loaded
// into registers during this tile's first tap column and written into the ring slots this
// tile has finished with (relative columns 0..TW-1, last read at tap column
// min(TW - 1, fw - 1)); the next tile's first filter column is loaded by this tile's last tap
// column.  A tile that does not continue its predecessor stages its whole feature tile
// first.  WPC waves share each of the TW output columns: wave w takes column w % TW and rows
// 64 (w / TW) / WPC .. + 64 / WPC, i.e. MT = 4 / WPC M-tiles x NT N-tiles of accumulators.
// The split forms' LDS tile leaves room for one workgroup per CU, which runs 8 waves: 8
// columns (TW = 8, half the filter traffic and LDS filter reads per output) where that tile
// fits the LDS, else 4 columns of two waves each (WPC = 2).  The per-output order of the
// products is the same for every TW and WPC.  nbuf = 2 double-buffers the filter columns;
// nbuf = 1 (a class whose double buffer would not fit the LDS) writes the next column after a
// barrier that ends this column's reads.
template <int KIND, int NT, bool SPLIT, int TW, int WPC>
__global__ __launch_bounds__(64 * TW * WPC)
__attribute__((amdgpu_waves_per_eu(2, 2))) void k_conv_mfma(
    const MfmaTile *__restrict__ tiles, int ntiles, int fh, int fw, int nbuf,
    const u32x4 *__restrict__ feat, const u32x4 *__restrict__ filt,
    const long long *__restrict__ map_off, const int2 *__restrict__ map_dims,
    float *__restrict__ out) {
  extern __shared__ u32x4 lds[];
  constexpr int NTH = 64 * TW * WPC, MT = 4 / WPC;
  constexpr int PN = ConvRegs<NT, SPLIT, NTH, TW>::PN, PRE = ConvRegs<NT, SPLIT, NTH, TW>::PRE;
  constexpr int CPC = SPLIT ? 8 : 4;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wcol = wave % TW, wrow = (wave / TW) * 16 * MT;  // the wave's column and first row
  const int AX = TW + fw - 1, AY = 64 + fh - 1, CS = AY * CPC, NA = AX * CS;
  u32x4 *As = lds;                               // [AX slots][AY][CPC chunks], swizzled
  u32x4 *Bs = lds + NA;                          // [nbuf][fh][NT][CPC][16]
  const int bstage = fh * NT * CPC * 16;         // 16-byte chunks per tap column
  const bool in_regs = bstage <= PRE * NTH;
  const bool ring = TW * CS <= PN * NTH;
  const int bmask = nbuf - 1;

  int bx = blockIdx.x;
  if ((int)gridDim.x == ntiles) {
    // one tile per workgroup: runs of 16 consecutive tiles (neighbouring columns, sharing
    // feature columns) go to one XCD (workgroup b runs on XCD b % 8)
    constexpr int X = 8, C = 16, GRP = X * C;
    if (bx < (ntiles / GRP) * GRP) {
      const int g = bx / GRP, l = bx - g * GRP;
      bx = g * GRP + (l % X) * C + l / X;
    }
  }
  const int tb = (int)((long long)bx * ntiles / gridDim.x);
  const int te = (int)((long long)(bx + 1) * ntiles / gridDim.x);
  if (tb >= te) return;
  const int r = lane & 15, chunk = lane >> 4;

  MfmaTile t = tiles[tb];
  bool have_feat = false, have_filt = false;
  int bpar = 0;                                  // filter buffer of this tile's tap column 0
  for (int ti = tb; ti < te; ti++) {
    const bool has_next = ti + 1 < te;
    const MfmaTile tn = tiles[has_next ? ti + 1 : ti];
    const bool cont = has_next && ring && tn.featb_off == t.featb_off && tn.y0 == t.y0 &&
                      tn.x0 == t.x0 + TW;
    const u32x4 *Wg = conv_filt_base<CPC>(filt, t);
    const int sx = t.x0 % AX;                  // slot of the tile's relative column 0
    // the previous tile ended with a barrier after its last reads of As
    if (!have_feat) conv_stage_tile<CPC, NTH>(As, feat, t, AX, AY, tid);
    if (!have_filt)
      for (int i = tid; i < bstage; i += NTH) Bs[bpar * bstage + i] = Wg[i];
    if (!have_feat || !have_filt) __syncthreads();

    f32x4 acc[MT][NT];
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
      for (int nt = 0; nt < NT; nt++) acc[mt][nt] = (f32x4){0.f, 0.f, 0.f, 0.f};

    u32x4 nf[PN];
    unsigned vmask = 0;
    auto step = [&](int xp, bool first) {
      const int buf = (bpar + xp) & bmask;
      // the next filter column (this tile's xp + 1, or the next tile's column 0) is loaded into
      // registers before this column's MFMAs and written to LDS after them (the register copy
      // runs even with nothing to copy -- last column of the last tile -- into the idle
      // buffer, and alongside the plain copy, writing the same values, when a column exceeds
      // the registers: branch-free, so the wait counts stay exact)
      const bool more = xp + 1 < fw;
      const u32x4 *src = more ? Wg + (long long)(xp + 1) * bstage : conv_filt_base<CPC>(filt, tn);
      u32x4 *dst = Bs + ((buf ^ 1) & bmask) * bstage;
      u32x4 pre[PRE];
#pragma unroll
      for (int c = 0; c < PRE; c++) pre[c] = src[(unsigned)min(tid + c * NTH, bstage - 1)];
      if (nbuf == 2 && !in_regs && (more || has_next))
        for (int i = tid; i < bstage; i += NTH) dst[i] = src[i];
      // issued after the filter loads, the next tile's feature loads may stay in flight past
      // this column's wait for the filters
      if (first) conv_fetch_new<CPC, PN, NTH, TW>(nf, vmask, feat, tn, cont, AX, CS, tid);
      int slot = sx + wcol + xp;
      if (slot >= AX) slot -= AX;
      conv_column<KIND, NT, SPLIT, MT>(acc, Bs + buf * bstage, As, slot * AY + wrow + r, fh, r,
                                       chunk);
      if (nbuf == 1) {
        __syncthreads();                       // every wave's reads of the one buffer done
        if (!in_regs && (more || has_next))
          for (int i = tid; i < bstage; i += NTH) dst[i] = src[i];
      }
#pragma unroll
      for (int c = 0; c < PRE; c++) dst[(unsigned)min(tid + c * NTH, bstage - 1)] = pre[c];
      __syncthreads();
    };
    // tap column xp reads relative columns xp..xp+TW-1, so from tap column TW on (or after the
    // last one) the slots of relative columns 0..TW-1 take the next tile's new columns.  With
    // 8-column tiles (fw <= 8, so always after the last one) the new columns are loaded only
    // then, when the tap columns' registers are free.
    step(0, TW == 4);
    for (int xp = 1; xp < fw; xp++) {
      if (TW == 4 && xp == TW && cont)
        conv_put_new<CPC, PN, NTH, TW>(As, nf, vmask, sx, AX, AY, tid);
      step(xp, false);
    }
    if (fw <= TW && cont) {
      if constexpr (TW == 8) conv_fetch_new<CPC, PN, NTH, TW>(nf, vmask, feat, tn, cont, AX, CS, tid);
      conv_put_new<CPC, PN, NTH, TW>(As, nf, vmask, sx, AX, AY, tid);
      __syncthreads();
    }
    conv_store<NT, MT>(acc, t, t.x0 + wcol, t.y0 + wrow, r, chunk, map_off, map_dims, out);
    have_feat = cont;
    have_filt = has_next;
    bpar = (bpar + fw) & bmask;
    t = tn;
  }
}

constexpr int kQueueMaxLds = 160 * 1024 - 1024;  // dynamic LDS cap of the queue kernel

// Work-queue form.  A persistent workgroup takes tile indices from *counter until the list
// is exhausted (every workgroup reaches that exit).  The tile carries its filter class's
// shape (fh, fw) and its number of 16-filter tiles nt (1..3), so one launch covers every
// level of every class; per output the taps are accumulated in the same order as the
// per-class kernel (split: hi*hi, hi*lo, lo*hi per tap), so both give the same bits.  Waves
// own one column of 16*MT rows each.
template <int KIND, int MT, bool SPLIT>
__global__ __launch_bounds__(256) void k_conv_mfma_q(const MfmaTile *__restrict__ tiles,
                                                     int ntiles, int *__restrict__ counter,
                                                     const u32x4 *__restrict__ feat,
                                                     const u32x4 *__restrict__ filt,
                                                     const long long *__restrict__ map_off,
                                                     const int2 *__restrict__ map_dims,
                                                     float *__restrict__ out) {
  extern __shared__ u32x4 lds[];
  __shared__ int s_tile;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r = lane & 15, chunk = lane >> 4;
  for (;;) {
    if (tid == 0) s_tile = atomicAdd(counter, 1);
    __syncthreads();
    const int ti = s_tile;
    __syncthreads();
    if (ti >= ntiles) break;
    const MfmaTile t = tiles[ti];
    constexpr int CPC = SPLIT ? 8 : 4;
    const int fh = t.fh, fw = t.fw, nt = t.nt;
    const int AX = 4 + fw - 1, AY = 16 * MT + fh - 1;
    u32x4 *As = lds;                               // [AX][AY][CPC chunks]
    u32x4 *Bs = lds + AX * AY * CPC;               // [2][fh][nt][CPC][16]
    const int bstage = fh * nt * CPC * 16;
    const u32x4 *F = conv_feat_base<CPC>(feat, t);
    const u32x4 *Wg = conv_filt_base<CPC>(filt, t);
    for (int i = tid; i < AX * AY * CPC; i += 256) {
      const int cell = i / CPC, part = i & (CPC - 1);
      const int ax = cell / AY, ay = cell - ax * AY;
      const int gx = t.x0 + ax, gy = t.y0 + ay;
      u32x4 v = {0u, 0u, 0u, 0u};
      if (gx < t.F1 && gy < t.F0) v = F[((long long)gx * t.F0 + gy) * CPC + part];
      As[i] = v;
    }
    for (int i = tid; i < bstage; i += 256) Bs[i] = Wg[i];
    __syncthreads();
    f32x4 acc[MT][3];
#pragma unroll
    for (int mt = 0; mt < MT; mt++)
#pragma unroll
      for (int n = 0; n < 3; n++) acc[mt][n] = (f32x4){0.f, 0.f, 0.f, 0.f};
    for (int xp = 0; xp < fw; xp++) {
      const int buf = xp & 1;
      if (xp + 1 < fw) {
        const u32x4 *src = Wg + (long long)(xp + 1) * bstage;
        u32x4 *dst = Bs + (buf ^ 1) * bstage;
        for (int i = tid; i < bstage; i += 256) dst[i] = src[i];
      }
      const u32x4 *Bb = Bs + buf * bstage;
      const u32x4 *Ab = As + ((wave + xp) * AY) * CPC + chunk;
      for (int yp = 0; yp < fh; yp++) {
        u32x4 b[2][3];
#pragma unroll
        for (int h = 0; h < (SPLIT ? 2 : 1); h++) {
          b[h][0] = Bb[((yp * nt + 0) * CPC + 4 * h + chunk) * 16 + r];
          b[h][1] = b[h][2] = b[h][0];
          if (nt > 1) b[h][1] = Bb[((yp * nt + 1) * CPC + 4 * h + chunk) * 16 + r];
          if (nt > 2) b[h][2] = Bb[((yp * nt + 2) * CPC + 4 * h + chunk) * 16 + r];
        }
#pragma unroll
        for (int mt = 0; mt < MT; mt++) {
          // the cell's hi chunk (its lo chunk lies 4 further): this kernel's ring is a plain
          // [AX][AY][CPC] array
          const u32x4 a = Ab[(mt * 16 + r + yp) * CPC];
#pragma unroll
          for (int n = 0; n < 3; n++)
            if (n < nt) acc[mt][n] = mfma16<KIND>(a, b[0][n], acc[mt][n]);
          if constexpr (SPLIT) {
            const u32x4 al = Ab[(mt * 16 + r + yp) * CPC + 4];
#pragma unroll
            for (int n = 0; n < 3; n++)
              if (n < nt) acc[mt][n] = mfma16<KIND>(a, b[1][n], acc[mt][n]);
#pragma unroll
            for (int n = 0; n < 3; n++)
              if (n < nt) acc[mt][n] = mfma16<KIND>(al, b[0][n], acc[mt][n]);
          }
        }
      }
      __syncthreads();
    }
    const int x = t.x0 + wave;
#pragma unroll
    for (int n = 0; n < 3; n++) {
      if (n >= nt) break;
      const int f = n * 16 + r;
      if (f >= t.nvalid) continue;
      const int slot = t.out_slot + f;
      const int2 hw = map_dims[slot];
      if (x >= hw.y) continue;
      float *o = out + map_off[slot] + (long long)x * hw.x;
#pragma unroll
      for (int mt = 0; mt < MT; mt++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int y = t.y0 + mt * 16 + chunk * 4 + i;
          if (y < hw.x) o[y] = acc[mt][n][i];
        }
    }
  }
}

size_t mfma_q_lds_bytes(int fh, int fw, int nt, int MT, bool split) {
  const size_t cpc = split ? 8 : 4;
  return (size_t)16 * cpc * ((size_t)(4 + fw - 1) * (16 * MT + fh - 1) + 2 * (size_t)fh * nt * 16);
}

template <int KIND, int MT, bool SPLIT>
static int launch_q(hipStream_t s, int blocks, size_t lds, const MfmaTile *tiles, int ntiles,
                    int *counter, const void *feat, const void *filt, const long long *map_off,
                    const int2 *map_dims, float *out) {
  // the kernel's static LDS (the tile index) comes out of the same 160 KB
  static const bool attr = hipFuncSetAttribute((const void *)k_conv_mfma_q<KIND, MT, SPLIT>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               kQueueMaxLds) == hipSuccess;
  if (!attr) return -1;
  hipLaunchKernelGGL((k_conv_mfma_q<KIND, MT, SPLIT>), dim3(blocks), dim3(256), lds, s, tiles,
                     ntiles, counter, (const u32x4 *)feat, (const u32x4 *)filt, map_off, map_dims,
                     out);
  return 0;
}

// kind: 1 = BF16, 2 = FP16; SPLIT: the three-product form on hi/lo operand pairs.
template <bool SPLIT>
static int launch_q_kind(hipStream_t s, int kind, int MT, int blocks, size_t lds,
                         const MfmaTile *tiles, int ntiles, int *counter, const void *feat,
                         const void *filt, const long long *map_off, const int2 *map_dims,
                         float *out) {
  if (kind == 1)
    return MT == 2 ? launch_q<1, 2, SPLIT>(s, blocks, lds, tiles, ntiles, counter, feat, filt, map_off, map_dims, out)
                   : launch_q<1, 4, SPLIT>(s, blocks, lds, tiles, ntiles, counter, feat, filt, map_off, map_dims, out);
  return MT == 2 ? launch_q<2, 2, SPLIT>(s, blocks, lds, tiles, ntiles, counter, feat, filt, map_off, map_dims, out)
                 : launch_q<2, 4, SPLIT>(s, blocks, lds, tiles, ntiles, counter, feat, filt, map_off, map_dims, out);
}

int launch_conv_mfma_queue(hipStream_t s, int kind, bool split, int MT, int blocks, size_t lds,
                           const MfmaTile *tiles, int ntiles, int *counter, const void *feat,
                           const void *filt, const long long *map_off, const int2 *map_dims,
                           float *out) {
  if (ntiles <= 0) return 0;
  if (lds > (size_t)kQueueMaxLds || blocks <= 0 || (kind != 1 && kind != 2)) return -1;
  return split ? launch_q_kind<true>(s, kind, MT, blocks, lds, tiles, ntiles, counter, feat, filt, map_off, map_dims, out)
               : launch_q_kind<false>(s, kind, MT, blocks, lds, tiles, ntiles, counter, feat, filt, map_off, map_dims, out);
}

size_t mfma_lds_bytes(int fh, int fw, int nt, bool split, int nbuf, int tw) {
  const size_t cpc = split ? 8 : 4;
  return (size_t)16 * cpc * ((size_t)(tw + fw - 1) * (64 + fh - 1) + (size_t)nbuf * fh * nt * 16);
}

// Filter buffers of the per-class kernel at tile width tw: 2, or 1 for a split class whose
// double buffer would not fit the 160 KB of LDS; 0 = unsupported (a split class whose 64-row
// feature tile and one tap column of filters exceed the LDS, e.g. 12x12 filters in groups of
// 48).
static int mfma_nbuf(int fh, int fw, int nt, bool split, int tw) {
  if (mfma_lds_bytes(fh, fw, nt, split, 2, tw) <= 160 * 1024) return 2;
  if (split && mfma_lds_bytes(fh, fw, nt, split, 1, tw) <= 160 * 1024) return 1;
  return 0;
}

bool mfma_wide_tiles(int fh, int fw, int nt, bool split) {
  return split && fw <= 8 && mfma_nbuf(fh, fw, nt, split, 8) > 0;
}

// Launches the persistent kernel with every workgroup resident at once (its tile ranges are
// static).  With fewer than 4 tiles per resident workgroup a static split would leave the
// last round uneven, so each tile gets its own workgroup and the hardware balances them.
template <int KIND, int NT, bool SPLIT, int TW>
static int launch_one(hipStream_t s, const MfmaTile *tiles, int ntiles, int fh, int fw,
                      const void *feat, const void *filt, const long long *map_off,
                      const int2 *map_dims, float *out, int eighths) {
  constexpr int WPC = SPLIT && TW == 4 ? 2 : 1;  // waves per output column
  const void *kern = (const void *)k_conv_mfma<KIND, NT, SPLIT, TW, WPC>;
  const int nbuf = mfma_nbuf(fh, fw, NT, SPLIT, TW);
  if (nbuf == 0) return -1;
  const size_t lds = mfma_lds_bytes(fh, fw, NT, SPLIT, nbuf, TW);
  // once per process (function-local statics are initialised once, also when serving
  // workers launch concurrently): the LDS attribute and the CU count, 0 on failure
  static const int ncu = [kern] {
    int dev = 0, n = 0;
    if (hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) !=
            hipSuccess ||
        hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      return 0;
    return n;
  }();
  if (ncu <= 0) return -1;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * TW * WPC, lds) !=
          hipSuccess ||
      per_cu < 1)
    return -1;
  const int full = std::max(1, per_cu * ncu * std::min(std::max(eighths, 1), 8) / 8);
  const int blocks = ntiles <= 4 * full ? ntiles : full;
  hipLaunchKernelGGL((k_conv_mfma<KIND, NT, SPLIT, TW, WPC>), dim3(blocks), dim3(64 * TW * WPC),
                     lds, s, tiles, ntiles, fh, fw, nbuf, (const u32x4 *)feat,
                     (const u32x4 *)filt, map_off, map_dims, out);
  return 0;
}

template <bool SPLIT, int TW>
static int launch_kind(hipStream_t s, int kind, int nt, const MfmaTile *tiles, int ntiles, int fh,
                       int fw, const void *feat, const void *filt, const long long *map_off,
                       const int2 *map_dims, float *out, int mb) {
  if (kind == 1) {
    if (nt == 3) return launch_one<1, 3, SPLIT, TW>(s, tiles, ntiles, fh, fw, feat, filt, map_off, map_dims, out, mb);
    if (nt == 2) return launch_one<1, 2, SPLIT, TW>(s, tiles, ntiles, fh, fw, feat, filt, map_off, map_dims, out, mb);
    return launch_one<1, 1, SPLIT, TW>(s, tiles, ntiles, fh, fw, feat, filt, map_off, map_dims, out, mb);
  }
  if (nt == 3) return launch_one<2, 3, SPLIT, TW>(s, tiles, ntiles, fh, fw, feat, filt, map_off, map_dims, out, mb);
  if (nt == 2) return launch_one<2, 2, SPLIT, TW>(s, tiles, ntiles, fh, fw, feat, filt, map_off, map_dims, out, mb);
  return launch_one<2, 1, SPLIT, TW>(s, tiles, ntiles, fh, fw, feat, filt, map_off, map_dims, out, mb);
}

int launch_conv_mfma(hipStream_t s, int kind, bool split, int nt, const MfmaTile *tiles4,
                     int ntiles4, const MfmaTile *tiles8, int ntiles8, int fh, int fw,
                     const void *feat, const void *filt, const long long *map_off,
                     const int2 *map_dims, float *out, int eighths) {
  if (ntiles4 <= 0) return 0;
  if ((kind != 1 && kind != 2) || nt < 1 || nt > 3) return -1;
  if (tiles8 && ntiles8 > 0 && mfma_wide_tiles(fh, fw, nt, split))
    return launch_kind<true, 8>(s, kind, nt, tiles8, ntiles8, fh, fw, feat, filt, map_off,
                                map_dims, out, eighths);
  if (mfma_nbuf(fh, fw, nt, split, 4) == 0) return -1;
  return split ? launch_kind<true, 4>(s, kind, nt, tiles4, ntiles4, fh, fw, feat, filt, map_off, map_dims, out, eighths)
               : launch_kind<false, 4>(s, kind, nt, tiles4, ntiles4, fh, fw, feat, filt, map_off, map_dims, out, eighths);
}

}  // namespace geye

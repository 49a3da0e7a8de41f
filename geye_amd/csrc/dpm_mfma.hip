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
#include "dpm_mfma.h"

namespace geye {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(8))) _Float16 f16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;
typedef __attribute__((ext_vector_type(4))) unsigned int u32x4;

template <int KIND>
__device__ inline f32x4 mfma16(u32x4 a, u32x4 b, f32x4 c) {
  if constexpr (KIND == 1)
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a),
                                                  __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
}

template <int KIND, int NT>
__global__ __launch_bounds__(256) void k_conv_mfma(const MfmaTile *__restrict__ tiles, int fh,
                                                   int fw, const u32x4 *__restrict__ feat,
                                                   const u32x4 *__restrict__ filt,
                                                   const long long *__restrict__ map_off,
                                                   const int2 *__restrict__ map_dims,
                                                   float *__restrict__ out) {
  extern __shared__ u32x4 lds[];
  // runs of 16 consecutive tiles (neighbouring columns of one level) per XCD
  constexpr int X = 8, C = 16, GRP = X * C;
  int bx = blockIdx.x;
  if (bx < ((int)gridDim.x / GRP) * GRP) {
    const int g = bx / GRP, l = bx - g * GRP;
    bx = g * GRP + (l % X) * C + l / X;
  }
  const MfmaTile t = tiles[bx];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int AX = 4 + fw - 1, AY = 64 + fh - 1;
  u32x4 *As = lds;                               // [AX][AY][4 chunks]
  u32x4 *Bs = lds + AX * AY * 4;                 // [2][fh][NT][4][16]
  const int bstage = fh * NT * 64;               // 16-byte chunks per tap column
  const u32x4 *F = feat + t.featb_off / 8;       // featb_off counts 16-bit elements
  const u32x4 *Wg = filt + t.filt_off / 8;

  // stage the feature tile (zero outside the level), 4 loads in flight per thread
  for (int i0 = tid; i0 < AX * AY * 4; i0 += 4 * 256) {
    u32x4 v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int i = i0 + u * 256;
      const int cell = i >> 2, part = i & 3;
      const int ax = cell / AY, ay = cell - ax * AY;
      const int gx = t.x0 + ax, gy = t.y0 + ay;
      v[u] = (u32x4){0u, 0u, 0u, 0u};
      if (i < AX * AY * 4 && gx < t.F1 && gy < t.F0)
        v[u] = F[((long long)gx * t.F0 + gy) * 4 + part];
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (i0 + u * 256 < AX * AY * 4) As[i0 + u * 256] = v[u];
  }
  for (int i = tid; i < bstage; i += 256) Bs[i] = Wg[i];
  __syncthreads();

  f32x4 acc[4][NT];
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int nt = 0; nt < NT; nt++) acc[mt][nt] = (f32x4){0.f, 0.f, 0.f, 0.f};

  const int r = lane & 15, chunk = lane >> 4;
  // the next tap column's filters are loaded into registers before this column's MFMAs and
  // written to LDS after them, so their latency hides behind the arithmetic
  constexpr int PRE = 8;
  const bool in_regs = bstage <= PRE * 256;
  for (int xp = 0; xp < fw; xp++) {
    const int buf = xp & 1;
    u32x4 pre[PRE];
    const bool next = xp + 1 < fw;
    const u32x4 *src = Wg + (long long)(xp + 1) * bstage;
    u32x4 *dst = Bs + (buf ^ 1) * bstage;
    if (next) {
      if (in_regs) {
#pragma unroll
        for (int c = 0; c < PRE; c++) {
          const int i = tid + c * 256;
          if (i < bstage) pre[c] = src[i];
        }
      } else {
        for (int i = tid; i < bstage; i += 256) dst[i] = src[i];
      }
    }
    const u32x4 *Bb = Bs + buf * bstage;
    const u32x4 *Ab = As + ((wave + xp) * AY) * 4 + chunk;
    // two register sets: tap yp + 1's fragments are read while tap yp's MFMAs run (the
    // scheduling barriers keep the compiler from sinking the reads next to their use)
    u32x4 b0[NT], a0[4], b1[NT], a1[4];
    auto load = [&](u32x4 *b, u32x4 *a, int yp) {
#pragma unroll
      for (int nt = 0; nt < NT; nt++) b[nt] = Bb[((yp * NT + nt) * 4 + chunk) * 16 + r];
#pragma unroll
      for (int mt = 0; mt < 4; mt++) a[mt] = Ab[(mt * 16 + r + yp) * 4];
    };
    auto mma = [&](const u32x4 *b, const u32x4 *a) {
#pragma unroll
      for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int nt = 0; nt < NT; nt++) acc[mt][nt] = mfma16<KIND>(a[mt], b[nt], acc[mt][nt]);
    };
    // the reads are unconditional (the last ones repeat tap fh - 1) so that the wait
    // counts before each set's MFMAs are fixed
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
    if (next && in_regs) {
#pragma unroll
      for (int c = 0; c < PRE; c++) {
        const int i = tid + c * 256;
        if (i < bstage) dst[i] = pre[c];
      }
    }
    __syncthreads();
  }

  const int x = t.x0 + wave;
#pragma unroll
  for (int nt = 0; nt < NT; nt++) {
    const int n = nt * 16 + r;
    if (n >= t.nvalid) continue;
    const int slot = t.out_slot + n;
    const int2 hw = map_dims[slot];
    if (x >= hw.y) continue;
    float *o = out + map_off[slot] + (long long)x * hw.x;
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int y = t.y0 + mt * 16 + chunk * 4 + i;
        if (y < hw.x) o[y] = acc[mt][nt][i];
      }
  }
}

constexpr int kQueueMaxLds = 160 * 1024 - 1024;  // dynamic LDS cap of the queue kernel

// Work-queue form.  A persistent workgroup takes tile indices from *counter until the list
// is exhausted (every workgroup reaches that exit).  The tile carries its filter class's
// shape (fh, fw) and its number of 16-filter tiles nt (1..3), so one launch covers every
// level of every class; per output the taps are accumulated in the same order as the
// per-class kernel, so both give the same bits.  Waves own one column of 16*MT rows each.
template <int KIND, int MT>
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
    const int fh = t.fh, fw = t.fw, nt = t.nt;
    const int AX = 4 + fw - 1, AY = 16 * MT + fh - 1;
    u32x4 *As = lds;                               // [AX][AY][4 chunks]
    u32x4 *Bs = lds + AX * AY * 4;                 // [2][fh][nt][4][16]
    const int bstage = fh * nt * 64;
    const u32x4 *F = feat + t.featb_off / 8;
    const u32x4 *Wg = filt + t.filt_off / 8;
    for (int i = tid; i < AX * AY * 4; i += 256) {
      const int cell = i >> 2, part = i & 3;
      const int ax = cell / AY, ay = cell - ax * AY;
      const int gx = t.x0 + ax, gy = t.y0 + ay;
      u32x4 v = {0u, 0u, 0u, 0u};
      if (gx < t.F1 && gy < t.F0) v = F[((long long)gx * t.F0 + gy) * 4 + part];
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
      const u32x4 *Ab = As + ((wave + xp) * AY) * 4 + chunk;
      for (int yp = 0; yp < fh; yp++) {
        u32x4 b0 = Bb[((yp * nt + 0) * 4 + chunk) * 16 + r], b1 = b0, b2 = b0;
        if (nt > 1) b1 = Bb[((yp * nt + 1) * 4 + chunk) * 16 + r];
        if (nt > 2) b2 = Bb[((yp * nt + 2) * 4 + chunk) * 16 + r];
#pragma unroll
        for (int mt = 0; mt < MT; mt++) {
          const u32x4 a = Ab[(mt * 16 + r + yp) * 4];
          acc[mt][0] = mfma16<KIND>(a, b0, acc[mt][0]);
          if (nt > 1) acc[mt][1] = mfma16<KIND>(a, b1, acc[mt][1]);
          if (nt > 2) acc[mt][2] = mfma16<KIND>(a, b2, acc[mt][2]);
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

size_t mfma_q_lds_bytes(int fh, int fw, int nt, int MT) {
  return (size_t)16 * ((size_t)(4 + fw - 1) * (16 * MT + fh - 1) * 4 + 2 * (size_t)fh * nt * 64);
}

template <int KIND, int MT>
static int launch_q(hipStream_t s, int blocks, size_t lds, const MfmaTile *tiles, int ntiles,
                    int *counter, const void *feat, const void *filt, const long long *map_off,
                    const int2 *map_dims, float *out) {
  // the kernel's static LDS (the tile index) comes out of the same 160 KB
  static const bool attr = hipFuncSetAttribute((const void *)k_conv_mfma_q<KIND, MT>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               kQueueMaxLds) == hipSuccess;
  if (!attr) return -1;
  hipLaunchKernelGGL((k_conv_mfma_q<KIND, MT>), dim3(blocks), dim3(256), lds, s, tiles, ntiles,
                     counter, (const u32x4 *)feat, (const u32x4 *)filt, map_off, map_dims, out);
  return 0;
}

int launch_conv_mfma_queue(hipStream_t s, int kind, int MT, int blocks, size_t lds,
                           const MfmaTile *tiles, int ntiles, int *counter, const void *feat,
                           const void *filt, const long long *map_off, const int2 *map_dims,
                           float *out) {
  if (ntiles <= 0) return 0;
  if (lds > (size_t)kQueueMaxLds || blocks <= 0) return -1;
  if (kind == 1)
    return MT == 2 ? launch_q<1, 2>(s, blocks, lds, tiles, ntiles, counter, feat, filt, map_off, map_dims, out)
                   : launch_q<1, 4>(s, blocks, lds, tiles, ntiles, counter, feat, filt, map_off, map_dims, out);
  return MT == 2 ? launch_q<2, 2>(s, blocks, lds, tiles, ntiles, counter, feat, filt, map_off, map_dims, out)
                 : launch_q<2, 4>(s, blocks, lds, tiles, ntiles, counter, feat, filt, map_off, map_dims, out);
}

size_t mfma_lds_bytes(int fh, int fw, int nt) {
  return (size_t)16 * ((size_t)(4 + fw - 1) * (64 + fh - 1) * 4 + 2 * (size_t)fh * nt * 64);
}

template <int KIND, int NT>
static int launch_one(hipStream_t s, const MfmaTile *tiles, int ntiles, int fh, int fw,
                      const void *feat, const void *filt, const long long *map_off,
                      const int2 *map_dims, float *out) {
  const size_t lds = mfma_lds_bytes(fh, fw, NT);
  static bool attr_set = false;
  if (!attr_set) {
    if (hipFuncSetAttribute((const void *)k_conv_mfma<KIND, NT>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) != hipSuccess)
      return -1;
    attr_set = true;
  }
  hipLaunchKernelGGL((k_conv_mfma<KIND, NT>), dim3(ntiles), dim3(256), lds, s, tiles, fh, fw,
                     (const u32x4 *)feat, (const u32x4 *)filt, map_off, map_dims, out);
  return 0;
}

int launch_conv_mfma(hipStream_t s, int kind, int nt, const MfmaTile *tiles, int ntiles, int fh,
                     int fw, const void *feat, const void *filt, const long long *map_off,
                     const int2 *map_dims, float *out) {
  if (ntiles <= 0) return 0;
  if (mfma_lds_bytes(fh, fw, nt) > 160 * 1024) return -1;
  if (kind == 1) {
    if (nt == 3) return launch_one<1, 3>(s, tiles, ntiles, fh, fw, feat, filt, map_off, map_dims, out);
    if (nt == 2) return launch_one<1, 2>(s, tiles, ntiles, fh, fw, feat, filt, map_off, map_dims, out);
    return launch_one<1, 1>(s, tiles, ntiles, fh, fw, feat, filt, map_off, map_dims, out);
  }
  if (nt == 3) return launch_one<2, 3>(s, tiles, ntiles, fh, fw, feat, filt, map_off, map_dims, out);
  if (nt == 2) return launch_one<2, 2>(s, tiles, ntiles, fh, fw, feat, filt, map_off, map_dims, out);
  return launch_one<2, 1>(s, tiles, ntiles, fh, fw, feat, filt, map_off, map_dims, out);
}

}  // namespace geye

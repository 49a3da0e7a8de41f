// dpm_oracle.cpp -- CPU oracle for the DPM detection path.
//
// TEST INFRASTRUCTURE ONLY: a restatement of the reference's CPU semantics, used by the
// tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.  The
// product library (geye_amd/) never links or calls it.
//
// Arithmetic contract.  The reference is FLOAT=float C++ compiled by g++ for x86-64 (SSE,
// FLT_EVAL_METHOD 0, no FMA contraction).  Every expression below spells out the type of
// each operation: float ops are written on float operands, and where the reference mixes in
// a double literal (eps, 0.5, 0.2357, 1e-3 ...) the operation is done in double and
// rounded to float exactly where the reference assigns to a FLOAT.  This file must be
// built with -ffp-contract=off.  The convolution is the exception that proves the rule:
// the reference has no CPU convolution (CT/get_boxes_func.h:28 declares fconvsMT, nothing
// defines it); its only convolution is the CUDA kernel CT/GPU_function.cu:302-358, and nvcc
// contracts `val += A*B` to a fused multiply-add by default, so the oracle uses fmaf there.
//
// Parity pins (what checks this file, see tests/):
//   resize / resize1dtran .... reference resize.cc compiled into oracle/_ref + golden fixtures
//   dt ....................... reference dt.cpp compiled into oracle/_ref + golden fixtures,
//                              brute-force O(n^2) minimum, closed forms
//   calc_feature ............. closed forms (constant image -> 0, single-orientation edges,
//                              truncation, texture = 0.2357*sum), FP64 brute force
//   conv ..................... brute force in numpy (float64), one-hot filters, linearity
//   nms ...................... hand-built box sets (min-area overlap, last-survivor drop,
//                              N=1 -> 0, original order); the reference nms.cpp needs
//                              OpenCV headers and cannot be built here
//   level plan / boxes ....... closed-form geometry, known CSV values
#include "dpm_oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------------------------------
// Small helpers
// ---------------------------------------------------------------------------------------
template <typename Fn>
void parallel_for(int n, int threads, Fn fn) {
  if (threads <= 1 || n <= 1) {
    for (int i = 0; i < n; ++i) fn(i);
    return;
  }
  std::atomic<int> next(0);
  int nt = std::min(threads, n);
  std::vector<std::thread> pool;
  pool.reserve(nt);
  for (int t = 0; t < nt; ++t)
    pool.emplace_back([&]() {
      for (int i = next.fetch_add(1); i < n; i = next.fetch_add(1)) fn(i);
    });
  for (auto &th : pool) th.join();
}

// Hcos/Hsin: the reference's 4-decimal orientation table (CT/featurepyramid.cpp.new2:76-77),
// stored as FLOAT.
const float kHcos[9] = {1.0000f, 0.9397f, 0.7660f, 0.5000f, 0.1736f,
                        -0.1736f, -0.5000f, -0.7660f, -0.9397f};
const float kHsin[9] = {0.0000f, 0.3420f, 0.6428f, 0.8660f, 0.9848f,
                        0.9848f, 0.8660f, 0.6428f, 0.3420f};

// min_2: `x <= 0.2 ? x : 0.2` with the comparison in double (CT/featurepyramid.cpp.new2:161).
inline float min_2(float x) { return ((double)x <= 0.2) ? x : (float)0.2; }

// ---------------------------------------------------------------------------------------
// Level plan: ini_scales (CT/featurepyramid.cpp.new2:174-239) and the size recursion of
// calc_f_pyramid (CT/featurepyramid.cpp.org:570-630).
// ---------------------------------------------------------------------------------------
struct LevelPlan {
  int max_scale = 0, LEN = 0;
  std::vector<float> scales;
  std::vector<int> img_h, img_w, sbin_l, feat_h, feat_w;
  std::vector<int> src_level;  // level whose image this level's image is resized from (-1 = frame)
  std::vector<float> resize_scale;
};

int compute_max_scale(const int *rsize, int ncomp, int sbin, int interval, int W, int H) {
  const float sc = (float)std::pow(2.0, (1 / (double)interval));
  int max_scale = 36;
  const int L_NUM = interval + max_scale;
  float MRY = (float)rsize[0];
  float MRX = (float)rsize[1];
  for (int kk = 1; kk < ncomp; kk++) {
    if ((float)rsize[kk * 2] < MRY) MRY = (float)rsize[kk * 2];
    if ((float)rsize[kk * 2 + 1] < MRX) MRX = (float)rsize[kk * 2 + 1];
  }
  MRY /= 2;
  MRX /= 2;
  float height = (float)H / (float)sbin;
  float width = (float)W / (float)sbin;
  float sc_step = 1 / sc;
  for (int kk = 0; kk < L_NUM; kk++) {
    height *= sc_step;
    width *= sc_step;
    if (height < MRY || width < MRX) {
      max_scale = kk - interval - 1;
      break;
    }
  }
  if (max_scale < interval) max_scale = interval;
  return max_scale;
}

int blocks_of(int n, int sbin) { return (int)std::floor(double(n) / double(sbin) + 0.5); }

LevelPlan make_plan(int max_scale, int sbin, int interval, int W, int H) {
  LevelPlan p;
  p.max_scale = max_scale;
  p.LEN = max_scale + interval;
  const int LEN = p.LEN;
  p.scales.assign(LEN, 0.f);
  p.img_h.assign(LEN, 0);
  p.img_w.assign(LEN, 0);
  p.sbin_l.assign(LEN, 0);
  p.feat_h.assign(LEN, 0);
  p.feat_w.assign(LEN, 0);
  p.src_level.assign(LEN, -1);
  p.resize_scale.assign(LEN, 1.f);
  const int sbin2 = (int)std::floor((double)sbin / 2.0);
  const float sc = (float)std::pow(2, (1.0 / (double)interval));
  for (int ii = 0; ii < interval && ii < LEN; ii++) {
    float st = (float)(1.0 / std::pow((double)sc, (double)ii));
    int h, w;
    if (st == 1.0f) {
      h = H;
      w = W;
    } else {
      h = (int)((double)((float)H * st) + 0.5);
      w = (int)((double)((float)W * st) + 0.5);
    }
    p.img_h[ii] = h; p.img_w[ii] = w; p.sbin_l[ii] = sbin2; p.scales[ii] = st * 2;
    p.src_level[ii] = -1; p.resize_scale[ii] = st;
    if (ii + interval < LEN) {
      int l = ii + interval;
      p.img_h[l] = h; p.img_w[l] = w; p.sbin_l[l] = sbin; p.scales[l] = st;
      p.src_level[l] = ii; p.resize_scale[l] = 1.f;  // same image as level ii
    }
    int ph = h, pw = w, prev = ii;
    for (int jj = ii + interval; jj < max_scale; jj += interval) {
      int l = jj + interval;
      int nh = (int)((double)((float)ph * 0.5f) + 0.5);
      int nw = (int)((double)((float)pw * 0.5f) + 0.5);
      p.img_h[l] = nh; p.img_w[l] = nw; p.sbin_l[l] = sbin;
      p.scales[l] = (float)(0.5 * (double)p.scales[jj]);
      p.src_level[l] = prev; p.resize_scale[l] = 0.5f;
      ph = nh; pw = nw; prev = l;
    }
  }
  for (int l = 0; l < LEN; l++) {
    int b0 = blocks_of(p.img_h[l], p.sbin_l[l]);
    int b1 = blocks_of(p.img_w[l], p.sbin_l[l]);
    p.feat_h[l] = std::max(b0 - 2, 0);
    p.feat_w[l] = std::max(b1 - 2, 0);
  }
  return p;
}

// ---------------------------------------------------------------------------------------
// resize (CT/resize.cc:52-160)
// ---------------------------------------------------------------------------------------
struct AlphaInfo {
  int si, di;
  float alpha;
};

std::vector<AlphaInfo> resize_table(int sheight, int dheight, int width) {
  const float scale = (float)dheight / (float)sheight;
  const float invscale = (float)sheight / (float)dheight;
  std::vector<AlphaInfo> ofs;
  ofs.reserve((size_t)(dheight * invscale + 0.99) + 2 * dheight + 4);
  for (int dy = 0; dy < dheight; dy++) {
    float fsy1 = (float)dy * invscale;
    float fsy2 = fsy1 + invscale;
    int sy1 = (int)((double)fsy1 + 0.99);
    int sy2 = (int)fsy2;
    int dyW = dy * width;
    float e1 = (float)sy1 - fsy1;
    if ((double)e1 > 1e-3) ofs.push_back({sy1 - 1, dyW, e1 * scale});
    for (int sy = sy1; sy < sy2; sy++) ofs.push_back({sy, dyW, scale});
    float e2 = fsy2 - (float)sy2;
    if ((double)e2 > 1e-3) ofs.push_back({sy2, dyW, e2 * scale});
  }
  for (auto &o : ofs) {
    if (o.si < 0 || o.si >= sheight) {
      std::fprintf(stderr, "oracle: resize table index out of range\n");
      std::abort();
    }
  }
  return ofs;
}

void resize1dtran(const float *src, int sheight, float *dst, int dheight, int width, int chan) {
  std::vector<AlphaInfo> ofs = resize_table(sheight, dheight, width);
  const long long WD = (long long)width * dheight;
  const long long WS = (long long)width * sheight;
  std::memset(dst, 0, sizeof(float) * chan * WD);
  for (int c = 0; c < chan; c++) {
    const float *s = src + c * WS;
    float *d = dst + c * WD;
    for (int x = 0; x < width; x++) {
      for (const AlphaInfo &o : ofs) d[o.di] += o.alpha * s[o.si];
      s += sheight;
      d++;
    }
  }
}

std::vector<float> resize_img(const std::vector<float> &src, const int *sdims, float scale,
                              int *odims) {
  if (scale == 1.0f) {
    odims[0] = sdims[0]; odims[1] = sdims[1]; odims[2] = sdims[2];
    return src;
  }
  odims[0] = (int)((double)((float)sdims[0] * scale) + 0.5);
  odims[1] = (int)((double)((float)sdims[1] * scale) + 0.5);
  odims[2] = sdims[2];
  std::vector<float> tmp((size_t)odims[0] * sdims[1] * sdims[2]);
  std::vector<float> dst((size_t)odims[0] * odims[1] * sdims[2]);
  resize1dtran(src.data(), sdims[0], tmp.data(), odims[0], sdims[1], sdims[2]);
  resize1dtran(tmp.data(), sdims[1], dst.data(), odims[1], odims[0], sdims[2]);
  return dst;
}

// ---------------------------------------------------------------------------------------
// calc_feature (CT/featurepyramid.cpp.new2:263-501)
// ---------------------------------------------------------------------------------------
std::vector<float> calc_feature(const float *SRC, int height, int width, int sbin, int *F0,
                                int *F1) {
  const int dims0 = height, dims1 = width;
  const int b0 = blocks_of(height, sbin), b1 = blocks_of(width, sbin);
  const int out0 = std::max(b0 - 2, 0), out1 = std::max(b1 - 2, 0);
  const int vis0 = b0 * sbin, vis1 = b1 * sbin;
  const long long BSQ = (long long)b0 * b1;
  std::vector<float> hist((size_t)BSQ * 18, 0.f), norm((size_t)BSQ, 0.f);
  std::vector<float> feat((size_t)out0 * out1 * 31, 0.f);
  const long long SQ = (long long)dims0 * dims1;

  for (int x = 1; x < vis1 - 1; x++) {
    for (int y = 1; y < vis0 - 1; y++) {
      const float *s = SRC + (long long)std::min(x, dims1 - 2) * dims0 + std::min(y, dims0 - 2);
      float dy = s[1] - s[-1];
      float dx = s[dims0] - s[-dims0];
      float v = dx * dx + dy * dy;
      s += SQ;
      float dy2 = s[1] - s[-1];
      float dx2 = s[dims0] - s[-dims0];
      float v2 = dx2 * dx2 + dy2 * dy2;
      s += SQ;
      float dy3 = s[1] - s[-1];
      float dx3 = s[dims0] - s[-dims0];
      float v3 = dx3 * dx3 + dy3 * dy3;
      if (v2 > v) { v = v2; dx = dx2; dy = dy2; }
      if (v3 > v) { v = v3; dx = dx3; dy = dy3; }
      float best_dot = 0;
      int best_o = 0;
      for (int o = 0; o < 9; o++) {
        float dot = kHcos[o] * dx + kHsin[o] * dy;
        if (dot > best_dot) {
          best_dot = dot;
          best_o = o;
        } else if (-dot > best_dot) {
          best_dot = -dot;
          best_o = o + 9;
        }
      }
      float xp = (float)(((double)(float)x + 0.5) / (double)(float)sbin - 0.5);
      float yp = (float)(((double)(float)y + 0.5) / (double)(float)sbin - 0.5);
      int ixp = (int)std::floor(xp);
      int iyp = (int)std::floor(yp);
      float vx0 = xp - (float)ixp;
      float vy0 = yp - (float)iyp;
      float vx1 = (float)(1.0 - (double)vx0);
      float vy1 = (float)(1.0 - (double)vy0);
      v = std::sqrt(v);
      float *H = hist.data() + (long long)best_o * BSQ;
      if (ixp >= 0 && iyp >= 0) H[(long long)ixp * b0 + iyp] += vx1 * vy1 * v;
      if (ixp + 1 < b1 && iyp >= 0) H[(long long)(ixp + 1) * b0 + iyp] += vx0 * vy1 * v;
      if (ixp >= 0 && iyp + 1 < b0) H[(long long)ixp * b0 + (iyp + 1)] += vx1 * vy0 * v;
      if (ixp + 1 < b1 && iyp + 1 < b0) H[(long long)(ixp + 1) * b0 + (iyp + 1)] += vx0 * vy0 * v;
    }
  }
  for (int o = 0; o < 9; o++) {
    const float *s1 = hist.data() + o * BSQ;
    const float *s2 = hist.data() + (o + 9) * BSQ;
    for (long long i = 0; i < BSQ; i++) {
      float t = s1[i] + s2[i];
      norm[i] += t * t;
    }
  }
  const double eps = 0.0001;
  const long long OSQ = (long long)out0 * out1;
  for (int x = 0; x < out1; x++) {
    for (int y = 0; y < out0; y++) {
      float *dst = feat.data() + (long long)x * out0 + y;
      const float *p;
      p = norm.data() + (long long)(x + 1) * b0 + y + 1;
      float n1 = (float)(1.0 / std::sqrt((double)(p[0] + p[1] + p[b0] + p[b0 + 1]) + eps));
      p = norm.data() + (long long)(x + 1) * b0 + y;
      float n2 = (float)(1.0 / std::sqrt((double)(p[0] + p[1] + p[b0] + p[b0 + 1]) + eps));
      p = norm.data() + (long long)x * b0 + y + 1;
      float n3 = (float)(1.0 / std::sqrt((double)(p[0] + p[1] + p[b0] + p[b0 + 1]) + eps));
      p = norm.data() + (long long)x * b0 + y;
      float n4 = (float)(1.0 / std::sqrt((double)(p[0] + p[1] + p[b0] + p[b0 + 1]) + eps));
      float t1 = 0, t2 = 0, t3 = 0, t4 = 0;
      const float *src = hist.data() + (long long)(x + 1) * b0 + (y + 1);
      for (int o = 0; o < 18; o++) {
        float h1 = min_2(*src * n1);
        float h2 = min_2(*src * n2);
        float h3 = min_2(*src * n3);
        float h4 = min_2(*src * n4);
        *dst = (float)(0.5 * (double)(h1 + h2 + h3 + h4));
        t1 += h1; t2 += h2; t3 += h3; t4 += h4;
        dst += OSQ;
        src += BSQ;
      }
      src = hist.data() + (long long)(x + 1) * b0 + (y + 1);
      for (int o = 0; o < 9; o++) {
        float sum = *src + *(src + 9 * BSQ);
        float h1 = min_2(sum * n1);
        float h2 = min_2(sum * n2);
        float h3 = min_2(sum * n3);
        float h4 = min_2(sum * n4);
        *dst = (float)(0.5 * (double)(h1 + h2 + h3 + h4));
        dst += OSQ;
        src += BSQ;
      }
      *dst = (float)(0.2357 * (double)t1); dst += OSQ;
      *dst = (float)(0.2357 * (double)t2); dst += OSQ;
      *dst = (float)(0.2357 * (double)t3); dst += OSQ;
      *dst = (float)(0.2357 * (double)t4);
    }
  }
  *F0 = out0;
  *F1 = out1;
  return feat;
}

// ---------------------------------------------------------------------------------------
// Ingest (CT/featurepyramid.cpp.new2:816-850)
// ---------------------------------------------------------------------------------------
void ipl_to_float(const unsigned char *bgr, int width, int height, int ws, float *out) {
  const long long SQ = (long long)height * width;
  float *P0 = out, *P1 = out + SQ, *P2 = out + 2 * SQ;
  for (int x = 0; x < width; x++) {
    int XT = x * 3;
    for (int y = 0; y < height; y++) {
      long long pp = (long long)ws * y + XT;
      *(P0++) = (float)bgr[pp];
      *(P1++) = (float)bgr[pp + 1];
      *(P2++) = (float)bgr[pp + 2];
    }
  }
}

// ---------------------------------------------------------------------------------------
// Pyramid (CT/featurepyramid.cpp.org:547-661 structure, new2 arithmetic)
// ---------------------------------------------------------------------------------------
struct Pyramid {
  LevelPlan plan;
  std::vector<std::vector<float>> feat;
};

Pyramid build_pyramid(const unsigned char *bgr, int W, int H, int ws, int sbin, int interval,
                      int max_scale, int threads) {
  Pyramid P;
  P.plan = make_plan(max_scale, sbin, interval, W, H);
  const int LEN = P.plan.LEN;
  std::vector<float> D((size_t)3 * W * H);
  ipl_to_float(bgr, W, H, ws, D.data());
  int in_dims[3] = {H, W, 3};
  std::vector<std::vector<float>> imgs(LEN);
  // Resizes run on the calling thread in the reference's order (org:578-628).
  for (int ii = 0; ii < interval && ii < LEN; ii++) {
    float st = P.plan.resize_scale[ii];
    int od[3];
    imgs[ii] = resize_img(D, in_dims, st, od);
    if (od[0] != P.plan.img_h[ii] || od[1] != P.plan.img_w[ii]) std::abort();
  }
  for (int ii = 0; ii < interval && ii < LEN; ii++) {
    int cur = ii;
    int dims[3] = {P.plan.img_h[ii], P.plan.img_w[ii], 3};
    for (int jj = ii + interval; jj < max_scale; jj += interval) {
      int l = jj + interval;
      int od[3];
      imgs[l] = resize_img(imgs[cur], dims, 0.5f, od);
      if (od[0] != P.plan.img_h[l] || od[1] != P.plan.img_w[l]) std::abort();
      dims[0] = od[0]; dims[1] = od[1];
      cur = l;
    }
  }
  P.feat.resize(LEN);
  parallel_for(LEN, threads, [&](int l) {
    const std::vector<float> &img = (l >= interval && l < 2 * interval) ? imgs[l - interval] : imgs[l];
    int F0, F1;
    P.feat[l] = calc_feature(img.data(), P.plan.img_h[l], P.plan.img_w[l], P.plan.sbin_l[l], &F0, &F1);
    if (F0 != P.plan.feat_h[l] || F1 != P.plan.feat_w[l]) std::abort();
  });
  return P;
}

// ---------------------------------------------------------------------------------------
// Convolution: per-output order of CT/GPU_function.cu:302-358.
// ---------------------------------------------------------------------------------------
void conv_f32(const float *A, int F0, int F1, const float *B, int h, int w, float *C) {
  const int H = F0 - h + 1, W = F1 - w + 1;
  const long long ASQ = (long long)F0 * F1;
  const int BSQ = h * w;
  for (int x = 0; x < W; x++) {
    for (int y = 0; y < H; y++) {
      float add_val = 0;
      for (int f = 0; f < 31; f++) {
        const float *a = A + f * ASQ + (long long)x * F0 + y;
        const float *b = B + f * BSQ;
        float val = 0;
        for (int xp = 0; xp < w; xp++) {
          for (int yp = 0; yp < h; yp++) val = std::fma(a[yp], b[yp], val);
          a += F0;
          b += h;
        }
        add_val += val;
      }
      C[(long long)x * H + y] = 0.0f + add_val;
    }
  }
}

void conv_f64(const float *A, int F0, int F1, const float *B, int h, int w, double *C) {
  const int H = F0 - h + 1, W = F1 - w + 1;
  const long long ASQ = (long long)F0 * F1;
  const int BSQ = h * w;
  for (int x = 0; x < W; x++) {
    for (int y = 0; y < H; y++) {
      double add_val = 0;
      for (int f = 0; f < 31; f++) {
        const float *a = A + f * ASQ + (long long)x * F0 + y;
        const float *b = B + f * BSQ;
        double val = 0;
        for (int xp = 0; xp < w; xp++) {
          for (int yp = 0; yp < h; yp++) val += (double)a[yp] * (double)b[yp];
          a += F0;
          b += h;
        }
        add_val += val;
      }
      C[(long long)x * H + y] = add_val;
    }
  }
}

// ---------------------------------------------------------------------------------------
// Distance transform (CT/dt.cpp:61-136).  The cost is evaluated exactly as the reference's
// expression parses: ((src + (a*t)*t) + b*t) with t an int converted to T.
// ---------------------------------------------------------------------------------------
template <typename T>
void dt_helper(const T *src, T *dst, int *ptr, long long step, int s1, int s2, int d1, int d2,
               T a, T b) {
  if (d2 >= d1) {
    int d = (d1 + d2) >> 1;
    long long ds = d * step;
    int s = s1;
    T src_ss = src[s * step];
    for (int p = s1 + 1; p <= s2; p++) {
      int t1 = d - s;
      int t2 = d - p;
      T lhs = src_ss + a * (T)t1 * (T)t1 + b * (T)t1;
      T rhs = src[p * step] + a * (T)t2 * (T)t2 + b * (T)t2;
      if (lhs > rhs) {
        s = p;
        src_ss = src[s * step];
      }
    }
    int D = d - s;
    dst[ds] = src[s * step] + a * (T)D * (T)D + b * (T)D;
    ptr[ds] = s;
    dt_helper(src, dst, ptr, step, s1, s, d1, d - 1, a, b);
    dt_helper(src, dst, ptr, step, s, s2, d + 1, d2, a, b);
  }
}

template <typename T>
void dt_passes(const T *vals, T ax, T bx, T ay, T by, const int *dims, T *tmpM, int *tmpIy,
               T *M, int *tmpIx) {
  const int d0 = dims[0], d1 = dims[1];
  for (int x = 0; x < d1; x++) {
    long long XD = (long long)x * d0;
    dt_helper<T>(vals + XD, tmpM + XD, tmpIy + XD, 1, 0, d0 - 1, 0, d0 - 1, ay, by);
  }
  for (int y = 0; y < d0; y++)
    dt_helper<T>(tmpM + y, M + y, tmpIx + y, d0, 0, d1 - 1, 0, d1 - 1, ax, bx);
}

template <typename T>
void dt_full(const T *vals, T ax, T bx, T ay, T by, const int *dims, T *M, int *Ix, int *Iy) {
  const long long SQ = (long long)dims[0] * dims[1];
  std::vector<T> tmpM(SQ);
  std::vector<int> tmpIx(SQ), tmpIy(SQ);
  dt_passes<T>(vals, ax, bx, ay, by, dims, tmpM.data(), tmpIy.data(), M, tmpIx.data());
  long long k = 0;
  for (int x = 0; x < dims[1]; x++)
    for (int y = 0; y < dims[0]; y++, k++) {
      Ix[k] = tmpIx[k];
      Iy[k] = tmpIy[(long long)tmpIx[k] * dims[0] + y];
    }
}

template <typename T>
void add_part(T *score, const T *M, const int *rootsize, const int *partsize, int ax, int ay) {
  T *S = score;
  int jj_L = ax + 2 * (rootsize[1] - 1) - 1;
  int ii_L = ay + 2 * (rootsize[0] - 1);
  int axm = ax - 1;
  for (int jj = axm; jj <= jj_L; jj += 2) {
    long long L = (long long)jj * partsize[0];
    for (int ii = ay; ii <= ii_L; ii += 2) {
      *S -= M[ii + L - 1];
      S++;
    }
  }
}

// ---------------------------------------------------------------------------------------
// NMS (CT/nms.cpp:70-217)
// ---------------------------------------------------------------------------------------
void quick_sort(float *ary, int *order, int first, int last) {
  int i = first, j = last;
  float key = ary[(first + last) / 2];
  while (true) {
    while (ary[i] > key) i++;
    while (ary[j] < key) j--;
    if (i >= j) break;
    std::swap(ary[i], ary[j]);
    std::swap(order[i], order[j]);
    i++;
    j--;
  }
  if (first < i - 1) quick_sort(ary, order, first, i - 1);
  if (last > j + 1) quick_sort(ary, order, j + 1, last);
}

int nms(const float *boxes, int NUM, int GL, float overlap, float *out) {
  if (NUM <= 0) return 0;
  std::vector<float> area(NUM), scores(NUM);
  std::vector<int> order(NUM), checked(NUM, 0);
  for (int i = 0; i < NUM; i++) {
    const float *P = boxes + (long long)GL * i;
    area[i] = (float)(((double)(P[3] - P[1]) + 1.0) * ((double)(P[2] - P[0]) + 1.0));
    scores[i] = P[GL - 2];
    order[i] = i;
  }
  quick_sort(scores.data(), order.data(), 0, NUM - 1);
  int cnum = NUM - 1, cur = 0, pi_num = 0;
  while (cnum > 0) {
    int A = order[cur];
    const float *P = boxes + (long long)GL * A;
    float Ay1 = P[0], Ax1 = P[1], Ay2 = P[2], Ax2 = P[3];
    checked[A] = 1;
    cnum--;
    cur++;
    pi_num++;
    for (int kk = cur; kk < NUM; kk++) {
      int B = order[kk];
      if (checked[B] != 0) continue;
      P = boxes + (long long)GL * B;
      float yy1 = (Ay1 >= P[0]) ? Ay1 : P[0];
      float xx1 = (Ax1 >= P[1]) ? Ax1 : P[1];
      float yy2 = (Ay2 <= P[2]) ? Ay2 : P[2];
      float xx2 = (Ax2 <= P[3]) ? Ax2 : P[3];
      float w = (float)((double)(xx2 - xx1) + 1.0);
      float h = (float)((double)(yy2 - yy1) + 1.0);
      float R_AREA = (area[A] <= area[B]) ? area[A] : area[B];
      if (w > 0 && h > 0) {
        float o = w * h / R_AREA;
        if (o > overlap) {
          checked[B] = -1;
          cnum--;
        }
      } else if (Ay1 < P[0] && Ax1 < P[1] && Ay2 > P[2] && Ax2 > P[3]) {
        checked[B] = -1;
        cnum--;
      } else if (Ay1 > P[0] && Ax1 > P[1] && Ay2 < P[2] && Ax2 < P[3]) {
        checked[B] = -1;
        cnum--;
      }
    }
    for (int kk = cur; kk < NUM; kk++)
      if (checked[order[kk]] == 0) {
        cur = kk;
        break;
      }
  }
  int count = 0;
  for (int ii = 0; count < pi_num; ii++)
    if (checked[ii] == 1) {
      std::memcpy(out + (long long)count * GL, boxes + (long long)GL * ii, sizeof(float) * GL);
      count++;
    }
  return pi_num;
}

// ---------------------------------------------------------------------------------------
// get_boxes (CT/get_boxes.cpp:583-1797)
// ---------------------------------------------------------------------------------------
struct ScoreMaps {
  LevelPlan plan;
  std::vector<int> valid;                    // per level
  // per (L, jj): root+bias-parts score map, sizes
  std::vector<std::vector<double>> score_d;  // precision 1
  std::vector<std::vector<float>> score_f;   // precision 0
  std::vector<int> rdims;                    // [L][jj][2]
  // per (L, jj, kk): Ix, Iy and part map dims
  std::vector<std::vector<int>> Ix, Iy;
  std::vector<int> pdims;                    // [L][p][2]
  int maxpart = 0;
};

template <typename T>
void compute_scores(const ora_model_desc *m, const Pyramid &P, int threads, ScoreMaps &S,
                    std::vector<std::vector<T>> &score) {
  const int interval = m->interval, LEN = P.plan.LEN, NoC = m->numcomponent;
  const int NoR = m->NoR, NoP = m->NoP, MP = m->maxpart;
  S.plan = P.plan;
  S.maxpart = MP;
  S.valid.assign(LEN, 0);
  for (int level = interval; level < LEN; level++) {
    bool skip = (P.plan.feat_h[level] + 2 * m->pady < m->max_Y) ||
                (P.plan.feat_w[level] + 2 * m->padx < m->max_X);
    S.valid[level] = skip ? 0 : 1;
  }
  const int NL = LEN - interval;
  // root and part convolution, one task per filter (the "MT" of fconvsMT)
  std::vector<std::vector<T>> rootm((size_t)LEN * NoR), partm((size_t)NL * NoP);
  S.rdims.assign((size_t)NL * NoC * 2, 0);
  S.pdims.assign((size_t)NL * NoP * 2, 0);
  parallel_for(NoR + NoP, threads, [&](int t) {
    for (int level = interval; level < LEN; level++) {
      if (!S.valid[level]) continue;
      int L = level - interval;
      if (t < NoR) {
        const int *bs = m->root_size + 3 * t;
        int lv = level;
        int H = P.plan.feat_h[lv] - bs[0] + 1, W = P.plan.feat_w[lv] - bs[1] + 1;
        if (H < 1 || W < 1) std::abort();
        auto &o = rootm[(size_t)level * NoR + t];
        o.resize((size_t)H * W);
        if constexpr (sizeof(T) == 4)
          conv_f32(P.feat[lv].data(), P.plan.feat_h[lv], P.plan.feat_w[lv],
                   m->rootfilter + m->root_off[t], bs[0], bs[1], o.data());
        else
          conv_f64(P.feat[lv].data(), P.plan.feat_h[lv], P.plan.feat_w[lv],
                   m->rootfilter + m->root_off[t], bs[0], bs[1], o.data());
      } else {
        int p = t - NoR;
        const int *bs = m->part_size + 3 * p;
        int lv = L;
        int H = P.plan.feat_h[lv] - bs[0] + 1, W = P.plan.feat_w[lv] - bs[1] + 1;
        if (H < 1 || W < 1) std::abort();
        S.pdims[((size_t)L * NoP + p) * 2] = H;
        S.pdims[((size_t)L * NoP + p) * 2 + 1] = W;
        auto &o = partm[(size_t)L * NoP + p];
        o.resize((size_t)H * W);
        if constexpr (sizeof(T) == 4)
          conv_f32(P.feat[lv].data(), P.plan.feat_h[lv], P.plan.feat_w[lv],
                   m->partfilter + m->part_off[p], bs[0], bs[1], o.data());
        else
          conv_f64(P.feat[lv].data(), P.plan.feat_h[lv], P.plan.feat_w[lv],
                   m->partfilter + m->part_off[p], bs[0], bs[1], o.data());
      }
    }
  });
  score.assign((size_t)NL * NoC, {});
  S.Ix.assign((size_t)NL * NoC * MP, {});
  S.Iy.assign((size_t)NL * NoC * MP, {});
  for (int level = interval; level < LEN; level++) {
    if (!S.valid[level]) continue;
    int L = level - interval;
    for (int jj = 0; jj < NoC; jj++) {
      const int *rs = m->root_size + 3 * jj;
      int R0 = P.plan.feat_h[level] - rs[0] + 1, R1 = P.plan.feat_w[level] - rs[1] + 1;
      S.rdims[((size_t)L * NoC + jj) * 2] = R0;
      S.rdims[((size_t)L * NoC + jj) * 2 + 1] = R1;
      auto &sc = score[(size_t)L * NoC + jj];
      sc = rootm[(size_t)level * NoR + jj];
      const T off = (T)m->offw[m->ridx[jj]];
      for (auto &v : sc) v += off;
      for (int kk = 0; kk < m->numpart[jj]; kk++) {
        int pid = m->pidx[jj * MP + kk];
        int did = m->didx[jj * MP + kk];
        int P0 = S.pdims[((size_t)L * NoP + pid) * 2], P1 = S.pdims[((size_t)L * NoP + pid) * 2 + 1];
        int ax = m->anchor[did * 2] + 1, ay = m->anchor[did * 2 + 1] + 1;
        // add_part reads M[(ay-1+2y) + (ax-1+2x)*P0]; reject models that would read outside.
        if (ay - 1 + 2 * (R0 - 1) >= P0 || ax - 1 + 2 * (R1 - 1) >= P1 || ay < 1 || ax < 1) {
          std::fprintf(stderr, "oracle: part anchor reads outside the part map\n");
          std::abort();
        }
        const auto &pm = partm[(size_t)L * NoP + pid];
        std::vector<T> neg(pm.size());
        for (size_t i = 0; i < pm.size(); i++) neg[i] = pm[i] * (T)-1;
        int dims[2] = {P0, P1};
        std::vector<T> M(pm.size());
        auto &Ix = S.Ix[((size_t)L * NoC + jj) * MP + kk];
        auto &Iy = S.Iy[((size_t)L * NoC + jj) * MP + kk];
        Ix.resize(pm.size());
        Iy.resize(pm.size());
        const float *df = m->def + did * 4;
        dt_full<T>(neg.data(), (T)df[0], (T)df[1], (T)df[2], (T)df[3], dims, M.data(), Ix.data(),
                   Iy.data());
        int rsz[2] = {R0, R1};
        add_part<T>(sc.data(), M.data(), rsz, dims, ax, ay);
      }
    }
  }
}

template <typename T>
std::vector<float> boxes_from_scores(const ora_model_desc *m, const ScoreMaps &S,
                                     const std::vector<std::vector<T>> &score, float thresh,
                                     int *GLout) {
  const int interval = m->interval, LEN = S.plan.LEN, NoC = m->numcomponent, NoP = m->NoP;
  const int MP = S.maxpart;
  const int GL0 = (m->numpart[0] + 1) * 4 + 3;
  *GLout = GL0;
  std::vector<float> out;
  for (int level = interval; level < LEN; level++) {
    if (!S.valid[level]) continue;
    int L = level - interval;
    const float scale = (float)m->sbin / S.plan.scales[level];
    for (int jj = 0; jj < NoC; jj++) {
      const int R0 = S.rdims[((size_t)L * NoC + jj) * 2];
      const int R1 = S.rdims[((size_t)L * NoC + jj) * 2 + 1];
      const auto &sc = score[(size_t)L * NoC + jj];
      const int GL = (m->numpart[jj] + 1) * 4 + 3;
      if (GL != GL0) std::abort();
      const int rsize0 = m->rsize[jj * 2], rsize1 = m->rsize[jj * 2 + 1];
      for (long long ii = 0; ii < (long long)R0 * R1; ii++) {
        if (!(sc[ii] > (T)thresh)) continue;
        int y = (int)(ii % R0), x = (int)(ii / R0);
        float row[4];
        row[0] = ((float)y - (float)m->pady + 1) * scale;
        row[1] = ((float)x - (float)m->padx + 1) * scale;
        row[2] = (float)((double)(row[0] + (float)rsize0 * scale) - 1.0);
        row[3] = (float)((double)(row[1] + (float)rsize1 * scale) - 1.0);
        out.insert(out.end(), row, row + 4);
        for (int pp = 0; pp < m->numpart[jj]; pp++) {
          int pid = m->pidx[jj * MP + pp], did = m->didx[jj * MP + pp];
          int ax = m->anchor[did * 2] + 1, ay = m->anchor[did * 2 + 1] + 1;
          int P0 = S.pdims[((size_t)L * NoP + pid) * 2];
          int probex = (x - 1) * 2 + ax, probey = (y - 1) * 2 + ay;
          long long PI = probey + (long long)probex * P0;
          const auto &Ix = S.Ix[((size_t)L * NoC + jj) * MP + pp];
          const auto &Iy = S.Iy[((size_t)L * NoC + jj) * MP + pp];
          if (PI < 0 || PI >= (long long)Ix.size()) {
            std::fprintf(stderr, "oracle: partbox index outside the part map\n");
            std::abort();
          }
          float px = (float)((double)(float)Ix[PI] + 1.0);
          float py = (float)((double)(float)Iy[PI] + 1.0);
          int ps0 = m->psize[(jj * MP + pp) * 2], ps1 = m->psize[(jj * MP + pp) * 2 + 1];
          float b[4];
          b[0] = (float)((((double)py - 2.0) / 2.0 + 1.0 - (double)(float)m->pady) * (double)scale);
          b[1] = (float)((((double)px - 2.0) / 2.0 + 1.0 - (double)(float)m->padx) * (double)scale);
          b[2] = (float)(((double)b[0] + (double)((float)ps0 * scale) / 2.0) - 1.0);
          b[3] = (float)(((double)b[1] + (double)((float)ps1 * scale) / 2.0) - 1.0);
          out.insert(out.end(), b, b + 4);
        }
        out.push_back((float)jj);
        out.push_back((float)sc[ii]);
        out.push_back(scale);
      }
    }
  }
  return out;
}

// ---------------------------------------------------------------------------------------
// accumulated score (CT/get_boxes.cpp:265-300, :1765-1787)
// ---------------------------------------------------------------------------------------
void calc_a_score(float *ac_score, const float *score, const int *ssize, const int *rsize,
                  int IHEI, int IWID, int padx_n, int pady_n, float scale) {
  const int block_pad = (int)(scale / 2.0);
  const int RY = (int)((float)rsize[0] * scale / 2.0 - 1.0 + block_pad);
  const int RX = (int)((float)rsize[1] * scale / 2.0 - 1.0 + block_pad);
  for (int ii = 0; ii < IWID; ii++) {
    const int Xn = (int)((float)ii / scale + padx_n);
    for (int jj = 0; jj < IHEI; jj++) {
      const int Yn = (int)((float)jj / scale + pady_n);
      if (Yn < ssize[0] && Xn < ssize[1]) {
        const float sc = score[Yn + Xn * ssize[0]];
        const int Im_Y = jj + RY, Im_X = ii + RX;
        if (Im_Y < IHEI && Im_X < IWID) {
          float *PP = ac_score + Im_Y + Im_X * IHEI;
          if (sc > *PP) *PP = sc;
        }
      }
    }
  }
}

// AS_OFF = |thresh| as a float (the reference's abs() on a FLOAT).
void ascore_finish(float *A, long long n, float thresh) {
  const float AS_OFF = std::fabs(thresh);
  float max_ac = 0.0f;
  for (long long ii = 0; ii < n; ii++) {
    if (A[ii] < thresh) {
      A[ii] = 0.0f;
    } else {
      A[ii] += AS_OFF;
      if (A[ii] > max_ac) max_ac = A[ii];
    }
  }
  if (max_ac > 0.0f) {
    const float ac_ratio = (float)(1.0 / max_ac);
    for (long long ii = 0; ii < n; ii++) A[ii] *= ac_ratio;
  }
}

// ---------------------------------------------------------------------------------------
// CSV model (CT/load_model.cpp:57-558)
// ---------------------------------------------------------------------------------------
struct OwnedModel {
  int numcomponent = 0, sbin = 0, interval = 0, max_Y = 0, max_X = 0, maxpart = 0;
  std::vector<int> ridx, oidx, rsize, numpart, pidx, didx, psize, anchor;
  std::vector<float> offw, def;
  int DefL = 0, NoR = 0, NoP = 0;
  std::vector<int> root_size, part_size, part_partner;
  std::vector<float> rootfilter, partfilter;
  std::vector<long long> root_off, part_off;
  std::vector<float> lsq;  // x1,y1,x2,y2 per component, kept for completeness
};

bool read_csv_floats(const char *path, std::vector<float> &v) {
  FILE *f = std::fopen(path, "r");
  if (!f) return false;
  std::string s;
  char buf[1 << 16];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, n);
  std::fclose(f);
  const char *p = s.c_str();
  while (*p) {
    while (*p && (*p == ',' || *p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) p++;
    if (!*p) break;
    char *e;
    float x = std::strtof(p, &e);
    if (e == p) return false;
    v.push_back(x);
    p = e;
  }
  return true;
}

}  // namespace

// =======================================================================================
// C API
// =======================================================================================
extern "C" {

void ora_ipl_to_float(const unsigned char *bgr, int width, int height, int widthStep,
                      float *out) {
  ipl_to_float(bgr, width, height, widthStep, out);
}

int ora_level_plan(const int *rsize, int numcomponent, int sbin, int interval, int width,
                   int height, int *max_scale, int *LEN, float *scales, int *img_h, int *img_w,
                   int *sbin_l, int *feat_h, int *feat_w) {
  int ms = compute_max_scale(rsize, numcomponent, sbin, interval, width, height);
  LevelPlan p = make_plan(ms, sbin, interval, width, height);
  *max_scale = ms;
  *LEN = p.LEN;
  for (int l = 0; l < p.LEN; l++) {
    if (scales) scales[l] = p.scales[l];
    if (img_h) img_h[l] = p.img_h[l];
    if (img_w) img_w[l] = p.img_w[l];
    if (sbin_l) sbin_l[l] = p.sbin_l[l];
    if (feat_h) feat_h[l] = p.feat_h[l];
    if (feat_w) feat_w[l] = p.feat_w[l];
  }
  return 0;
}

long long ora_resize(const float *src, const int *sdims, float scale, float *dst, int *odims) {
  if (scale == 1.0f) {
    odims[0] = sdims[0]; odims[1] = sdims[1]; odims[2] = sdims[2];
  } else {
    odims[0] = (int)((double)((float)sdims[0] * scale) + 0.5);
    odims[1] = (int)((double)((float)sdims[1] * scale) + 0.5);
    odims[2] = sdims[2];
  }
  long long n = (long long)odims[0] * odims[1] * odims[2];
  if (!dst) return n;
  std::vector<float> s(src, src + (size_t)sdims[0] * sdims[1] * sdims[2]);
  int od[3];
  std::vector<float> r = resize_img(s, sdims, scale, od);
  std::memcpy(dst, r.data(), sizeof(float) * n);
  return n;
}

int ora_resize_table(int sheight, int dheight, int width, int *si, int *di, float *alpha,
                     int cap) {
  std::vector<AlphaInfo> t = resize_table(sheight, dheight, width);
  int n = std::min<int>((int)t.size(), cap);
  for (int i = 0; i < n; i++) {
    si[i] = t[i].si; di[i] = t[i].di; alpha[i] = t[i].alpha;
  }
  return (int)t.size();
}

int ora_calc_feature(const float *img, int height, int width, int sbin, float *feat, int *F0,
                     int *F1) {
  if (!feat) {
    *F0 = std::max(blocks_of(height, sbin) - 2, 0);
    *F1 = std::max(blocks_of(width, sbin) - 2, 0);
    return 0;
  }
  std::vector<float> f = calc_feature(img, height, width, sbin, F0, F1);
  std::memcpy(feat, f.data(), sizeof(float) * f.size());
  return 0;
}

void ora_conv_f32(const float *feat, int F0, int F1, const float *filt, int h, int w,
                  float *out) {
  conv_f32(feat, F0, F1, filt, h, w, out);
}
void ora_conv_f64(const float *feat, int F0, int F1, const float *filt, int h, int w,
                  double *out) {
  conv_f64(feat, F0, F1, filt, h, w, out);
}

void ora_dt_f32(const float *vals, float ax, float bx, float ay, float by, const int *dims,
                float *M, int *Ix, int *Iy) {
  dt_full<float>(vals, ax, bx, ay, by, dims, M, Ix, Iy);
}
void ora_dt_f64(const double *vals, double ax, double bx, double ay, double by,
                const int *dims, double *M, int *Ix, int *Iy) {
  dt_full<double>(vals, ax, bx, ay, by, dims, M, Ix, Iy);
}
void ora_dt_passes_f32(const float *vals, float ax, float bx, float ay, float by,
                       const int *dims, float *tmpM, int *tmpIy, float *M, int *tmpIx) {
  dt_passes<float>(vals, ax, bx, ay, by, dims, tmpM, tmpIy, M, tmpIx);
}

void ora_add_part_f32(float *score, const float *M, const int *rootsize, const int *partsize,
                      int ax, int ay) {
  add_part<float>(score, M, rootsize, partsize, ax, ay);
}

int ora_nms(const float *boxes, int n, int GL, float overlap, float *out) {
  return nms(boxes, n, GL, overlap, out);
}

void ora_quicksort(float *scores, int *order, int n) {
  if (n > 0) quick_sort(scores, order, 0, n - 1);
}

long long ora_pyramid(const unsigned char *bgr, int width, int height, int widthStep, int sbin,
                      int interval, int max_scale, int threads, float **feat, int *feat_h,
                      int *feat_w, float *scales) {
  Pyramid P = build_pyramid(bgr, width, height, widthStep, sbin, interval, max_scale, threads);
  long long total = 0;
  for (auto &f : P.feat) total += (long long)f.size();
  float *buf = (float *)std::malloc(sizeof(float) * std::max<long long>(total, 1));
  long long off = 0;
  for (int l = 0; l < P.plan.LEN; l++) {
    std::memcpy(buf + off, P.feat[l].data(), sizeof(float) * P.feat[l].size());
    off += (long long)P.feat[l].size();
    feat_h[l] = P.plan.feat_h[l];
    feat_w[l] = P.plan.feat_w[l];
    scales[l] = P.plan.scales[l];
  }
  *feat = buf;
  return total;
}

long long ora_detect(const ora_model_desc *m, const unsigned char *bgr, int width, int height,
                     int widthStep, float thresh, int precision, int threads, float **boxes,
                     int *GL) {
  int ms = compute_max_scale(m->rsize, m->numcomponent, m->sbin, m->interval, width, height);
  Pyramid P = build_pyramid(bgr, width, height, widthStep, m->sbin, m->interval, ms, threads);
  ScoreMaps S;
  std::vector<float> rows;
  if (precision == 1) {
    std::vector<std::vector<double>> sc;
    compute_scores<double>(m, P, threads, S, sc);
    rows = boxes_from_scores<double>(m, S, sc, thresh, GL);
  } else {
    std::vector<std::vector<float>> sc;
    compute_scores<float>(m, P, threads, S, sc);
    rows = boxes_from_scores<float>(m, S, sc, thresh, GL);
  }
  long long n = (long long)rows.size() / *GL;
  float *buf = (float *)std::malloc(sizeof(float) * std::max<size_t>(rows.size(), 1));
  std::memcpy(buf, rows.data(), sizeof(float) * rows.size());
  *boxes = buf;
  return n;
}

long long ora_rows_one_map(const ora_model_desc *m, int jj, float scale, const float *score,
                           int R0, int R1, const int *Ix, const int *Iy, const int *pdims,
                           float thresh, float *out, long long cap, int *GL) {
  // a one-level plan: the level `interval` is the only valid one, with L = 0 and
  // sbin / scales[level] == scale; component jj's maps are filled, the others stay empty
  const int NoC = m->numcomponent, MP = m->maxpart, NoP = m->NoP, np = m->numpart[jj];
  ScoreMaps S;
  S.maxpart = MP;
  S.plan.LEN = m->interval + 1;
  S.plan.scales.assign(S.plan.LEN, 1.0f);
  S.plan.scales[m->interval] = (float)m->sbin / scale;
  S.valid.assign(S.plan.LEN, 0);
  S.valid[m->interval] = 1;
  S.rdims.assign((size_t)NoC * 2, 0);
  S.rdims[jj * 2] = R0;
  S.rdims[jj * 2 + 1] = R1;
  S.pdims.assign((size_t)NoP * 2, 0);
  S.Ix.assign((size_t)NoC * MP, {});
  S.Iy.assign((size_t)NoC * MP, {});
  long long off = 0;
  for (int kk = 0; kk < np; kk++) {
    const int pid = m->pidx[jj * MP + kk];
    const int P0 = pdims[2 * kk], P1 = pdims[2 * kk + 1];
    S.pdims[pid * 2] = P0;
    S.pdims[pid * 2 + 1] = P1;
    S.Ix[(size_t)jj * MP + kk].assign(Ix + off, Ix + off + (long long)P0 * P1);
    S.Iy[(size_t)jj * MP + kk].assign(Iy + off, Iy + off + (long long)P0 * P1);
    off += (long long)P0 * P1;
  }
  std::vector<std::vector<float>> sc((size_t)NoC);
  sc[jj].assign(score, score + (long long)R0 * R1);
  // only component jj has candidates; boxes_from_scores checks every component's GL
  ora_model_desc one = *m;
  std::vector<int> npart(m->numpart, m->numpart + NoC);
  for (int c = 0; c < NoC; c++) npart[c] = np;
  one.numpart = npart.data();
  std::vector<float> rows = boxes_from_scores<float>(&one, S, sc, thresh, GL);
  const long long n = (long long)rows.size() / *GL;
  if (out) std::memcpy(out, rows.data(), sizeof(float) * (size_t)(std::min(n, cap) * *GL));
  return n;
}

long long ora_detect_maps(const ora_model_desc *m, const unsigned char *bgr, int width,
                          int height, int widthStep, int precision, int threads, void **maps,
                          int **map_dims, int *nmaps) {
  int ms = compute_max_scale(m->rsize, m->numcomponent, m->sbin, m->interval, width, height);
  Pyramid P = build_pyramid(bgr, width, height, widthStep, m->sbin, m->interval, ms, threads);
  ScoreMaps S;
  const int NoC = m->numcomponent, NL = P.plan.LEN - m->interval;
  long long total = 0;
  std::vector<int> dims;
  int count = 0;
  auto emit = [&](auto &sc, auto elt) {
    using E = decltype(elt);
    for (int L = 0; L < NL; L++) {
      if (!S.valid[L + m->interval]) continue;
      for (int jj = 0; jj < NoC; jj++) total += (long long)sc[(size_t)L * NoC + jj].size();
    }
    E *buf = (E *)std::malloc(sizeof(E) * std::max<long long>(total, 1));
    long long off = 0;
    for (int L = 0; L < NL; L++) {
      if (!S.valid[L + m->interval]) continue;
      for (int jj = 0; jj < NoC; jj++) {
        auto &v = sc[(size_t)L * NoC + jj];
        std::memcpy(buf + off, v.data(), sizeof(E) * v.size());
        off += (long long)v.size();
        dims.push_back(L + m->interval);
        dims.push_back(jj);
        dims.push_back(S.rdims[((size_t)L * NoC + jj) * 2]);
        dims.push_back(S.rdims[((size_t)L * NoC + jj) * 2 + 1]);
        count++;
      }
    }
    *maps = buf;
  };
  if (precision == 1) {
    std::vector<std::vector<double>> sc;
    compute_scores<double>(m, P, threads, S, sc);
    emit(sc, double(0));
  } else {
    std::vector<std::vector<float>> sc;
    compute_scores<float>(m, P, threads, S, sc);
    emit(sc, float(0));
  }
  int *d = (int *)std::malloc(sizeof(int) * std::max<size_t>(dims.size(), 1));
  std::memcpy(d, dims.data(), sizeof(int) * dims.size());
  *map_dims = d;
  *nmaps = count;
  return total;
}

int ora_detect_ascore(const ora_model_desc *m, const unsigned char *bgr, int width, int height,
                      int widthStep, float thresh, int threads, float *ac) {
  int ms = compute_max_scale(m->rsize, m->numcomponent, m->sbin, m->interval, width, height);
  Pyramid P = build_pyramid(bgr, width, height, widthStep, m->sbin, m->interval, ms, threads);
  ScoreMaps S;
  std::vector<std::vector<float>> sc;
  compute_scores<float>(m, P, threads, S, sc);
  const int NoC = m->numcomponent;
  for (int level = m->interval; level < P.plan.LEN; level++) {
    if (!S.valid[level]) continue;
    const int L = level - m->interval;
    const float scale = (float)m->sbin / S.plan.scales[level];
    for (int jj = 0; jj < NoC; jj++) {
      const int ssize[2] = {S.rdims[((size_t)L * NoC + jj) * 2],
                            S.rdims[((size_t)L * NoC + jj) * 2 + 1]};
      calc_a_score(ac, sc[(size_t)L * NoC + jj].data(), ssize, m->rsize + 2 * jj, height, width,
                   m->padx, m->pady, scale);
    }
  }
  ascore_finish(ac, (long long)width * height, thresh);
  return 0;
}

void ora_calc_a_score(float *ac, int IHEI, int IWID, const float *score, int R0, int R1,
                      const int *rsize, int padx, int pady, float scale) {
  const int ssize[2] = {R0, R1};
  calc_a_score(ac, score, ssize, rsize, IHEI, IWID, padx, pady, scale);
}

void ora_ascore_finish(float *ac, long long n, float thresh) { ascore_finish(ac, n, thresh); }

void ora_new_rects(const ora_model_desc *m, const float *boxes, int num, int GL, int width,
                   int height, float ratio, int *point, int *or_point, int *type, float *scale,
                   float *score) {
  const int np0 = m->numpart[0];
  const int LL = GL - 3;
  const int ML = 1 + 2 * (1 + np0);
  std::vector<long long> loff(m->numcomponent + 1, 0);
  for (int c = 0; c < m->numcomponent; c++)
    loff[c + 1] = loff[c] + 4LL * (1 + 2 * (1 + m->numpart[c]));
  std::vector<float> Avec(ML);
  for (int ii = 0; ii < num; ii++) {
    const float *P = boxes + (long long)GL * ii;
    const int CNUM = (int)P[GL - 3];
    int k = 0;
    Avec[k++] = P[3] - P[1];
    for (int kk = 0; kk < LL; kk += 4) {
      Avec[k++] = P[kk + 1];
      Avec[k++] = P[kk];
    }
    const int MLc = 1 + 2 * (1 + m->numpart[CNUM]);
    const float *x1 = m->lsq + loff[CNUM], *y1 = x1 + MLc, *x2 = y1 + MLc, *y2 = x2 + MLc;
    float XP1 = 0, XP2 = 0, YP1 = 0, YP2 = 0;
    for (int kk = 0; kk < ML; kk++) {
      YP1 += Avec[kk] * y1[kk];
      YP2 += Avec[kk] * y2[kk];
      XP1 += Avec[kk] * x1[kk];
      XP2 += Avec[kk] * x2[kk];
    }
    int PP[4];
    PP[0] = XP1 > 0 ? (int)XP1 : 0;
    PP[1] = YP1 > 0 ? (int)YP1 : 0;
    PP[2] = XP2 < width ? (int)XP2 : width;
    PP[3] = YP2 < height ? (int)YP2 : height;
    for (int q = 0; q < 4; q++) {
      point[ii * 4 + q] = PP[q];
      or_point[ii * 4 + q] = (int)((float)PP[q] / ratio);
    }
    scale[ii] = P[GL - 1];
    score[ii] = P[GL - 2];
    type[ii] = CNUM;
  }
}

void ora_free(void *p) { std::free(p); }

void *ora_load_csv(const char *comp, const char *root, const char *part, ora_model_desc *desc) {
  std::vector<float> c, r, p;
  if (!read_csv_floats(comp, c) || !read_csv_floats(root, r) || !read_csv_floats(part, p))
    return nullptr;
  OwnedModel *M = new OwnedModel();
  size_t k = 0;
  auto next = [&](std::vector<float> &v) -> float {
    if (k >= v.size()) throw 1;
    return v[k++];
  };
  try {
    M->numcomponent = (int)next(c);
    M->sbin = (int)next(c);
    M->interval = (int)next(c);
    M->max_Y = (int)next(c);
    M->max_X = (int)next(c);
    const int NC = M->numcomponent;
    std::vector<std::vector<int>> pidx(NC), didx(NC), psize(NC);
    for (int ii = 0; ii < NC; ii++) {
      M->ridx.push_back((int)next(c) - 1);
      M->oidx.push_back((int)next(c) - 1);
      M->offw.push_back(next(c));
      M->rsize.push_back((int)next(c));
      M->rsize.push_back((int)next(c));
      int np = (int)next(c);
      M->numpart.push_back(np);
      M->maxpart = std::max(M->maxpart, np);
      for (int jj = 0; jj < np; jj++) {
        pidx[ii].push_back((int)next(c) - 1);
        didx[ii].push_back((int)next(c) - 1);
        psize[ii].push_back((int)next(c));
        psize[ii].push_back((int)next(c));
      }
    }
    M->DefL = (int)next(c);
    for (int kk = 0; kk < M->DefL; kk++) {
      for (int q = 0; q < 4; q++) M->def.push_back(next(c));
      M->anchor.push_back((int)next(c));
      M->anchor.push_back((int)next(c));
    }
    for (int ii = 0; ii < NC; ii++) {
      int GLs = 1 + 2 * (1 + M->numpart[ii]);
      for (int q = 0; q < 4 * GLs; q++) M->lsq.push_back(next(c));
    }
    const int MP = std::max(M->maxpart, 1);
    M->maxpart = MP;
    M->pidx.assign((size_t)NC * MP, 0);
    M->didx.assign((size_t)NC * MP, 0);
    M->psize.assign((size_t)NC * MP * 2, 0);
    for (int ii = 0; ii < NC; ii++)
      for (int jj = 0; jj < M->numpart[ii]; jj++) {
        M->pidx[ii * MP + jj] = pidx[ii][jj];
        M->didx[ii * MP + jj] = didx[ii][jj];
        M->psize[(ii * MP + jj) * 2] = psize[ii][jj * 2];
        M->psize[(ii * MP + jj) * 2 + 1] = psize[ii][jj * 2 + 1];
      }
    k = 0;
    M->NoR = (int)next(r);
    for (int ii = 0; ii < M->NoR; ii++) {
      int h = (int)next(r), w = (int)next(r), ch = (int)next(r);
      M->root_size.insert(M->root_size.end(), {h, w, ch});
      M->root_off.push_back((long long)M->rootfilter.size());
      for (int q = 0; q < h * w * ch; q++) M->rootfilter.push_back(next(r));
    }
    k = 0;
    M->NoP = (int)next(p);
    for (int ii = 0; ii < M->NoP; ii++) {
      int h = (int)next(p), w = (int)next(p), ch = (int)next(p);
      M->part_size.insert(M->part_size.end(), {h, w, ch});
      M->part_off.push_back((long long)M->partfilter.size());
      for (int q = 0; q < h * w * ch; q++) M->partfilter.push_back(next(p));
      M->part_partner.push_back((int)next(p));
    }
  } catch (...) {
    delete M;
    return nullptr;
  }
  desc->numcomponent = M->numcomponent;
  desc->sbin = M->sbin;
  desc->interval = M->interval;
  desc->max_Y = M->max_Y;
  desc->max_X = M->max_X;
  desc->padx = 0;  // load_model overrides the computed pads (CT/load_model.cpp:552-553)
  desc->pady = 0;
  desc->maxpart = M->maxpart;
  desc->ridx = M->ridx.data();
  desc->oidx = M->oidx.data();
  desc->offw = M->offw.data();
  desc->rsize = M->rsize.data();
  desc->numpart = M->numpart.data();
  desc->pidx = M->pidx.data();
  desc->didx = M->didx.data();
  desc->psize = M->psize.data();
  desc->DefL = M->DefL;
  desc->def = M->def.data();
  desc->anchor = M->anchor.data();
  desc->NoR = M->NoR;
  desc->root_size = M->root_size.data();
  desc->rootfilter = M->rootfilter.data();
  desc->root_off = M->root_off.data();
  desc->NoP = M->NoP;
  desc->part_size = M->part_size.data();
  desc->partfilter = M->partfilter.data();
  desc->part_off = M->part_off.data();
  desc->lsq = M->lsq.data();
  return M;
}

void ora_model_free(void *h) { delete (OwnedModel *)h; }

}  // extern "C"

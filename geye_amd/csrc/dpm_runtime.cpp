// dpm_runtime.cpp -- context, model residency, the per-frame pipeline and the C ABI.
//
// Replaces the reference's host orchestration: init_cuda/clean_cuda (CT/GPU_init.cpp:50-356),
// load_model (CT/load_model.cpp:541), calc_f_pyramid (CT/featurepyramid.cpp.new2:952),
// fconvsMT_GPU (CT/fconvsMT.cpp:310), dt_GPU (CT/dt_GPU.cpp:763) and get_boxes
// (CT/get_boxes.cpp:583).  Unlike the reference, filters, plans and every intermediate stay
// resident in HBM: a frame crosses PCIe once (in) and only its boxes come back.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <tuple>
#include <memory>
#include <string>

#include "dpm_ascore.h"
#include "dpm_kernels.h"
#include "dpm_mfma.h"
#include "dpm_nms.h"
#include "dpm_rects.h"
#include "dpm_runtime.h"

using namespace geye;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(call)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(GEYE_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));     \
  } while (0)

#define CHK(expr)            \
  do {                       \
    int r_ = (expr);         \
    if (r_ != GEYE_OK) return r_; \
  } while (0)

int ensure(DeviceBuffer &b, size_t bytes) {
  if (bytes == 0) bytes = 256;
  if (b.p && b.bytes >= bytes) return GEYE_OK;
  if (b.p) {
    HIPCHK(hipFree(b.p));
    b.p = nullptr;
    b.bytes = 0;
  }
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    return fail(GEYE_ERR_NOMEM, std::string("hipMalloc(") + std::to_string(bytes) +
                                    "): " + hipGetErrorString(e));
  }
  b.bytes = bytes;
  return GEYE_OK;
}

void release(DeviceBuffer &b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

// Host-to-device copies are issued on the stream the kernels run on and waited for: the
// context's stream does not synchronise with the null stream, and a pageable-memory
// hipMemcpy may return before its DMA has landed.
template <typename T>
int upload(DeviceBuffer &b, const std::vector<T> &v, hipStream_t s) {
  CHK(ensure(b, sizeof(T) * v.size()));
  if (!v.empty()) {
    HIPCHK(hipMemcpyAsync(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  return GEYE_OK;
}

template <typename T>
T *dp(DeviceBuffer &b) {
  return (T *)b.p;
}

// ---------------------------------------------------------------------------------------
// Per-kernel timing: events around launches on the context's stream.
// ---------------------------------------------------------------------------------------
struct KTimer {
  geye_ctx *c;
  std::string name;
  hipStream_t s;
  hipEvent_t a = nullptr, b = nullptr;
  // s: the stream the timed launches go to (the context's by default)
  KTimer(geye_ctx *ctx, std::string n, hipStream_t st = nullptr)
      : c(ctx), name(std::move(n)), s(st ? st : ctx->stream) {
    if (!c->profile || name.empty()) return;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    (void)hipEventRecord(a, s);
  }
  ~KTimer() {
    if (!c->profile || name.empty()) return;
    (void)hipEventRecord(b, s);
    c->pending.push_back({name, {a, b}});
  }
};

void flush_stats(geye_ctx *c) {
  for (auto &p : c->pending) {
    float ms = 0;
    (void)hipEventSynchronize(p.second.second);
    (void)hipEventElapsedTime(&ms, p.second.first, p.second.second);
    auto &st = c->stats[p.first];
    st.ms += ms;
    st.count += 1;
    (void)hipEventDestroy(p.second.first);
    (void)hipEventDestroy(p.second.second);
  }
  c->pending.clear();
}

// ---------------------------------------------------------------------------------------
// Model
// ---------------------------------------------------------------------------------------
int pick_nf(int count) {
  if (count >= 8) return 8;
  if (count >= 4) return 4;
  if (count >= 2) return 2;
  return 1;
}

// float -> bf16 with round-to-nearest-even (finite inputs)
unsigned short to_bf16(float f) {
  unsigned int u;
  std::memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

unsigned short to_f16(float f) {
  _Float16 h = (_Float16)f;
  unsigned short u;
  std::memcpy(&u, &h, 2);
  return u;
}

float from_bf16(unsigned short b) {
  const unsigned int u = (unsigned int)b << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

float from_f16(unsigned short b) {
  _Float16 h;
  std::memcpy(&h, &b, 2);
  return (float)h;
}

// The 16-bit operand of an MFMA form for v: its rounding (BF16, FP16), or for the split forms
// the pair hi = round16(v), lo = round16(v - hi) (v - hi is exact in FP32).
void mfma_operand(float v, int precision, unsigned short &hi, unsigned short &lo) {
  const bool bf = mfma_kind(precision) == 1;
  hi = bf ? to_bf16(v) : to_f16(v);
  lo = 0;
  if (split_precision(precision)) {
    const float h = bf ? from_bf16(hi) : from_f16(hi);
    lo = bf ? to_bf16(v - h) : to_f16(v - h);
  }
}

}  // namespace

// The host side of a model: fields, checks, filter classes and the packed filters (FP32 and
// one 16-bit pack per MFMA form) the device copies are made from.  No device work.
int geye::model_pack(const geye_model_desc *d, Model &m, ModelPacks &pk) {
  if (!d || d->numcomponent <= 0 || d->NoR <= 0 || d->sbin <= 0 || d->interval <= 0)
    return fail(GEYE_ERR_ARG, "model: empty or malformed description");
  m.numcomponent = d->numcomponent;
  m.sbin = d->sbin;
  m.interval = d->interval;
  m.max_Y = d->max_Y;
  m.max_X = d->max_X;
  m.padx = d->padx;
  m.pady = d->pady;
  m.maxpart = std::max(d->maxpart, 1);
  m.NoR = d->NoR;
  m.NoP = d->NoP;
  m.DefL = d->DefL;
  const int C = m.numcomponent, MP = m.maxpart;
  m.ridx.assign(d->ridx, d->ridx + C);
  m.oidx.assign(d->oidx, d->oidx + C);
  m.offw.assign(d->offw, d->offw + C);
  m.rsize.assign(d->rsize, d->rsize + 2 * C);
  m.numpart.assign(d->numpart, d->numpart + C);
  m.pidx.assign(C * MP, 0);
  m.didx.assign(C * MP, 0);
  m.psize.assign(C * MP * 2, 0);
  for (int i = 0; i < C; i++) {
    if (m.numpart[i] > MP || m.numpart[i] < 0) return fail(GEYE_ERR_ARG, "model: numpart > maxpart");
    if (m.ridx[i] < 0 || m.ridx[i] >= C) return fail(GEYE_ERR_ARG, "model: ridx out of range");
    for (int k = 0; k < m.numpart[i]; k++) {
      m.pidx[i * MP + k] = d->pidx[i * d->maxpart + k];
      m.didx[i * MP + k] = d->didx[i * d->maxpart + k];
      m.psize[(i * MP + k) * 2] = d->psize[(i * d->maxpart + k) * 2];
      m.psize[(i * MP + k) * 2 + 1] = d->psize[(i * d->maxpart + k) * 2 + 1];
      if (m.pidx[i * MP + k] < 0 || m.pidx[i * MP + k] >= m.NoP)
        return fail(GEYE_ERR_ARG, "model: pidx out of range");
      if (m.didx[i * MP + k] < 0 || m.didx[i * MP + k] >= m.DefL)
        return fail(GEYE_ERR_ARG, "model: didx out of range");
    }
  }
  // each part filter may be used by one (component, part) pair: the DT negates per use
  std::vector<int> uses(m.NoP, 0);
  for (int i = 0; i < C; i++)
    for (int k = 0; k < m.numpart[i]; k++)
      if (++uses[m.pidx[i * MP + k]] > 1)
        return fail(GEYE_ERR_UNSUPPORTED, "model: a part filter shared by two parts");
  m.def.assign(d->def, d->def + 4 * m.DefL);
  m.anchor.assign(d->anchor, d->anchor + 2 * m.DefL);
  m.root_size.assign(d->root_size, d->root_size + 3 * m.NoR);
  m.part_size.assign(d->part_size, d->part_size + 3 * m.NoP);
  m.rootf.resize(m.NoR);
  m.partf.resize(m.NoP);
  for (int r = 0; r < m.NoR; r++) {
    const int *s = &m.root_size[3 * r];
    if (s[2] != 31 || s[0] < 1 || s[1] < 1 || s[0] > 16 || s[1] > 16)
      return fail(GEYE_ERR_UNSUPPORTED, "model: root filters must be h,w <= 16 with 31 channels");
    m.rootf[r].assign(d->rootfilter + d->root_off[r], d->rootfilter + d->root_off[r] + 31LL * s[0] * s[1]);
  }
  for (int p = 0; p < m.NoP; p++) {
    const int *s = &m.part_size[3 * p];
    if (s[2] != 31 || s[0] < 1 || s[1] < 1 || s[0] > 16 || s[1] > 16)
      return fail(GEYE_ERR_UNSUPPORTED, "model: part filters must be h,w <= 16 with 31 channels");
    m.partf[p].assign(d->partfilter + d->part_off[p], d->partfilter + d->part_off[p] + 31LL * s[0] * s[1]);
  }
  for (int i = 0; i < C; i++)
    if (m.rsize[2 * i] != m.root_size[3 * i] || m.rsize[2 * i + 1] != m.root_size[3 * i + 1])
      return fail(GEYE_ERR_ARG, "model: rsize disagrees with the root filter size");
  // filter classes: all roots together at their bounding size, parts grouped by shape
  m.classes.clear();
  {
    Model::FilterClass rc;
    rc.root = true;
    rc.h = rc.w = 0;
    for (int r = 0; r < m.NoR; r++) {
      rc.members.push_back(r);
      rc.msize.push_back(int2{m.root_size[3 * r], m.root_size[3 * r + 1]});
      rc.h = std::max(rc.h, m.root_size[3 * r]);
      rc.w = std::max(rc.w, m.root_size[3 * r + 1]);
    }
    m.classes.push_back(rc);
  }
  {
    std::vector<bool> done(m.NoP, false);
    for (int i = 0; i < m.NoP; i++) {
      if (done[i]) continue;
      Model::FilterClass pc;
      pc.root = false;
      pc.h = m.part_size[3 * i];
      pc.w = m.part_size[3 * i + 1];
      for (int j = i; j < m.NoP; j++)
        if (!done[j] && m.part_size[3 * j] == pc.h && m.part_size[3 * j + 1] == pc.w) {
          pc.members.push_back(j);
          pc.msize.push_back(int2{pc.h, pc.w});
          done[j] = true;
        }
      m.classes.push_back(pc);
    }
  }
  std::vector<float> &packed = pk.f32;
  packed.clear();
  for (auto &h : pk.h16) h.clear();
  for (Model::FilterClass &fc : m.classes) {
    const int nmem = (int)fc.members.size();
    fc.nf = pick_nf(nmem);
    fc.nt = std::min(3, (nmem + 15) / 16);
    const int g32 = (nmem + fc.nf - 1) / fc.nf, g16 = (nmem + 16 * fc.nt - 1) / (16 * fc.nt);
    fc.nslots = std::max(g32 * fc.nf, g16 * 16 * fc.nt);
    auto weight = [&](int member, int c, int xp, int yp) -> float {
      if (member >= nmem || c >= 31) return 0.0f;
      const int fid = fc.members[member];
      const int mh = fc.msize[member].x, mw = fc.msize[member].y;
      if (xp >= mw || yp >= mh) return 0.0f;
      const std::vector<float> &src = fc.root ? m.rootf[fid] : m.partf[fid];
      return src[(size_t)c * mh * mw + (size_t)xp * mh + yp];
    };
    fc.off = (long long)packed.size();
    for (int g = 0; g < g32; g++)
      for (int f = 0; f < 31; f++)
        for (int xp = 0; xp < fc.w; xp++)
          for (int yp = 0; yp < fc.h; yp++)
            for (int n = 0; n < fc.nf; n++) packed.push_back(weight(g * fc.nf + n, f, xp, yp));
    // [group][xp][yp][nt][chunk][16 filters][8]: 4 chunks of 8 channels, or 8 in the split
    // forms (the hi halves, then the lo halves); the split packs' offsets are twice hoff
    fc.hoff = (long long)pk.h16[GEYE_BF16].size();
    for (int p = GEYE_BF16; p <= GEYE_FP16X3; p++) {
      const int nh = split_precision(p) ? 2 : 1;
      for (int g = 0; g < g16; g++)
        for (int xp = 0; xp < fc.w; xp++)
          for (int yp = 0; yp < fc.h; yp++)
            for (int t = 0; t < fc.nt; t++)
              for (int h = 0; h < nh; h++)
                for (int ch = 0; ch < 4; ch++)
                  for (int n = 0; n < 16; n++)
                    for (int j = 0; j < 8; j++) {
                      unsigned short hi, lo;
                      mfma_operand(weight(g * 16 * fc.nt + t * 16 + n, ch * 8 + j, xp, yp), p,
                                   hi, lo);
                      pk.h16[p].push_back(h ? lo : hi);
                    }
    }
  }
  m.lsq.clear();
  m.lsq_idx.clear();
  if (d->lsq) {
    m.lsq_idx.assign(2 * C, 0);
    for (int i = 0; i < C; i++) {
      const int ml = 1 + 2 * (1 + m.numpart[i]);
      m.lsq_idx[i] = (int)m.lsq.size();
      m.lsq_idx[C + i] = ml;
      m.lsq.insert(m.lsq.end(), d->lsq + m.lsq.size(), d->lsq + m.lsq.size() + 4 * ml);
    }
  }
  return GEYE_OK;
}

namespace {

int model_from_desc(geye_ctx *ctx, const geye_model_desc *d, Model &m) {
  ModelPacks pk;
  CHK(model_pack(d, m, pk));
  m.ctx = ctx;
  for (int p = GEYE_BF16; p <= GEYE_FP16X3; p++) CHK(upload(m.d_filt16[p], pk.h16[p], ctx->stream));
  CHK(upload(m.d_filters, pk.f32, ctx->stream));
  CHK(upload(m.d_rsize, m.rsize, ctx->stream));
  if (!m.lsq.empty()) {
    CHK(upload(m.d_lsq, m.lsq, ctx->stream));
    CHK(upload(m.d_lsq_idx, m.lsq_idx, ctx->stream));
  }
  return GEYE_OK;
}

// Every number of a model CSV file (CT/load_model.cpp:57-530 reads them with fscanf, commas
// and white space between); false when the file cannot be read or holds a non-number.
bool read_csv(const char *path, std::vector<float> &v) {
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

// CSV model format of CT/load_model.cpp:57-558; the pads are zero as load_model forces
// (CT/load_model.cpp:552-553).
int load_csv(geye_ctx *ctx, const char *comp, const char *root, const char *part, Model &m) {
  std::vector<float> c, r, p;
  if (!read_csv(comp, c)) return fail(GEYE_ERR_IO, std::string("cannot read ") + comp);
  if (!read_csv(root, r)) return fail(GEYE_ERR_IO, std::string("cannot read ") + root);
  if (!read_csv(part, p)) return fail(GEYE_ERR_IO, std::string("cannot read ") + part);
  size_t k = 0;
  bool bad = false;
  auto next = [&](const std::vector<float> &v) -> float {
    if (k >= v.size()) { bad = true; return 0.f; }
    return v[k++];
  };
  geye_model_desc d{};
  d.numcomponent = (int)next(c);
  d.sbin = (int)next(c);
  d.interval = (int)next(c);
  d.max_Y = (int)next(c);
  d.max_X = (int)next(c);
  const int C = d.numcomponent;
  if (bad || C <= 0 || C > 1024) return fail(GEYE_ERR_IO, "component file malformed");
  std::vector<int> ridx(C), oidx(C), rsize(2 * C), numpart(C);
  std::vector<float> offw(C);
  std::vector<std::vector<int>> pi(C), di(C), ps(C);
  int MP = 1;
  for (int i = 0; i < C && !bad; i++) {
    ridx[i] = (int)next(c) - 1;
    oidx[i] = (int)next(c) - 1;
    offw[i] = next(c);
    rsize[2 * i] = (int)next(c);
    rsize[2 * i + 1] = (int)next(c);
    numpart[i] = (int)next(c);
    if (numpart[i] < 0 || numpart[i] > 4096) { bad = true; break; }
    MP = std::max(MP, numpart[i]);
    for (int j = 0; j < numpart[i]; j++) {
      pi[i].push_back((int)next(c) - 1);
      di[i].push_back((int)next(c) - 1);
      ps[i].push_back((int)next(c));
      ps[i].push_back((int)next(c));
    }
  }
  const int DefL = (int)next(c);
  if (bad || DefL < 0 || DefL > 1 << 20) return fail(GEYE_ERR_IO, "component file malformed");
  std::vector<float> def(4 * DefL);
  std::vector<int> anchor(2 * DefL);
  for (int q = 0; q < DefL; q++) {
    for (int t = 0; t < 4; t++) def[4 * q + t] = next(c);
    anchor[2 * q] = (int)next(c);
    anchor[2 * q + 1] = (int)next(c);
  }
  // least-squares box coefficients: per component x1, y1, x2, y2 of 1 + 2(1 + np) values
  std::vector<float> lsq;
  for (int i = 0; i < C && !bad; i++)
    for (int q = 0; q < 4 * (1 + 2 * (1 + numpart[i])); q++) lsq.push_back(next(c));
  if (bad) return fail(GEYE_ERR_IO, "component file truncated");
  std::vector<int> pidx(C * MP, 0), didx(C * MP, 0), psize(C * MP * 2, 0);
  for (int i = 0; i < C; i++)
    for (int j = 0; j < numpart[i]; j++) {
      pidx[i * MP + j] = pi[i][j];
      didx[i * MP + j] = di[i][j];
      psize[(i * MP + j) * 2] = ps[i][2 * j];
      psize[(i * MP + j) * 2 + 1] = ps[i][2 * j + 1];
    }
  auto filters = [&](const std::vector<float> &v, bool partner, int &N, std::vector<int> &size,
                     std::vector<float> &w, std::vector<long long> &off) -> bool {
    k = 0;
    N = (int)next(v);
    if (bad || N < 0 || N > 1 << 16) return false;
    for (int i = 0; i < N; i++) {
      int h = (int)next(v), ww = (int)next(v), ch = (int)next(v);
      if (bad || h <= 0 || ww <= 0 || ch <= 0) return false;
      size.insert(size.end(), {h, ww, ch});
      off.push_back((long long)w.size());
      for (long long q = 0; q < (long long)h * ww * ch; q++) w.push_back(next(v));
      if (partner) (void)next(v);
      if (bad) return false;
    }
    return true;
  };
  int NoR = 0, NoP = 0;
  std::vector<int> rs, pss;
  std::vector<float> rw, pw;
  std::vector<long long> ro, po;
  if (!filters(r, false, NoR, rs, rw, ro)) return fail(GEYE_ERR_IO, "root filter file malformed");
  if (!filters(p, true, NoP, pss, pw, po)) return fail(GEYE_ERR_IO, "part filter file malformed");
  d.padx = 0;
  d.pady = 0;
  d.maxpart = MP;
  d.ridx = ridx.data();
  d.oidx = oidx.data();
  d.offw = offw.data();
  d.rsize = rsize.data();
  d.numpart = numpart.data();
  d.pidx = pidx.data();
  d.didx = didx.data();
  d.psize = psize.data();
  d.DefL = DefL;
  d.def = def.data();
  d.anchor = anchor.data();
  d.NoR = NoR;
  d.root_size = rs.data();
  d.rootfilter = rw.data();
  d.root_off = ro.data();
  d.NoP = NoP;
  d.part_size = pss.data();
  d.partfilter = pw.data();
  d.part_off = po.data();
  d.lsq = lsq.data();
  return model_from_desc(ctx, &d, m);
}

// ---------------------------------------------------------------------------------------
// Plans: built once per frame size and kept with the model.
// ---------------------------------------------------------------------------------------
int plan_upload(Plan &P, hipStream_t st) {
  std::vector<ResizeOp> ops;
  for (auto &R : P.rounds) {
    ops.insert(ops.end(), R.rows.begin(), R.rows.end());
    ops.insert(ops.end(), R.cols.begin(), R.cols.end());
  }
  CHK(upload(P.d_resize_ops, ops, st));
  CHK(upload(P.d_resize_tab, P.resize_tab, st));
  CHK(upload(P.d_resize_si, P.resize_si, st));
  CHK(upload(P.d_resize_alpha, P.resize_alpha, st));
  // per destination index, its (at most 4) taps in one 16-byte load each
  P.resize_si4.assign(P.resize_tab.size(), int4{0, 0, 0, 0});
  P.resize_al4.assign(P.resize_tab.size(), float4{0.f, 0.f, 0.f, 0.f});
  for (size_t k = 0; k < P.resize_tab.size(); k++) {
    const int2 t = P.resize_tab[k];
    int *si4 = &P.resize_si4[k].x;
    float *al4 = &P.resize_al4[k].x;
    for (int q = 0; q < std::min(t.y, 4); q++) {
      si4[q] = P.resize_si[t.x + q];
      al4[q] = P.resize_alpha[t.x + q];
    }
  }
  CHK(upload(P.d_resize_si4, P.resize_si4, st));
  CHK(upload(P.d_resize_al4, P.resize_al4, st));
  CHK(upload(P.d_hog, P.hog, st));
  CHK(upload(P.d_conv_tiles, P.conv_tiles, st));
  CHK(upload(P.d_mfma_tiles, P.mfma_tiles, st));
  CHK(upload(P.d_mfma_tiles8, P.mfma_tiles8, st));
  CHK(upload(P.d_map_dims, P.map_dims, st));
  CHK(upload(P.d_map_off, P.map_off, st));
  CHK(upload(P.d_dt1, P.dt_pass1, st));
  CHK(upload(P.d_dt2, P.dt_pass2, st));
  CHK(upload(P.d_dt1w, P.dt_pass1w, st));
  CHK(upload(P.d_dt2w, P.dt_pass2w, st));
  CHK(upload(P.d_nodes, P.dt_nodes, st));
  CHK(upload(P.d_depths, P.dt_depths, st));
  CHK(upload(P.d_smaps, P.smaps, st));
  CHK(upload(P.d_sparts, P.sparts, st));
  CHK(upload(P.d_ascore_tab, P.ascore_tab, st));
  CHK(upload(P.d_grad, P.grad, st));
  {
    std::vector<FusedResize> fl;
    for (auto &v : P.fused) fl.insert(fl.end(), v.begin(), v.end());
    CHK(upload(P.d_fused, fl, st));
  }
  CHK(upload(P.d_qtiles0, P.q_tiles[0], st));
  CHK(upload(P.d_qtiles1, P.q_tiles[1], st));
  CHK(ensure(P.d_qcounter, 256));
  // the pyramid's arenas (images, gradients, histograms, features) are allocated by the
  // first pyramid run: a plan that convolves another plan's pyramid never needs them; the
  // fused resize keeps the row pass in LDS, so there is no intermediate image
  CHK(ensure(P.d_conv, sizeof(float) * P.conv_total));
  CHK(ensure(P.d_tmpM, sizeof(float) * P.dt_total));
  CHK(ensure(P.d_tmpIy, sizeof(int) * P.dt_total));
  CHK(ensure(P.d_M, sizeof(float) * P.dt_total));
  CHK(ensure(P.d_tmpIx, sizeof(int) * P.dt_total));
  CHK(ensure(P.d_score, sizeof(float) * P.score_total));
  CHK(ensure(P.d_flags, sizeof(int) * P.score_total));
  CHK(ensure(P.d_blocksum, sizeof(int) * (compact_blocks(P.score_total) + 1)));
  CHK(ensure(P.d_cand, sizeof(int) * P.score_total));
  CHK(ensure(P.d_count, sizeof(int) * 4));
  return GEYE_OK;
}

std::atomic<unsigned long long> g_plan_gen{0};

// Drops the plan's captured graphs (a buffer they use moved, or the plan goes); the new
// generation also retires every batch graph captured against the plan.
void drop_graphs(Plan &P) {
  for (auto &kv : P.graphs) (void)hipGraphExecDestroy(kv.second);
  P.graphs.clear();
  P.gen = ++g_plan_gen;
}

void plan_release(Plan &P) {
  drop_graphs(P);
  for (auto *v : {&P.b_small, &P.b_mask, &P.b_blocksum, &P.b_sort, &P.b_grid, &P.b_acount})
    for (DeviceBuffer &b : *v) release(b);
  for (hipEvent_t e : P.b_ev) (void)hipEventDestroy(e);
  P.b_ev.clear();
  if (P.b_host) (void)hipHostFree(P.b_host);
  P.b_host = P.b_dev = nullptr;
  for (DeviceBuffer *b : {&P.d_resize_ops, &P.d_resize_tab, &P.d_resize_si, &P.d_resize_alpha,
                          &P.d_resize_si4, &P.d_resize_al4,
                          &P.d_hog, &P.d_conv_tiles, &P.d_mfma_tiles, &P.d_mfma_tiles8, &P.d_map_dims,
                          &P.d_map_off, &P.d_dt1, &P.d_dt2, &P.d_dt1w, &P.d_dt2w,
                          &P.d_nodes, &P.d_depths, &P.d_smaps, &P.d_sparts, &P.d_frame,
                          &P.d_img, &P.d_tmp, &P.d_hist, &P.d_norm, &P.d_feat, &P.d_featb,
                          &P.d_conv, &P.d_tmpM, &P.d_tmpIy, &P.d_M, &P.d_tmpIx, &P.d_score,
                          &P.d_flags, &P.d_blocksum, &P.d_cand, &P.d_count, &P.d_rows,
                          &P.n_box, &P.n_keys, &P.n_keys2, &P.n_sbox, &P.n_alive0, &P.n_alive1,
                          &P.n_status, &P.n_fsup, &P.n_picks, &P.n_counters, &P.n_mask,
                          &P.n_sort, &P.n_flags, &P.n_keep_orig,
                          &P.n_kept, &P.n_total, &P.n_vals, &P.n_ord,
                          &P.n_rank, &P.n_dec, &P.n_alist, &P.n_acount, &P.n_grid, &P.n_tie,
                          &P.d_ascore_tab,
                          &P.d_ascore,
                          &P.d_grad, &P.d_grid,
                          &P.d_qtiles0, &P.d_qtiles1, &P.d_qcounter, &P.d_fused})
    release(*b);
}

}  // namespace

// ---------------------------------------------------------------------------------------
// Plan registry.  A plan (tables + every per-frame arena) belongs to one (context, model,
// frame geometry): contexts that run concurrently (serving slots) never share arenas,
// while the model's filters are shared read-only.
// ---------------------------------------------------------------------------------------
namespace geye {
namespace {
std::mutex g_plan_mu;
std::map<std::tuple<const geye_ctx *, const Model *, long long>, Plan *> g_plans;
}  // namespace

int set_error(int code, const std::string &msg) { return fail(code, msg); }

long long plan_key(int W, int H, int max_scale) {
  return ((long long)std::max(max_scale, 0) << 40) | ((long long)W << 20) | (long long)H;
}

// max_scale > 0 pins the pyramid depth (ini_scales computes it once per model from the
// first frame, CT/featurepyramid.cpp.new2:179-226); otherwise it follows this frame.
int get_plan(geye_ctx *c, Model &m, int W, int H, Plan **out, int max_scale) {
  const auto key = std::make_tuple((const geye_ctx *)c, (const Model *)&m, plan_key(W, H, max_scale));
  {
    std::lock_guard<std::mutex> lk(g_plan_mu);
    auto it = g_plans.find(key);
    if (it != g_plans.end()) {
      *out = it->second;
      return GEYE_OK;
    }
  }
  std::unique_ptr<Plan> P(new Plan());
  P->gen = ++g_plan_gen;
  std::string err;
  int rc = build_plan(m, W, H, *P, err, max_scale > 0 ? max_scale : -1);
  if (rc != GEYE_OK) return fail(rc, err);
  rc = plan_upload(*P, c->stream);
  if (rc != GEYE_OK) {
    plan_release(*P);
    return rc;
  }
  *out = P.release();
  std::lock_guard<std::mutex> lk(g_plan_mu);
  g_plans[key] = *out;
  return GEYE_OK;
}

// The plan of a batch of nb frames for model m (plan_key with nb and the feature stride's
// owner depth in its high bits): the one-frame plan's descriptors repeated per frame, plus
// the resources of nb concurrent NMS jobs.  feat_len > 0 lays each frame's features out as a
// model with feat_len levels would (the deepest model of a set sharing one pyramid).
int get_batch_plan(geye_ctx *c, Model &m, int W, int H, int nb, int feat_len, Plan **out) {
  const long long key = plan_key(W, H, 0) | ((long long)nb << 40) | ((long long)feat_len << 52) |
                        (1LL << 62);
  const auto k = std::make_tuple((const geye_ctx *)c, (const Model *)&m, key);
  {
    std::lock_guard<std::mutex> lk(g_plan_mu);
    auto it = g_plans.find(k);
    if (it != g_plans.end()) {
      *out = it->second;
      return GEYE_OK;
    }
  }
  Plan one;
  std::string err;
  int rc = build_plan(m, W, H, one, err);
  if (rc != GEYE_OK) return fail(rc, err);
  long long fs = 0, fbs = 0;
  if (feat_len > 0 && feat_len != one.LEN) {
    // the deeper model's one-frame feature sizes (this model's levels are a prefix of its):
    // every level with cells, in level order, as the plan lays them out
    std::vector<Level> lv;
    std::vector<int> ih, iw, isrc;
    std::vector<float> isc;
    level_geometry(feat_len - m.interval, m.sbin, m.interval, W, H, lv, ih, iw, isrc, isc);
    for (const Level &L : lv)
      if (L.F0 > 0 && L.F1 > 0) {
        fs += 31LL * L.F0 * L.F1;
        fbs += 32LL * L.F0 * L.F1;
      }
  }
  std::unique_ptr<Plan> P(new Plan());
  P->gen = ++g_plan_gen;
  replicate_plan(one, nb, fs, fbs, *P);
  rc = plan_upload(*P, c->stream);
  const long long cells = std::max<long long>(one.score_total, 1);
  P->b_small.resize(nb);
  P->b_mask.resize(nb);
  P->b_blocksum.resize(nb);
  P->b_sort.resize(nb);
  P->b_grid.resize(nb);
  P->b_acount.resize(nb);
  for (int f = 0; f < nb && rc == GEYE_OK; f++) {
    rc = ensure(P->b_small[f], 64 + sizeof(float) * (kNmsTieCap + kNmsResolvedCap));
    if (rc == GEYE_OK)
      rc = ensure(P->b_mask[f], sizeof(unsigned long long) * kNmsWindow * (kNmsWindow / 64));
    if (rc == GEYE_OK) rc = ensure(P->b_blocksum[f], sizeof(int) * (compact_blocks(cells) + 1));
    if (rc == GEYE_OK) rc = ensure(P->b_sort[f], nms_sort_temp_bytes((int)cells));
    for (int e = 0; e < 2 && rc == GEYE_OK; e++) {
      hipEvent_t ev = nullptr;
      if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess)
        rc = fail(GEYE_ERR_HIP, "batch events");
      else
        P->b_ev.push_back(ev);
    }
  }
  if (rc == GEYE_OK &&
      (hipHostMalloc((void **)&P->b_host, sizeof(int) * (4 * nb + nb + 1),
                     hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess ||
       hipHostGetDevicePointer((void **)&P->b_dev, P->b_host, 0) != hipSuccess))
    rc = fail(GEYE_ERR_HIP, "batch host slots");
  if (rc != GEYE_OK) {
    plan_release(*P);
    return rc;
  }
  *out = P.release();
  std::lock_guard<std::mutex> lk(g_plan_mu);
  g_plans[k] = *out;
  return GEYE_OK;
}

// Device bytes a plan for this geometry holds once the frame has run, with the NMS buffers
// at their largest (every score cell a candidate).  Host-side: builds the tables only.
int plan_bytes(const Model &m, int W, int H, long long *bytes) {
  Plan P;
  std::string err;
  int rc = build_plan(const_cast<Model &>(m), W, H, P, err);
  if (rc != GEYE_OK) return fail(rc, err);
  long long tmp = 0;
  for (auto &R : P.rounds) tmp = std::max(tmp, R.tmp_total);
  const long long N = std::max<long long>(P.score_total, 1);
  (void)tmp;
  long long b = 4 * (P.img_total + P.hist_total + P.norm_total + P.feat_total + P.conv_total +
                     4 * P.dt_total + 3 * P.score_total) +
                4 * P.featb_total + 5 * P.grid_total + (long long)P.W * P.H * (3 + 4) +
                N * (2 * (long long)sizeof(NmsBox) + 16 * 4) + (long long)P.GL * 4 * N;
  *bytes = b;
  return GEYE_OK;
}

// Releases the plans matching (c, m, key); a null c, m or key < 0 matches any.
void release_plans(const geye_ctx *c, const Model *m, long long key) {
  std::vector<Plan *> gone;
  {
    std::lock_guard<std::mutex> lk(g_plan_mu);
    for (auto it = g_plans.begin(); it != g_plans.end();) {
      const bool hit = (!c || std::get<0>(it->first) == c) && (!m || std::get<1>(it->first) == m) &&
                       (key < 0 || std::get<2>(it->first) == key);
      if (hit) {
        geye_ctx *owner = const_cast<geye_ctx *>(std::get<0>(it->first));
        if (owner->last_plan == it->second) owner->last_plan = nullptr;
        gone.push_back(it->second);
        it = g_plans.erase(it);
      } else {
        ++it;
      }
    }
  }
  for (Plan *P : gone) {
    plan_release(*P);
    delete P;
  }
}
}  // namespace geye

namespace {

// ---------------------------------------------------------------------------------------
// Pipeline stages
// ---------------------------------------------------------------------------------------
// The DT form a detection runs: the divide and conquer (the reference's order of work) for
// FP32, the windowed minimum for the 16-bit forms, unless geye_dt_form chose one for all.
// Both give the same outputs; the windowed one writes them as half grids.
bool dt_windowed(const geye_ctx *c, int precision) {
  return c->dt_form == GEYE_DT_WINDOW || (c->dt_form == GEYE_DT_AUTO && precision != GEYE_FP32);
}

int run_pyramid(geye_ctx *c, Plan &P, const unsigned char *d_bgr, int ws, int hkind) {
  hipStream_t s = c->stream;
  CHK(ensure(P.d_img, sizeof(float) * std::max<long long>(P.img_total, 1)));
  CHK(ensure(P.d_grid, 5 * std::max<long long>(P.grid_total, 1)));
  CHK(ensure(P.d_hist, sizeof(float) * P.hist_total));
  CHK(ensure(P.d_norm, sizeof(float) * P.norm_total));
  if (!hkind) CHK(ensure(P.d_feat, sizeof(float) * P.feat_total));
  // the frame itself is never copied to float planes: the first resize round and its
  // gradients read its BGR bytes
  int fbase = 0;
  for (size_t r = 0; r < P.fused.size(); r++) {
    KTimer t(c, "resize");
    // round 0 resizes the frame itself: it reads the BGR bytes
    launch_resize_fused(s, dp<FusedResize>(P.d_fused) + fbase, (int)P.fused[r].size(),
                        P.fused_blocks[r], dp<int2>(P.d_resize_tab), dp<int4>(P.d_resize_si4),
                        dp<float4>(P.d_resize_al4), dp<float>(P.d_img), dp<float>(P.d_img),
                        r == 0 ? d_bgr : nullptr, ws, (long long)ws * P.H);
    fbase += (int)P.fused[r].size();
  }
  const HogLevel *hog = dp<HogLevel>(P.d_hog);
  {
    KTimer t(c, "hog_grad");
    launch_hog_grad(s, dp<GradImg>(P.d_grad), (int)P.grad.size(), P.grad_blocks,
                    dp<float>(P.d_img), P.d_grid.p, P.grid_total, d_bgr, ws,
                    (long long)ws * P.H);
  }
  if (P.hog16_n + P.hog8_n > 0) {
    KTimer t(c, "hog_hist");
    launch_hog_votes(s, hog, P.hog16_n + P.hog8_n, P.hog_hist_blocks + P.hog8_blocks,
                     P.d_grid.p, P.grid_total, dp<float>(P.d_hist), dp<float>(P.d_norm));
  }
  // the MFMA forms' channel-last copy: 32 16-bit channels per cell, 64 in the split forms.
  // Sized for the split forms whatever the form, so that switching forms never moves it
  // under a captured graph; should it move all the same, the plan's graphs are dropped.
  if (hkind) {
    const void *old = P.d_featb.p;
    CHK(ensure(P.d_featb, 4 * P.featb_total));
    if (old && P.d_featb.p != old) drop_graphs(P);
  }
  if (P.hogf_n > 0) {
    KTimer t(c, "hog_feat");
    // the 16-bit forms convolve the channel-last copy only, so the FP32 features are not
    // written for them
    launch_hog_feat(s, hog + P.hog16_n + P.hog8_n, P.hogf_n, P.hog_feat_blocks,
                    dp<float>(P.d_hist), dp<float>(P.d_norm),
                    hkind ? nullptr : dp<float>(P.d_feat), hkind ? P.d_featb.p : nullptr,
                    hkind);
  }
  HIPCHK(hipGetLastError());
  return GEYE_OK;
}

int read_count(geye_ctx *c, Plan &P, int *ncand, int slot = kHostCandidates);

std::mutex g_tune_mu;
std::map<std::tuple<const Model *, long long, int>, int> g_tuned;  // autotuned MFMA schedules
// held while a schedule is tuned: one context tunes, the others (a server's workers starting
// together) wait and take its result, so the candidates are timed without each other's
// kernels on the device
std::mutex g_tune_run_mu;

// MFMA convolution schedule `choice`: 0 = one launch per filter class (4 x 64 tiles), else
// the work-queue launch over every class with 16 * MT row tiles and `bpc` workgroups per CU.
int run_mfma(geye_ctx *c, Model &m, Plan &P, const Plan &pyr, int precision, int choice,
             bool timed) {
  hipStream_t s = c->stream;
  const void *filt = m.d_filt16[precision].p;
  static const char *const tags[5] = {"", "conv_bf16_", "conv_fp16_", "conv_bf16x3_", "conv_fp16x3_"};
  const std::string tag = tags[precision];
  const int kind = mfma_kind(precision);
  const bool split = split_precision(precision);
  if (choice == 0) {
    for (auto &mc : P.mfma_classes) {
      if (mc.tile_count == 0) continue;
      KTimer t(c, timed ? tag + std::to_string(mc.h) + "x" + std::to_string(mc.w) : std::string());
      if (launch_conv_mfma(s, kind, split, mc.nt, dp<MfmaTile>(P.d_mfma_tiles) + mc.tile_begin,
                           mc.tile_count, dp<MfmaTile>(P.d_mfma_tiles8) + mc.tile8_begin,
                           mc.tile8_count, mc.h, mc.w, pyr.d_featb.p, filt,
                           dp<long long>(P.d_map_off), dp<int2>(P.d_map_dims),
                           dp<float>(P.d_conv), c->conv_eighths) != 0)
        return fail(GEYE_ERR_UNSUPPORTED, "filter class too large for the MFMA kernel's LDS");
    }
    return GEYE_OK;
  }
  const int mi = (choice - 1) / 8, bpc = (choice - 1) % 8 + 1;
  const int MT = mi == 0 ? 2 : 4;
  const DeviceBuffer &qt = mi == 0 ? P.d_qtiles0 : P.d_qtiles1;
  launch_fill(s, dp<int>(P.d_qcounter), 1, 0);
  KTimer t(c, timed ? tag + "queue" : std::string());
  if (launch_conv_mfma_queue(s, kind, split, MT, c->num_cu * bpc, P.q_lds[mi] * (split ? 2 : 1),
                             (const MfmaTile *)qt.p,
                             (int)P.q_tiles[mi].size(), dp<int>(P.d_qcounter), pyr.d_featb.p, filt,
                             dp<long long>(P.d_map_off), dp<int2>(P.d_map_dims),
                             dp<float>(P.d_conv)) != 0)
    return fail(GEYE_ERR_UNSUPPORTED, "filter classes too large for the work-queue kernel's LDS");
  return GEYE_OK;
}

// Autotuner: times every admissible schedule on this plan's own features (the pyramid of
// the frame being processed is already in HBM) and keeps the fastest.  All schedules give
// the same bits, so the choice only changes speed.  Runs once per plan and precision,
// before any graph capture.
int tune_mfma(geye_ctx *c, Model &m, Plan &P, const Plan &pyr, int precision) {
  std::vector<int> cands{0};
  const size_t qscale = split_precision(precision) ? 2 : 1;  // the split forms' LDS per tile
  for (int mi = 0; mi < 2; mi++)
    for (int bpc = 1; bpc <= 4; bpc++)
      if (!P.q_tiles[mi].empty() && P.q_lds[mi] * qscale * bpc <= 159 * 1024)
        cands.push_back(1 + mi * 8 + bpc - 1);
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  const bool prof = c->profile;
  c->profile = false;
  float best = 1e30f;
  int pick = 0;
  P.mfma_tuning[precision].clear();
  for (int cand : cands) {
    int rc = run_mfma(c, m, P, pyr, precision, cand, false);  // warm-up (attributes, caches)
    // the fastest of three runs: a run that shared the device with other work (another
    // context's kernels still in flight) is not the schedule's own time
    float ms = 1e30f;
    for (int k = 0; k < 3 && rc == GEYE_OK; k++) {
      float t = 0;
      if (hipEventRecord(e0, c->stream) != hipSuccess) {
        rc = GEYE_ERR_HIP;
        break;
      }
      rc = run_mfma(c, m, P, pyr, precision, cand, false);
      (void)hipEventRecord(e1, c->stream);
      (void)hipEventSynchronize(e1);
      (void)hipEventElapsedTime(&t, e0, e1);
      ms = std::min(ms, t);
    }
    if (rc != GEYE_OK) continue;
    const int mi = (cand - 1) / 8, bpc = (cand - 1) % 8 + 1;
    P.mfma_tuning[precision].push_back(
        {cand == 0 ? std::string("per-class 4x64") :
                     "queue 4x" + std::to_string(mi == 0 ? 32 : 64) + " bpc" + std::to_string(bpc),
         ms});
    if (ms < best) {
      best = ms;
      pick = cand;
    }
  }
  c->profile = prof;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipGetLastError();
  P.mfma_choice[precision] = pick;
  return GEYE_OK;
}

// Launches the convolution, DT, scoring and candidate compaction; with ncand non-null also
// waits for the candidate count.  pyr (default P) is the plan whose pyramid holds the
// features: another model's plan over the same frame when the pyramid is shared (its levels
// are a prefix of pyr's, at the same arena offsets).  slot is the mapped host slot that
// receives the candidate count.
int run_scores(geye_ctx *c, Model &m, Plan &P, float thresh, int precision, int *ncand,
               const Plan *pyr = nullptr, int slot = kHostCandidates) {
  hipStream_t s = c->stream;
  const Plan &F = pyr ? *pyr : P;
  if (precision == GEYE_FP32) {
    for (auto &cc : P.conv_classes) {
      if (cc.tile_count == 0) continue;
      KTimer t(c, "conv_f32_" + std::to_string(cc.h) + "x" + std::to_string(cc.w));
      launch_conv_f32(s, cc.nf, dp<ConvTile>(P.d_conv_tiles) + cc.tile_begin, cc.tile_count,
                      cc.h, cc.w, (const float *)F.d_feat.p, dp<float>(m.d_filters),
                      dp<long long>(P.d_map_off), dp<int2>(P.d_map_dims), dp<float>(P.d_conv));
    }
  } else if (mfma_precision(precision)) {
    if (P.mfma_choice[precision] < 0) {
      // a schedule tuned for the same model and frame geometry on another context is reused
      const auto key = std::make_tuple((const Model *)&m, plan_key(P.W, P.H, P.max_scale), precision);
      auto lookup = [&] {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) P.mfma_choice[precision] = it->second;
      };
      lookup();
      if (P.mfma_choice[precision] < 0) {
        std::lock_guard<std::mutex> run(g_tune_run_mu);
        lookup();  // another context may have tuned it meanwhile
        if (P.mfma_choice[precision] < 0) {
          CHK(tune_mfma(c, m, P, F, precision));
          std::lock_guard<std::mutex> lk(g_tune_mu);
          g_tuned[key] = P.mfma_choice[precision];
        }
      }
    }
    CHK(run_mfma(c, m, P, F, precision, P.mfma_choice[precision], true));
  } else {
    return fail(GEYE_ERR_ARG, "unknown precision");
  }
  // FP32 runs the reference's divide and conquer; the 16-bit forms take the windowed
  // minimum, which gives the same outputs and writes only the cells the scoring and the part
  // boxes read
  const bool exact_dt = !dt_windowed(c, precision);
  auto *dt_launch = exact_dt ? launch_dt : launch_dt_window;
  const auto &jobs1 = exact_dt ? P.dt_pass1 : P.dt_pass1w;
  const auto &jobs2 = exact_dt ? P.dt_pass2 : P.dt_pass2w;
  if (!jobs1.empty()) {
    KTimer t(c, "dt_cols");
    if (dt_launch(s, dp<DtJob>(exact_dt ? P.d_dt1 : P.d_dt1w), (int)jobs1.size(),
                  exact_dt ? P.dt_lds1 : P.dt_lds1w, dp<int4>(P.d_nodes), dp<int>(P.d_depths),
                  dp<float>(P.d_conv), dp<float>(P.d_tmpM), dp<int>(P.d_tmpIy)) != 0)
      return fail(GEYE_ERR_HIP, "DT kernel: dynamic LDS attribute not set");
  }
  if (!jobs2.empty()) {
    KTimer t(c, "dt_rows");
    if (dt_launch(s, dp<DtJob>(exact_dt ? P.d_dt2 : P.d_dt2w), (int)jobs2.size(),
                  exact_dt ? P.dt_lds2 : P.dt_lds2w, dp<int4>(P.d_nodes), dp<int>(P.d_depths),
                  dp<float>(P.d_tmpM), dp<float>(P.d_M), dp<int>(P.d_tmpIx)) != 0)
      return fail(GEYE_ERR_HIP, "DT kernel: dynamic LDS attribute not set");
  }
  if (P.score_total > 0) {
    {
      KTimer t(c, "score");
      launch_score(s, dp<ScoreMap>(P.d_smaps), (int)P.smaps.size(), dp<ScorePart>(P.d_sparts),
                   dp<float>(P.d_conv), dp<float>(P.d_M), thresh, exact_dt ? 0 : 1,
                   dp<float>(P.d_score), dp<int>(P.d_flags), P.score_blocks);
    }
    {
      KTimer t(c, "compact");
      launch_compact(s, dp<int>(P.d_flags), P.score_total, dp<int>(P.d_blocksum),
                     dp<int>(P.d_count), dp<int>(P.d_cand), c->d_hcount + slot);
    }
  }
  HIPCHK(hipGetLastError());
  return ncand ? read_count(c, P, ncand, slot) : GEYE_OK;
}

// number of candidates the last scoring pass left in P.d_count (waits for it)
int read_count(geye_ctx *c, Plan &P, int *ncand, int slot) {
  *ncand = 0;
  if (P.score_total <= 0) return GEYE_OK;
  // the compaction's last workgroup wrote the count into the context's mapped host slot
  HIPCHK(hipStreamSynchronize(c->stream));
  *ncand = ((volatile int *)c->h_count)[slot];
  return GEYE_OK;
}

// Rows of the candidates of the last scoring pass: all of them, or (sel non-null) the
// candidate positions sel[0..ncand).  half: the DT outputs are the windowed DT's half grids
// (the 16-bit forms).
int run_rows(geye_ctx *c, Model &m, Plan &P, const int *sel, int ncand, bool half,
             std::vector<float> &rows) {
  rows.assign((size_t)ncand * P.GL, 0.0f);
  if (ncand == 0) return GEYE_OK;
  CHK(ensure(P.d_rows, sizeof(float) * (size_t)ncand * P.GL));
  {
    KTimer t(c, "rows");
    launch_rows(c->stream, dp<ScoreMap>(P.d_smaps), (int)P.smaps.size(), dp<ScorePart>(P.d_sparts),
                dp<float>(P.d_score), dp<int>(P.d_cand), sel, ncand, dp<int>(P.d_tmpIx),
                dp<int>(P.d_tmpIy), half ? 1 : 0, P.GL, m.padx, m.pady, dp<int>(m.d_rsize),
                dp<float>(P.d_rows));
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(rows.data(), P.d_rows.p, sizeof(float) * rows.size(),
                        hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return GEYE_OK;
}

// Buffers of the device NMS for n boxes (held by a plan).
int nms_buffers(Plan &P, int n) {
  const size_t N = (size_t)n;
  CHK(ensure(P.n_box, sizeof(NmsBox) * N));
  CHK(ensure(P.n_sbox, sizeof(NmsBox) * N));
  CHK(ensure(P.n_keys, 4 * N));
  CHK(ensure(P.n_keys2, 4 * N));
  CHK(ensure(P.n_vals, 4 * N));
  CHK(ensure(P.n_ord, 4 * N));
  CHK(ensure(P.n_alive0, 4 * N));
  CHK(ensure(P.n_alive1, 4 * N));
  CHK(ensure(P.n_status, 4 * N));
  CHK(ensure(P.n_fsup, 4 * N));
  CHK(ensure(P.n_picks, 4 * N));
  CHK(ensure(P.n_keep_orig, 4 * N));
  CHK(ensure(P.n_kept, 4 * N));
  CHK(ensure(P.n_counters, 16));
  CHK(ensure(P.n_flags, 16));
  CHK(ensure(P.n_total, 16));
  CHK(ensure(P.n_tie, sizeof(float) * (kNmsTieCap + kNmsResolvedCap)));
  CHK(ensure(P.n_mask, sizeof(unsigned long long) * kNmsWindow * (kNmsWindow / 64)));
  CHK(ensure(P.d_blocksum, sizeof(int) * (compact_blocks((long long)N) + 1)));
  CHK(ensure(P.n_sort, nms_sort_temp_bytes(n)));
  return GEYE_OK;
}

// One greedy NMS over a segment of a plan's NMS arrays: candidates [off, off + n) of the box,
// key and rank arrays (the whole candidate set, or one frame of a batch).  A job has its own
// stream, events, mapped host slots and small buffers, so several jobs run at once; their
// rounds are driven together (nms_run), each one round ahead of the host.
struct NmsJob {
  Plan *P = nullptr;
  long long off = 0;
  int n = 0;
  hipStream_t s = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  volatile int *h_round[2] = {nullptr, nullptr}, *h_surv = nullptr, *h_tie = nullptr;
  int *d_round[2] = {nullptr, nullptr}, *d_surv = nullptr, *d_tie = nullptr;
  int *counters = nullptr, *flags = nullptr, *total = nullptr, *blocksum = nullptr;
  // tie records ([kNmsTieCap]) then the resolved scores ([kNmsResolvedCap]) on the device;
  // the resolved scores on the host
  float *tie = nullptr;
  std::vector<float> resolved;
  unsigned long long *mask = nullptr;
  void *sort_temp = nullptr;
  size_t sort_bytes = 0;
  bool use_grid = false;  // spatial index of each round's picks (overlap >= 0.5)
  NmsGrid grid{};
  const int *cand = nullptr;  // candidate -> flat score index, segment base (batch rows)
  // state of the rounds
  int na = 0, cur = 0, slot = 0, nk = 0, round = 0;
  int *alive = nullptr, *alive_next = nullptr;
  bool done = true, waiting = false;
  template <typename T>
  T *seg(DeviceBuffer &b) const {
    return (T *)b.p + off;
  }
  int *na_buf(int k) const { return total + 2 * k; }
};

// Sorts the job's keys (unless sort is false: the order is already in place), gathers the
// boxes in that order and enqueues the first round.
int nms_job_start(geye_ctx *c, NmsJob &J, float overlap, bool sort = true) {
  Plan &P = *J.P;
  J.done = J.n <= 0;
  J.waiting = false;
  J.nk = 0;
  if (J.done) return GEYE_OK;
  hipStream_t s = J.s;
  if (sort) {
    KTimer t(c, "nms_sort", s);
    if (launch_nms_sort(s, J.sort_temp, J.sort_bytes, J.seg<unsigned int>(P.n_keys),
                        J.seg<unsigned int>(P.n_keys2), J.seg<int>(P.n_vals),
                        J.seg<int>(P.n_ord), J.n) != 0)
      return fail(GEYE_ERR_HIP, "rocprim radix sort failed");
  }
  J.cur = 0;
  J.slot = 0;
  launch_nms_gather(s, J.seg<int>(P.n_ord), J.seg<NmsBox>(P.n_box), J.n, J.seg<NmsBox>(P.n_sbox),
                    J.seg<int>(P.n_alive0), J.seg<int>(P.n_status), J.seg<int>(P.n_fsup),
                    J.counters, J.na_buf(0), J.flags, J.seg<int>(P.n_keep_orig),
                    J.use_grid ? &J.grid : nullptr, J.tie);
  J.alive = J.seg<int>(P.n_alive0);
  J.alive_next = J.seg<int>(P.n_alive1);
  J.na = J.n;
  J.round = 0;
  KTimer t(c, "nms_round", s);
  if (launch_nms_round(s, J.seg<NmsBox>(P.n_sbox), J.alive, J.na_buf(J.cur), J.na, overlap, J.mask,
                       J.seg<int>(P.n_status), J.seg<int>(P.n_fsup), J.seg<int>(P.n_picks),
                       J.counters, J.blocksum, J.use_grid ? &J.grid : nullptr, 0) != 0)
    return fail(GEYE_ERR_HIP, "NMS greedy kernel: dynamic LDS attribute not set");
  return GEYE_OK;
}

// Rounds run one ahead of the host: after round k's compaction, round k+1 is enqueued with a
// bound of round k's boxes after its window (its alive count can only be smaller) before the
// host waits for round k's survivor count, so the GPU never idles on the read-back.  The
// kernels read the true counts from device memory and do nothing past them.
int nms_job_enqueue(geye_ctx *c, NmsJob &J, float overlap) {
  if (J.done) return GEYE_OK;
  Plan &P = *J.P;
  const int rest = J.na - std::min(J.na, kNmsWindow);  // boxes after that round's window
  if (rest <= 0) {                                      // it held every alive box
    J.done = true;
    return GEYE_OK;
  }
  launch_nms_compact(J.s, J.alive, J.na_buf(J.cur), J.seg<int>(P.n_fsup), J.blocksum, J.na,
                     J.na_buf(J.cur ^ 1), J.alive_next, J.d_round[J.slot]);
  HIPCHK(hipEventRecord(J.ev[J.slot], J.s));
  std::swap(J.alive, J.alive_next);
  J.cur ^= 1;
  J.round++;
  KTimer t(c, "nms_round", J.s);
  if (launch_nms_round(J.s, J.seg<NmsBox>(P.n_sbox), J.alive, J.na_buf(J.cur), rest, overlap,
                       J.mask, J.seg<int>(P.n_status), J.seg<int>(P.n_fsup),
                       J.seg<int>(P.n_picks), J.counters, J.blocksum,
                       J.use_grid ? &J.grid : nullptr, J.round) != 0)
    return fail(GEYE_ERR_HIP, "NMS greedy kernel: dynamic LDS attribute not set");
  J.waiting = true;
  return GEYE_OK;
}

int nms_job_wait(NmsJob &J) {
  if (!J.waiting) return GEYE_OK;
  HIPCHK(hipEventSynchronize(J.ev[J.slot]));
  J.na = *J.h_round[J.slot];  // the exact alive count of the round enqueued last
  J.slot ^= 1;
  J.waiting = false;
  if (J.na == 0) J.done = true;
  return GEYE_OK;
}

// Finalisation: the tie check and the survivor compaction are both enqueued (the pick count
// stays on the device); the tie flag and the count arrive in the job's mapped host slots.
void nms_job_finish(NmsJob &J) {
  if (J.n <= 0) return;
  Plan &P = *J.P;
  launch_nms_check(J.s, J.seg<NmsBox>(P.n_sbox), J.n, J.seg<int>(P.n_status),
                   J.seg<int>(P.n_fsup), J.seg<int>(P.n_picks), J.counters, J.flags,
                   J.tie + kNmsTieCap, (int)J.resolved.size(), J.tie);
  launch_nms_keep(J.s, J.seg<NmsBox>(P.n_sbox), J.seg<int>(P.n_picks), J.counters, J.n, J.flags,
                  J.seg<int>(P.n_keep_orig), J.d_tie);
  launch_compact(J.s, J.seg<int>(P.n_keep_orig), (long long)J.n, J.blocksum, J.total + 1,
                 J.seg<int>(P.n_kept), J.d_surv);
}

// After the job's stream has finished: its survivor count (returns 0), or 1 when equal scores
// decided an outcome and the order of their runs has been set to the reference's (the job
// must run again from its gather), or 2 when that cannot be done here (the host replay must
// run).  The reference's quicksort is unstable: equal scores end in an order fixed by the
// whole array.  The sort of the device keeps them in candidate order, which is the
// reference's order wherever no tie decides anything; the finalisation check records each
// score whose run's order did decide something.  For those scores the host runs the
// reference's quicksort on the scores in candidate order, restricted to the subarrays that
// can hold them (O(N) work rather than the host replay's O(N^2)), and writes each run in the
// order found into the device order; the check of the next run skips those scores.
int nms_job_collect(geye_ctx *c, NmsJob &J) {
  J.nk = 0;
  if (J.n <= 0) return 0;
  Plan &P = *J.P;
  const int tie = *J.h_tie;
  if (tie == 0) {
    J.nk = *J.h_surv;
    return 0;
  }
  if ((tie & 4) || (int)J.resolved.size() >= kNmsResolvedCap) return 2;
  int fl[2] = {0, 0};
  std::vector<unsigned int> tab(kNmsTieCap);
  HIPCHK(hipMemcpyAsync(fl, J.flags + 2, sizeof(fl), hipMemcpyDeviceToHost, J.s));
  HIPCHK(hipMemcpyAsync(tab.data(), J.tie, sizeof(float) * kNmsTieCap, hipMemcpyDeviceToHost,
                        J.s));
  HIPCHK(hipStreamSynchronize(J.s));
  if (fl[0] <= 0 || fl[1] != 0) return 2;  // nothing recorded, or the set overflowed
  std::vector<float> vals;
  for (unsigned int b : tab)
    if (b != kNmsTieEmpty) {
      float v;
      std::memcpy(&v, &b, 4);
      vals.push_back(v);
    }
  for (float v : vals)
    if (v == 0.0f || v != v) return 2;  // +0 and -0 are one run for the reference, two here
  for (float v : vals)
    if (std::find(J.resolved.begin(), J.resolved.end(), v) == J.resolved.end())
      J.resolved.push_back(v);
  if ((int)J.resolved.size() > kNmsResolvedCap) return 2;
  // the scores in candidate order, from the sort keys (descending order of the float bits)
  const size_t N = (size_t)J.n;
  std::vector<unsigned int> keys(N);
  HIPCHK(hipMemcpyAsync(keys.data(), J.seg<unsigned int>(P.n_keys), 4 * N, hipMemcpyDeviceToHost,
                        J.s));
  HIPCHK(hipStreamSynchronize(J.s));
  std::vector<float> sc(N);
  std::vector<int> order(N);
  for (size_t i = 0; i < N; i++) {
    unsigned int u = ~keys[i];
    u = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    std::memcpy(&sc[i], &u, 4);
    order[i] = (int)i;
  }
  // each resolved score's run starts after every greater score in the device order
  std::vector<int> start(J.resolved.size(), 0), count(J.resolved.size(), 0);
  for (size_t i = 0; i < N; i++)
    for (size_t k = 0; k < J.resolved.size(); k++) {
      start[k] += sc[i] > J.resolved[k];
      count[k] += sc[i] == J.resolved[k];
    }
  quicksort_host(sc.data(), order.data(), 0, (int)N - 1, &J.resolved);
  for (size_t k = 0; k < J.resolved.size(); k++) {
    const float v = J.resolved[k];
    const int a = start[k], b = a + count[k];
    for (int p = a; p < b; p++)
      if (p >= (int)N || sc[p] != v) return 2;  // the run is not where the counts put it
    HIPCHK(hipMemcpyAsync(J.seg<int>(P.n_ord) + a, order.data() + a, sizeof(int) * (b - a),
                          hipMemcpyHostToDevice, J.s));
  }
  HIPCHK(hipMemcpyAsync(J.tie + kNmsTieCap, J.resolved.data(), sizeof(float) * J.resolved.size(),
                        hipMemcpyHostToDevice, J.s));
  HIPCHK(hipStreamSynchronize(J.s));
  return 1;
}

// The host replay of the reference's NMS (its exact quicksort order); the survivors are
// written where the device's would be.
int nms_job_host_replay(geye_ctx *c, NmsJob &J, float overlap) {
  Plan &P = *J.P;
  P.nms_fallbacks++;
  c->nms_fallbacks++;
  c->stats["nms_host_fallback"].count++;
  const size_t N = (size_t)J.n;
  std::vector<NmsBox> hb(N);
  HIPCHK(hipMemcpyAsync(hb.data(), J.seg<NmsBox>(P.n_box), sizeof(NmsBox) * N,
                        hipMemcpyDeviceToHost, J.s));
  HIPCHK(hipStreamSynchronize(J.s));
  std::vector<float> rows(N * 7), out(N * 7);
  for (size_t i = 0; i < N; i++) {
    float *r = &rows[i * 7];
    r[0] = hb[i].y1; r[1] = hb[i].x1; r[2] = hb[i].y2; r[3] = hb[i].x2;
    r[4] = (float)i; r[5] = hb[i].score; r[6] = 0.0f;
  }
  const int nk = nms_host(rows.data(), J.n, 7, overlap, out.data());
  std::vector<int> kept(nk);
  for (int k = 0; k < nk; k++) kept[k] = (int)out[(size_t)k * 7 + 4];
  if (nk > 0) {
    HIPCHK(hipMemcpyAsync(J.seg<int>(P.n_kept), kept.data(), sizeof(int) * nk,
                          hipMemcpyHostToDevice, J.s));
    HIPCHK(hipStreamSynchronize(J.s));
  }
  J.nk = nk;
  return GEYE_OK;
}

// Drives started jobs' rounds together to their finalisation and waits for their streams.
int nms_drive(geye_ctx *c, std::vector<NmsJob *> &jobs, float overlap) {
  for (;;) {
    bool any = false;
    for (NmsJob *J : jobs) {
      CHK(nms_job_enqueue(c, *J, overlap));
      any = any || J->waiting;
    }
    if (!any) break;
    for (NmsJob *J : jobs) CHK(nms_job_wait(*J));
  }
  for (NmsJob *J : jobs) nms_job_finish(*J);
  for (NmsJob *J : jobs) HIPCHK(hipStreamSynchronize(J->s));
  return GEYE_OK;
}

// Runs the jobs' NMS together: each job's boxes, keys and values are in place (a prep
// kernel on its stream); leaves each job's survivors' local candidate positions, ascending,
// in its segment of P.n_kept and their count in nk.  A job whose outcome a tie decided runs
// again with the reference's order of the tied runs (up to 4 times), else on the host.
int nms_run(geye_ctx *c, std::vector<NmsJob> &jobs, float overlap) {
  std::vector<NmsJob *> all;
  for (NmsJob &J : jobs) {
    J.resolved.clear();
    CHK(nms_job_start(c, J, overlap));
    all.push_back(&J);
  }
  CHK(nms_drive(c, all, overlap));
  for (NmsJob &J : jobs) {
    int st = nms_job_collect(c, J);
    for (int attempt = 0; st == 1 && attempt < 4; attempt++) {
      if (attempt == 0) {
        c->nms_tie_repairs++;
        c->stats["nms_tie_repair"].count++;
      }
      std::vector<NmsJob *> one{&J};
      CHK(nms_job_start(c, J, overlap, false));
      CHK(nms_drive(c, one, overlap));
      st = nms_job_collect(c, J);
    }
    if (st != 0) CHK(nms_job_host_replay(c, J, overlap));
  }
  HIPCHK(hipGetLastError());
  return GEYE_OK;
}

// The whole candidate set of a one-frame plan as one job on the context's stream, with the
// context's events and mapped host slots.
NmsJob single_job(geye_ctx *c, Plan &P, int n, const NmsGrid *grid) {
  NmsJob J;
  J.P = &P;
  J.off = 0;
  J.n = n;
  J.s = c->stream;
  J.ev[0] = c->ev_count[0];
  J.ev[1] = c->ev_count[1];
  volatile int *h = (volatile int *)c->h_count;
  J.h_round[0] = h;
  J.h_round[1] = h + 1;
  J.h_surv = h + kHostSurvivors;
  J.h_tie = h + kHostTie;
  J.d_round[0] = c->d_hcount;
  J.d_round[1] = c->d_hcount + 1;
  J.d_surv = c->d_hcount + kHostSurvivors;
  J.d_tie = c->d_hcount + kHostTie;
  J.counters = dp<int>(P.n_counters);
  J.flags = dp<int>(P.n_flags);
  J.total = dp<int>(P.n_total);
  J.tie = dp<float>(P.n_tie);
  J.blocksum = dp<int>(P.d_blocksum);
  J.mask = dp<unsigned long long>(P.n_mask);
  J.sort_temp = P.n_sort.p;
  J.sort_bytes = P.n_sort.bytes;
  J.use_grid = grid != nullptr;
  if (grid) J.grid = *grid;
  return J;
}

// The indexed suppression pass's buffers for the n candidates of a one-frame plan (P.n_box
// filled) over a W x H frame: each candidate's rank and decided flag, and the tile lists.
// Only for overlap >= 0.5, where the index is complete.
int nms_grid_buffers(Plan &P, int n, int W, int H, NmsGrid *g) {
  nms_grid_shape(W, H, &g->gx, &g->gy, &g->shift);
  const size_t N = (size_t)std::max(n, 1);
  CHK(ensure(P.n_rank, sizeof(int) * N));
  CHK(ensure(P.n_dec, N));
  const size_t NB = (N + 1023) / 1024;
  CHK(ensure(P.n_alist, sizeof(int) * (N + 1024 * NB)));       // list, then slot
  CHK(ensure(P.n_acount, sizeof(int) * (4 * NB + 64)));         // lcount, then bsum
  CHK(ensure(P.n_grid, nms_grid_bytes(g->gx, g->gy)));
  nms_grid_bind(*g, P.n_grid.p);
  g->rank_of = dp<int>(P.n_rank);
  g->dec = (unsigned char *)P.n_dec.p;
  g->box = dp<NmsBox>(P.n_box);
  g->list = dp<int>(P.n_alist);
  g->slot = g->list + N;
  g->lcount = dp<int>(P.n_acount);
  g->bsum = g->lcount + 64;
  g->n = n;
  return GEYE_OK;
}

// Greedy NMS over P.n_box / P.n_keys / P.n_vals (filled by a prep kernel, original order);
// leaves the survivors' original indices, ascending, in P.n_kept.  grid (optional) selects
// the indexed suppression pass.
int nms_core(geye_ctx *c, Plan &P, int n, float overlap, const NmsGrid *grid, int *nkept) {
  *nkept = 0;
  if (n <= 0) return GEYE_OK;
  std::vector<NmsJob> jobs{single_job(c, P, n, grid)};
  CHK(nms_run(c, jobs, overlap));
  *nkept = jobs[0].nk;
  return GEYE_OK;
}

// Device NMS over the ncand candidates of the last scoring pass; leaves the survivors'
// candidate positions (original order) in P.n_kept.
int run_nms(geye_ctx *c, Model &m, Plan &P, int n, float overlap, int *nkept) {
  *nkept = 0;
  if (n <= 0) return GEYE_OK;
  CHK(nms_buffers(P, n));
  {
    KTimer t(c, "nms_prep");
    launch_nms_prep(c->stream, dp<ScoreMap>(P.d_smaps), (int)P.smaps.size(),
                    dp<float>(P.d_score), dp<int>(P.d_cand), n, m.padx, m.pady,
                    dp<int>(m.d_rsize), dp<NmsBox>(P.n_box), dp<unsigned int>(P.n_keys),
                    dp<int>(P.n_vals));
  }
  NmsGrid g;
  const bool use_grid = overlap >= 0.5f;
  if (use_grid) CHK(nms_grid_buffers(P, n, P.W, P.H, &g));
  int nk = 0;
  CHK(nms_core(c, P, n, overlap, use_grid ? &g : nullptr, &nk));
  HIPCHK(hipGetLastError());
  *nkept = nk;
  return GEYE_OK;
}

// The plan's frame buffer; graphs captured against an older buffer are dropped when it grows.
int ensure_frame(Plan &P, size_t bytes) {
  const void *old = P.d_frame.p;
  CHK(ensure(P.d_frame, bytes));
  if (P.d_frame.p != old) drop_graphs(P);
  return GEYE_OK;
}

// The per-frame work before the first host wait -- resize rounds, HOG, filter
// convolution, DT, scoring and candidate compaction, about 30 launches -- runs as one HIP
// graph per (plan, precision, threshold, widthStep): captured on the first frame after a
// warm-up run (which makes every allocation and one-time attribute call), then replayed.
// The graph reads the frame from the plan's own frame buffer, so a caller's device frame is
// copied there first.  With profiling on, or graphs off, the launches run one by one.
int run_front(geye_ctx *c, Model &m, Plan &P, const unsigned char *d_bgr, int ws, float thresh,
              int precision, int *ncand) {
  const int hkind = precision == GEYE_FP32 ? 0 : precision;
  if (c->profile || !c->graphs) {
    CHK(run_pyramid(c, P, d_bgr, ws, hkind));
    return run_scores(c, m, P, thresh, precision, ncand);
  }
  unsigned int tb;
  std::memcpy(&tb, &thresh, 4);
  const long long key = ((long long)precision << 48) ^ ((long long)dt_windowed(c, precision) << 52) ^
                        ((long long)ws << 32) ^ (long long)tb;
  CHK(ensure_frame(P, (size_t)ws * P.H));
  if (d_bgr != P.d_frame.p)
    HIPCHK(hipMemcpyAsync(P.d_frame.p, d_bgr, (size_t)ws * P.H, hipMemcpyDeviceToDevice, c->stream));
  const unsigned char *fr = (const unsigned char *)P.d_frame.p;
  auto it = P.graphs.find(key);
  if (it == P.graphs.end()) {
    // warm-up: allocations and attributes happen outside the capture
    CHK(run_pyramid(c, P, fr, ws, hkind));
    CHK(run_scores(c, m, P, thresh, precision, nullptr));
    HIPCHK(hipStreamSynchronize(c->stream));
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = run_pyramid(c, P, fr, ws, hkind);
    if (rc == GEYE_OK) rc = run_scores(c, m, P, thresh, precision, nullptr);
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc != GEYE_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(GEYE_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    hipGraphExec_t ex = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) return fail(GEYE_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ei));
    it = P.graphs.emplace(key, ex).first;
  }
  HIPCHK(hipGraphLaunch(it->second, c->stream));
  return read_count(c, P, ncand);
}

int frame_in(geye_ctx *c, Plan &P, const unsigned char *bgr, int W, int H, int ws,
             const unsigned char **d_bgr) {
  if (!bgr || ws < 3 * W) return fail(GEYE_ERR_ARG, "frame: null pointer or widthStep < 3*width");
  CHK(ensure_frame(P, (size_t)ws * H));
  HIPCHK(hipMemcpyAsync(P.d_frame.p, bgr, (size_t)ws * H, hipMemcpyHostToDevice, c->stream));
  *d_bgr = (const unsigned char *)P.d_frame.p;
  return GEYE_OK;
}

float *to_malloc(const std::vector<float> &v) {
  float *p = (float *)std::malloc(sizeof(float) * std::max<size_t>(v.size(), 1));
  if (p && !v.empty()) std::memcpy(p, v.data(), sizeof(float) * v.size());
  return p;
}

int detect_common(geye_ctx *c, const geye_model *model, const unsigned char *bgr, bool dev,
                  int W, int H, int ws, float thresh, float overlap, int precision, bool do_nms,
                  float **boxes, int *num, int *GL) {
  if (!c || !model || !boxes || !num || !GL) return fail(GEYE_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  HIPCHK(hipSetDevice(c->device));
  (void)hipGetLastError();
  Model &m = const_cast<Model &>(model->m);
  Plan *P = nullptr;
  CHK(get_plan(c, m, W, H, &P, -1));
  const unsigned char *d_bgr = bgr;
  if (!dev) CHK(frame_in(c, *P, bgr, W, H, ws, &d_bgr));
  int ncand = 0;
  CHK(run_front(c, m, *P, d_bgr, ws, thresh, precision, &ncand));
  std::vector<float> rows;
  c->last_plan = P;
  if (do_nms) {
    int nk = 0;
    CHK(run_nms(c, m, *P, ncand, overlap, &nk));
    CHK(run_rows(c, m, *P, dp<int>(P->n_kept), nk, dt_windowed(c, precision), rows));
    *num = nk;
  } else {
    CHK(run_rows(c, m, *P, nullptr, ncand, dt_windowed(c, precision), rows));
    *num = ncand;
  }
  *boxes = to_malloc(rows);
  *GL = P->GL;
  flush_stats(c);
  return *boxes ? GEYE_OK : fail(GEYE_ERR_NOMEM, "malloc");
}

// Levels of B are a prefix of A's at the same feature arena offsets (the same frame size,
// sbin and interval, and B's pyramid no deeper): B's convolution can read A's features.
bool pyramid_prefix(const Plan &A, const Plan &B) {
  if (A.W != B.W || A.H != B.H || A.sbin != B.sbin || A.interval != B.interval || B.LEN > A.LEN)
    return false;
  for (int l = 0; l < B.LEN; l++) {
    const Level &a = A.lv[l], &b = B.lv[l];
    if (a.img_h != b.img_h || a.img_w != b.img_w || a.sbin != b.sbin || a.F0 != b.F0 ||
        a.F1 != b.F1 || a.scale != b.scale || A.feat_off[l] != B.feat_off[l])
      return false;
  }
  return true;
}

// The level geometry and feature offsets of model m's plan for a W x H frame, nothing else
// (what pyramid_prefix compares).
void level_only_plan(const Model &m, int W, int H, Plan &P) {
  P.W = W;
  P.H = H;
  P.sbin = m.sbin;
  P.interval = m.interval;
  P.max_scale = compute_max_scale(m, W, H);
  P.LEN = P.max_scale + m.interval;
  std::vector<int> isrc;
  std::vector<float> isc;
  level_geometry(P.max_scale, m.sbin, m.interval, W, H, P.lv, P.img_h, P.img_w, isrc, isc);
  P.feat_off.assign(P.LEN, -1);
  long long fo = 0;
  for (int l = 0; l < P.LEN; l++)
    if (P.lv[l].F0 > 0 && P.lv[l].F1 > 0) {
      P.feat_off[l] = fo;
      fo += 31LL * P.lv[l].F0 * P.lv[l].F1;
    }
}

// mapped host slot of model i's candidate count in a several-model detection
constexpr int kMaxModels = 8, kHostModelCounts = 8;
int count_slot(int i) { return i == 0 ? kHostCandidates : kHostModelCounts + i; }

// Several models on one frame: the deepest model's plan builds the pyramid once and every
// model whose levels are a prefix of it convolves those features; each model then scores,
// runs its NMS and assembles its rows exactly as a detection of it alone does.
int detect_models(geye_ctx *c, const geye_model *const *models, int n, const unsigned char *bgr,
                  bool dev, int W, int H, int ws, float thresh, float overlap, int precision,
                  float **boxes, int *num, int *GL) {
  if (!c || !models || n < 1 || n > kMaxModels || !boxes || !num || !GL)
    return fail(GEYE_ERR_ARG, "detect_models: null argument or model count outside 1..8");
  for (int i = 0; i < n; i++) {
    if (!models[i]) return fail(GEYE_ERR_ARG, "detect_models: null model");
    boxes[i] = nullptr;
    num[i] = GL[i] = 0;
  }
  if (!valid_precision(precision)) return fail(GEYE_ERR_ARG, "unknown precision");
  std::lock_guard<std::mutex> lk(c->mu);
  HIPCHK(hipSetDevice(c->device));
  (void)hipGetLastError();
  std::vector<Plan *> P(n, nullptr);
  for (int i = 0; i < n; i++)
    CHK(get_plan(c, const_cast<Model &>(models[i]->m), W, H, &P[i], -1));
  int owner = 0;
  for (int i = 1; i < n; i++)
    if (P[i]->LEN > P[owner]->LEN) owner = i;
  const unsigned char *d_bgr = bgr;
  if (!dev) CHK(frame_in(c, *P[owner], bgr, W, H, ws, &d_bgr));
  const int hkind = precision == GEYE_FP32 ? 0 : precision;
  CHK(run_pyramid(c, *P[owner], d_bgr, ws, hkind));
  for (int i = 0; i < n; i++) {
    Model &m = const_cast<Model &>(models[i]->m);
    const bool shared = i == owner || pyramid_prefix(*P[owner], *P[i]);
    if (!shared) CHK(run_pyramid(c, *P[i], d_bgr, ws, hkind));
    CHK(run_scores(c, m, *P[i], thresh, precision, nullptr, shared ? P[owner] : nullptr,
                   count_slot(i)));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  int rc = GEYE_OK;
  for (int i = 0; i < n && rc == GEYE_OK; i++) {
    Model &m = const_cast<Model &>(models[i]->m);
    const int ncand = P[i]->score_total > 0 ? ((volatile int *)c->h_count)[count_slot(i)] : 0;
    int nk = 0;
    std::vector<float> rows;
    rc = run_nms(c, m, *P[i], ncand, overlap, &nk);
    if (rc == GEYE_OK) rc = run_rows(c, m, *P[i], dp<int>(P[i]->n_kept), nk, dt_windowed(c, precision), rows);
    if (rc == GEYE_OK) {
      boxes[i] = to_malloc(rows);
      num[i] = nk;
      GL[i] = P[i]->GL;
      if (!boxes[i]) rc = fail(GEYE_ERR_NOMEM, "malloc");
    }
  }
  if (rc != GEYE_OK) {
    for (int i = 0; i < n; i++) {
      std::free(boxes[i]);
      boxes[i] = nullptr;
      num[i] = GL[i] = 0;
    }
    return rc;
  }
  c->last_plan = P[n - 1];
  flush_stats(c);
  return GEYE_OK;
}

// nf frames of one size through nm models: per model a batch plan (its one-frame plan's
// descriptors repeated per frame), the deepest model's pyramid built once for every frame
// and read by every model whose levels are a prefix of it, each model's convolution, DT and
// scoring in one launch sequence for all frames; then every (frame, model) NMS as a job, the
// jobs running together on the context's job streams.  boxes/num/GL: nf * nm entries, frame
// major.  Every result equals that frame's detection alone.
int detect_batch(geye_ctx *c, const geye_model *const *models, int nm,
                 const unsigned char *const *frames, bool dev, int nf, int W, int H, int ws,
                 float thresh, float overlap, int precision, float **boxes, int *num, int *GL) {
  if (!c || !models || nm < 1 || nm > kMaxModels || !frames || nf < 1 || nf > 64 || !boxes ||
      !num || !GL)
    return fail(GEYE_ERR_ARG, "detect_batch: null argument, 1..8 models or 1..64 frames");
  if (ws < 3 * W || W < 8 || H < 8) return fail(GEYE_ERR_ARG, "detect_batch: frame geometry");
  for (int i = 0; i < nm; i++)
    if (!models[i]) return fail(GEYE_ERR_ARG, "detect_batch: null model");
  for (int f = 0; f < nf; f++)
    if (!frames[f]) return fail(GEYE_ERR_ARG, "detect_batch: null frame");
  if (!valid_precision(precision)) return fail(GEYE_ERR_ARG, "unknown precision");
  for (int k = 0; k < nf * nm; k++) {
    boxes[k] = nullptr;
    num[k] = GL[k] = 0;
  }
  std::lock_guard<std::mutex> lk(c->mu);
  HIPCHK(hipSetDevice(c->device));
  (void)hipGetLastError();
  for (hipStream_t &js : c->job_streams)
    if (!js) HIPCHK(hipStreamCreateWithFlags(&js, hipStreamNonBlocking));
  // the level plans decide which models share the deepest model's pyramid
  std::vector<Plan> lvl(nm);
  for (int i = 0; i < nm; i++) level_only_plan(models[i]->m, W, H, lvl[i]);
  int owner = 0;
  for (int i = 1; i < nm; i++)
    if (lvl[i].LEN > lvl[owner].LEN) owner = i;
  std::vector<bool> shared(nm);
  std::vector<Plan *> P(nm, nullptr);
  for (int i = 0; i < nm; i++) {
    shared[i] = i == owner || pyramid_prefix(lvl[owner], lvl[i]);
    CHK(get_batch_plan(c, const_cast<Model &>(models[i]->m), W, H, nf,
                       shared[i] ? lvl[owner].LEN : 0, &P[i]));
  }
  // frames into the owner's frame buffer, one after another
  const size_t fbytes = (size_t)ws * H;
  CHK(ensure_frame(*P[owner], fbytes * nf));
  unsigned char *fr = (unsigned char *)P[owner]->d_frame.p;
  for (int f = 0; f < nf; f++)
    HIPCHK(hipMemcpyAsync(fr + fbytes * f, frames[f], fbytes,
                          dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
  const int hkind = precision == GEYE_FP32 ? 0 : precision;
  // every launch before the NMS, for every frame and model
  auto front = [&]() -> int {
    CHK(run_pyramid(c, *P[owner], fr, ws, hkind));
    for (int i = 0; i < nm; i++) {
      Model &m = const_cast<Model &>(models[i]->m);
      if (!shared[i]) {
        CHK(ensure_frame(*P[i], fbytes * nf));
        HIPCHK(hipMemcpyAsync(P[i]->d_frame.p, fr, fbytes * nf, hipMemcpyDeviceToDevice,
                              c->stream));
        CHK(run_pyramid(c, *P[i], (const unsigned char *)P[i]->d_frame.p, ws, hkind));
      }
      CHK(run_scores(c, m, *P[i], thresh, precision, nullptr, shared[i] ? P[owner] : nullptr,
                     count_slot(i)));
      if (P[i]->score_total > 0)
        launch_frame_offsets(c->stream, dp<int>(P[i]->d_cand), dp<int>(P[i]->d_count),
                             P[i]->score_one, nf, P[i]->b_dev + 4 * nf);
    }
    return GEYE_OK;
  };
  if (c->profile || !c->graphs) {
    CHK(front());
  } else {
    // one HIP graph for the whole batch front, per (form, threshold, frames, models' plans),
    // captured after a warm-up run (allocations and one-time attributes) and replayed while
    // none of the plans has dropped its graphs since
    unsigned int tb;
    std::memcpy(&tb, &thresh, 4);
    std::vector<long long> key{precision, dt_windowed(c, precision) ? 1 : 0, ws, (long long)tb,
                               nf, nm, W, H};
    std::vector<unsigned long long> gens;
    for (int i = 0; i < nm; i++) {
      key.push_back((long long)(uintptr_t)P[i]);
      gens.push_back(P[i]->gen);
    }
    // graphs captured against plans that are gone (freed models, evicted plans) or whose
    // buffers have moved since are never replayed again: destroy them here
    {
      std::vector<std::pair<const Plan *, unsigned long long>> live;
      {
        std::lock_guard<std::mutex> lk(g_plan_mu);
        for (const auto &kv : g_plans)
          if (std::get<0>(kv.first) == c) live.emplace_back(kv.second, kv.second->gen);
      }
      for (auto g = c->batch_graphs.begin(); g != c->batch_graphs.end();) {
        const std::vector<long long> &k = g->first;  // the plans from position 8 on
        bool stale = false;
        for (size_t j = 8; j < k.size() && !stale; j++) {
          const std::pair<const Plan *, unsigned long long> pg(
              (const Plan *)(uintptr_t)k[j], g->second.second[j - 8]);
          stale = std::find(live.begin(), live.end(), pg) == live.end();
        }
        if (stale) {
          (void)hipGraphExecDestroy(g->second.first);
          g = c->batch_graphs.erase(g);
        } else {
          ++g;
        }
      }
    }
    auto it = c->batch_graphs.find(key);
    if (it != c->batch_graphs.end() && it->second.second != gens) {
      (void)hipGraphExecDestroy(it->second.first);
      c->batch_graphs.erase(it);
      it = c->batch_graphs.end();
    }
    if (it == c->batch_graphs.end()) {
      CHK(front());
      HIPCHK(hipStreamSynchronize(c->stream));
      for (int i = 0; i < nm; i++) gens[i] = P[i]->gen;  // the warm-up may have grown buffers
      hipGraph_t g = nullptr;
      HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      const int rc = front();
      const hipError_t e = hipStreamEndCapture(c->stream, &g);
      if (rc != GEYE_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess)
        return fail(GEYE_ERR_HIP, std::string("batch graph capture: ") + hipGetErrorString(e));
      hipGraphExec_t ex = nullptr;
      const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ei != hipSuccess)
        return fail(GEYE_ERR_HIP, std::string("batch graph instantiate: ") + hipGetErrorString(ei));
      it = c->batch_graphs.emplace(key, std::make_pair(ex, gens)).first;
    }
    HIPCHK(hipGraphLaunch(it->second.first, c->stream));
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  // every (frame, model) NMS as one job
  const bool window = dt_windowed(c, precision);
  std::vector<NmsJob> jobs;
  std::vector<std::pair<int, int>> who;  // (frame, model) of each job
  for (int i = 0; i < nm; i++) {
    Model &m = const_cast<Model &>(models[i]->m);
    Plan &B = *P[i];
    const int ncand = B.score_total > 0 ? ((volatile int *)c->h_count)[count_slot(i)] : 0;
    CHK(nms_buffers(B, std::max(ncand, 1)));
    NmsGrid g0;
    const bool use_grid = overlap >= 0.5f;
    if (use_grid) CHK(nms_grid_buffers(B, std::max(ncand, 1), W, H, &g0));
    const volatile int *off = (const volatile int *)B.b_host + 4 * nf;
    for (int f = 0; f < nf; f++) {
      NmsJob J;
      J.P = &B;
      J.off = B.score_total > 0 ? off[f] : 0;
      J.n = B.score_total > 0 ? off[f + 1] - off[f] : 0;
      J.s = c->job_streams[jobs.size() % 4];
      J.ev[0] = B.b_ev[2 * f];
      J.ev[1] = B.b_ev[2 * f + 1];
      volatile int *h = (volatile int *)B.b_host + 4 * f;
      int *d = B.b_dev + 4 * f;
      J.h_round[0] = h;
      J.h_round[1] = h + 1;
      J.h_surv = h + 2;
      J.h_tie = h + 3;
      J.d_round[0] = d;
      J.d_round[1] = d + 1;
      J.d_surv = d + 2;
      J.d_tie = d + 3;
      int *small = dp<int>(B.b_small[f]);
      J.counters = small;
      J.flags = small + 4;
      J.total = small + 8;
      J.tie = (float *)(small + 16);
      J.blocksum = dp<int>(B.b_blocksum[f]);
      J.mask = dp<unsigned long long>(B.b_mask[f]);
      J.sort_temp = B.b_sort[f].p;
      J.sort_bytes = B.b_sort[f].bytes;
      J.cand = dp<int>(B.d_cand) + J.off;
      // frame f's grid: its own tile lists, its segment of the rank and decided arrays
      J.use_grid = use_grid;
      if (use_grid) {
        J.grid = g0;
        CHK(ensure(B.b_grid[f], nms_grid_bytes(g0.gx, g0.gy)));
        nms_grid_bind(J.grid, B.b_grid[f].p);
        J.grid.rank_of = g0.rank_of + J.off;
        J.grid.dec = g0.dec + J.off;
        J.grid.box = g0.box + J.off;
        // the job's own list, slots and counts
        const size_t nbj = ((size_t)std::max(J.n, 1) + 1023) / 1024;
        CHK(ensure(B.b_acount[f], sizeof(int) * (2 * 1024 * nbj + 4 * nbj + 64)));
        J.grid.list = dp<int>(B.b_acount[f]);
        J.grid.slot = J.grid.list + 1024 * nbj;
        J.grid.lcount = J.grid.slot + 1024 * nbj;
        J.grid.bsum = J.grid.lcount + 64;
        J.grid.n = J.n;
      }
      if (J.n > 0) {
        KTimer t(c, "nms_prep", J.s);
        launch_nms_prep(J.s, dp<ScoreMap>(B.d_smaps), (int)B.smaps.size(), dp<float>(B.d_score),
                        J.cand, J.n, m.padx, m.pady, dp<int>(m.d_rsize),
                        J.seg<NmsBox>(B.n_box), J.seg<unsigned int>(B.n_keys),
                        J.seg<int>(B.n_vals));
      }
      jobs.push_back(J);
      who.push_back({f, i});
    }
  }
  CHK(nms_run(c, jobs, overlap));
  // rows of every job's survivors, then one copy back per model
  int rc = GEYE_OK;
  for (int i = 0; i < nm && rc == GEYE_OK; i++) {
    Model &m = const_cast<Model &>(models[i]->m);
    Plan &B = *P[i];
    long long total = 0;
    for (size_t j = 0; j < jobs.size(); j++)
      if (who[j].second == i) total += jobs[j].nk;
    rc = ensure(B.d_rows, sizeof(float) * (size_t)std::max<long long>(total, 1) * B.GL);
    std::vector<float> rows((size_t)total * B.GL);
    long long at = 0;
    for (size_t j = 0; j < jobs.size() && rc == GEYE_OK; j++) {
      if (who[j].second != i || jobs[j].nk == 0) continue;
      NmsJob &J = jobs[j];
      launch_rows(J.s, dp<ScoreMap>(B.d_smaps), (int)B.smaps.size(), dp<ScorePart>(B.d_sparts),
                  dp<float>(B.d_score), J.cand, J.seg<int>(B.n_kept), J.nk, dp<int>(B.d_tmpIx),
                  dp<int>(B.d_tmpIy), window ? 1 : 0, B.GL, m.padx, m.pady, dp<int>(m.d_rsize),
                  dp<float>(B.d_rows) + at * B.GL);
      at += J.nk;
    }
    for (hipStream_t js : c->job_streams)
      if (rc == GEYE_OK && hipStreamSynchronize(js) != hipSuccess)
        rc = fail(GEYE_ERR_HIP, "batch rows");
    if (rc == GEYE_OK && total > 0 &&
        (hipMemcpyAsync(rows.data(), B.d_rows.p, sizeof(float) * rows.size(),
                        hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
         hipStreamSynchronize(c->stream) != hipSuccess))
      rc = fail(GEYE_ERR_HIP, "batch rows copy");
    at = 0;
    for (size_t j = 0; j < jobs.size() && rc == GEYE_OK; j++) {
      if (who[j].second != i) continue;
      const int k = who[j].first * nm + i;
      std::vector<float> r(rows.begin() + at * B.GL, rows.begin() + (at + jobs[j].nk) * B.GL);
      boxes[k] = to_malloc(r);
      num[k] = jobs[j].nk;
      GL[k] = B.GL;
      if (!boxes[k]) rc = fail(GEYE_ERR_NOMEM, "malloc");
      at += jobs[j].nk;
    }
  }
  if (rc != GEYE_OK) {
    for (int k = 0; k < nf * nm; k++) {
      std::free(boxes[k]);
      boxes[k] = nullptr;
      num[k] = GL[k] = 0;
    }
    return rc;
  }
  c->last_plan = P[nm - 1];
  flush_stats(c);
  return GEYE_OK;
}

// single-level plan pieces used by the stage entry points
struct Scratch {
  std::deque<DeviceBuffer> bufs;  // references stay valid as buffers are added
  ~Scratch() {
    for (auto &b : bufs) release(b);
  }
  DeviceBuffer &get() {
    bufs.emplace_back();
    return bufs.back();
  }
};

}  // namespace

// =======================================================================================
// Host NMS (CT/nms.cpp:70-217), exact: the reference's middle-pivot quicksort, the
// min-area overlap, the `while (cnum > 0)` loop that never emits the last survivor, and the
// output in original candidate order.
// =======================================================================================
namespace geye {

void quicksort_host(float *ary, int *order, int first, int last, const std::vector<float> *only) {
  while (true) {
    int i = first, j = last;
    const float key = ary[(first + last) / 2];
    while (true) {
      while (ary[i] > key) i++;
      while (ary[j] < key) j--;
      if (i >= j) break;
      std::swap(ary[i], ary[j]);
      std::swap(order[i], order[j]);
      i++;
      j--;
    }
    // [first, i-1] holds values >= key and [j+1, last] values <= key (i <= j + 1, so the two
    // ranges are disjoint); with `only`, a side is sorted further only if one of those
    // values can lie in it
    bool left = first < i - 1, right = last > j + 1;
    if (only) {
      bool ge = false, le = false;
      for (float v : *only) {
        ge = ge || v >= key;
        le = le || v <= key;
      }
      left = left && ge;
      right = right && le;
    }
    // recurse into the smaller side first to bound the stack; the order of the two
    // recursive calls does not change the result (they touch disjoint ranges)
    if (left && right) {
      if ((i - 1 - first) < (last - j - 1)) {
        quicksort_host(ary, order, first, i - 1, only);
        first = j + 1;
      } else {
        quicksort_host(ary, order, j + 1, last, only);
        last = i - 1;
      }
    } else if (left) {
      last = i - 1;
    } else if (right) {
      first = j + 1;
    } else {
      return;
    }
  }
}

int nms_host(const float *boxes, int NUM, int GL, float overlap, float *out) {
  if (NUM <= 0) return 0;
  std::vector<float> area(NUM), scores(NUM);
  std::vector<int> order(NUM), checked(NUM, 0);
  for (int i = 0; i < NUM; i++) {
    const float *P = boxes + (size_t)GL * i;
    area[i] = (float)(((double)(P[3] - P[1]) + 1.0) * ((double)(P[2] - P[0]) + 1.0));
    scores[i] = P[GL - 2];
    order[i] = i;
  }
  quicksort_host(scores.data(), order.data(), 0, NUM - 1);
  int cnum = NUM - 1, cur = 0, picked = 0;
  while (cnum > 0) {
    const int A = order[cur];
    const float *P = boxes + (size_t)GL * A;
    const float Ay1 = P[0], Ax1 = P[1], Ay2 = P[2], Ax2 = P[3];
    checked[A] = 1;
    cnum--;
    cur++;
    picked++;
    for (int kk = cur; kk < NUM; kk++) {
      const int B = order[kk];
      if (checked[B] != 0) continue;
      P = boxes + (size_t)GL * B;
      const float yy1 = Ay1 >= P[0] ? Ay1 : P[0];
      const float xx1 = Ax1 >= P[1] ? Ax1 : P[1];
      const float yy2 = Ay2 <= P[2] ? Ay2 : P[2];
      const float xx2 = Ax2 <= P[3] ? Ax2 : P[3];
      const float w = (float)((double)(xx2 - xx1) + 1.0);
      const float h = (float)((double)(yy2 - yy1) + 1.0);
      const float R = area[A] <= area[B] ? area[A] : area[B];
      if (w > 0 && h > 0) {
        if (w * h / R > overlap) {
          checked[B] = -1;
          cnum--;
        }
      } else if ((Ay1 < P[0] && Ax1 < P[1] && Ay2 > P[2] && Ax2 > P[3]) ||
                 (Ay1 > P[0] && Ax1 > P[1] && Ay2 < P[2] && Ax2 < P[3])) {
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
  for (int ii = 0; count < picked; ii++)
    if (checked[ii] == 1) {
      std::memcpy(out + (size_t)count * GL, boxes + (size_t)GL * ii, sizeof(float) * GL);
      count++;
    }
  return picked;
}

}  // namespace geye

int geye::ctx_create(int device, bool own_queue, geye_ctx **ctx) {
  if (!ctx) return fail(GEYE_ERR_ARG, "null ctx");
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(GEYE_ERR_ARG, "no such device");
  HIPCHK(hipSetDevice(device));
  geye_ctx *c = new geye_ctx();
  c->device = device;
  if (hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      c->num_cu <= 0)
    c->num_cu = 256;
  // own_queue (a serving worker's context): the stream is created with a CU mask (every CU
  // enabled), which gives it a hardware queue of its own.  Plain streams share the process's
  // pool of GPU_MAX_HW_QUEUES queues (4 unless set): a server's four workers beside the
  // caller's context were five streams on four queues, and the two sharing one ran one after
  // the other (1920x1080 bf16x3 at 4 slots: 697 frames/s with plain streams, 795 with masked
  // ones, every process).  Other contexts keep a plain stream, which leaves the device's
  // queues to the workers (4 slots: 829-833 frames/s with every context masked, 847-855
  // with the workers' alone).  A plain stream is also the fallback.
  hipError_t e = hipErrorNotSupported;
  if (own_queue) {
    std::vector<uint32_t> cu_all((c->num_cu + 31) / 32, 0xffffffffu);
    e = hipExtStreamCreateWithCUMask(&c->stream, (uint32_t)cu_all.size(), cu_all.data());
    if (e != hipSuccess) (void)hipGetLastError();
  }
  if (e != hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return fail(GEYE_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  // the NMS rounds' alive counts: written by the compaction kernel into coherent pinned
  // memory, read by the host after the round's event
  if (hipHostMalloc((void **)&c->h_count, 64, hipHostMallocCoherent | hipHostMallocMapped) !=
          hipSuccess ||
      hipHostGetDevicePointer((void **)&c->d_hcount, c->h_count, 0) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_count[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_count[1], hipEventDisableTiming) != hipSuccess) {
    if (c->h_count) (void)hipHostFree(c->h_count);
    for (hipEvent_t ev : c->ev_count)
      if (ev) (void)hipEventDestroy(ev);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return fail(GEYE_ERR_HIP, "pinned count buffer / events");
  }
  // the allocation-free entry point's scratch array on this device (its address is
  // resolved here, once per device, with the device current); a failure is reported by
  // that entry point, not here
  (void)noalloc_setup(device);
  *ctx = c;
  return GEYE_OK;
}

// =======================================================================================
// C ABI
// =======================================================================================
extern "C" {

const char *geye_last_error(void) { return g_err.c_str(); }

int geye_device_count(int *n) {
  if (!n) return fail(GEYE_ERR_ARG, "null");
  HIPCHK(hipGetDeviceCount(n));
  return GEYE_OK;
}

int geye_init(int device, geye_ctx **ctx) { return geye::ctx_create(device, false, ctx); }

void geye_shutdown(geye_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  flush_stats(c);
  for (auto &kv : c->batch_graphs) (void)hipGraphExecDestroy(kv.second.first);
  c->batch_graphs.clear();
  release_plans(c, nullptr, -1);
  for (auto &b : c->scratch_bufs) release(b);
  if (c->h_count) (void)hipHostFree(c->h_count);
  for (hipEvent_t e : c->ev_count)
    if (e) (void)hipEventDestroy(e);
  for (hipStream_t js : c->job_streams)
    if (js) (void)hipStreamDestroy(js);
  (void)hipStreamDestroy(c->stream);
  delete c;
}

int geye_model_create(geye_ctx *c, const geye_model_desc *desc, geye_model **model) {
  if (!c || !desc || !model) return fail(GEYE_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  HIPCHK(hipSetDevice(c->device));
  (void)hipGetLastError();
  geye_model *gm = new geye_model();
  int rc = model_from_desc(c, desc, gm->m);
  if (rc != GEYE_OK) {
    geye_model_free(gm);
    return rc;
  }
  *model = gm;
  return GEYE_OK;
}

int geye_model_load_csv(geye_ctx *c, const char *comp, const char *root, const char *part,
                        geye_model **model) {
  if (!c || !comp || !root || !part || !model) return fail(GEYE_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  HIPCHK(hipSetDevice(c->device));
  (void)hipGetLastError();
  geye_model *gm = new geye_model();
  int rc = load_csv(c, comp, root, part, gm->m);
  if (rc != GEYE_OK) {
    geye_model_free(gm);
    return rc;
  }
  *model = gm;
  return GEYE_OK;
}

void geye_model_free(geye_model *gm) {
  if (!gm) return;
  Model &m = gm->m;
  release_plans(nullptr, &m, -1);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (auto it = g_tuned.begin(); it != g_tuned.end();)
      it = std::get<0>(it->first) == &m ? g_tuned.erase(it) : std::next(it);
  }
  release(m.d_filters);
  release(m.d_rsize);
  for (auto &b : m.d_filt16) release(b);
  release(m.d_lsq);
  release(m.d_lsq_idx);
  delete gm;
}

int geye_model_info(const geye_model *gm, int *numcomponent, int *sbin, int *interval,
                    int *max_Y, int *max_X, int *NoR, int *NoP) {
  if (!gm) return fail(GEYE_ERR_ARG, "null model");
  const Model &m = gm->m;
  if (numcomponent) *numcomponent = m.numcomponent;
  if (sbin) *sbin = m.sbin;
  if (interval) *interval = m.interval;
  if (max_Y) *max_Y = m.max_Y;
  if (max_X) *max_X = m.max_X;
  if (NoR) *NoR = m.NoR;
  if (NoP) *NoP = m.NoP;
  return GEYE_OK;
}

int geye_level_plan(const geye_model *gm, int width, int height, int *max_scale, int *LEN,
                    float *scales, int *feat_h, int *feat_w) {
  if (!gm || !max_scale || !LEN) return fail(GEYE_ERR_ARG, "null argument");
  const Model &m = gm->m;
  int ms = compute_max_scale(m, width, height);
  std::vector<Level> lv;
  std::vector<int> ih, iw, is;
  std::vector<float> isc;
  level_geometry(ms, m.sbin, m.interval, width, height, lv, ih, iw, is, isc);
  *max_scale = ms;
  *LEN = (int)lv.size();
  for (size_t l = 0; l < lv.size(); l++) {
    if (scales) scales[l] = lv[l].scale;
    if (feat_h) feat_h[l] = lv[l].F0;
    if (feat_w) feat_w[l] = lv[l].F1;
  }
  return GEYE_OK;
}

int geye_level_plan_dims(const int *rsize, int numcomponent, int sbin, int interval, int width,
                         int height, int *max_scale, int *LEN, float *scales, int *feat_h,
                         int *feat_w) {
  if (!rsize || numcomponent <= 0 || sbin <= 0 || interval <= 0 || !max_scale || !LEN)
    return fail(GEYE_ERR_ARG, "bad argument");
  Model m;
  m.numcomponent = numcomponent;
  m.sbin = sbin;
  m.interval = interval;
  m.rsize.assign(rsize, rsize + 2 * numcomponent);
  const int ms = compute_max_scale(m, width, height);
  std::vector<Level> lv;
  std::vector<int> ih, iw, is;
  std::vector<float> isc;
  level_geometry(ms, sbin, interval, width, height, lv, ih, iw, is, isc);
  *max_scale = ms;
  *LEN = (int)lv.size();
  for (size_t l = 0; l < lv.size(); l++) {
    if (scales) scales[l] = lv[l].scale;
    if (feat_h) feat_h[l] = lv[l].F0;
    if (feat_w) feat_w[l] = lv[l].F1;
  }
  return GEYE_OK;
}

int geye_nms_host(const float *boxes, int num, int GL, float overlap, float *out, int *out_num) {
  if (!out_num || (num > 0 && (!boxes || !out)) || GL < 7) return fail(GEYE_ERR_ARG, "bad argument");
  *out_num = nms_host(boxes, num, GL, overlap, out);
  return GEYE_OK;
}

int geye_get_boxes(geye_ctx *c, const geye_model *model, const unsigned char *bgr, int width,
                   int height, int widthStep, float thresh, int precision, float **boxes,
                   int *num, int *GL) {
  return detect_common(c, model, bgr, false, width, height, widthStep, thresh, 0.0f, precision,
                       false, boxes, num, GL);
}

int geye_detect(geye_ctx *c, const geye_model *model, const unsigned char *bgr, int width,
                int height, int widthStep, float thresh, float overlap, int precision,
                float **boxes, int *num, int *GL) {
  return detect_common(c, model, bgr, false, width, height, widthStep, thresh, overlap,
                       precision, true, boxes, num, GL);
}

int geye_detect_dev(geye_ctx *c, const geye_model *model, const void *d_bgr, int width,
                    int height, int widthStep, float thresh, float overlap, int precision,
                    float **boxes, int *num, int *GL) {
  return detect_common(c, model, (const unsigned char *)d_bgr, true, width, height, widthStep,
                       thresh, overlap, precision, true, boxes, num, GL);
}

int geye_detect_models(geye_ctx *c, const geye_model *const *models, int nmodels,
                       const unsigned char *bgr, int width, int height, int widthStep,
                       float thresh, float overlap, int precision, float **boxes, int *num,
                       int *GL) {
  return detect_models(c, models, nmodels, bgr, false, width, height, widthStep, thresh, overlap,
                       precision, boxes, num, GL);
}

int geye_detect_models_dev(geye_ctx *c, const geye_model *const *models, int nmodels,
                           const void *d_bgr, int width, int height, int widthStep, float thresh,
                           float overlap, int precision, float **boxes, int *num, int *GL) {
  if (!d_bgr) return fail(GEYE_ERR_ARG, "null frame");
  return detect_models(c, models, nmodels, (const unsigned char *)d_bgr, true, width, height,
                       widthStep, thresh, overlap, precision, boxes, num, GL);
}

int geye_detect_batch(geye_ctx *c, const geye_model *const *models, int nmodels,
                      const unsigned char *const *bgr, int nframes, int width, int height,
                      int widthStep, float thresh, float overlap, int precision, float **boxes,
                      int *num, int *GL) {
  return detect_batch(c, models, nmodels, bgr, false, nframes, width, height, widthStep, thresh,
                      overlap, precision, boxes, num, GL);
}

int geye_detect_batch_dev(geye_ctx *c, const geye_model *const *models, int nmodels,
                          const void *const *d_bgr, int nframes, int width, int height,
                          int widthStep, float thresh, float overlap, int precision,
                          float **boxes, int *num, int *GL) {
  return detect_batch(c, models, nmodels, (const unsigned char *const *)d_bgr, true, nframes,
                      width, height, widthStep, thresh, overlap, precision, boxes, num, GL);
}

int geye_last_scores(geye_ctx *c, float *out, long long cap, long long *total) {
  if (!c || !total) return fail(GEYE_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  Plan *P = c->last_plan;
  if (!P) return fail(GEYE_ERR_ARG, "no detection has run on this context");
  *total = P->score_total;
  if (out && cap > 0) {
    long long n = std::min(cap, P->score_total);
    HIPCHK(hipMemcpyAsync(out, P->d_score.p, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return GEYE_OK;
}

namespace {
// A_SCORE of the last detection into device memory d_ac (IHEI x IWID column-major, read and
// overwritten).
int ascore_run(geye_ctx *c, Plan &P, float thresh, float *d_ac) {
  CHK(ensure(P.d_ascore, sizeof(float) * ((size_t)P.W * P.H + 64)));
  unsigned int *maxbits = (unsigned int *)(dp<float>(P.d_ascore) + (size_t)P.W * P.H);
  launch_fill(c->stream, (int *)maxbits, 1, 0);
  {
    KTimer t(c, "ascore");
    launch_ascore(c->stream, dp<ScoreMap>(P.d_smaps), (int)P.smaps.size(), dp<int>(P.d_ascore_tab),
                  dp<float>(P.d_score), P.W, P.H, thresh, d_ac, maxbits);
  }
  HIPCHK(hipGetLastError());
  return GEYE_OK;
}
}  // namespace

int geye_accumulated_score(geye_ctx *c, float thresh, float *a_score) {
  if (!c || !a_score) return fail(GEYE_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  HIPCHK(hipSetDevice(c->device));
  (void)hipGetLastError();
  Plan *P = c->last_plan;
  if (!P) return fail(GEYE_ERR_ARG, "no detection has run on this context");
  const size_t n = (size_t)P->W * P->H;
  CHK(ensure(P->d_ascore, sizeof(float) * (n + 64)));
  float *d_ac = dp<float>(P->d_ascore);
  HIPCHK(hipMemcpyAsync(d_ac, a_score, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
  CHK(ascore_run(c, *P, thresh, d_ac));
  HIPCHK(hipMemcpyAsync(a_score, d_ac, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  flush_stats(c);
  return GEYE_OK;
}

int geye_accumulated_score_dev(geye_ctx *c, float thresh, void *d_a_score) {
  if (!c || !d_a_score) return fail(GEYE_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  HIPCHK(hipSetDevice(c->device));
  (void)hipGetLastError();
  Plan *P = c->last_plan;
  if (!P) return fail(GEYE_ERR_ARG, "no detection has run on this context");
  CHK(ascore_run(c, *P, thresh, (float *)d_a_score));
  HIPCHK(hipStreamSynchronize(c->stream));
  flush_stats(c);
  return GEYE_OK;
}

void geye_free(void *p) { std::free(p); }

int geye_new_rects(geye_ctx *c, const geye_model *model, const float *boxes, int num, int GL,
                   int width, int height, float ratio, int *point, int *or_point, int *type,
                   float *scale, float *score) {
  if (!c || !model || num < 0 || (num > 0 && (!boxes || !point || !or_point || !type || !scale ||
                                              !score)))
    return fail(GEYE_ERR_ARG, "null argument");
  if (num == 0) return GEYE_OK;
  const Model &m = model->m;
  if (m.lsq.empty()) return fail(GEYE_ERR_ARG, "model has no least-squares box coefficients");
  const int ML = 1 + 2 * (1 + m.numpart[0]);
  if (GL != (m.numpart[0] + 1) * 4 + 3) return fail(GEYE_ERR_ARG, "GL does not match the model");
  std::lock_guard<std::mutex> lk(c->mu);
  HIPCHK(hipSetDevice(c->device));
  (void)hipGetLastError();
  Scratch S;
  DeviceBuffer &drows = S.get(), &dint = S.get(), &dflt = S.get();
  const size_t n = (size_t)num;
  CHK(ensure(drows, sizeof(float) * n * GL));
  CHK(ensure(dint, sizeof(int) * (n * 9 + 1)));
  CHK(ensure(dflt, sizeof(float) * n * 2));
  int *d_point = dp<int>(dint), *d_or = d_point + 4 * n, *d_type = d_or + 4 * n, *d_bad = d_type + n;
  float *d_scale = dp<float>(dflt), *d_score = d_scale + n;
  HIPCHK(hipMemcpyAsync(drows.p, boxes, sizeof(float) * n * GL, hipMemcpyHostToDevice, c->stream));
  launch_fill(c->stream, d_bad, 1, 0);
  {
    KTimer t(c, "new_rects");
    launch_new_rects(c->stream, dp<float>(drows), num, GL, ML, (const float *)m.d_lsq.p,
                     (const int *)m.d_lsq_idx.p, (const int *)m.d_lsq_idx.p + m.numcomponent,
                     m.numcomponent, width, height, ratio, d_point, d_or, d_type, d_scale,
                     d_score, d_bad);
  }
  HIPCHK(hipGetLastError());
  int bad = 0;
  HIPCHK(hipMemcpyAsync(point, d_point, sizeof(int) * 4 * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(or_point, d_or, sizeof(int) * 4 * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(type, d_type, sizeof(int) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(scale, d_scale, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(score, d_score, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  flush_stats(c);
  if (bad) return fail(GEYE_ERR_ARG, "a row's component index is outside the model");
  return GEYE_OK;
}

int geye_add_part(geye_ctx *c, float *score, const float *M, const int *rootsize,
                  const int *partsize, int ax, int ay) {
  if (!c || !score || !M || !rootsize || !partsize) return fail(GEYE_ERR_ARG, "null argument");
  const int R0 = rootsize[0], R1 = rootsize[1], P0 = partsize[0], P1 = partsize[1];
  if (R0 <= 0 || R1 <= 0 || P0 <= 0 || P1 <= 0) return fail(GEYE_ERR_ARG, "empty map");
  if (ay < 1 || ax < 1 || ay - 1 + 2 * (R0 - 1) >= P0 || ax - 1 + 2 * (R1 - 1) >= P1)
    return fail(GEYE_ERR_ARG, "anchor reads outside the part map");
  std::lock_guard<std::mutex> lk(c->mu);
  HIPCHK(hipSetDevice(c->device));
  (void)hipGetLastError();
  Scratch S;
  DeviceBuffer &ds = S.get(), &dm = S.get();
  const size_t nr = (size_t)R0 * R1, np = (size_t)P0 * P1;
  CHK(ensure(ds, sizeof(float) * nr));
  CHK(ensure(dm, sizeof(float) * np));
  HIPCHK(hipMemcpyAsync(ds.p, score, sizeof(float) * nr, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(dm.p, M, sizeof(float) * np, hipMemcpyHostToDevice, c->stream));
  {
    KTimer t(c, "add_part");
    launch_add_part(c->stream, dp<float>(ds), dp<float>(dm), R0, R1, P0, ax, ay);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(score, ds.p, sizeof(float) * nr, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  flush_stats(c);
  return GEYE_OK;
}

int geye_get_boxes_features(geye_ctx *c, const geye_model *model, int width, int height,
                            int max_scale, const float *feat, long long nfeat, float thresh,
                            float **boxes, int *num, int *GL) {
  if (!c || !model || !feat || !boxes || !num || !GL) return fail(GEYE_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  HIPCHK(hipSetDevice(c->device));
  (void)hipGetLastError();
  Model &m = const_cast<Model &>(model->m);
  Plan *P = nullptr;
  CHK(get_plan(c, m, width, height, &P, max_scale));
  long long total = 0;
  auto lvl_n = [&](int l) { return 31LL * std::max(0, P->lv[l].F0) * std::max(0, P->lv[l].F1); };
  for (int l = 0; l < P->LEN; l++) total += lvl_n(l);
  if (nfeat != total) return fail(GEYE_ERR_ARG, "feature count does not match the level plan");
  CHK(ensure(P->d_feat, sizeof(float) * std::max<long long>(P->feat_total, 1)));
  long long off = 0;
  for (int l = 0; l < P->LEN; l++) {
    const long long n = lvl_n(l);
    if (n > 0 && P->feat_off[l] >= 0)
      HIPCHK(hipMemcpyAsync(dp<float>(P->d_feat) + P->feat_off[l], feat + off, sizeof(float) * n,
                            hipMemcpyHostToDevice, c->stream));
    off += n;
  }
  int ncand = 0;
  CHK(run_scores(c, m, *P, thresh, GEYE_FP32, &ncand));
  std::vector<float> rows;
  c->last_plan = P;
  CHK(run_rows(c, m, *P, nullptr, ncand, dt_windowed(c, GEYE_FP32), rows));
  *num = ncand;
  *boxes = to_malloc(rows);
  *GL = P->GL;
  flush_stats(c);
  return *boxes ? GEYE_OK : fail(GEYE_ERR_NOMEM, "malloc");
}

int geye_nms(geye_ctx *c, const float *boxes, int num, int GL, float overlap, float *out,
             int *out_num) {
  if (!c || !out_num || (num > 0 && (!boxes || !out)) || GL < 7)
    return fail(GEYE_ERR_ARG, "bad argument");
  *out_num = 0;
  if (num <= 0) return GEYE_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  HIPCHK(hipSetDevice(c->device));
  (void)hipGetLastError();
  Plan B;
  int rc = nms_buffers(B, num);
  DeviceBuffer rows;
  if (rc == GEYE_OK) rc = ensure(rows, sizeof(float) * (size_t)num * GL);
  if (rc == GEYE_OK) {
    hipError_t e = hipMemcpyAsync(rows.p, boxes, sizeof(float) * (size_t)num * GL,
                                  hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) rc = fail(GEYE_ERR_HIP, hipGetErrorString(e));
  }
  int nk = 0;
  if (rc == GEYE_OK) {
    launch_nms_prep_rows(c->stream, dp<float>(rows), num, GL, dp<NmsBox>(B.n_box),
                         dp<unsigned int>(B.n_keys), dp<int>(B.n_vals));
    // the indexed pass needs overlap >= 0.5 and boxes with x2 >= x1, y2 >= y1 (then a box
    // that contains another also intersects it); the frame is the boxes' extent
    bool grid_ok = overlap >= 0.5f;
    float ex = 0.0f, ey = 0.0f;
    for (int i = 0; i < num && grid_ok; i++) {
      const float *r = boxes + (size_t)i * GL;
      grid_ok = r[2] >= r[0] && r[3] >= r[1] && std::isfinite(r[2]) && std::isfinite(r[3]) &&
                std::isfinite(r[0]) && std::isfinite(r[1]);
      ey = std::max(ey, r[2]);
      ex = std::max(ex, r[3]);
    }
    NmsGrid g;
    if (grid_ok)
      rc = nms_grid_buffers(B, num, (int)std::min(ex + 2.0f, 1e6f), (int)std::min(ey + 2.0f, 1e6f),
                            &g);
    if (rc == GEYE_OK) rc = nms_core(c, B, num, overlap, grid_ok ? &g : nullptr, &nk);
  }
  std::vector<int> kept(nk);
  if (rc == GEYE_OK && nk > 0) {
    hipError_t e = hipMemcpyAsync(kept.data(), B.n_kept.p, sizeof(int) * nk,
                                  hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = fail(GEYE_ERR_HIP, hipGetErrorString(e));
  }
  if (rc == GEYE_OK) {
    for (int k = 0; k < nk; k++)
      std::memcpy(out + (size_t)k * GL, boxes + (size_t)kept[k] * GL, sizeof(float) * GL);
    *out_num = nk;
  }
  release(rows);
  plan_release(B);
  return rc;
}

// ---- stage entry points ----------------------------------------------------------------

int geye_resize(geye_ctx *c, const float *src, const int *sdims, float scale, float *dst,
                int *odims) {
  if (!c || !src || !sdims || !odims) return fail(GEYE_ERR_ARG, "null argument");
  const int h = sdims[0], w = sdims[1], ch = sdims[2];
  if (h < 1 || w < 1 || ch != 3) return fail(GEYE_ERR_ARG, "resize: sdims must be (h, w, 3)");
  if (scale == 1.0f) {
    odims[0] = h; odims[1] = w; odims[2] = ch;
    if (dst) std::memcpy(dst, src, sizeof(float) * 3 * h * w);
    return GEYE_OK;
  }
  const int oh = (int)((double)((float)h * scale) + 0.5);
  const int ow = (int)((double)((float)w * scale) + 0.5);
  odims[0] = oh; odims[1] = ow; odims[2] = ch;
  if (!dst) return GEYE_OK;
  if (oh < 1 || ow < 1) return fail(GEYE_ERR_ARG, "resize: empty output");
  std::lock_guard<std::mutex> lk(c->mu);
  HIPCHK(hipSetDevice(c->device));
  (void)hipGetLastError();
  // column-major [c][x][y] -> row-major [c][y][x]
  std::vector<float> in((size_t)3 * h * w);
  for (int cc = 0; cc < 3; cc++)
    for (int x = 0; x < w; x++)
      for (int y = 0; y < h; y++) in[((size_t)cc * h + y) * w + x] = src[((size_t)cc * w + x) * h + y];
  std::vector<int2> tab;
  std::vector<int> si;
  std::vector<float> al;
  resize_tables(h, oh, tab, si, al);
  const int tab2 = (int)tab.size();
  resize_tables(w, ow, tab, si, al);
  ResizeOp ops[2] = {{0, 0, h, w, oh, w, 0, 0}, {0, 3LL * h * w, oh, w, oh, ow, tab2, 0}};
  Scratch S;
  DeviceBuffer &dops = S.get(), &dtab = S.get(), &dsi = S.get(), &dal = S.get(),
               &dimg = S.get(), &dtmp = S.get();
  CHK(upload(dops, std::vector<ResizeOp>(ops, ops + 2), c->stream));
  CHK(upload(dtab, tab, c->stream));
  CHK(upload(dsi, si, c->stream));
  CHK(upload(dal, al, c->stream));
  CHK(ensure(dimg, sizeof(float) * (3LL * h * w + 3LL * oh * ow)));
  CHK(ensure(dtmp, sizeof(float) * 3LL * oh * w));
  HIPCHK(hipMemcpyAsync(dimg.p, in.data(), sizeof(float) * in.size(), hipMemcpyHostToDevice,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  launch_resize_rows(c->stream, dp<ResizeOp>(dops), 1, (int)((3LL * oh * w + 255) / 256),
                     dp<int2>(dtab), dp<int>(dsi), dp<float>(dal), dp<float>(dimg), dp<float>(dtmp));
  launch_resize_cols(c->stream, dp<ResizeOp>(dops) + 1, 1, (int)((3LL * oh * ow + 255) / 256),
                     dp<int2>(dtab), dp<int>(dsi), dp<float>(dal), dp<float>(dtmp), dp<float>(dimg));
  HIPCHK(hipGetLastError());
  std::vector<float> outr((size_t)3 * oh * ow);
  HIPCHK(hipMemcpyAsync(outr.data(), dp<float>(dimg) + 3LL * h * w, sizeof(float) * outr.size(),
                        hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int cc = 0; cc < 3; cc++)
    for (int x = 0; x < ow; x++)
      for (int y = 0; y < oh; y++) dst[((size_t)cc * ow + x) * oh + y] = outr[((size_t)cc * oh + y) * ow + x];
  return GEYE_OK;
}

int geye_calc_feature(geye_ctx *c, const float *img, int height, int width, int sbin,
                      float *feat, int *F0, int *F1) {
  if (!c || !img || !F0 || !F1) return fail(GEYE_ERR_ARG, "null argument");
  if (sbin < 2 || sbin > 8) return fail(GEYE_ERR_UNSUPPORTED, "sbin outside [2, 8]");
  if (height < 3 || width < 3) return fail(GEYE_ERR_ARG, "image smaller than 3x3");
  const int b0 = (int)std::floor(double(height) / double(sbin) + 0.5);
  const int b1 = (int)std::floor(double(width) / double(sbin) + 0.5);
  *F0 = std::max(b0 - 2, 0);
  *F1 = std::max(b1 - 2, 0);
  if (!feat || *F0 == 0 || *F1 == 0) return GEYE_OK;
  std::lock_guard<std::mutex> lk(c->mu);
  HIPCHK(hipSetDevice(c->device));
  (void)hipGetLastError();
  const int h = height, w = width;
  std::vector<float> in((size_t)3 * h * w);
  for (int cc = 0; cc < 3; cc++)
    for (int x = 0; x < w; x++)
      for (int y = 0; y < h; y++) in[((size_t)cc * h + y) * w + x] = img[((size_t)cc * w + x) * h + y];
  HogLevel L{};
  L.img_off = 0;
  L.h = h; L.w = w; L.sbin = sbin; L.b0 = b0; L.b1 = b1; L.F0 = *F0; L.F1 = *F1;
  L.T = kHogVoteRows;
  L.tiles_x = (b1 + kHogVoteCols - 1) / kHogVoteCols;
  L.tiles_y = (b0 + L.T - 1) / L.T;
  L.blocks_begin = 0;
  L.grid_off = 0;
  GradImg G{};
  G.h = h;
  G.w = w;
  G.tiles_x = (w + 63) / 64;
  G.tiles_y = (h + kGradTY - 1) / kGradTY;
  G.frame = -1;
  Scratch S;
  DeviceBuffer &dl = S.get(), &dimg = S.get(), &dh = S.get(), &dn = S.get(), &df = S.get(),
               &dg = S.get(), &dgrid = S.get();
  CHK(upload(dl, std::vector<HogLevel>{L}, c->stream));
  CHK(upload(dg, std::vector<GradImg>{G}, c->stream));
  CHK(upload(dimg, in, c->stream));
  const long long grid_total = (long long)grid_stride(h) * w;
  CHK(ensure(dgrid, 5 * grid_total));
  CHK(ensure(dh, sizeof(float) * 18LL * b0 * b1));
  CHK(ensure(dn, sizeof(float) * (size_t)b0 * b1));
  CHK(ensure(df, sizeof(float) * 31LL * *F0 * *F1));
  launch_hog_grad(c->stream, dp<GradImg>(dg), 1, G.tiles_x * G.tiles_y, dp<float>(dimg),
                  dgrid.p, grid_total);
  launch_hog_votes(c->stream, dp<HogLevel>(dl), 1, L.tiles_x * L.tiles_y, dgrid.p, grid_total,
                   dp<float>(dh), dp<float>(dn));
  launch_hog_feat(c->stream, dp<HogLevel>(dl), 1, (int)((1LL * *F0 * *F1 + 255) / 256),
                  dp<float>(dh), dp<float>(dn), dp<float>(df), nullptr, 0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(feat, df.p, sizeof(float) * 31LL * *F0 * *F1, hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return GEYE_OK;
}

int geye_pyramid(geye_ctx *c, const unsigned char *bgr, int width, int height, int widthStep,
                 int sbin, int interval, int max_scale, float *feat, long long cap, int *feat_h,
                 int *feat_w, float *scales) {
  if (!c || !bgr || !feat_h || !feat_w || !scales) return fail(GEYE_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  HIPCHK(hipSetDevice(c->device));
  (void)hipGetLastError();
  Model m;
  m.sbin = sbin;
  m.interval = interval;
  m.numcomponent = 0;
  Plan P;
  std::string err;
  int rc = build_plan(m, width, height, P, err, max_scale);
  if (rc != GEYE_OK) return fail(rc, err);
  rc = plan_upload(P, c->stream);
  if (rc == GEYE_OK) {
    const unsigned char *d_bgr = nullptr;
    rc = frame_in(c, P, bgr, width, height, widthStep, &d_bgr);
    if (rc == GEYE_OK) rc = run_pyramid(c, P, d_bgr, widthStep, 0);
  }
  if (rc == GEYE_OK) {
    for (int l = 0; l < P.LEN; l++) {
      feat_h[l] = P.lv[l].F0;
      feat_w[l] = P.lv[l].F1;
      scales[l] = P.lv[l].scale;
    }
    if (feat) {
      // the feature arena holds the levels with F0*F1 > 0 in level order, i.e. exactly the
      // reference's concatenation of all levels
      long long n = std::min(cap, P.feat_total);
      hipError_t e = hipMemcpyAsync(feat, P.d_feat.p, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      if (e != hipSuccess) rc = fail(GEYE_ERR_HIP, hipGetErrorString(e));
    }
  }
  plan_release(P);
  return rc;
}

int geye_fconv(geye_ctx *c, const float *feat, int F0, int F1, const float *filters, int n,
               int h, int w, int precision, float *out) {
  if (!c || !feat || !filters || !out || n <= 0) return fail(GEYE_ERR_ARG, "null argument");
  if (h < 1 || w < 1 || h > 16 || w > 16) return fail(GEYE_ERR_UNSUPPORTED, "filter size outside 1..16");
  const int H = F0 - h + 1, W = F1 - w + 1;
  if (H < 1 || W < 1) return fail(GEYE_ERR_ARG, "filter larger than the feature map");
  if (!valid_precision(precision)) return fail(GEYE_ERR_ARG, "unknown precision");
  std::lock_guard<std::mutex> lk(c->mu);
  HIPCHK(hipSetDevice(c->device));
  (void)hipGetLastError();
  auto wt = [&](int fid, int ch, int xp, int yp) -> float {
    if (fid >= n || ch >= 31) return 0.0f;
    return filters[(size_t)fid * 31 * h * w + (size_t)ch * h * w + (size_t)xp * h + yp];
  };
  Scratch S;
  DeviceBuffer &dfe = S.get(), &dfi = S.get(), &dt = S.get(), &dm = S.get(), &dd = S.get(),
               &dout = S.get();
  std::vector<long long> moff;
  std::vector<int2> mdims;
  if (precision == GEYE_FP32) {
    const int nf = pick_nf(n);
    const int ng = (n + nf - 1) / nf;
    std::vector<float> packed;
    for (int g = 0; g < ng; g++)
      for (int f = 0; f < 31; f++)
        for (int xp = 0; xp < w; xp++)
          for (int yp = 0; yp < h; yp++)
            for (int k = 0; k < nf; k++) packed.push_back(wt(g * nf + k, f, xp, yp));
    std::vector<ConvTile> tiles;
    for (int k = 0; k < ng * nf; k++) {
      moff.push_back((long long)std::min(k, n - 1) * H * W);
      mdims.push_back(k < n ? int2{H, W} : int2{0, 0});
    }
    for (int g = 0; g < ng; g++)
      for (int x0 = 0; x0 < W; x0 += 16)
        for (int y0 = 0; y0 < H; y0 += 64)
          tiles.push_back(ConvTile{0, F0, F1, H, W, x0, y0, (long long)g * 31 * w * h * nf,
                                   g * nf, std::min(nf, n - g * nf)});
    CHK(upload(dfe, std::vector<float>(feat, feat + 31LL * F0 * F1), c->stream));
    CHK(upload(dfi, packed, c->stream));
    CHK(upload(dt, tiles, c->stream));
    CHK(upload(dm, moff, c->stream));
    CHK(upload(dd, mdims, c->stream));
    CHK(ensure(dout, sizeof(float) * (size_t)n * H * W));
    launch_conv_f32(c->stream, nf, dp<ConvTile>(dt), (int)tiles.size(), h, w, dp<float>(dfe),
                    dp<float>(dfi), dp<long long>(dm), dp<int2>(dd), dp<float>(dout));
  } else {
    const int nt = std::min(3, (n + 15) / 16), gsz = 16 * nt, ng = (n + gsz - 1) / gsz;
    const bool split = split_precision(precision);
    const int nh = split ? 2 : 1, cw = 32 * nh;   // channels per cell: [hi 32][lo 32] if split
    std::vector<unsigned short> packed, fb((size_t)F0 * F1 * cw, 0);
    unsigned short hi, lo;
    for (int g = 0; g < ng; g++)
      for (int xp = 0; xp < w; xp++)
        for (int yp = 0; yp < h; yp++)
          for (int t = 0; t < nt; t++)
            for (int hh = 0; hh < nh; hh++)
              for (int ch = 0; ch < 4; ch++)
                for (int k = 0; k < 16; k++)
                  for (int j = 0; j < 8; j++) {
                    mfma_operand(wt(g * gsz + t * 16 + k, ch * 8 + j, xp, yp), precision, hi, lo);
                    packed.push_back(hh ? lo : hi);
                  }
    for (int x = 0; x < F1; x++)
      for (int y = 0; y < F0; y++)
        for (int f = 0; f < 31; f++) {
          mfma_operand(feat[(size_t)f * F0 * F1 + (size_t)x * F0 + y], precision, hi, lo);
          fb[((size_t)x * F0 + y) * cw + f] = hi;
          if (split) fb[((size_t)x * F0 + y) * cw + 32 + f] = lo;
        }
    std::vector<MfmaTile> tiles;  // 4-column tiles, then the same outputs as 8-column tiles
    for (int k = 0; k < ng * gsz; k++) {
      moff.push_back((long long)std::min(k, n - 1) * H * W);
      mdims.push_back(k < n ? int2{H, W} : int2{0, 0});
    }
    int n4 = 0;
    for (int tw = 4; tw <= 8; tw += 4) {
      for (int g = 0; g < ng; g++)
        for (int y0 = 0; y0 < H; y0 += 64)
          for (int x0 = 0; x0 < W; x0 += tw)
            tiles.push_back(MfmaTile{0, F0, F1, x0, y0, (long long)g * w * h * nt * 512,
                                     g * gsz, std::min(gsz, n - g * gsz)});
      if (tw == 4) n4 = (int)tiles.size();
    }
    CHK(upload(dfe, fb, c->stream));
    CHK(upload(dfi, packed, c->stream));
    CHK(upload(dt, tiles, c->stream));
    CHK(upload(dm, moff, c->stream));
    CHK(upload(dd, mdims, c->stream));
    CHK(ensure(dout, sizeof(float) * (size_t)n * H * W));
    if (launch_conv_mfma(c->stream, mfma_kind(precision), split, nt, dp<MfmaTile>(dt), n4,
                         dp<MfmaTile>(dt) + n4, (int)tiles.size() - n4, h, w,
                         dfe.p, dfi.p, dp<long long>(dm), dp<int2>(dd), dp<float>(dout)) != 0)
      return fail(GEYE_ERR_UNSUPPORTED, "filter too large for the MFMA kernel's LDS");
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, dout.p, sizeof(float) * (size_t)n * H * W, hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return GEYE_OK;
}

static int dt_standalone(geye_ctx *c, const float *vals, float ax, float bx, float ay, float by,
                         const int *dims, float *M, int *Ix, int *Iy, bool window) {
  if (!c || !vals || !dims || !M || !Ix || !Iy) return fail(GEYE_ERR_ARG, "null argument");
  const int d0 = dims[0], d1 = dims[1];
  if (d0 < 1 || d1 < 1) return fail(GEYE_ERR_ARG, "dt: empty map");
  std::lock_guard<std::mutex> lk(c->mu);
  HIPCHK(hipSetDevice(c->device));
  (void)hipGetLastError();
  std::vector<DtJob> p1, p2;
  std::vector<int4> nodes;
  std::vector<int> depths;
  size_t lds1 = 0, lds2 = 0;
  if (dt_single_jobs(d0, d1, ax, bx, ay, by, p1, p2, nodes, depths, lds1, lds2, window) !=
      GEYE_OK)
    return fail(GEYE_ERR_UNSUPPORTED, "dt: map dimension too large");
  const long long SQ = (long long)d0 * d1;
  Scratch S;
  DeviceBuffer &dv = S.get(), &dtm = S.get(), &diy = S.get(), &dm = S.get(), &dix = S.get(),
               &dj1 = S.get(), &dj2 = S.get(), &dn = S.get(), &dd = S.get();
  CHK(upload(dv, std::vector<float>(vals, vals + SQ), c->stream));
  CHK(ensure(dtm, sizeof(float) * SQ));
  CHK(ensure(diy, sizeof(int) * SQ));
  CHK(ensure(dm, sizeof(float) * SQ));
  CHK(ensure(dix, sizeof(int) * SQ));
  CHK(upload(dj1, p1, c->stream));
  CHK(upload(dj2, p2, c->stream));
  CHK(upload(dn, nodes, c->stream));
  CHK(upload(dd, depths, c->stream));
  auto *dt_launch = window ? launch_dt_window : launch_dt;
  if (dt_launch(c->stream, dp<DtJob>(dj1), (int)p1.size(), lds1, dp<int4>(dn), dp<int>(dd),
                dp<float>(dv), dp<float>(dtm), dp<int>(diy)) != 0 ||
      dt_launch(c->stream, dp<DtJob>(dj2), (int)p2.size(), lds2, dp<int4>(dn), dp<int>(dd),
                dp<float>(dtm), dp<float>(dm), dp<int>(dix)) != 0)
    return fail(GEYE_ERR_HIP, "DT kernel: dynamic LDS attribute not set");
  HIPCHK(hipGetLastError());
  std::vector<int> tIx(SQ), tIy(SQ);
  HIPCHK(hipMemcpyAsync(M, dm.p, sizeof(float) * SQ, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(tIx.data(), dix.p, sizeof(int) * SQ, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(tIy.data(), diy.p, sizeof(int) * SQ, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  // Ix = tmpIx, Iy = tmpIy[tmpIx*dims0 + y] (CT/dt.cpp:120-131)
  long long k = 0;
  for (int x = 0; x < d1; x++)
    for (int y = 0; y < d0; y++, k++) {
      Ix[k] = tIx[k];
      Iy[k] = tIy[(long long)tIx[k] * d0 + y];
    }
  return GEYE_OK;
}

int geye_dt(geye_ctx *c, const float *vals, float ax, float bx, float ay, float by,
            const int *dims, float *M, int *Ix, int *Iy) {
  return dt_standalone(c, vals, ax, bx, ay, by, dims, M, Ix, Iy, false);
}

int geye_dt_window(geye_ctx *c, const float *vals, float ax, float bx, float ay, float by,
                   const int *dims, float *M, int *Ix, int *Iy) {
  return dt_standalone(c, vals, ax, bx, ay, by, dims, M, Ix, Iy, true);
}

int geye_dev_alloc(geye_ctx *c, size_t bytes, void **ptr) {
  if (!c || !ptr) return fail(GEYE_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(c->device));
  hipError_t e = hipMalloc(ptr, std::max<size_t>(bytes, 1));
  if (e != hipSuccess) return fail(GEYE_ERR_NOMEM, hipGetErrorString(e));
  return GEYE_OK;
}

int geye_dev_free(geye_ctx *c, void *ptr) {
  if (!c) return fail(GEYE_ERR_ARG, "null ctx");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipFree(ptr));
  return GEYE_OK;
}

int geye_memcpy_h2d(geye_ctx *c, void *dst, const void *src, size_t bytes) {
  if (!c || !dst || !src) return fail(GEYE_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return GEYE_OK;
}

int geye_synchronize(geye_ctx *c) {
  if (!c) return fail(GEYE_ERR_ARG, "null ctx");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  flush_stats(c);
  return GEYE_OK;
}

int geye_mfma_tuning(geye_ctx *c, int precision, int *choice, char *text, int cap) {
  if (!c || !mfma_precision(precision)) return fail(GEYE_ERR_ARG, "bad argument");
  std::lock_guard<std::mutex> lk(c->mu);
  Plan *P = c->last_plan;
  if (!P) return fail(GEYE_ERR_ARG, "no detection has run on this context");
  if (choice) *choice = P->mfma_choice[precision];
  std::string t;
  for (auto &kv : P->mfma_tuning[precision]) {
    char b[96];
    std::snprintf(b, sizeof b, "%s=%.4f;", kv.first.c_str(), kv.second);
    t += b;
  }
  if (text && cap > 0) {
    std::strncpy(text, t.c_str(), (size_t)cap - 1);
    text[cap - 1] = 0;
  }
  return GEYE_OK;
}

int geye_mfma_schedule(geye_ctx *c, int precision, int choice) {
  if (!c || !mfma_precision(precision)) return fail(GEYE_ERR_ARG, "bad argument");
  std::lock_guard<std::mutex> lk(c->mu);
  Plan *P = c->last_plan;
  if (!P) return fail(GEYE_ERR_ARG, "no detection has run on this context");
  if (choice != 0) {
    const int mi = (choice - 1) / 8, bpc = (choice - 1) % 8 + 1;
    const size_t lds = P->q_lds[mi] * (split_precision(precision) ? 2 : 1);
    if (choice < 0 || mi > 1 || P->q_tiles[mi].empty() || lds * bpc > 159 * 1024)
      return fail(GEYE_ERR_ARG, "no such MFMA schedule for this plan");
  }
  P->mfma_choice[precision] = choice;
  drop_graphs(*P);
  return GEYE_OK;
}

int geye_graph_enable(geye_ctx *c, int on) {
  if (!c) return fail(GEYE_ERR_ARG, "null ctx");
  std::lock_guard<std::mutex> lk(c->mu);
  c->graphs = on != 0;
  return GEYE_OK;
}

int geye_profile_enable(geye_ctx *c, int on) {
  if (!c) return fail(GEYE_ERR_ARG, "null ctx");
  std::lock_guard<std::mutex> lk(c->mu);
  flush_stats(c);
  c->profile = on != 0;
  c->stats.clear();
  return GEYE_OK;
}

int geye_dt_form(geye_ctx *c, int form) {
  if (!c) return fail(GEYE_ERR_ARG, "null ctx");
  if (form != GEYE_DT_AUTO && form != GEYE_DT_DIVIDE && form != GEYE_DT_WINDOW)
    return fail(GEYE_ERR_ARG, "unknown DT form");
  std::lock_guard<std::mutex> lk(c->mu);
  c->dt_form = form;
  return GEYE_OK;
}

int geye_plan_bytes_desc(const geye_model_desc *desc, int width, int height, long long *bytes) {
  if (!desc || !bytes || width < 8 || height < 8) return fail(GEYE_ERR_ARG, "bad argument");
  Model m;
  ModelPacks pk;
  CHK(model_pack(desc, m, pk));
  return plan_bytes(m, width, height, bytes);
}

int geye_csv_values(const char *path, float **values, long long *count) {
  if (!path || !values || !count) return fail(GEYE_ERR_ARG, "null argument");
  *values = nullptr;
  *count = 0;
  std::vector<float> v;
  if (!read_csv(path, v)) return fail(GEYE_ERR_IO, std::string("cannot read ") + path);
  *values = to_malloc(v);
  if (!*values) return fail(GEYE_ERR_NOMEM, "malloc");
  *count = (long long)v.size();
  return GEYE_OK;
}

int geye_nms_tie_repairs(geye_ctx *c, long long *count) {
  if (!c || !count) return fail(GEYE_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  *count = c->nms_tie_repairs;
  return GEYE_OK;
}

int geye_nms_fallbacks(geye_ctx *c, long long *count) {
  if (!c || !count) return fail(GEYE_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  *count = c->nms_fallbacks;
  return GEYE_OK;
}

int geye_kernel_stats(geye_ctx *c, char *names, int names_cap, double *ms, long long *count,
                      int cap, int *n) {
  if (!c || !n) return fail(GEYE_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  flush_stats(c);
  std::string all;
  int i = 0;
  for (auto &kv : c->stats) {
    if (i < cap) {
      if (ms) ms[i] = kv.second.ms;
      if (count) count[i] = kv.second.count;
    }
    all += kv.first;
    all += '\n';
    i++;
  }
  *n = i;
  if (names && names_cap > 0) {
    std::strncpy(names, all.c_str(), names_cap - 1);
    names[names_cap - 1] = 0;
  }
  return GEYE_OK;
}

}  // extern "C"

// dpm_compat.cpp -- the reference's C++ entry points (include/geye_dpm_compat.h) on the C ABI.
//
// Host code only: every computation goes through include/geye_dpm.h, which runs it on the
// device.  This file holds what the reference keeps in host structures: the MODEL <->
// geye_model association, the CSV model loader that fills a MODEL, ini_scales' once-per-
// model pyramid depth, and the malloc'd output conventions.
#include "../../include/geye_dpm_compat.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/geye_dpm.h"

namespace {

std::mutex g_mu;
geye_ctx *g_ctx = nullptr;

struct Registered {
  geye_model *model = nullptr;
};
std::map<const MODEL *, Registered> g_models;

[[noreturn]] void die(const char *what) {
  std::fprintf(stderr, "%s failed: %s\n", what, geye_last_error());
  std::exit(1);
}

void check(int rc, const char *what) {
  if (rc != GEYE_OK) die(what);
}

geye_ctx *ctx() {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ctx) {
    const char *d = std::getenv("GEYE_DEVICE");
    check(geye_init(d ? std::atoi(d) : 0, &g_ctx), "init_cuda");
  }
  return g_ctx;
}

template <typename T>
T *alloc(size_t n) {
  T *p = (T *)std::calloc(std::max<size_t>(n, 1), sizeof(T));
  if (!p) {
    std::fprintf(stderr, "out of host memory\n");
    std::exit(1);
  }
  return p;
}

// Flat description of a MODEL (the reference's structures hold per-component and
// per-filter pointers).
struct FlatModel {
  std::vector<int> pidx, didx, psize, root_size, part_size;
  std::vector<float> rootf, partf, lsq;
  std::vector<long long> roff, poff;
  geye_model_desc d{};
};

void flatten(const MODEL *MO, FlatModel &F) {
  const Model_info *MI = MO->MI;
  const int C = MI->numcomponent;
  int MP = 1, DefL = 0;
  for (int i = 0; i < C; i++) MP = std::max(MP, MI->numpart[i]);
  F.pidx.assign((size_t)C * MP, 0);
  F.didx.assign((size_t)C * MP, 0);
  F.psize.assign((size_t)C * MP * 2, 0);
  for (int i = 0; i < C; i++)
    for (int k = 0; k < MI->numpart[i]; k++) {
      F.pidx[i * MP + k] = MI->pidx[i][k];
      F.didx[i * MP + k] = MI->didx[i][k];
      F.psize[(i * MP + k) * 2] = MI->psize[i][k * 2];
      F.psize[(i * MP + k) * 2 + 1] = MI->psize[i][k * 2 + 1];
      DefL = std::max(DefL, MI->didx[i][k] + 1);
    }
  for (int r = 0; r < MO->RF->NoR; r++) {
    const int *s = MO->RF->root_size[r];
    F.root_size.insert(F.root_size.end(), s, s + 3);
    F.roff.push_back((long long)F.rootf.size());
    F.rootf.insert(F.rootf.end(), MO->RF->rootfilter[r], MO->RF->rootfilter[r] + (size_t)s[0] * s[1] * s[2]);
  }
  for (int p = 0; p < MO->PF->NoP; p++) {
    const int *s = MO->PF->part_size[p];
    F.part_size.insert(F.part_size.end(), s, s + 3);
    F.poff.push_back((long long)F.partf.size());
    F.partf.insert(F.partf.end(), MO->PF->partfilter[p], MO->PF->partfilter[p] + (size_t)s[0] * s[1] * s[2]);
  }
  if (MI->x1 && MI->y1 && MI->x2 && MI->y2)
    for (int i = 0; i < C; i++) {
      const int ml = 1 + 2 * (1 + MI->numpart[i]);
      for (FLOAT **a : {MI->x1, MI->y1, MI->x2, MI->y2}) F.lsq.insert(F.lsq.end(), a[i], a[i] + ml);
    }
  geye_model_desc &d = F.d;
  d.numcomponent = C;
  d.sbin = MI->sbin;
  d.interval = MI->interval;
  d.max_Y = MI->max_Y;
  d.max_X = MI->max_X;
  d.padx = MI->padx;
  d.pady = MI->pady;
  d.maxpart = MP;
  d.ridx = MI->ridx;
  d.oidx = MI->oidx;
  d.offw = MI->offw;
  d.rsize = MI->rsize;
  d.numpart = MI->numpart;
  d.pidx = F.pidx.data();
  d.didx = F.didx.data();
  d.psize = F.psize.data();
  d.DefL = DefL;
  d.def = MI->def;
  d.anchor = MI->anchor;
  d.NoR = MO->RF->NoR;
  d.root_size = F.root_size.data();
  d.rootfilter = F.rootf.data();
  d.root_off = F.roff.data();
  d.NoP = MO->PF->NoP;
  d.part_size = F.part_size.data();
  d.partfilter = F.partf.data();
  d.part_off = F.poff.data();
  d.lsq = F.lsq.empty() ? nullptr : F.lsq.data();
}

// The device copy of a MODEL: made on first use, kept until free_model.
geye_model *device_model(const MODEL *MO) {
  geye_ctx *c = ctx();
  std::lock_guard<std::mutex> lk(g_mu);
  Registered &r = g_models[MO];
  if (!r.model) {
    FlatModel F;
    flatten(MO, F);
    check(geye_model_create(c, &F.d, &r.model), "load_model");
  }
  return r.model;
}

int GL_of(const MODEL *MO) { return (MO->MI->numpart[0] + 1) * 4 + 3; }

// ---- CSV loader (formats of CT/load_model.cpp:57-530) -----------------------------------
bool read_values(const char *path, std::vector<float> &v) {
  float *vals = nullptr;
  long long n = 0;
  if (geye_csv_values(path, &vals, &n) != GEYE_OK) return false;
  v.assign(vals, vals + n);
  geye_free(vals);
  return true;
}

struct Reader {
  const std::vector<float> &v;
  size_t k = 0;
  const char *file;
  float next() {
    if (k >= v.size()) {
      std::fprintf(stderr, "load_model: %s is truncated\n", file);
      std::exit(1);
    }
    return v[k++];
  }
  int inext() { return (int)next(); }
};

const char *env_or(const char *name, const char *dflt) {
  const char *e = std::getenv(name);
  return e ? e : dflt;
}

}  // namespace

// =======================================================================================
extern "C" void init_cuda(void) { (void)ctx(); }

extern "C" void clean_cuda(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto &kv : g_models) geye_model_free(kv.second.model);
  g_models.clear();
  if (g_ctx) geye_shutdown(g_ctx);
  g_ctx = nullptr;
}

MODEL *load_model(FLOAT ratio) {
  const char *fc = env_or("GEYE_F_NAME_COM", "./CAR_TRACKING/car_comp.csv");
  const char *fr = env_or("GEYE_F_NAME_ROOT", "./CAR_TRACKING/car_root.csv");
  const char *fp = env_or("GEYE_F_NAME_PART", "./CAR_TRACKING/car_part.csv");
  std::vector<float> cv, rv, pv;
  for (auto [path, vec] : {std::pair<const char *, std::vector<float> *>{fc, &cv}, {fr, &rv}, {fp, &pv}})
    if (!read_values(path, *vec)) {
      std::fprintf(stderr, "load_model: cannot read %s\n", path);
      std::exit(1);
    }
  Model_info *MI = alloc<Model_info>(1);
  Reader c{cv, 0, fc};
  MI->numcomponent = c.inext();
  MI->sbin = c.inext();
  MI->interval = c.inext();
  MI->max_Y = c.inext();
  MI->max_X = c.inext();
  const int C = MI->numcomponent;
  if (C <= 0) {
    std::fprintf(stderr, "load_model: %s has no components\n", fc);
    std::exit(1);
  }
  MI->ridx = alloc<int>(C);
  MI->oidx = alloc<int>(C);
  MI->offw = alloc<FLOAT>(C);
  MI->rsize = alloc<int>(2 * C);
  MI->numpart = alloc<int>(C);
  MI->pidx = alloc<int *>(C);
  MI->didx = alloc<int *>(C);
  MI->psize = alloc<int *>(C);
  for (int i = 0; i < C; i++) {
    MI->ridx[i] = c.inext() - 1;
    MI->oidx[i] = c.inext() - 1;
    MI->offw[i] = c.next();
    MI->rsize[2 * i] = c.inext();
    MI->rsize[2 * i + 1] = c.inext();
    const int np = MI->numpart[i] = c.inext();
    MI->pidx[i] = alloc<int>(np);
    MI->didx[i] = alloc<int>(np);
    MI->psize[i] = alloc<int>(2 * np);
    for (int k = 0; k < np; k++) {
      MI->pidx[i][k] = c.inext() - 1;
      MI->didx[i][k] = c.inext() - 1;
      MI->psize[i][2 * k] = c.inext();
      MI->psize[i][2 * k + 1] = c.inext();
    }
  }
  const int DefL = c.inext();
  MI->def = alloc<FLOAT>(4 * DefL);
  MI->anchor = alloc<int>(2 * DefL);
  for (int q = 0; q < DefL; q++) {
    for (int t = 0; t < 4; t++) MI->def[4 * q + t] = c.next();
    MI->anchor[2 * q] = c.inext();
    MI->anchor[2 * q + 1] = c.inext();
  }
  MI->x1 = alloc<FLOAT *>(C);
  MI->y1 = alloc<FLOAT *>(C);
  MI->x2 = alloc<FLOAT *>(C);
  MI->y2 = alloc<FLOAT *>(C);
  for (int i = 0; i < C; i++) {
    const int ml = 1 + 2 * (1 + MI->numpart[i]);
    for (FLOAT **a : {MI->x1, MI->y1, MI->x2, MI->y2}) {
      a[i] = alloc<FLOAT>(ml);
      for (int j = 0; j < ml; j++) a[i][j] = c.next();
    }
  }
  MI->ini = true;
  MI->ratio = ratio;
  MI->padx = 0;  // load_model's override of the computed pads (CT/load_model.cpp:552-553)
  MI->pady = 0;
  auto filters = [](const std::vector<float> &v, const char *file, bool partner, int &N,
                    int **&size, FLOAT **&w, int *&extra) {
    Reader r{v, 0, file};
    N = r.inext();
    size = alloc<int *>(N);
    w = alloc<FLOAT *>(N);
    extra = alloc<int>(N);
    for (int i = 0; i < N; i++) {
      size[i] = alloc<int>(3);
      for (int t = 0; t < 3; t++) size[i][t] = r.inext();
      const size_t n = (size_t)size[i][0] * size[i][1] * size[i][2];
      w[i] = alloc<FLOAT>(n);
      for (size_t q = 0; q < n; q++) w[i][q] = r.next();
      if (partner) extra[i] = r.inext();
    }
  };
  Rootfilters *RF = alloc<Rootfilters>(1);
  filters(rv, fr, false, RF->NoR, RF->root_size, RF->rootfilter, RF->rootsym);
  Partfilters *PF = alloc<Partfilters>(1);
  filters(pv, fp, true, PF->NoP, PF->part_size, PF->partfilter, PF->part_partner);
  PF->part_sym = alloc<int>(PF->NoP);
  MODEL *MO = alloc<MODEL>(1);
  MO->MI = MI;
  MO->RF = RF;
  MO->PF = PF;
  return MO;
}

void free_model(MODEL *MO) {
  if (!MO) return;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_models.find(MO);
    if (it != g_models.end()) {
      geye_model_free(it->second.model);
      g_models.erase(it);
    }
  }
  Model_info *MI = MO->MI;
  for (int i = 0; i < MI->numcomponent; i++) {
    std::free(MI->pidx[i]);
    std::free(MI->didx[i]);
    std::free(MI->psize[i]);
    for (FLOAT **a : {MI->x1, MI->y1, MI->x2, MI->y2})
      if (a) std::free(a[i]);
  }
  for (void *p : {(void *)MI->ridx, (void *)MI->oidx, (void *)MI->offw, (void *)MI->rsize,
                  (void *)MI->numpart, (void *)MI->pidx, (void *)MI->didx, (void *)MI->psize,
                  (void *)MI->def, (void *)MI->anchor, (void *)MI->x1, (void *)MI->y1,
                  (void *)MI->x2, (void *)MI->y2})
    std::free(p);
  std::free(MI);
  for (int r = 0; r < MO->RF->NoR; r++) {
    std::free(MO->RF->root_size[r]);
    std::free(MO->RF->rootfilter[r]);
  }
  std::free(MO->RF->root_size);
  std::free(MO->RF->rootfilter);
  std::free(MO->RF->rootsym);
  std::free(MO->RF);
  for (int p = 0; p < MO->PF->NoP; p++) {
    std::free(MO->PF->part_size[p]);
    std::free(MO->PF->partfilter[p]);
  }
  std::free(MO->PF->part_size);
  std::free(MO->PF->partfilter);
  std::free(MO->PF->part_partner);
  std::free(MO->PF->part_sym);
  std::free(MO->PF);
  std::free(MO);
}

// ---- pyramid ---------------------------------------------------------------------------
// ini_scales (CT/featurepyramid.cpp.new2:174-239): the pyramid depth is fixed on the first
// call for a model (MI->ini) and reused for every later frame.
// The depth loop reads the image's own size; X and Y are unused there as well.
FLOAT *ini_scales(Model_info *MI, IplImage *IM, int, int) {
  if (MI->ini) {
    int ms = 0, LEN = 0;
    check(geye_level_plan_dims(MI->rsize, MI->numcomponent, MI->sbin, MI->interval, IM->width,
                               IM->height, &ms, &LEN, nullptr, nullptr, nullptr),
          "ini_scales");
    MI->max_scale = ms;
    MI->ini = false;
  }
  MI->IM_HEIGHT = IM->height;
  MI->IM_WIDTH = IM->width;
  return alloc<FLOAT>((size_t)MI->max_scale + MI->interval);
}

int *ini_featsize(Model_info *MI) { return alloc<int>(2 * ((size_t)MI->max_scale + MI->interval)); }

FLOAT **calc_f_pyramid(IplImage *Image, Model_info *MI, int *FTSIZE, FLOAT *scale) {
  if (Image->nChannels != 3) {
    std::fprintf(stderr, "calc_f_pyramid: 3-channel 8-bit images only\n");
    std::exit(1);
  }
  const int LEN = MI->max_scale + MI->interval;
  std::vector<int> fh(LEN), fw(LEN);
  std::vector<float> sc(LEN);
  // sizes first (no feature buffer), then the features into one block
  check(geye_pyramid(ctx(), (const unsigned char *)Image->imageData, Image->width, Image->height,
                     Image->widthStep, MI->sbin, MI->interval, MI->max_scale, nullptr, 0,
                     fh.data(), fw.data(), sc.data()),
        "calc_f_pyramid");
  long long total = 0;
  for (int l = 0; l < LEN; l++) total += 31LL * std::max(fh[l], 0) * std::max(fw[l], 0);
  FLOAT *block = alloc<FLOAT>((size_t)total);
  check(geye_pyramid(ctx(), (const unsigned char *)Image->imageData, Image->width, Image->height,
                     Image->widthStep, MI->sbin, MI->interval, MI->max_scale, block, total,
                     fh.data(), fw.data(), sc.data()),
        "calc_f_pyramid");
  FLOAT **feat = alloc<FLOAT *>(LEN);
  long long off = 0;
  for (int l = 0; l < LEN; l++) {
    feat[l] = block + off;
    off += 31LL * std::max(fh[l], 0) * std::max(fw[l], 0);
    FTSIZE[l * 2] = fh[l];
    FTSIZE[l * 2 + 1] = fw[l];
    scale[l] = sc[l];
  }
  return feat;
}

// Features are one block (free_features frees &features[0][0], CT/featurepyramid.cpp.new2:1656).
void free_features(FLOAT **features, Model_info *) {
  if (!features) return;
  std::free(features[0]);
  std::free(features);
}

// ---- detection -------------------------------------------------------------------------
FLOAT *get_boxes(FLOAT **features, FLOAT *scales, int *FSIZE, MODEL *MO, int *Dnum,
                 FLOAT *A_SCORE, FLOAT thresh) {
  (void)scales;  // recomputed by the level plan (the same values calc_f_pyramid wrote)
  Model_info *MI = MO->MI;
  geye_model *gm = device_model(MO);
  const int LEN = MI->max_scale + MI->interval;
  long long total = 0;
  bool contiguous = true;
  for (int l = 0; l < LEN; l++) {
    if (features[l] != features[0] + total) contiguous = false;
    total += 31LL * std::max(FSIZE[l * 2], 0) * std::max(FSIZE[l * 2 + 1], 0);
  }
  std::vector<float> gathered;
  const float *feat = features[0];
  if (!contiguous) {
    gathered.reserve((size_t)total);
    for (int l = 0; l < LEN; l++)
      gathered.insert(gathered.end(), features[l],
                      features[l] + 31LL * std::max(FSIZE[l * 2], 0) * std::max(FSIZE[l * 2 + 1], 0));
    feat = gathered.data();
  }
  float *boxes = nullptr;
  int n = 0, GL = 0;
  check(geye_get_boxes_features(ctx(), gm, MI->IM_WIDTH, MI->IM_HEIGHT, MI->max_scale, feat, total,
                                thresh, &boxes, &n, &GL),
        "get_boxes");
  if (A_SCORE) check(geye_accumulated_score(ctx(), thresh, A_SCORE), "get_boxes (A_SCORE)");
  *Dnum = n;
  return boxes;
}

FLOAT *nms(FLOAT *boxes, FLOAT overlap, int *num, MODEL *MO) {
  const int NUM = *num;
  if (NUM <= 0) return nullptr;
  const int GL = GL_of(MO);
  FLOAT *out = alloc<FLOAT>((size_t)NUM * GL);
  int k = 0;
  check(geye_nms(ctx(), boxes, NUM, GL, overlap, out, &k), "nms");
  *num = k;
  return out;
}

// create_result's allocation (CT/tracking.cpp:265-284) for get_new_rects' output.
RESULT *get_new_rects(IplImage *Image, MODEL *MO, FLOAT *boxes, int *NUM) {
  const int n = *NUM;
  RESULT *R = alloc<RESULT>(1);
  R->num = n;
  if (n == 0) return R;
  R->point = alloc<int>(4 * (size_t)n);
  R->OR_point = alloc<int>(4 * (size_t)n);
  R->type = alloc<int>(n);
  R->score = alloc<FLOAT>(n);
  R->scale = alloc<FLOAT>(n);
  R->IM = alloc<IplImage *>(n);
  check(geye_new_rects(ctx(), device_model(MO), boxes, n, GL_of(MO), Image->width, Image->height,
                       MO->MI->ratio, R->point, R->OR_point, R->type, R->scale, R->score),
        "get_new_rects");
  return R;
}

// ---- stage functions -------------------------------------------------------------------
FLOAT **fconvsMT(FLOAT *feat, FLOAT *, FLOAT **filter, int *, int start, int end, int *A_SIZE,
                 int **B_SIZE, int *M_size) {
  const int n = end - start + 1;
  if (n <= 0) return nullptr;
  const int F0 = A_SIZE[0], F1 = A_SIZE[1];
  long long total = 0;
  for (int i = 0; i < n; i++) {
    const int H = F0 - B_SIZE[start + i][0] + 1, W = F1 - B_SIZE[start + i][1] + 1;
    M_size[i * 2] = H;
    M_size[i * 2 + 1] = W;
    total += (long long)std::max(H, 0) * std::max(W, 0);
  }
  FLOAT *block = alloc<FLOAT>((size_t)total);
  FLOAT **out = alloc<FLOAT *>(n);
  long long off = 0;
  for (int i = 0; i < n; i++) {
    const int h = B_SIZE[start + i][0], w = B_SIZE[start + i][1];
    out[i] = block + off;
    if (M_size[i * 2] > 0 && M_size[i * 2 + 1] > 0)
      check(geye_fconv(ctx(), feat, F0, F1, filter[start + i], 1, h, w, GEYE_FP32, out[i]), "fconvsMT");
    off += (long long)std::max(M_size[i * 2], 0) * std::max(M_size[i * 2 + 1], 0);
  }
  return out;
}

FLOAT *dt(FLOAT *vals, FLOAT ax, FLOAT bx, FLOAT ay, FLOAT by, int *dims, int *Ix, int *Iy) {
  FLOAT *M = alloc<FLOAT>((size_t)dims[0] * dims[1]);
  check(geye_dt(ctx(), vals, ax, bx, ay, by, dims, M, Ix, Iy), "dt");
  return M;
}

void add_part_calculation(FLOAT *score, FLOAT *M, int *rootsize, int *partsize, int ax, int ay) {
  check(geye_add_part(ctx(), score, M, rootsize, partsize, ax, ay), "add_part_calculation");
}

FLOAT *resize(FLOAT *src, int *sdims, int *odims, FLOAT scale) {
  check(geye_resize(ctx(), src, sdims, scale, nullptr, odims), "resize");
  FLOAT *dst = alloc<FLOAT>((size_t)odims[0] * odims[1] * odims[2]);
  check(geye_resize(ctx(), src, sdims, scale, dst, odims), "resize");
  return dst;
}

FLOAT *calc_feature(FLOAT *SRC, int *ISIZE, int *FTSIZE, int sbin) {
  int F0 = 0, F1 = 0;
  check(geye_calc_feature(ctx(), SRC, ISIZE[0], ISIZE[1], sbin, nullptr, &F0, &F1), "calc_feature");
  FLOAT *feat = alloc<FLOAT>(31 * (size_t)F0 * F1);
  if (F0 > 0 && F1 > 0)
    check(geye_calc_feature(ctx(), SRC, ISIZE[0], ISIZE[1], sbin, feat, &F0, &F1), "calc_feature");
  FTSIZE[0] = F0;
  FTSIZE[1] = F1;
  return feat;
}

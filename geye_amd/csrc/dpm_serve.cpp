// dpm_serve.cpp -- the serving loop: several camera streams, frames admitted by HBM budget,
// several frames on the device at once.
//
// The reference runs one frame at a time from main() (CT/main.cpp:240-300); this server is
// the device-side equivalent for many cameras.  A server owns `slots` workers; each worker
// has its own device context (its own HIP stream and its own per-frame arenas), so frames
// of different workers overlap on the GPU while every model's filters stay resident once.
// Requests queue in arrival order.  A worker takes the oldest request whose working memory
// (the plan of its model and frame size on that worker) is already resident or fits in the
// HBM budget, evicting plans of idle workers (least recently used first) when needed; a
// request that cannot fit waits until memory frees up.  Each frame runs the same kernels
// as geye_detect, so a result does not depend on what else is in flight.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "dpm_internal.h"
#include "dpm_runtime.h"

using namespace geye;

namespace {

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
  std::vector<long long> tickets;          // one per model
  std::vector<const geye_model *> models;  // several share the frame's pyramid
  std::vector<unsigned char> frame;        // host copy, or empty for a device frame
  const void *dev = nullptr;               // device-resident frame (caller-owned)
  int W = 0, H = 0, ws = 0;
  double t_submit = 0;
};

struct Result {
  int rc = GEYE_OK;
  std::string err;
  float *rows = nullptr;
  int num = 0, GL = 0, slot = -1;
  double t_submit = 0, t_start = 0, t_done = 0;
};

struct Resident {             // a plan held by one worker
  int slot;
  const Model *model;
  long long key;
  long long bytes;
  double last_use;
};

}  // namespace

struct geye_server {
  geye_serve_config cfg{};
  int device = 0;
  std::vector<geye_ctx *> ctx;
  std::vector<unsigned char *> pinned;        // per worker: frame staging (pinned host memory)
  std::vector<size_t> pinned_bytes;
  std::vector<std::thread> workers;
  std::vector<bool> busy;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request> queue;
  std::unordered_map<long long, Result> done;
  std::vector<Resident> resident;
  std::map<std::pair<const Model *, long long>, long long> bytes_cache;
  long long held = 0, peak = 0, next_ticket = 1, completed = 0, evictions = 0;
  std::vector<double> latencies;
  double t_first = -1, t_last = 0;
  bool stop = false;
};

namespace {

// Memory for (model, frame) on worker s: 0 when resident there, else its estimate.
long long need_bytes(geye_server *S, int s, const Model *m, int W, int H, long long key,
                     bool *resident_here) {
  *resident_here = false;
  for (auto &r : S->resident)
    if (r.slot == s && r.model == m && r.key == key) {
      *resident_here = true;
      return 0;
    }
  auto ck = std::make_pair(m, key);
  auto it = S->bytes_cache.find(ck);
  if (it != S->bytes_cache.end()) return it->second;
  long long b = 0;
  if (plan_bytes(*m, W, H, &b) != GEYE_OK) b = 0;  // the run reports the error
  S->bytes_cache[ck] = b;
  return b;
}

// Frees plans of idle workers (oldest use first) until `need` more bytes fit.  Called with
// S->mu held; the evicted workers are idle, so none of their kernels is running.
bool make_room(geye_server *S, long long need) {
  if (S->cfg.hbm_budget <= 0) return true;
  while (S->held + need > S->cfg.hbm_budget) {
    int victim = -1;
    for (int i = 0; i < (int)S->resident.size(); i++) {
      const Resident &r = S->resident[i];
      if (S->busy[r.slot]) continue;
      if (victim < 0 || r.last_use < S->resident[victim].last_use) victim = i;
    }
    if (victim < 0) return false;
    const Resident r = S->resident[victim];
    release_plans(S->ctx[r.slot], r.model, r.key);
    S->held -= r.bytes;
    S->resident.erase(S->resident.begin() + victim);
    S->evictions++;
  }
  return true;
}

void worker(geye_server *S, int s) {
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    // the oldest request, once this worker can hold its working memory (arrival order is
    // kept: the head waits for memory rather than being overtaken)
    auto pick = S->queue.end();
    std::vector<long long> need;
    std::vector<bool> here;
    while (!S->stop) {
      pick = S->queue.end();
      if (!S->queue.empty()) {
        auto it = S->queue.begin();
        const long long key = plan_key(it->W, it->H, -1);
        const size_t nm = it->models.size();
        need.assign(nm, 0);
        here.assign(nm, false);
        long long total = 0;
        for (size_t k = 0; k < nm; k++) {
          bool h = false;
          need[k] = need_bytes(S, s, &it->models[k]->m, it->W, it->H, key, &h);
          here[k] = h;
          if (!h) total += need[k];
        }
        if (S->cfg.hbm_budget > 0 && total > S->cfg.hbm_budget) {
          // can never fit: answer it now instead of blocking the queue
          Result res;
          res.rc = GEYE_ERR_NOMEM;
          res.err = "frame working memory exceeds the server's HBM budget";
          res.t_submit = it->t_submit;
          res.t_start = res.t_done = now_ms();
          const std::vector<long long> tickets = it->tickets;
          S->queue.erase(it);
          for (long long t : tickets) S->done[t] = res;
          S->cv_done.notify_all();
          continue;
        }
        if (total == 0 || make_room(S, total)) pick = it;
      }
      if (pick != S->queue.end()) break;
      S->cv_work.wait(lk);
    }
    if (S->stop) return;
    Request req = std::move(*pick);
    S->queue.erase(pick);
    S->busy[s] = true;
    const long long key = plan_key(req.W, req.H, -1);
    for (size_t k = 0; k < req.models.size(); k++)
      if (!here[k]) {
        S->resident.push_back(Resident{s, &req.models[k]->m, key, need[k], now_ms()});
        S->held += need[k];
      }
    S->peak = std::max(S->peak, S->held);
    lk.unlock();

    const int nm = (int)req.models.size();
    std::vector<Result> res(nm);
    const double t_start = now_ms();
    std::vector<float *> rows(nm, nullptr);
    std::vector<int> num(nm, 0), GL(nm, 0);
    int rc = GEYE_OK;
    if (req.dev) {
      rc = geye_detect_models_dev(S->ctx[s], req.models.data(), nm, req.dev, req.W, req.H, req.ws,
                                  S->cfg.thresh, S->cfg.overlap, S->cfg.precision, rows.data(),
                                  num.data(), GL.data());
    } else {
      const size_t fb = req.frame.size();
      if (S->pinned_bytes[s] < fb) {
        if (S->pinned[s]) (void)hipHostFree(S->pinned[s]);
        S->pinned[s] = nullptr;
        S->pinned_bytes[s] = 0;
        if (hipHostMalloc((void **)&S->pinned[s], fb, hipHostMallocDefault) == hipSuccess)
          S->pinned_bytes[s] = fb;
      }
      const unsigned char *src = req.frame.data();
      if (S->pinned[s]) {
        std::memcpy(S->pinned[s], req.frame.data(), fb);
        src = S->pinned[s];
      }
      rc = geye_detect_models(S->ctx[s], req.models.data(), nm, src, req.W, req.H, 3 * req.W,
                              S->cfg.thresh, S->cfg.overlap, S->cfg.precision, rows.data(),
                              num.data(), GL.data());
    }
    const std::string err = rc != GEYE_OK ? std::string(geye_last_error()) : std::string();
    const double t_done = now_ms();
    for (int k = 0; k < nm; k++) {
      res[k].rc = rc;
      res[k].err = err;
      res[k].rows = rows[k];
      res[k].num = num[k];
      res[k].GL = GL[k];
      res[k].slot = s;
      res[k].t_submit = req.t_submit;
      res[k].t_start = t_start;
      res[k].t_done = t_done;
    }

    lk.lock();
    for (auto &r : S->resident)
      if (r.slot == s && r.key == key)
        for (const geye_model *m : req.models)
          if (r.model == &m->m) r.last_use = t_done;
    S->busy[s] = false;
    for (int k = 0; k < nm; k++) {
      S->completed++;
      S->latencies.push_back(t_done - req.t_submit);
      S->done[req.tickets[k]] = std::move(res[k]);
    }
    S->t_last = std::max(S->t_last, t_done);
    S->cv_done.notify_all();
    S->cv_work.notify_all();
  }
}

// Queues a request for nm models on one frame; one ticket per model.
int enqueue(geye_server *S, Request &&r, int nm, long long *tickets) {
  std::lock_guard<std::mutex> lk(S->mu);
  if ((int)S->queue.size() >= S->cfg.queue_cap) return GEYE_ERR_BUSY;
  r.t_submit = now_ms();
  if (S->t_first < 0) S->t_first = r.t_submit;
  for (int k = 0; k < nm; k++) {
    r.tickets.push_back(S->next_ticket++);
    tickets[k] = r.tickets.back();
  }
  S->queue.push_back(std::move(r));
  S->cv_work.notify_all();
  return GEYE_OK;
}

bool models_ok(const geye_model *const *models, int nm) {
  if (!models || nm < 1 || nm > 8) return false;
  for (int k = 0; k < nm; k++)
    if (!models[k]) return false;
  return true;
}

}  // namespace

extern "C" {

int geye_serve_create(int device, const geye_serve_config *cfg, geye_server **out) {
  if (!cfg || !out) return GEYE_ERR_ARG;
  if (cfg->slots < 1 || cfg->slots > 16 || cfg->queue_cap < 1 ||
      !valid_precision(cfg->precision))
    return GEYE_ERR_ARG;
  geye_server *S = new geye_server();
  S->cfg = *cfg;
  S->device = device;
  for (int s = 0; s < cfg->slots; s++) {
    geye_ctx *c = nullptr;
    const int rc = ctx_create(device, true, &c);
    if (rc != GEYE_OK) {
      for (geye_ctx *x : S->ctx) geye_shutdown(x);
      delete S;
      return rc;
    }
    // with other frames in flight, each frame's MFMA convolution keeps part of the CUs:
    // 1920x1080 bf16x3 at 4 slots gave 798-801 frames/s with every CU, 841-855 with 6/8,
    // 813-825 with 5/8, 808-820 with 7/8; at 3 slots 5/8 was best (814-831 against 797-806
    // with 6/8).  One frame alone is faster with every CU.
    if (cfg->slots >= 2) c->conv_eighths = cfg->slots >= 4 ? 6 : 5;
    S->ctx.push_back(c);
  }
  S->pinned.assign(cfg->slots, nullptr);
  S->pinned_bytes.assign(cfg->slots, 0);
  S->busy.assign(cfg->slots, false);
  for (int s = 0; s < cfg->slots; s++) S->workers.emplace_back(worker, S, s);
  *out = S;
  return GEYE_OK;
}

int geye_serve_submit_models(geye_server *S, const geye_model *const *models, int nmodels,
                             const unsigned char *bgr, int width, int height, int widthStep,
                             long long *tickets) {
  if (!S || !models_ok(models, nmodels) || !bgr || !tickets || width < 8 || height < 8 ||
      widthStep < 3 * width)
    return GEYE_ERR_ARG;
  Request r;
  r.models.assign(models, models + nmodels);
  r.W = width;
  r.H = height;
  r.frame.resize((size_t)3 * width * height);
  for (int y = 0; y < height; y++)
    std::memcpy(&r.frame[(size_t)y * 3 * width], bgr + (size_t)y * widthStep, (size_t)3 * width);
  return enqueue(S, std::move(r), nmodels, tickets);
}

int geye_serve_submit_models_dev(geye_server *S, const geye_model *const *models, int nmodels,
                                 const void *d_bgr, int width, int height, int widthStep,
                                 long long *tickets) {
  if (!S || !models_ok(models, nmodels) || !d_bgr || !tickets || width < 8 || height < 8 ||
      widthStep < 3 * width)
    return GEYE_ERR_ARG;
  Request r;
  r.models.assign(models, models + nmodels);
  r.dev = d_bgr;
  r.W = width;
  r.H = height;
  r.ws = widthStep;
  return enqueue(S, std::move(r), nmodels, tickets);
}

int geye_serve_submit(geye_server *S, const geye_model *model, const unsigned char *bgr,
                      int width, int height, int widthStep, long long *ticket) {
  return geye_serve_submit_models(S, &model, 1, bgr, width, height, widthStep, ticket);
}

int geye_serve_submit_dev(geye_server *S, const geye_model *model, const void *d_bgr, int width,
                          int height, int widthStep, long long *ticket) {
  return geye_serve_submit_models_dev(S, &model, 1, d_bgr, width, height, widthStep, ticket);
}

int geye_serve_wait(geye_server *S, long long ticket, float **boxes, int *num, int *GL,
                    double *latency_ms) {
  if (!S || !boxes || !num || !GL) return GEYE_ERR_ARG;
  std::unique_lock<std::mutex> lk(S->mu);
  if (ticket <= 0 || ticket >= S->next_ticket) return GEYE_ERR_ARG;
  S->cv_done.wait(lk, [&] { return S->done.count(ticket) > 0; });
  Result r = std::move(S->done[ticket]);
  S->done.erase(ticket);
  lk.unlock();
  if (latency_ms) *latency_ms = r.t_done - r.t_submit;
  if (r.rc != GEYE_OK) return set_error(r.rc, r.err);
  *boxes = r.rows;
  *num = r.num;
  *GL = r.GL;
  return GEYE_OK;
}

int geye_serve_stats(geye_server *S, long long *completed, double *p50_ms, double *p99_ms,
                     double *fps, long long *peak_bytes, long long *evictions) {
  if (!S) return GEYE_ERR_ARG;
  std::lock_guard<std::mutex> lk(S->mu);
  std::vector<double> l = S->latencies;
  std::sort(l.begin(), l.end());
  auto pct = [&](double q) {
    if (l.empty()) return 0.0;
    const size_t i = (size_t)std::min<double>(l.size() - 1, q * (l.size() - 1) + 0.5);
    return l[i];
  };
  if (completed) *completed = S->completed;
  if (p50_ms) *p50_ms = pct(0.50);
  if (p99_ms) *p99_ms = pct(0.99);
  if (fps) *fps = (S->completed > 0 && S->t_last > S->t_first)
                      ? 1000.0 * S->completed / (S->t_last - S->t_first) : 0.0;
  if (peak_bytes) *peak_bytes = S->peak;
  if (evictions) *evictions = S->evictions;
  return GEYE_OK;
}

int geye_serve_nms_fallbacks(geye_server *S, long long *count) {
  if (!S || !count) return GEYE_ERR_ARG;
  long long total = 0;
  for (geye_ctx *c : S->ctx) {
    long long k = 0;
    if (geye_nms_fallbacks(c, &k) == GEYE_OK) total += k;
  }
  *count = total;
  return GEYE_OK;
}

int geye_serve_nms_tie_repairs(geye_server *S, long long *count) {
  if (!S || !count) return GEYE_ERR_ARG;
  long long total = 0;
  for (geye_ctx *c : S->ctx) {
    long long k = 0;
    if (geye_nms_tie_repairs(c, &k) == GEYE_OK) total += k;
  }
  *count = total;
  return GEYE_OK;
}

int geye_plan_bytes(const geye_model *model, int width, int height, long long *bytes) {
  if (!model || !bytes || width < 8 || height < 8) return GEYE_ERR_ARG;
  return plan_bytes(model->m, width, height, bytes);
}

int geye_serve_reset_stats(geye_server *S) {
  if (!S) return GEYE_ERR_ARG;
  std::lock_guard<std::mutex> lk(S->mu);
  S->latencies.clear();
  S->completed = 0;
  S->t_first = -1;
  S->t_last = 0;
  S->peak = S->held;
  S->evictions = 0;
  return GEYE_OK;
}

void geye_serve_destroy(geye_server *S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
    S->cv_work.notify_all();
  }
  for (auto &t : S->workers) t.join();
  for (auto &kv : S->done) std::free(kv.second.rows);
  for (size_t s = 0; s < S->ctx.size(); s++) {
    geye_shutdown(S->ctx[s]);
    if (S->pinned[s]) (void)hipHostFree(S->pinned[s]);
  }
  delete S;
}

}  // extern "C"

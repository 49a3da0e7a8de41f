/*
 * geye_dpm.h -- C ABI of the MI355X (gfx950) DPM detection path.
 *
 * The entry points replace the reference's detection hot path (CT = CAR_TRACKING/ of the
 * reference repository): every function returns a status code (GEYE_OK = 0) instead of
 * calling exit(), and geye_last_error() returns the message of the calling thread's last
 * failure.  Arrays crossing the boundary are plain host pointers in the reference's
 * layouts (column-major planes, FLOAT = float); device-resident variants take pointers
 * obtained from geye_dev_alloc().  No torch or HIP types appear in the signatures.
 *
 * Threading: one geye_ctx per device; a context serialises the calls made on it (a mutex),
 * so two threads may share one.  Concurrent detection on one device is done by the serving
 * loop (geye_serve_*), which owns its own streams.
 */
#ifndef GEYE_DPM_H
#define GEYE_DPM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  GEYE_OK = 0,
  GEYE_ERR_ARG = -1,       /* invalid argument (sizes, null pointers, model shape) */
  GEYE_ERR_HIP = -2,       /* a HIP runtime call failed */
  GEYE_ERR_NOMEM = -3,     /* device or host allocation failed, or the budget is exceeded */
  GEYE_ERR_IO = -4,        /* model file missing or malformed */
  GEYE_ERR_UNSUPPORTED = -5,
  GEYE_ERR_BUSY = -6       /* serving: no slot free for the request */
};

/* Arithmetic of the filter convolution; every other step is the same in all forms.
 *   GEYE_FP32:   FP32 FMA chains in the reference's order (bit-exact).
 *   GEYE_BF16, GEYE_FP16: one MFMA per filter tap on 16-bit operands, FP32 accumulation.
 *   GEYE_BF16X3, GEYE_FP16X3: split operands x = hi + lo (hi the 16-bit rounding of x, lo
 *     that of x - hi), three MFMAs per tap (hi*hi + hi*lo + lo*hi) with FP32 accumulation:
 *     16 (BF16) or 22 (FP16) significant bits per operand. */
enum { GEYE_FP32 = 0, GEYE_BF16 = 1, GEYE_FP16 = 2, GEYE_BF16X3 = 3, GEYE_FP16X3 = 4 };

typedef struct geye_ctx geye_ctx;
typedef struct geye_model geye_model;

/* Flat model description; field meaning follows Model_info / Rootfilters / Partfilters
 * (CT/MODEL_info.h:75-142), 0-based indices.  Filters are concatenated in the reference
 * layout [31][w][h]. */
typedef struct {
  int numcomponent, sbin, interval, max_Y, max_X, padx, pady;
  int maxpart;
  const int *ridx, *oidx;
  const float *offw;
  const int *rsize;        /* [numcomponent][2] (Y, X) */
  const int *numpart;
  const int *pidx, *didx;  /* [numcomponent][maxpart] */
  const int *psize;        /* [numcomponent][maxpart][2] */
  int DefL;
  const float *def;        /* [DefL][4] = (ax, bx, ay, by) */
  const int *anchor;       /* [DefL][2] = (x, y) */
  int NoR;
  const int *root_size;    /* [NoR][3] */
  const float *rootfilter;
  const long long *root_off;
  int NoP;
  const int *part_size;
  const float *partfilter;
  const long long *part_off;
  /* Least-squares box coefficients (Model_info x1/y1/x2/y2, CT/load_model.cpp:207-229):
   * per component c, four arrays (x1, y1, x2, y2) of 1 + 2*(1 + numpart[c]) floats,
   * components concatenated.  May be null (then geye_new_rects is unavailable). */
  const float *lsq;
} geye_model_desc;

/* ---- context / model (replaces init_cuda CT/GPU_init.cpp:50, clean_cuda :310,
 *      load_model CT/load_model.cpp:541, free_model :567) ---------------------------- */
const char *geye_last_error(void);
int geye_init(int device, geye_ctx **ctx);
void geye_shutdown(geye_ctx *ctx);
int geye_device_count(int *n);
int geye_model_load_csv(geye_ctx *ctx, const char *comp, const char *root, const char *part,
                        geye_model **model);
int geye_model_create(geye_ctx *ctx, const geye_model_desc *desc, geye_model **model);
/* The numbers of one model CSV file in order (the format of CT/load_model.cpp:57-530:
 * numbers separated by commas and white space), malloc'd (free with geye_free);
 * GEYE_ERR_IO when the file cannot be read or holds something else.  Host only. */
int geye_csv_values(const char *path, float **values, long long *count);
void geye_model_free(geye_model *model);
/* Reads back the loaded model's scalar fields and the first filter value (loader checks). */
int geye_model_info(const geye_model *model, int *numcomponent, int *sbin, int *interval,
                    int *max_Y, int *max_X, int *NoR, int *NoP);

/* ---- level plan (replaces ini_scales CT/featurepyramid.cpp.new2:174, ini_featsize :247) */
int geye_level_plan(const geye_model *model, int width, int height, int *max_scale, int *LEN,
                    float *scales, int *feat_h, int *feat_w);

/* Host-only level plan from the root sizes (no device work): same outputs as
 * geye_level_plan (ini_scales CT/featurepyramid.cpp.new2:174-239, calc_f_pyramid sizes
 * CT/featurepyramid.cpp.org:570-630).  Arrays may be null; capacity >= LEN. */
int geye_level_plan_dims(const int *rsize, int numcomponent, int sbin, int interval, int width,
                         int height, int *max_scale, int *LEN, float *scales, int *feat_h,
                         int *feat_w);

/* Host-only NMS with the reference's exact quicksort order (CT/nms.cpp:70-217); the device
 * path falls back to it when equal scores could change the outcome. */
int geye_nms_host(const float *boxes, int num, int GL, float overlap, float *out, int *out_num);

/* ---- whole-frame detection -------------------------------------------------------------
 * geye_get_boxes: replaces detect() + get_boxes() (CT/detect.cpp:151-226,
 *   CT/get_boxes.cpp:583-1797): every candidate above thresh, rows of GL floats
 *   [root box][part boxes][component][score][scale], in the reference's order.
 * geye_detect: replaces car_detection()'s detect + nms (CT/detect.cpp:232-276,
 *   CT/nms.cpp:94-217): the NMS survivors in original candidate order.
 * Output buffers are malloc'd; free with geye_free.  precision: GEYE_FP32 / BF16 / FP16 /
 * BF16X3 / FP16X3 (the MFMA forms change only the filter convolution). */
int geye_get_boxes(geye_ctx *ctx, const geye_model *model, const unsigned char *bgr, int width,
                   int height, int widthStep, float thresh, int precision, float **boxes,
                   int *num, int *GL);
int geye_detect(geye_ctx *ctx, const geye_model *model, const unsigned char *bgr, int width,
                int height, int widthStep, float thresh, float overlap, int precision,
                float **boxes, int *num, int *GL);
/* Device-resident frame (d_bgr from geye_dev_alloc + geye_memcpy_h2d). */
int geye_detect_dev(geye_ctx *ctx, const geye_model *model, const void *d_bgr, int width,
                    int height, int widthStep, float thresh, float overlap, int precision,
                    float **boxes, int *num, int *GL);
/* Several models on one frame (car_detection once per model, CT/detect.cpp:232-276): the
 * models whose sbin and interval match share one HOG pyramid, built once to the deepest
 * model's level count, and each model's detections are geye_detect's for it alone, bit for
 * bit.  boxes, num and GL are arrays of nmodels entries (free each boxes[i] with geye_free);
 * 1 <= nmodels <= 8.  The _dev form takes a device-resident frame. */
int geye_detect_models(geye_ctx *ctx, const geye_model *const *models, int nmodels,
                       const unsigned char *bgr, int width, int height, int widthStep,
                       float thresh, float overlap, int precision, float **boxes, int *num,
                       int *GL);
int geye_detect_models_dev(geye_ctx *ctx, const geye_model *const *models, int nmodels,
                           const void *d_bgr, int width, int height, int widthStep, float thresh,
                           float overlap, int precision, float **boxes, int *num, int *GL);
/* A batch of nframes frames of one size through nmodels models (geye_detect_models for each
 * frame, batched): every stage before the NMS runs once for all frames -- each launch covers
 * every frame's tiles, lines and maps -- the models share the deepest one's pyramid as in
 * geye_detect_models, and the frames' NMS passes run concurrently.  Every result is that
 * frame's geye_detect for that model, bit for bit.  boxes, num and GL hold nframes * nmodels
 * entries, frame major (entry f * nmodels + m); free each boxes[k] with geye_free.
 * 1 <= nframes <= 64, 1 <= nmodels <= 8.  The _dev form takes device-resident frames. */
int geye_detect_batch(geye_ctx *ctx, const geye_model *const *models, int nmodels,
                      const unsigned char *const *bgr, int nframes, int width, int height,
                      int widthStep, float thresh, float overlap, int precision, float **boxes,
                      int *num, int *GL);
int geye_detect_batch_dev(geye_ctx *ctx, const geye_model *const *models, int nmodels,
                          const void *const *d_bgr, int nframes, int width, int height,
                          int widthStep, float thresh, float overlap, int precision,
                          float **boxes, int *num, int *GL);
/* Root+bias-parts score maps of the last geye_get_boxes/geye_detect* call on ctx,
 * concatenated per (valid level, component) in the reference's order, column-major.
 * Writes at most cap floats; returns the total count through *total. */
int geye_last_scores(geye_ctx *ctx, float *out, long long cap, long long *total);
/* Accumulated-score map of the last geye_get_boxes/geye_detect* call on ctx: the A_SCORE
 * side output of get_boxes (calc_a_score CT/get_boxes.cpp:265-300, normalisation
 * :1765-1787).  a_score is the caller's height x width map, column-major, as ini_ac_score
 * (CT/detect.cpp:136-145) allocates it; it is read and updated in place.  thresh is the
 * detection threshold.  The _dev form takes device memory (geye_dev_alloc). */
int geye_accumulated_score(geye_ctx *ctx, float thresh, float *a_score);
int geye_accumulated_score_dev(geye_ctx *ctx, float thresh, void *d_a_score);
void geye_free(void *p);

/* get_new_rects (CT/tracking.cpp:90-166): the least-squares refinement of NMS survivors
 * (rows as geye_detect returns them) into integer boxes clamped to the width x height
 * image, plus the same boxes divided by ratio (OR_point).  Outputs hold num*4 (point,
 * or_point) and num (type, scale, score) entries.  Computed on the device. */
int geye_new_rects(geye_ctx *ctx, const geye_model *model, const float *boxes, int num, int GL,
                   int width, int height, float ratio, int *point, int *or_point, int *type,
                   float *scale, float *score);
/* get_boxes() (CT/get_boxes.cpp:583-1797) on caller-supplied features: all levels of a
 * width x height frame's pyramid concatenated as calc_f_pyramid returns them
 * ([level][31][F1][F0]).  max_scale > 0 fixes the pyramid depth (Model_info.max_scale as
 * ini_scales left it); <= 0 derives it from the frame size.  FP32; uploads the features. */
int geye_get_boxes_features(geye_ctx *ctx, const geye_model *model, int width, int height,
                            int max_scale, const float *feat, long long nfeat, float thresh,
                            float **boxes, int *num, int *GL);

/* ---- stage entry points (host arrays, reference layouts) --------------------------------
 * geye_resize        replaces resize()          CT/resize.cc:135
 * geye_calc_feature  replaces calc_feature()    CT/featurepyramid.cpp.new2:263
 * geye_pyramid       replaces calc_f_pyramid()  CT/featurepyramid.cpp.new2:952 (features of
 *                    all levels concatenated [level][31][F1][F0])
 * geye_fconv         replaces fconvsMT()        CT/get_boxes_func.h:28, CT/fconvsMT.cpp:310
 *                    (one feature map against n filters of one size)
 * geye_dt            replaces dt()              CT/dt.cpp:101
 * geye_add_part      replaces add_part_calculation()  CT/dt.cpp:38
 * geye_nms           replaces nms()             CT/nms.cpp:94 (runs on the device) */
int geye_resize(geye_ctx *ctx, const float *src, const int *sdims, float scale, float *dst,
                int *odims);
int geye_calc_feature(geye_ctx *ctx, const float *img, int height, int width, int sbin,
                      float *feat, int *F0, int *F1);
int geye_pyramid(geye_ctx *ctx, const unsigned char *bgr, int width, int height, int widthStep,
                 int sbin, int interval, int max_scale, float *feat, long long cap,
                 int *feat_h, int *feat_w, float *scales);
int geye_fconv(geye_ctx *ctx, const float *feat, int F0, int F1, const float *filters, int n,
               int h, int w, int precision, float *out);
int geye_dt(geye_ctx *ctx, const float *vals, float ax, float bx, float ay, float by,
            const int *dims, float *M, int *Ix, int *Iy);
/* The DT of the MFMA forms (GEYE_BF16 ... GEYE_FP16X3), same arguments and the same outputs,
 * bit for bit, as geye_dt (CT/dt.cpp:101).  Each output takes the first least cost in a window
 * bounded by the line block's value range, which holds dt()'s argmin when ax, ay > 0 and the
 * costs' rounding cannot reorder the argmins along a line; where either fails, or the window
 * would pass 16, the block runs geye_dt's divide and conquer. */
int geye_dt_window(geye_ctx *ctx, const float *vals, float ax, float bx, float ay, float by,
                   const int *dims, float *M, int *Ix, int *Iy);
/* add_part_calculation (CT/dt.cpp:38-55): score[y + x*rootsize[0]] -=
 * M[(ay-1+2y) + (ax-1+2x)*partsize[0]] over the root map, in place; rejects anchors that
 * would read outside M (the reference does not check). */
int geye_add_part(geye_ctx *ctx, float *score, const float *M, const int *rootsize,
                  const int *partsize, int ax, int ay);
int geye_nms(geye_ctx *ctx, const float *boxes, int num, int GL, float overlap, float *out,
             int *out_num);
/* geye_dt on device memory, allocating nothing.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  d_vals, d_M, d_Ix and d_Iy are device arrays of
 * dims[0] x dims[1] cells (column-major, as geye_dt's host arrays); dims is a host array.
 * The outputs are geye_dt's, bit for bit, and are complete when the call returns.  The
 * working memory is a static device array of about 8.4 MB per GPU, held until the process
 * ends, which limits each dimension to 2048 and the map to 2^20 cells (larger maps return
 * GEYE_ERR_ARG).  Calls that overlap on one GPU share that array and race; calls on one
 * context are serialised. */
int geye_dt_noalloc(geye_ctx *ctx, const float *d_vals, float ax, float bx, float ay, float by,
                    const int *dims, float *d_M, int *d_Ix, int *d_Iy);

/* ---- serving: several camera streams on one device ---------------------------------------
 * The reference detects one frame at a time (CT/main.cpp:240-300).  A server runs `slots`
 * workers, each with its own stream and per-frame working memory, so several frames are on
 * the device at once; models (geye_model, created on any context of the same device) stay
 * resident and shared.  Requests are taken in arrival order as soon as a worker is free and
 * the frame's working memory is resident on it or fits in hbm_budget (bytes; 0 = no limit),
 * evicting idle workers' least recently used plans.  geye_serve_submit copies the frame and
 * returns at once (GEYE_ERR_BUSY when queue_cap requests are waiting); geye_serve_wait blocks
 * for one ticket and returns its NMS survivors like geye_detect (free with geye_free) and
 * its latency from submission.  Results do not depend on the other frames in flight.
 * Each worker's stream has a hardware queue of its own, so workers never wait on one
 * another's queue; with 2 or more slots a worker's MFMA convolution keeps 5/8 to 3/4 of the
 * CUs, so that the other frames' kernels find room beside it. */
typedef struct geye_server geye_server;
typedef struct {
  int slots;               /* 1..16 frames on the device at once */
  long long hbm_budget;    /* bytes of per-frame working memory held, 0 = unlimited */
  int queue_cap;           /* requests waiting before submit returns GEYE_ERR_BUSY */
  int precision;           /* GEYE_FP32 / GEYE_BF16 / GEYE_FP16 / GEYE_BF16X3 / GEYE_FP16X3 */
  float thresh, overlap;
} geye_serve_config;

int geye_serve_create(int device, const geye_serve_config *cfg, geye_server **srv);
int geye_serve_submit(geye_server *srv, const geye_model *model, const unsigned char *bgr,
                      int width, int height, int widthStep, long long *ticket);
/* Device-resident frame (geye_dev_alloc on the same device); it must stay valid and unchanged
 * until its ticket has been waited for. */
int geye_serve_submit_dev(geye_server *srv, const geye_model *model, const void *d_bgr,
                          int width, int height, int widthStep, long long *ticket);
/* One frame for several models (geye_detect_models on the worker that takes it: the frame's
 * pyramid is built once); tickets receives one ticket per model, in order, each waited for
 * with geye_serve_wait.  1 <= nmodels <= 8. */
int geye_serve_submit_models(geye_server *srv, const geye_model *const *models, int nmodels,
                             const unsigned char *bgr, int width, int height, int widthStep,
                             long long *tickets);
int geye_serve_submit_models_dev(geye_server *srv, const geye_model *const *models, int nmodels,
                                 const void *d_bgr, int width, int height, int widthStep,
                                 long long *tickets);
int geye_serve_wait(geye_server *srv, long long ticket, float **boxes, int *num, int *GL,
                    double *latency_ms);
/* Completed requests, latency percentiles (ms, submission to result), throughput since the
 * first submission, peak bytes of working memory held, and plan evictions. */
int geye_serve_stats(geye_server *srv, long long *completed, double *p50_ms, double *p99_ms,
                     double *fps, long long *peak_bytes, long long *evictions);
int geye_serve_reset_stats(geye_server *srv);
/* NMS calls of the server's workers replayed on the host (geye_nms_fallbacks) and with tie
 * runs reordered (geye_nms_tie_repairs), since the server was created. */
int geye_serve_nms_fallbacks(geye_server *srv, long long *count);
int geye_serve_nms_tie_repairs(geye_server *srv, long long *count);
/* Device bytes of one frame's working memory for this model and frame size, as the server
 * accounts it against hbm_budget (host-only; no device work). */
int geye_plan_bytes(const geye_model *model, int width, int height, long long *bytes);
/* The same from a model description, with no device at all (capacity planning on a host). */
int geye_plan_bytes_desc(const geye_model_desc *desc, int width, int height, long long *bytes);
void geye_serve_destroy(geye_server *srv);

/* ---- device memory and timing (bench / serving plumbing) ------------------------------- */
int geye_dev_alloc(geye_ctx *ctx, size_t bytes, void **ptr);
int geye_dev_free(geye_ctx *ctx, void *ptr);
int geye_memcpy_h2d(geye_ctx *ctx, void *dst, const void *src, size_t bytes);
int geye_synchronize(geye_ctx *ctx);
/* Per-kernel timing with HIP events on the stream the kernels run on.  While enabled,
 * every launch is bracketed by events; geye_kernel_stats returns, per kernel class, the
 * total milliseconds and launch count since the last reset. */
int geye_profile_enable(geye_ctx *ctx, int on);
/* The per-frame work up to the candidate count (ingest .. compaction) is captured once per
 * plan as a HIP graph and replayed (default on); off runs every launch individually.
 * Profiling always runs the launches individually. */
int geye_graph_enable(geye_ctx *ctx, int on);
/* MFMA convolution schedule of the last plan run on ctx.  On its first BF16/FP16 frame a
 * plan times every admissible schedule on its own features and keeps the fastest: 0 = one
 * launch per filter class (4 x 64 tiles), 1 + 8*m + (b - 1) = one persistent work-queue
 * launch over every level and filter class with 4 x (32, 64)[m] tiles and b workgroups per
 * CU.  geye_mfma_tuning reports the choice and "name=ms;" per schedule tried;
 * geye_mfma_schedule overrides it (every schedule gives the same bits). */
int geye_mfma_tuning(geye_ctx *ctx, int precision, int *choice, char *text, int cap);
/* The DT form whole-frame detection runs on ctx: GEYE_DT_AUTO (default) takes the
 * divide and conquer of dt() (CT/dt.cpp:61-136) for GEYE_FP32 and the windowed minimum of
 * geye_dt_window for the MFMA forms; GEYE_DT_DIVIDE / GEYE_DT_WINDOW take one of them for
 * every precision.  Both give the same bits. */
enum { GEYE_DT_AUTO = 0, GEYE_DT_DIVIDE = 1, GEYE_DT_WINDOW = 2 };
int geye_dt_form(geye_ctx *ctx, int form);
int geye_mfma_schedule(geye_ctx *ctx, int precision, int choice);
int geye_kernel_stats(geye_ctx *ctx, char *names, int names_cap, double *ms, long long *count,
                      int cap, int *n);
/* NMS runs on the device.  The reference's quicksort (CT/nms.cpp:70-84) orders equal scores
 * in its own way; where that order decides an outcome, the host runs the quicksort on the
 * scores restricted to the subarrays holding those scores, sets the device order of their
 * runs to the result and the device NMS runs again (geye_nms_tie_repairs counts such calls).
 * When that cannot settle it (a zero score, more than 32 tied scores), the call is replayed
 * on the host with the reference's exact order: geye_nms_fallbacks counts those.  Both
 * count NMS calls on ctx since the context was created. */
int geye_nms_fallbacks(geye_ctx *ctx, long long *count);
int geye_nms_tie_repairs(geye_ctx *ctx, long long *count);

#ifdef __cplusplus
}
#endif
#endif

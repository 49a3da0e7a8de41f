/*
 * dpm_oracle.h -- C API of the CPU oracle for the DPM detection path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, on the CPU, the arithmetic of the
 * reference's detection path (CT = CAR_TRACKING/ of the reference repository) so that
 * tests can check the HIP product against it.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product library never links it.
 *
 * Every stage is restated from the reference source it cites:
 *   ora_ipl_to_float     CT/featurepyramid.cpp.new2:816-850
 *   ora_level_plan       CT/featurepyramid.cpp.new2:174-239, CT/featurepyramid.cpp.org:547-661
 *   ora_resize           CT/resize.cc:52-160
 *   ora_calc_feature     CT/featurepyramid.cpp.new2:263-501
 *   ora_conv             CT/GPU_function.cu:302-358 (per-output order, FMA as nvcc emits it)
 *   ora_dt               CT/dt.cpp:61-136
 *   ora_add_part         CT/dt.cpp:38-55
 *   ora_detect           CT/get_boxes.cpp:583-1797, CT/detect.cpp:151-226
 *   ora_nms              CT/nms.cpp:70-217
 *   ora_detect_ascore    CT/get_boxes.cpp:265-300 (calc_a_score), :1765-1787 (normalisation)
 *   ora_new_rects        CT/tracking.cpp:90-166
 *   ora_load_csv         CT/load_model.cpp:57-558
 */
#ifndef GEYE_DPM_ORACLE_H
#define GEYE_DPM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Flat model description.  Field meaning follows Model_info / Rootfilters / Partfilters
 * (CT/MODEL_info.h:75-142).  Per-part arrays are [numcomponent][maxpart]; filters are
 * concatenated in the reference layout [31][w][h] (column-major, h fastest). */
typedef struct {
  int numcomponent, sbin, interval, max_Y, max_X, padx, pady;
  int maxpart;               /* row stride of pidx/didx and half the stride of psize */
  const int *ridx, *oidx;    /* [numcomponent], 0-based */
  const float *offw;         /* [numcomponent] */
  const int *rsize;          /* [numcomponent][2] (Y, X) */
  const int *numpart;        /* [numcomponent] */
  const int *pidx, *didx;    /* [numcomponent][maxpart], 0-based */
  const int *psize;          /* [numcomponent][maxpart][2] */
  int DefL;
  const float *def;          /* [DefL][4] = (ax, bx, ay, by) */
  const int *anchor;         /* [DefL][2] = (x, y), 0-based */
  int NoR;
  const int *root_size;      /* [NoR][3] (h, w, c) */
  const float *rootfilter;   /* concatenated */
  const long long *root_off; /* [NoR] element offsets into rootfilter */
  int NoP;
  const int *part_size;      /* [NoP][3] */
  const float *partfilter;
  const long long *part_off; /* [NoP] */
  const float *lsq;          /* per component: x1, y1, x2, y2 of 1+2(1+numpart) floats */
} ora_model_desc;

/* BGR interleaved uint8 -> planar float, column-major [c][x][y], plane 0 = B. */
void ora_ipl_to_float(const unsigned char *bgr, int width, int height, int widthStep, float *out);

/* Level plan.  Writes max_scale and LEN; if the arrays are non-null (capacity >= LEN)
 * fills scales[LEN], img_h/img_w[LEN] (resized image size feeding each level),
 * sbin_l[LEN], feat_h/feat_w[LEN]. */
int ora_level_plan(const int *rsize, int numcomponent, int sbin, int interval, int width,
                   int height, int *max_scale, int *LEN, float *scales, int *img_h,
                   int *img_w, int *sbin_l, int *feat_h, int *feat_w);

/* resize(): src column-major [c][x][y] with sdims=(h,w,c).  Returns output element count,
 * writes odims.  dst may be null to query the size. */
long long ora_resize(const float *src, const int *sdims, float scale, float *dst, int *odims);

/* One resize1dtran pass (CT/resize.cc:68-124): interpolation table only, for inspection.
 * Returns the number of (si, di, alpha) entries written (capacity cap). */
int ora_resize_table(int sheight, int dheight, int width, int *si, int *di, float *alpha, int cap);

/* calc_feature(): image column-major [3][w][h] (h,w); out [31][F1][F0].  Returns 0 and writes
 * F0/F1.  feat may be null to query the size. */
int ora_calc_feature(const float *img, int height, int width, int sbin, float *feat, int *F0,
                     int *F1);

/* Valid cross-correlation of one feature level with one filter (31 channels read). */
void ora_conv_f32(const float *feat, int F0, int F1, const float *filt, int h, int w,
                  float *out);
void ora_conv_f64(const float *feat, int F0, int F1, const float *filt, int h, int w,
                  double *out);

/* Generalized distance transform with argmax, dims = (dims0, dims1) column-major. */
void ora_dt_f32(const float *vals, float ax, float bx, float ay, float by, const int *dims,
                float *M, int *Ix, int *Iy);
void ora_dt_f64(const double *vals, double ax, double bx, double ay, double by,
                const int *dims, double *M, int *Ix, int *Iy);
/* Two separate passes, exposing the intermediate arrays. */
void ora_dt_passes_f32(const float *vals, float ax, float bx, float ay, float by,
                       const int *dims, float *tmpM, int *tmpIy, float *M, int *tmpIx);

void ora_add_part_f32(float *score, const float *M, const int *rootsize, const int *partsize,
                      int ax, int ay);

/* NMS over rows of GL floats.  out capacity n*GL.  Returns the number of rows kept. */
int ora_nms(const float *boxes, int n, int GL, float overlap, float *out);

/* quickSort on (scores, order) exactly as CT/nms.cpp:70-84. */
void ora_quicksort(float *scores, int *order, int n);

/* Whole-frame detection.  precision: 0 = FP32 (bit-exact contract), 1 = FP64 scoring
 * (features FP32, convolution/DT/part-add in double).  threads: 1 = serial, >1 = the
 * reference's threading (thread per level for HOG, thread per filter for convolution).
 * Boxes are written to a malloc'd buffer owned by the oracle (ora_free).  Rows follow
 * CT/get_boxes.cpp:1519-1546. Returns the candidate count or a negative error. */
long long ora_detect(const ora_model_desc *m, const unsigned char *bgr, int width, int height,
                     int widthStep, float thresh, int precision, int threads, float **boxes,
                     int *GL);

/* Candidate rows of one component's score map at one level (CT/get_boxes.cpp:185-258,
 * 1497-1546): get_gmpc's column-major scan over the R0 x R1 map, rootbox and partbox per
 * candidate, with the part argmin maps Ix/Iy (numpart[jj] maps of P0 x P1 each, column-major,
 * concatenated) as the DT left them.  scale = sbin / scales[level].  Exposed for hand-built
 * checks of the box geometry.  Writes at most cap rows of GL floats; returns the row count. */
long long ora_rows_one_map(const ora_model_desc *m, int jj, float scale, const float *score,
                           int R0, int R1, const int *Ix, const int *Iy, const int *pdims,
                           float thresh, float *out, long long cap, int *GL);

/* Raw score maps of the last ora_detect_maps call: root+bias-parts per (level, component)
 * concatenated in level-major order, in double for precision 1 and float otherwise. */
long long ora_detect_maps(const ora_model_desc *m, const unsigned char *bgr, int width,
                          int height, int widthStep, int precision, int threads, void **maps,
                          int **map_dims, int *nmaps);

/* The HOG pyramid only: concatenated features [level][31][F1][F0]. */
long long ora_pyramid(const unsigned char *bgr, int width, int height, int widthStep,
                      int sbin, int interval, int max_scale, int threads, float **feat,
                      int *feat_h, int *feat_w, float *scales);

/* Accumulated-score map (A_SCORE side output of get_boxes) of one detection: ac is the
 * caller's height x width map (column-major), read and updated in place; FP32 scores. */
int ora_detect_ascore(const ora_model_desc *m, const unsigned char *bgr, int width, int height,
                      int widthStep, float thresh, int threads, float *ac);
/* One calc_a_score call (CT/get_boxes.cpp:265-300) and the final normalisation
 * (CT/get_boxes.cpp:1765-1787), exposed for hand-built checks. */
void ora_calc_a_score(float *ac, int IHEI, int IWID, const float *score, int R0, int R1,
                      const int *rsize, int padx, int pady, float scale);
void ora_ascore_finish(float *ac, long long n, float thresh);
/* get_new_rects (CT/tracking.cpp:90-166) over num NMS survivors. */
void ora_new_rects(const ora_model_desc *m, const float *boxes, int num, int GL, int width,
                   int height, float ratio, int *point, int *or_point, int *type, float *scale,
                   float *score);
void ora_free(void *p);

/* CSV model loader (CT/load_model.cpp:57-558).  Returns an opaque handle; the desc view
 * stays valid until ora_model_free. */
void *ora_load_csv(const char *comp, const char *root, const char *part,
                   ora_model_desc *desc);
void ora_model_free(void *h);

#ifdef __cplusplus
}
#endif
#endif

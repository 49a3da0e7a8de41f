/*
 * geye_dpm_compat.h -- the reference's own C++ entry points of the DPM detection path,
 * implemented on the MI355X C ABI (geye_dpm.h) by libgeye_dpm_compat.so.
 *
 * A caller written against CT/detect_func.h, CT/get_boxes_func.h, CT/for_use_GPU.h and
 * CT/MODEL_info.h (CT = CAR_TRACKING/ of the reference repository), such as
 * CT/detect.cpp:151-276, compiles against this header unchanged and links
 * libgeye_dpm_compat.so + libgeye_amd.so in place of featurepyramid.cpp, resize.cc,
 * get_boxes.cpp, fconvsMT.cpp, dt.cpp, nms.cpp, load_model.cpp, GPU_init.cpp and
 * tracking.cpp's get_new_rects.
 *
 * Types.  FLOAT, Model_info, Rootfilters, Partfilters, MODEL and RESULT keep the field order
 * and types of CT/switch_float.h:2-8 and CT/MODEL_info.h:75-166, so structures built by the
 * reference's own loader can be passed in.  IplImage is OpenCV's C image header (only width,
 * height, nChannels, widthStep and imageData are read; 8-bit BGR, 3 channels).  A build that
 * has OpenCV and the reference's headers defines GEYE_HOST_TYPES and includes them first.
 *
 * Behaviour.  Every function runs on one process-wide device context (device 0, or
 * $GEYE_DEVICE), created by init_cuda() or by the first call that needs it and destroyed by
 * clean_cuda(), like the reference's global CUDA context (CT/GPU_init.cpp:23-29).  Results
 * are bit-identical to the reference's CPU path in FP32.  Outputs are malloc'd and freed by
 * the caller with free(), as in the reference.  On failure the library's message is printed
 * to stderr and the process exits with status 1, as the reference does; the C ABI returns
 * status codes instead.
 */
#ifndef GEYE_DPM_COMPAT_H
#define GEYE_DPM_COMPAT_H

#ifndef GEYE_HOST_TYPES
typedef float FLOAT;

/* OpenCV's IplImage header (cxcore types_c.h), field for field. */
typedef struct _IplImage {
  int nSize;
  int ID;
  int nChannels;
  int alphaChannel;
  int depth;
  char colorModel[4];
  char channelSeq[4];
  int dataOrder;
  int origin;
  int align;
  int width;
  int height;
  struct _IplROI *roi;
  struct _IplImage *maskROI;
  void *imageId;
  struct _IplTileInfo *tileInfo;
  int imageSize;
  char *imageData;
  int widthStep;
  int BorderMode[4];
  int BorderConst[4];
  char *imageDataOrigin;
} IplImage;

/* CT/MODEL_info.h:75-118 */
typedef struct {
  int numcomponent;
  int sbin;
  int interval;
  int max_X;
  int max_Y;
  int padx;
  int pady;
  int max_scale;
  int IM_WIDTH;
  int IM_HEIGHT;
  int *ridx;
  int *oidx;
  FLOAT *offw;
  int *rsize;
  int *numpart;
  int **pidx;
  int **didx;
  int **psize;
  FLOAT *def;
  int *anchor;
  FLOAT **x1;
  FLOAT **y1;
  FLOAT **x2;
  FLOAT **y2;
  bool ini;
  FLOAT ratio;
} Model_info;

/* CT/MODEL_info.h:120-135 */
typedef struct {
  int NoR;
  int **root_size;
  FLOAT **rootfilter;
  int *rootsym;
} Rootfilters;

typedef struct {
  int NoP;
  int **part_size;
  FLOAT **partfilter;
  int *part_partner;
  int *part_sym;
} Partfilters;

/* CT/MODEL_info.h:138-142 */
typedef struct {
  Model_info *MI;
  Rootfilters *RF;
  Partfilters *PF;
} MODEL;

/* CT/MODEL_info.h:158-166 */
typedef struct {
  int num;
  int *point;
  int *OR_point;
  IplImage **IM;
  int *type;
  FLOAT *scale;
  FLOAT *score;
} RESULT;
#endif /* GEYE_HOST_TYPES */

/* ---- device context (CT/for_use_GPU.h:128-130, CT/GPU_init.cpp:50-356) ------------------- */
extern "C" void init_cuda(void);
extern "C" void clean_cuda(void);

/* ---- model (CT/load_model.cpp:541-672).  Reads $GEYE_F_NAME_COM / _ROOT / _PART, or the
 *      reference's default paths ./CAR_TRACKING/car_{comp,root,part}.csv. ----------------- */
MODEL *load_model(FLOAT ratio);
void free_model(MODEL *MO);

/* ---- CT/detect_func.h:23-45 ------------------------------------------------------------- */
FLOAT *ini_scales(Model_info *MI, IplImage *IM, int X, int Y);
int *ini_featsize(Model_info *MI);
FLOAT **calc_f_pyramid(IplImage *Image, Model_info *MI, int *FTSIZE, FLOAT *scale);
FLOAT *get_boxes(FLOAT **features, FLOAT *scales, int *FSIZE, MODEL *MO, int *Dnum,
                 FLOAT *A_SCORE, FLOAT thresh);
void free_features(FLOAT **features, Model_info *MI);
FLOAT *nms(FLOAT *boxes, FLOAT overlap, int *num, MODEL *MO);
RESULT *get_new_rects(IplImage *Image, MODEL *MO, FLOAT *boxes, int *NUM);

/* ---- CT/get_boxes_func.h:28-35 ----------------------------------------------------------
 * fconvsMT: the reference declares it but has no CPU definition.  Here: features feat
 * (A_SIZE = {F0, F1, 31}, [31][F1][F0]) against filter[start..end] (B_SIZE[i] = {h, w}),
 * returning end-start+1 malloc'd maps [F1-w+1][F0-h+1] in one block (free(out[0]) then
 * free(out)); M_size receives {H, W} per map.  flfeat and sym_info are not read (the
 * symmetric-filter variant is unused by the model format). */
FLOAT **fconvsMT(FLOAT *feat, FLOAT *flfeat, FLOAT **filter, int *sym_info, int start, int end,
                 int *A_SIZE, int **B_SIZE, int *M_size);
FLOAT *dt(FLOAT *vals, FLOAT ax, FLOAT bx, FLOAT ay, FLOAT by, int *dims, int *Ix, int *Iy);
void add_part_calculation(FLOAT *score, FLOAT *M, int *rootsize, int *partsize, int ax, int ay);

/* ---- CT/featurepyramid.cpp.new2:114, :135 ----------------------------------------------- */
FLOAT *resize(FLOAT *src, int *sdims, int *odims, FLOAT scale);
FLOAT *calc_feature(FLOAT *SRC, int *ISIZE, int *FTSIZE, int sbin);

#endif /* GEYE_DPM_COMPAT_H */

// compat_detect.cpp -- test driver: the reference's detect() / car_detection() call sequence
// (CT/detect.cpp:136-276, CT/main.cpp:253-257) written against include/geye_dpm_compat.h
// with the reference's own names and types, run on one raw BGR frame.
//
// usage: compat_detect frame.bgr width height thresh overlap out.bin
// The model comes from load_model() ($GEYE_F_NAME_COM / _ROOT / _PART).  out.bin holds, as
// int32 / float32 arrays: ncand, GL, boxes[ncand*GL], nkept, rects[nkept*GL],
// A_SCORE[width*height], RESULT point[nkept*4], OR_point[nkept*4], type[nkept],
// scale[nkept], score[nkept], then the level count and features' total size of the pyramid.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "geye_dpm_compat.h"

// ini_ac_score (CT/detect.cpp:136-145) belongs to the caller.
static FLOAT *ini_ac_score(IplImage *IM) {
  const int L = IM->height * IM->width;
  FLOAT *A_SCORE = (FLOAT *)calloc(L, sizeof(FLOAT));
  for (int ii = 0; ii < L; ii++) A_SCORE[ii] = -100.0;
  return A_SCORE;
}

// detect() (CT/detect.cpp:151-226) without its timers
static FLOAT *detect(IplImage *IM, MODEL *MO, FLOAT thresh, int *D_NUMS, FLOAT *A_SCORE,
                     long long *featsum) {
  FLOAT *scales = ini_scales(MO->MI, IM, IM->width, IM->height);
  int *featsize = ini_featsize(MO->MI);
  FLOAT **feature = calc_f_pyramid(IM, MO->MI, featsize, scales);
  *featsum = 0;
  for (int l = 0; l < MO->MI->max_scale + MO->MI->interval; l++)
    *featsum += 31LL * featsize[l * 2] * featsize[l * 2 + 1];
  FLOAT *boxes = get_boxes(feature, scales, featsize, MO, D_NUMS, A_SCORE, thresh);
  free(scales);
  free(featsize);
  free_features(feature, MO->MI);
  return boxes;
}

int main(int argc, char **argv) {
  if (argc != 7) {
    std::fprintf(stderr, "usage: %s frame.bgr width height thresh overlap out.bin\n", argv[0]);
    return 2;
  }
  const int W = std::atoi(argv[2]), H = std::atoi(argv[3]);
  const FLOAT thresh = (FLOAT)std::atof(argv[4]), overlap = (FLOAT)std::atof(argv[5]);
  std::vector<char> pix((size_t)W * H * 3);
  FILE *f = std::fopen(argv[1], "rb");
  if (!f || std::fread(pix.data(), 1, pix.size(), f) != pix.size()) {
    std::fprintf(stderr, "cannot read %s\n", argv[1]);
    return 2;
  }
  std::fclose(f);
  init_cuda();
  MODEL *MO = load_model(1.0f);
  IplImage IM = {};
  IM.nChannels = 3;
  IM.depth = 8;
  IM.width = W;
  IM.height = H;
  IM.widthStep = 3 * W;
  IM.imageData = pix.data();
  FLOAT *A_SCORE = ini_ac_score(&IM);
  int D_NUMS = 0;
  long long featsum = 0;
  // car_detection() (CT/detect.cpp:232-276)
  FLOAT *boxes = detect(&IM, MO, thresh, &D_NUMS, A_SCORE, &featsum);
  const int ncand = D_NUMS, GL = (MO->MI->numpart[0] + 1) * 4 + 3;
  FLOAT *rects = nms(boxes, overlap, &D_NUMS, MO);
  const int nkept = rects ? D_NUMS : 0;
  RESULT *CUR = get_new_rects(&IM, MO, rects, &D_NUMS);

  FILE *o = std::fopen(argv[6], "wb");
  if (!o) return 2;
  const int LEN = MO->MI->max_scale + MO->MI->interval;
  std::fwrite(&ncand, 4, 1, o);
  std::fwrite(&GL, 4, 1, o);
  std::fwrite(boxes, 4, (size_t)ncand * GL, o);
  std::fwrite(&nkept, 4, 1, o);
  if (nkept) std::fwrite(rects, 4, (size_t)nkept * GL, o);
  std::fwrite(A_SCORE, 4, (size_t)W * H, o);
  if (nkept) {
    std::fwrite(CUR->point, 4, (size_t)nkept * 4, o);
    std::fwrite(CUR->OR_point, 4, (size_t)nkept * 4, o);
    std::fwrite(CUR->type, 4, nkept, o);
    std::fwrite(CUR->scale, 4, nkept, o);
    std::fwrite(CUR->score, 4, nkept, o);
  }
  std::fwrite(&LEN, 4, 1, o);
  std::fwrite(&featsum, 8, 1, o);
  std::fclose(o);
  free(boxes);
  free(rects);
  free(A_SCORE);
  free_model(MO);
  clean_cuda();
  std::printf("compat_detect: %d candidates, %d kept\n", ncand, nkept);
  return 0;
}

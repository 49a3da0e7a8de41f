// dpm_runtime.h -- internal interfaces between the plan, the pipeline and the C ABI.
#pragma once

#include <string>

#include "dpm_internal.h"

namespace geye {

int compute_max_scale(const Model &m, int W, int H);
// geye_init; own_queue: the context's stream gets a hardware queue of its own (serving
// workers, whose streams run concurrently)
int ctx_create(int device, bool own_queue, geye_ctx **ctx);
void level_geometry(int max_scale, int sbin, int interval, int W, int H, std::vector<Level> &lv,
                    std::vector<int> &img_h, std::vector<int> &img_w,
                    std::vector<int> &img_src, std::vector<float> &img_scale);
int build_plan(Model &m, int W, int H, Plan &P, std::string &err, int max_scale_override = -1);
// The plan of nb frames from a one-frame plan: every descriptor repeated per frame with the
// frame's arena offsets; the features of frame f at f * feat_stride (FP32) / featb_stride
// (16-bit) elements (0: one's own sizes).
void replicate_plan(const Plan &one, int nb, long long feat_stride, long long featb_stride,
                    Plan &P);

void resize_tables(int sheight, int dheight, std::vector<int2> &tab, std::vector<int> &si,
                   std::vector<float> &alpha);
// DT jobs for one standalone map (dims0 = n0, dims1 = n1): pass 1 over columns of a
// negation-free source, pass 2 over rows; lds1/lds2 sized for the windowed kernel if window.
int dt_single_jobs(int n0, int n1, float ax, float bx, float ay, float by,
                   std::vector<DtJob> &p1, std::vector<DtJob> &p2, std::vector<int4> &nodes,
                   std::vector<int> &depths, size_t &lds1, size_t &lds2, bool window = false);

// lanes per line and dynamic LDS of a DT job over lines of length n (host only, allocates
// nothing); GEYE_ERR_UNSUPPORTED when no job shape fits
int dt_line_geometry(int n, bool window, int *G, size_t *lds);
// resolves the allocation-free entry point's scratch array on the current device, once per
// device (dpm_noalloc.hip)
int noalloc_setup(int device);

// the packed filters of a model (FP32 [group][31][w][h][NF]; 16-bit for the MFMA kernels,
// indexed by precision GEYE_BF16 .. GEYE_FP16X3)
struct ModelPacks {
  std::vector<float> f32;
  std::vector<unsigned short> h16[5];
};

// precision codes: valid at all, an MFMA form, and its 16-bit kind (1 = BF16, 2 = FP16) and
// whether it splits every operand into a hi/lo pair
inline bool valid_precision(int p) { return p >= GEYE_FP32 && p <= GEYE_FP16X3; }
inline bool mfma_precision(int p) { return p >= GEYE_BF16 && p <= GEYE_FP16X3; }
inline int mfma_kind(int p) { return p == GEYE_BF16 || p == GEYE_BF16X3 ? 1 : 2; }
inline bool split_precision(int p) { return p == GEYE_BF16X3 || p == GEYE_FP16X3; }
// the host side of a model from its description (no device work)
int model_pack(const geye_model_desc *d, Model &m, ModelPacks &pk);

// sets the calling thread's geye_last_error() message; returns code
int set_error(int code, const std::string &msg);

// plan registry (dpm_runtime.cpp)
long long plan_key(int W, int H, int max_scale);
int get_plan(geye_ctx *c, Model &m, int W, int H, Plan **out, int max_scale);
int plan_bytes(const Model &m, int W, int H, long long *bytes);
void release_plans(const geye_ctx *c, const Model *m, long long key);

// host NMS with the reference's exact semantics (CT/nms.cpp:70-217)
int nms_host(const float *boxes, int n, int GL, float overlap, float *out);
// the reference's quicksort (CT/nms.cpp:70-84); with `only`, restricted to the subarrays that
// can hold one of those values (their elements end where the whole sort puts them)
void quicksort_host(float *ary, int *order, int first, int last,
                    const std::vector<float> *only = nullptr);

}  // namespace geye

// dpm_noalloc.hip -- geye_dt_noalloc: the DT of one map (dt(), CT/dt.cpp:101-136) on device
// memory the caller owns, with no device allocation.  For callers that cannot allocate
// (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// Its working memory -- the first pass's M and argmins, the job and node tables the DT
// kernels read -- lives in one file-scope __device__ array, `scratch`, of kNoAllocBytes
// (about 8.4 MB), which every GPU the library runs on holds until the process ends.  The
// job and node tables are built on the device from by-value kernel arguments, so nothing
// is copied from host memory either.  The passes are geye_dt's own kernels (launch_dt), so
// the outputs are geye_dt's, bit for bit.
//
// In host code `scratch` names a host-side stand-in, not device memory: its device address
// is resolved once per device with hipGetSymbolAddress (noalloc_setup, called by geye_init
// with that device current) and only that address is used.  Every call on a GPU shares that
// GPU's array, so calls that overlap on one GPU (two contexts, or two threads with
// different contexts) race; calls on one context are serialised by its mutex.  Nothing in
// the array needs to start at zero: every word a call reads was written by that call.
#include <mutex>

#include "dpm_kernels.h"
#include "dpm_runtime.h"

namespace geye {

namespace {

// limits: each map dimension up to 2048 cells, up to 2^20 cells in all (a 3840x2160 frame's
// largest part map has about 5 x 10^5)
constexpr int kNoAllocMaxLine = 2048;
constexpr long long kNoAllocMaxCells = 1LL << 20;
// jobs of the two passes: ceil(n1 / NL1) + ceil(n0 / NL2) with NL >= 4 lines per job
constexpr int kNoAllocMaxJobs = 2 * (kNoAllocMaxLine / 4) + 16;
constexpr int kNoAllocMaxDepth = 32;  // levels of the divide and conquer (12 at 2048)

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kOffTmpM = 0;
constexpr size_t kOffTmpIy = kOffTmpM + align256(sizeof(float) * kNoAllocMaxCells);
constexpr size_t kOffJobs = kOffTmpIy + align256(sizeof(int) * kNoAllocMaxCells);
constexpr size_t kOffNodes = kOffJobs + align256(sizeof(DtJob) * kNoAllocMaxJobs);
constexpr size_t kOffDepths = kOffNodes + align256(sizeof(int4) * 2 * kNoAllocMaxLine);
constexpr size_t kNoAllocBytes = kOffDepths + align256(sizeof(int) * 2 * (kNoAllocMaxDepth + 1));

__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_setup_mu;
void *g_scratch[kMaxDevices] = {};     // device address of `scratch` per device
hipError_t g_setup_rc[kMaxDevices] = {};
bool g_setup_done[kMaxDevices] = {};

// Breadth-first node table of the divide and conquer over [0, n) (the order the DT kernels
// walk, one depth at a time): node (d, d1 - 1, d2 + 1) with d = (d1 + d2) >> 1, its children
// [d1, d - 1] and [d + 1, d2]; depths[k] is the first node of depth k, depths[count] = n.
// One thread, in LDS; returns the depth count.
__device__ int dt_build_nodes(int n, int4 *nodes, int *depths) {
  nodes[0] = make_int4((n - 1) >> 1, -1, n, 0);
  int cnt = 1, begin = 0, level = 0;
  depths[0] = 0;
  while (begin < cnt) {
    const int end = cnt;
    for (int i = begin; i < end; i++) {
      const int4 nd = nodes[i];
      const int d = nd.x, d1 = nd.y + 1, d2 = nd.z - 1;
      if (d - 1 >= d1) nodes[cnt++] = make_int4((d1 + d - 1) >> 1, d1 - 1, d, 0);
      if (d2 >= d + 1) nodes[cnt++] = make_int4((d + 1 + d2) >> 1, d, d2 + 1, 0);
    }
    depths[++level] = end;
    begin = end;
  }
  return level;
}

// The jobs of both passes (as dt_single_jobs lays them out on the host) and the node tables
// of the two line lengths, in scratch.  One workgroup of 256 threads.
__global__ __launch_bounds__(256) void k_dt_noalloc_setup(int n0, int n1, float ax, float bx,
                                                           float ay, float by, int G1, int G2,
                                                           DtJob *jobs, int4 *nodes,
                                                           int *depths) {
  __shared__ int4 s_nodes[kNoAllocMaxLine];
  __shared__ int s_dep[kNoAllocMaxDepth + 1];
  __shared__ int s_count[2];
  for (int t = 0; t < 2; t++) {
    const int n = t == 0 ? n0 : n1;
    if (threadIdx.x == 0) s_count[t] = dt_build_nodes(n, s_nodes, s_dep);
    __syncthreads();
    int4 *gn = nodes + (t == 0 ? 0 : n0);
    int *gd = depths + (t == 0 ? 0 : kNoAllocMaxDepth + 1);
    for (int i = threadIdx.x; i < n; i += 256) gn[i] = s_nodes[i];
    for (int i = threadIdx.x; i <= s_count[t]; i += 256) gd[i] = s_dep[i];
    __syncthreads();
  }
  const int NL1 = 256 / G1, NL2 = 256 / G2;
  const int j1 = (n1 + NL1 - 1) / NL1, j2 = (n0 + NL2 - 1) / NL2;
  for (int k = threadIdx.x; k < j1 + j2; k += 256) {
    const bool p1 = k < j1;  // pass 1: columns (length n0) with (ay, by); pass 2: rows
    const int q = p1 ? k : k - j1, NL = p1 ? NL1 : NL2, lines = p1 ? n1 : n0;
    DtJob J{};
    J.src_off = 0;
    J.dst_off = 0;
    J.n = p1 ? n0 : n1;
    J.line0 = q * NL;
    J.nlines = min(NL, lines - J.line0);
    J.line_stride = p1 ? n0 : 1;
    J.elem_stride = p1 ? 1 : n0;
    J.node_off = p1 ? 0 : n0;
    J.depth_off = p1 ? 0 : kNoAllocMaxDepth + 1;
    J.depth_count = s_count[p1 ? 0 : 1];
    J.G = p1 ? G1 : G2;
    J.neg = 0;
    J.a = p1 ? ay : ax;
    J.b = p1 ? by : bx;
    J.sel = 0;
    J.ax = J.ay = 0;
    J.fb_lines = 0;
    jobs[k] = J;
  }
}

// Ix = tmpIx (the second pass wrote it into Ix), Iy = tmpIy[Ix * dims0 + y] (CT/dt.cpp:120-131)
__global__ __launch_bounds__(256) void k_dt_compose(const int *__restrict__ Ix,
                                                     const int *__restrict__ tmpIy, int n0,
                                                     long long cells, int *__restrict__ Iy) {
  const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
  if (k >= cells) return;
  const int y = (int)(k % n0);
  Iy[k] = tmpIy[(long long)Ix[k] * n0 + y];
}

}  // namespace

int noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return GEYE_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_setup_mu);
  if (!g_setup_done[device]) {
    void *p = nullptr;
    g_setup_rc[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch[device] = g_setup_rc[device] == hipSuccess ? p : nullptr;
    g_setup_done[device] = true;
  }
  return g_scratch[device] ? GEYE_OK : GEYE_ERR_HIP;
}

}  // namespace geye

using namespace geye;

extern "C" int geye_dt_noalloc(geye_ctx *c, const float *d_vals, float ax, float bx, float ay,
                               float by, const int *dims, float *d_M, int *d_Ix, int *d_Iy) {
  if (!c || !d_vals || !dims || !d_M || !d_Ix || !d_Iy)
    return set_error(GEYE_ERR_ARG, "dt_noalloc: null argument");
  const int n0 = dims[0], n1 = dims[1];
  if (n0 < 1 || n1 < 1) return set_error(GEYE_ERR_ARG, "dt_noalloc: empty map");
  const long long cells = (long long)n0 * n1;
  if (n0 > kNoAllocMaxLine || n1 > kNoAllocMaxLine || cells > kNoAllocMaxCells)
    return set_error(GEYE_ERR_ARG,
                     "dt_noalloc: map " + std::to_string(n0) + " x " + std::to_string(n1) +
                         " exceeds the scratch limits (each dimension <= 2048, cells <= 2^20)");
  int G1 = 0, G2 = 0;
  size_t lds1 = 0, lds2 = 0;
  if (dt_line_geometry(n0, false, &G1, &lds1) != GEYE_OK ||
      dt_line_geometry(n1, false, &G2, &lds2) != GEYE_OK)
    return set_error(GEYE_ERR_UNSUPPORTED, "dt_noalloc: map dimension too large for the DT kernel");
  const int j1 = (n1 + 256 / G1 - 1) / (256 / G1), j2 = (n0 + 256 / G2 - 1) / (256 / G2);
  if (j1 + j2 > kNoAllocMaxJobs) return set_error(GEYE_ERR_ARG, "dt_noalloc: too many DT jobs");
  std::lock_guard<std::mutex> lk(c->mu);
  if (hipSetDevice(c->device) != hipSuccess)
    return set_error(GEYE_ERR_HIP, "dt_noalloc: hipSetDevice failed");
  (void)hipGetLastError();
  if (noalloc_setup(c->device) != GEYE_OK)
    return set_error(GEYE_ERR_HIP, std::string("dt_noalloc: scratch symbol address: ") +
                                       hipGetErrorString(g_setup_rc[c->device]));
  unsigned char *base = (unsigned char *)g_scratch[c->device];
  float *tmpM = (float *)(base + kOffTmpM);
  int *tmpIy = (int *)(base + kOffTmpIy);
  DtJob *jobs = (DtJob *)(base + kOffJobs);
  int4 *nodes = (int4 *)(base + kOffNodes);
  int *depths = (int *)(base + kOffDepths);
  hipStream_t s = c->stream;
  hipLaunchKernelGGL(k_dt_noalloc_setup, dim3(1), dim3(256), 0, s, n0, n1, ax, bx, ay, by, G1,
                     G2, jobs, nodes, depths);
  if (launch_dt(s, jobs, j1, lds1, nodes, depths, d_vals, tmpM, tmpIy) != 0 ||
      launch_dt(s, jobs + j1, j2, lds2, nodes, depths, tmpM, d_M, d_Ix) != 0)
    return set_error(GEYE_ERR_HIP, "dt_noalloc: DT kernel's dynamic LDS attribute not set");
  hipLaunchKernelGGL(k_dt_compose, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s, d_Ix,
                     tmpIy, n0, cells, d_Iy);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess)
    return set_error(GEYE_ERR_HIP, std::string("dt_noalloc: ") + hipGetErrorString(e));
  return GEYE_OK;
}

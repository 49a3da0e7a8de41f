// host_checks.cpp -- the host code of the library and the CPU oracle run under AddressSanitizer
// and UndefinedBehaviorSanitizer (built by tests/sanitize/Makefile; driven by
// tests/test_host_sanitizers.py).  Only host-side work: the level plan, the plan builder
// (through geye_plan_bytes_desc, which packs the model and lays out every descriptor and arena
// of a frame size), the host NMS, the CSV reader, and the oracle's whole detection of small
// frames.  No device is needed.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/geye_dpm.h"
#include "../../oracle/dpm_oracle.h"

namespace {

uint64_t g_state = 0x9E3779B97F4A7C15ull;
uint64_t next_u64() {  // SplitMix64
  uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
float uniform() { return (float)((next_u64() >> 11) * (1.0 / 9007199254740992.0)); }

int failures = 0;
void expect(bool ok, const char *what) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s (%s)\n", what, geye_last_error());
    failures++;
  }
}

// A model of `C` components with roots (h, w), `np` parts of ps x ps each, anchors inside the
// 2x root window, the arrays owned by the struct.
struct HostModel {
  int C, np, ps;
  std::vector<int> ridx, oidx, rsize, numpart, pidx, didx, psize, anchor, root_size, part_size;
  std::vector<float> offw, def, rootf, partf, lsq;
  std::vector<long long> root_off, part_off;
  geye_model_desc d{};
  HostModel(const std::vector<std::pair<int, int>> &roots, int nparts, int psz)
      : C((int)roots.size()), np(nparts), ps(psz) {
    for (int c = 0; c < C; c++) {
      const int h = roots[c].first, w = roots[c].second;
      ridx.push_back(c);
      oidx.push_back(c);
      offw.push_back(uniform() - 0.5f);
      rsize.push_back(h);
      rsize.push_back(w);
      numpart.push_back(np);
      root_size.insert(root_size.end(), {h, w, 31});
      root_off.push_back((long long)rootf.size());
      for (int k = 0; k < 31 * h * w; k++) rootf.push_back((uniform() - 0.5f) * 0.1f);
      for (int k = 0; k < np; k++) {
        const int p = c * np + k;
        pidx.push_back(p);
        didx.push_back(p);
        psize.push_back(ps);
        psize.push_back(ps);
        part_size.insert(part_size.end(), {ps, ps, 31});
        part_off.push_back((long long)partf.size());
        for (int q = 0; q < 31 * ps * ps; q++) partf.push_back((uniform() - 0.5f) * 0.1f);
        def.insert(def.end(), {0.01f, 0.0f, 0.01f, 0.0f});
        anchor.push_back(1 + (k % 4) * (w - 2) / 3);
        anchor.push_back(1 + (k / 4) * (h - 2));
      }
      for (int q = 0; q < 4 * (1 + 2 * (1 + np)); q++) lsq.push_back(q % 3 == 0 ? 1.0f : 0.0f);
    }
    int maxY = 0, maxX = 0;
    for (int c = 0; c < C; c++) {
      maxY = std::max(maxY, rsize[2 * c]);
      maxX = std::max(maxX, rsize[2 * c + 1]);
    }
    d.numcomponent = C;
    d.sbin = 8;
    d.interval = 10;
    d.max_Y = maxY;
    d.max_X = maxX;
    d.maxpart = np;
    d.ridx = ridx.data();
    d.oidx = oidx.data();
    d.offw = offw.data();
    d.rsize = rsize.data();
    d.numpart = numpart.data();
    d.pidx = pidx.data();
    d.didx = didx.data();
    d.psize = psize.data();
    d.DefL = C * np;
    d.def = def.data();
    d.anchor = anchor.data();
    d.NoR = C;
    d.root_size = root_size.data();
    d.rootfilter = rootf.data();
    d.root_off = root_off.data();
    d.NoP = C * np;
    d.part_size = part_size.data();
    d.partfilter = partf.data();
    d.part_off = part_off.data();
    d.lsq = lsq.data();
  }
  // the same arrays as the oracle's description (identical field order and meaning)
  ora_model_desc oracle_desc() const {
    ora_model_desc o{};
    static_assert(sizeof(ora_model_desc) == sizeof(geye_model_desc), "description layouts");
    std::memcpy(&o, &d, sizeof o);
    return o;
  }
};

std::vector<unsigned char> frame(int W, int H) {
  std::vector<unsigned char> f((size_t)3 * W * H);
  for (auto &b : f) b = (unsigned char)(next_u64() >> 56);
  for (int r = 0; r < 8; r++) {  // constant rectangles
    const int x0 = (int)(uniform() * W), y0 = (int)(uniform() * H);
    const int w = 20 + (int)(uniform() * 100), h = 20 + (int)(uniform() * 100);
    const unsigned char c = (unsigned char)(next_u64() >> 56);
    for (int y = y0; y < std::min(H, y0 + h); y++)
      for (int x = x0; x < std::min(W, x0 + w); x++)
        for (int k = 0; k < 3; k++) f[((size_t)y * W + x) * 3 + k] = c;
  }
  return f;
}

}  // namespace

int main() {
  HostModel car({{6, 11}, {6, 11}, {7, 10}, {7, 10}, {8, 9}, {8, 9}}, 8, 6);
  HostModel tiny({{3, 4}, {4, 3}}, 2, 3);
  // plan builder and level plan over frame sizes, odd ones included
  const int sizes[][2] = {{640, 480}, {1280, 720}, {1920, 1080}, {3840, 2160}, {97, 131},
                          {200, 150}, {90, 400}, {1023, 17}};
  for (auto &sz : sizes) {
    long long b1 = 0, b2 = 0;
    expect(geye_plan_bytes_desc(&car.d, sz[0], sz[1], &b1) == GEYE_OK || sz[1] < 40,
           "plan bytes (car)");
    expect(geye_plan_bytes_desc(&tiny.d, sz[0], sz[1], &b2) == GEYE_OK, "plan bytes (tiny)");
    int ms = 0, len = 0;
    std::vector<float> scales(128);
    std::vector<int> fh(128), fw(128);
    expect(geye_level_plan_dims(car.rsize.data(), car.C, 8, 10, sz[0], sz[1], &ms, &len,
                                scales.data(), fh.data(), fw.data()) == GEYE_OK,
           "level plan");
  }
  // rejected descriptions leave no trace
  {
    HostModel bad({{6, 11}}, 2, 6);
    bad.pidx[1] = 99;  // part index out of range
    long long b = 0;
    expect(geye_plan_bytes_desc(&bad.d, 640, 480, &b) == GEYE_ERR_ARG, "bad pidx rejected");
  }
  // host NMS on random and tied boxes
  for (int n : {0, 1, 2, 50, 3000}) {
    const int GL = 11;
    std::vector<float> rows((size_t)std::max(n, 1) * GL), out(rows.size());
    for (int i = 0; i < n; i++) {
      float *r = &rows[(size_t)i * GL];
      r[0] = (float)(int)(uniform() * 400);
      r[1] = (float)(int)(uniform() * 600);
      r[2] = r[0] + 20 + (float)(int)(uniform() * 100);
      r[3] = r[1] + 20 + (float)(int)(uniform() * 140);
      r[GL - 2] = std::round((uniform() - 0.5f) * 8) / 4;  // many ties
      r[GL - 1] = 8;
    }
    int k = 0;
    expect(geye_nms_host(rows.data(), n, GL, 0.5f, out.data(), &k) == GEYE_OK, "host nms");
    std::vector<float> o2(rows.size());
    const int k2 = ora_nms(rows.data(), n, GL, 0.5f, o2.data());
    expect(k == k2 && std::memcmp(out.data(), o2.data(), sizeof(float) * (size_t)k * GL) == 0,
           "host nms equals the oracle");
  }
  // CSV reader: a file written here, then a missing one
  {
    const char *path = "tests/_bin/host_checks.csv";
    FILE *f = std::fopen(path, "w");
    if (f) {
      std::fprintf(f, "2,8,10,6,13\n-4.058444, 1e-3\n");
      std::fclose(f);
      float *v = nullptr;
      long long n = 0;
      expect(geye_csv_values(path, &v, &n) == GEYE_OK && n == 7, "csv values");
      geye_free(v);
    }
    float *v = nullptr;
    long long n = 0;
    expect(geye_csv_values("tests/_bin/none.csv", &v, &n) == GEYE_ERR_IO, "missing csv");
  }
  // the oracle's whole detection of small frames, threaded and not, FP32 and FP64
  for (HostModel *m : {&tiny, &car}) {
    const ora_model_desc od = m->oracle_desc();
    for (int t : {1, 4}) {
      const int W = m == &tiny ? 200 : 320, H = m == &tiny ? 150 : 240;
      std::vector<unsigned char> fr = frame(W, H);
      for (int prec : {0, 1}) {
        float *boxes = nullptr;
        int GL = 0;
        const long long n = ora_detect(&od, fr.data(), W, H, 3 * W, -0.5f, prec, t, &boxes, &GL);
        expect(n >= 0, "oracle detect");
        if (n > 0) {
          std::vector<float> out((size_t)n * GL);
          ora_nms(boxes, (int)n, GL, 0.5f, out.data());
        }
        ora_free(boxes);
      }
    }
  }
  // oracle DT and resize on odd shapes
  {
    const int d0 = 37, d1 = 29;
    std::vector<float> v((size_t)d0 * d1), M(v.size());
    std::vector<int> Ix(v.size()), Iy(v.size());
    for (auto &x : v) x = uniform() - 0.5f;
    const int dims[2] = {d0, d1};
    ora_dt_f32(v.data(), 0.05f, 0.01f, 0.03f, -0.01f, dims, M.data(), Ix.data(), Iy.data());
    const int sd[3] = {31, 23, 3};
    std::vector<float> img((size_t)31 * 23 * 3);
    for (auto &x : img) x = uniform() * 255;
    int od[3];
    const long long n = ora_resize(img.data(), sd, 0.707107f, nullptr, od);
    std::vector<float> out((size_t)std::max(n, 1LL));
    ora_resize(img.data(), sd, 0.707107f, out.data(), od);
  }
  std::printf("host checks: %d failures\n", failures);
  return failures ? 1 : 0;
}

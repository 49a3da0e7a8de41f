// ref_harness.cpp -- C entry points over the reference's own CPU functions, compiled from
// the reference tree (CT/resize.cc, CT/dt.cpp) into oracle/_ref/libdpm_ref.so.  Test
// infrastructure only: it pins the oracle's resize and dt against the reference's code.
// Declarations restate the reference signatures (CT/resize.cc:43, CT/dt.cpp:30-32) for
// FLOAT = float (CT/switch_float.h:2-5).
#include <cstdlib>
#include <cstring>

float *resize(float *src, int *sdims, int *odims, float scale);
float *dt(float *vals, float ax, float bx, float ay, float by, int *dims, int *Ix, int *Iy);
void add_part_calculation(float *score, float *M, int *rootsize, int *partsize, int ax, int ay);

extern "C" {
long long ref_resize(const float *src, const int *sdims, float scale, float *dst, int *odims) {
  int sd[3] = {sdims[0], sdims[1], sdims[2]};
  long long n = (long long)sd[0] * sd[1] * sd[2];
  float *s = (float *)malloc(sizeof(float) * n);
  memcpy(s, src, sizeof(float) * n);
  float *r = resize(s, sd, odims, scale);
  long long m = (long long)odims[0] * odims[1] * odims[2];
  if (dst) memcpy(dst, r, sizeof(float) * m);
  free(r);
  free(s);
  return m;
}

void ref_dt(const float *vals, float ax, float bx, float ay, float by, const int *dims,
            float *M, int *Ix, int *Iy) {
  int d[2] = {dims[0], dims[1]};
  long long n = (long long)d[0] * d[1];
  float *v = (float *)malloc(sizeof(float) * n);
  memcpy(v, vals, sizeof(float) * n);
  float *m = dt(v, ax, bx, ay, by, d, Ix, Iy);
  memcpy(M, m, sizeof(float) * n);
  free(m);
  free(v);
}

void ref_add_part(float *score, const float *M, const int *rootsize, const int *partsize,
                  int ax, int ay) {
  int rs[2] = {rootsize[0], rootsize[1]}, ps[2] = {partsize[0], partsize[1]};
  add_part_calculation(score, (float *)M, rs, ps, ax, ay);
}
}

// bh_host_util.cpp — host-side helpers of the C ABI that need no GPU:
// camera focal length, Object pose rotations, look-at framing, the Annulus
// constructor's derived normal, and the seeded input generators shared by
// every test and benchmark.
#include <cmath>
#include <cstring>

#include "bh_api.h"

namespace {

constexpr double kPi = 3.141592653589793238462643383279;

inline double dot3(const double* a, const double* b) {
  double s = 0;
  s += a[0] * b[0];
  s += a[1] * b[1];
  s += a[2] * b[2];
  return s;
}

// cv::normalize semantics: v * (|v| ? 1/|v| : 0)
inline void normalize3(double* v) {
  double n = std::sqrt(dot3(v, v));
  double k = n != 0.0 ? 1. / n : 0.;
  v[0] *= k;
  v[1] *= k;
  v[2] *= k;
}

inline void cross3(const double* a, const double* b, double* o) {
  double r0 = a[1] * b[2] - a[2] * b[1];
  double r1 = a[2] * b[0] - a[0] * b[2];
  double r2 = a[0] * b[1] - a[1] * b[0];
  o[0] = r0;
  o[1] = r1;
  o[2] = r2;
}

// RotationMatrixForAxis (matrix.h:342-356): Rodrigues matrix about the unit
// axis u = normalize(v) by theta, row-major.
void rotation_for_axis(const double* v, double theta, double* m) {
  double u[3] = {v[0], v[1], v[2]};
  normalize3(u);
  const double ux = u[0], uy = u[1], uz = u[2];
  const double c = std::cos(theta);
  const double c2 = 1 - c;
  const double s = std::sin(theta);
  m[0] = c + ux * ux * c2;
  m[1] = ux * uy * c2 - uz * s;
  m[2] = ux * uz * c2 + uy * s;
  m[3] = uy * ux * c2 + uz * s;
  m[4] = c + uy * uy * c2;
  m[5] = uy * uz * c2 - ux * s;
  m[6] = uz * ux * c2 - uy * s;
  m[7] = uz * uy * c2 + ux * s;
  m[8] = c + uz * uz * c2;
}

inline void matvec3(const double* m, const double* x, double* o) {
  double r[3];
  for (int i = 0; i < 3; ++i) {
    double s = 0;
    s += m[3 * i + 0] * x[0];
    s += m[3 * i + 1] * x[1];
    s += m[3 * i + 2] * x[2];
    r[i] = s;
  }
  o[0] = r[0];
  o[1] = r[1];
  o[2] = r[2];
}

}  // namespace

extern "C" {

int bh_version(void) { return BH_API_VERSION; }

// Camera::fov clamp and focal length (camera.h:37-41).
double bh_focal_length(int width, double fov) {
  double hi = std::nextafter(kPi, 0.0);
  double f = fov;
  if (!(0.0 < f)) f = 0.0;
  if (hi < f) f = hi;
  return width / (2 * std::tan(f / 2.0));
}

// Object::RotateX/Y/Z (object.h:58-80): rotate the two other axes (then
// renormalise) and every vertex about vertex[0].
void bh_object_rotate(double axes[9], double* verts, int n_verts, int which, double rad) {
  if (!axes || which < 0 || which > 2) return;
  double m[9];
  rotation_for_axis(axes + 3 * which, rad, m);
  for (int k = 0; k < 3; ++k) {
    if (k == which) continue;
    double* a = axes + 3 * k;
    matvec3(m, a, a);
    normalize3(a);
  }
  if (verts && n_verts > 0) {
    const double p0[3] = {verts[0], verts[1], verts[2]};
    for (int i = 0; i < n_verts; ++i) {
      double* v = verts + 3 * i;
      double d[3] = {v[0] - p0[0], v[1] - p0[1], v[2] - p0[2]};
      double r[3];
      matvec3(m, d, r);
      v[0] = r[0] + p0[0];
      v[1] = r[1] + p0[1];
      v[2] = r[2] + p0[2];
    }
  }
}

// Look-at framing (not in the reference, whose camera only has Move*/Rotate*):
// axis_x toward the target, axis_z = -(up orthogonalised), axis_y = z cross x,
// which reproduces the default frame (1,0,0),(0,-1,0),(0,0,-1) for a camera
// on the -x axis looking at the origin with +z up.
void bh_camera_look_at(bh_camera* cam, const double target[3], const double up[3]) {
  if (!cam || !target || !up) return;
  double x[3] = {target[0] - cam->position[0], target[1] - cam->position[1],
                 target[2] - cam->position[2]};
  normalize3(x);
  double k = dot3(up, x);
  double z[3] = {-(up[0] - k * x[0]), -(up[1] - k * x[1]), -(up[2] - k * x[2])};
  normalize3(z);
  double y[3];
  cross3(z, x, y);
  normalize3(y);
  std::memcpy(cam->axis_x, x, sizeof x);
  std::memcpy(cam->axis_y, y, sizeof y);
  std::memcpy(cam->axis_z, z, sizeof z);
}

// Annulus(p1, p2, p3, p4, r_outer, r_inner) (vector_object.h:320-326):
// vertex list {0, p1, p2, p3, p4}, normal = normalize((p3-p1) x (p2-p1)).
void bh_annulus_init(bh_annulus* a, const double p1[3], const double p2[3],
                     const double p3[3], const double p4[3], double r_outer,
                     double r_inner, int texture) {
  if (!a) return;
  std::memset(a, 0, sizeof *a);
  const double* p[4] = {p1, p2, p3, p4};
  for (int i = 0; i < 4; ++i) std::memcpy(a->vertex[i + 1], p[i], 3 * sizeof(double));
  double e1[3] = {p3[0] - p1[0], p3[1] - p1[1], p3[2] - p1[2]};
  double e2[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
  cross3(e1, e2, a->normal);
  normalize3(a->normal);
  a->r_outer = r_outer;
  a->r_inner = r_inner;
  a->texture = texture;
}

// SplitMix64 (Steele, Lea, Flood 2014; constants as published by Vigna).
uint64_t bh_splitmix64(uint64_t* state) {
  uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double bh_uniform(uint64_t* state, double lo, double hi) {
  double u = (double)(bh_splitmix64(state) >> 11) * (1.0 / 9007199254740992.0);
  return lo + (hi - lo) * u;
}

void bh_gen_noise(uint8_t* dst, size_t n, uint64_t seed) {
  uint64_t st = seed;
  size_t i = 0;
  while (i < n) {
    uint64_t r = bh_splitmix64(&st);
    for (int k = 0; k < 8 && i < n; ++k, ++i) dst[i] = (uint8_t)(r >> (8 * k));
  }
}

// Background: per texel three noise bytes from one SplitMix64 draw; cells of
// (cols/cells_x) x (rows/cells_y) texels alternate dark (n/4) and light
// (192 + n/4).
void bh_gen_background(uint8_t* dst, int rows, int cols, int cells_x, int cells_y,
                       uint64_t seed) {
  if (!dst || rows <= 0 || cols <= 0) return;
  if (cells_x < 1) cells_x = 1;
  if (cells_y < 1) cells_y = 1;
  const int cw = cols / cells_x > 0 ? cols / cells_x : 1;
  const int ch = rows / cells_y > 0 ? rows / cells_y : 1;
  uint64_t st = seed;
  for (int y = 0; y < rows; ++y) {
    for (int x = 0; x < cols; ++x) {
      uint64_t r = bh_splitmix64(&st);
      int light = ((x / cw) + (y / ch)) & 1;
      uint8_t* p = dst + ((size_t)y * cols + x) * 3;
      for (int c = 0; c < 3; ++c) {
        int n = (int)((r >> (8 * c)) & 0xff);
        p[c] = (uint8_t)(light ? 192 + (n >> 2) : (n >> 2));
      }
    }
  }
}

namespace {
uint32_t isqrt64(uint64_t v) {
  uint64_t r = (uint64_t)std::sqrt((double)v);
  while (r * r > v) --r;
  while ((r + 1) * (r + 1) <= v) ++r;
  return (uint32_t)r;
}
}  // namespace

// Disc texture, RGBA: band = floor(rho / band_px) of the texel's distance from
// the image centre (integer arithmetic on doubled coordinates); even bands
// warm (R high), odd bands cool (B high); each channel gets uniform noise
// in [0, 63]; alpha 255.
void bh_gen_disc_texture(uint8_t* dst, int rows, int cols, int band_px, uint64_t seed) {
  if (!dst || rows <= 0 || cols <= 0) return;
  if (band_px < 1) band_px = 1;
  uint64_t st = seed;
  static const uint8_t warm[3] = {192, 120, 40};
  static const uint8_t cool[3] = {60, 110, 190};
  for (int y = 0; y < rows; ++y) {
    for (int x = 0; x < cols; ++x) {
      int64_t dx = 2 * (int64_t)x + 1 - cols;
      int64_t dy = 2 * (int64_t)y + 1 - rows;
      uint32_t rho2 = isqrt64((uint64_t)(dx * dx + dy * dy));  // 2 * rho
      int band = (int)(rho2 / (uint32_t)(2 * band_px)) & 1;
      const uint8_t* base = band ? cool : warm;
      uint64_t r = bh_splitmix64(&st);
      uint8_t* p = dst + ((size_t)y * cols + x) * 4;
      for (int c = 0; c < 3; ++c) p[c] = (uint8_t)(base[c] + ((r >> (8 * c)) & 63));
      p[3] = 255;
    }
  }
}

}  // extern "C"

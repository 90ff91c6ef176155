// bh_flat.hip — the flat-space march of the ray_tracer_test entry point
// (include/blackhole/ray_tracer.h:52-96): from the camera through the pixel
// point, straight chords of a fixed vector (BasicLinearRayRecurrence,
// ray_tracer.h:17-35), each chord tested against every object for the nearest
// hit (object_manager.h:69-85).  FP64 in the reference's operation order, one
// lane per pixel; the work is a few dozen flops per chord, so the kernel is
// bound by its output stores.
#include "bh_internal.h"
#include "bh_fp64.h"

namespace bh {
namespace {

// Rectangle::Collide (vector_object.h:107-127): sign change of n.(P - p1)
// across the chord, crossing by linear interpolation, inside the
// parallelogram p1 + [0,1] s1 + [0,1] s3.
__device__ inline bool rect_collide(const double (*v)[3], d3 aa, d3 bb, d3* inter) {
  const d3 p0 = ld3(v[1]);
  const d3 p1 = ld3(v[2]) - p0;
  const d3 p3 = ld3(v[4]) - p0;
  const d3 n = normalize(cross(p1, p3));
  const d3 Q1 = aa - p0;
  const d3 Q2 = bb - p0;
  const double q1 = dot(n, Q1);
  const double q2 = dot(n, Q2);
  if (q1 * q2 >= 0) return false;
  const d3 Q = (Q1 * fabs(q2) + Q2 * fabs(q1)) * (1.0 / (fabs(q1) + fabs(q2)));
  if (dot(p1, Q) > 0 && dot(p1, p1) > dot(p1, Q) && dot(p3, Q) > 0 && dot(p3, p3) > dot(p3, Q)) {
    *inter = Q + p0;
    return true;
  }
  return false;
}

// InfinitePlane::Collide (vector_object.h:210-225); a chord lying in the
// plane (t1 = t2 = 0) gives the reference's 0/0 crossing.
__device__ inline bool plane_collide(const DevFlat& s, int i, d3 p1, d3 p2, d3* inter) {
  const double* z = s.pz[i];
  const double* o = s.pp[i];
  const double t1 = z[0] * (p1.x - o[0]) + z[1] * (p1.y - o[1]) + z[2] * (p1.z - o[2]);
  const double t2 = z[0] * (p2.x - o[0]) + z[1] * (p2.y - o[1]) + z[2] * (p2.z - o[2]);
  if (t1 * t2 > 0) return false;
  const double d1 = fabs(t1), d2 = fabs(t2);
  *inter = (p1 * d2 + p2 * d1) * (1.0 / (d1 + d2));
  return true;
}

// ChessPattern2D::mod (pattern.h:44-47): |a| minus whole multiples of
// 2 size, with C++ integer division of truncated values; size is a whole
// number >= 1 (checked on the host).  Beyond the int range, where the
// reference's (int) cast is undefined, the same truncated quotient is taken
// in double arithmetic.
__device__ inline double chess_mod(double a, double ps) {
  const double b = fabs(a);
  if (!(b < 2147483648.0)) return b - floor(floor(b) / (ps * 2)) * (ps * 2);
  return b - (double)((int)b / (int)(ps * 2)) * (ps * 2);
}

// InfinitePlane::color (vector_object.h:227-232) with ChessPattern2D
// (pattern.h:22-42): plane coordinates (a, b) of (x, y), then the checker.
__device__ inline uint32_t plane_color(const DevFlat& s, int i, d3 q) {
  if (!(s.psize[i] > 0)) return 0u;
  const double* vx = s.px[i];
  const double* vy = s.py[i];
  const double b = (vx[0] * q.y - vx[1] * q.x) / (vx[0] * vy[1] - vx[1] * vy[0]);
  const double a = (q.x - b * vy[0]) / vx[0];
  const double ps = s.psize[i];
  const double x2 = chess_mod(a, ps), y2 = chess_mod(b, ps);
  const bool in = (x2 <= ps && y2 <= ps) || (ps <= x2 && ps <= y2);
  const bool white = (a * b > 0) ? in : !in;
  return white ? 0xffffffu : 0u;
}

// Rectangle::color (vector_object.h:159-179), nearest texel, clamped.
__device__ inline uint32_t rect_color(const DevFlat& s, int i, d3 p) {
  const DevTex& t = s.rtex[i];
  const d3 v1 = ld3(s.rv[i][1]);
  const d3 s1 = ld3(s.rv[i][2]) - v1;
  const d3 s2 = ld3(s.rv[i][4]) - v1;
  const d3 v = p - v1;
  const double r = sqrt(dot(v, v));
  double theta = acos(dot(s1, v) / (sqrt(dot(s1, s1)) * r));
  theta = isnan(theta) ? 0 : theta;
  const double a = (r * cos(theta) / sqrt(dot(s1, s1))) * t.cols;
  const double b = (r * sin(theta) / sqrt(dot(s2, s2))) * t.rows;
  return texel(t, tex_index(a, t.cols), tex_index(b, t.rows));
}

__global__ __launch_bounds__(256) void flat_kernel(DevFlat s, DevOut out) {
  const int x = blockIdx.x * 16 + (threadIdx.x & 15);
  const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= s.W || y >= s.H) return;
  const size_t idx = (size_t)y * s.W + x;
  // Camera::PixelVector (camera.h:55-63), treated as a point (the chord from
  // the camera to it is the first one)
  const double sx = s.W / 2.0 - x, sy = s.H / 2.0 - y;
  d3 old = ld3(s.C);
  d3 present = (ld3(s.fv) - ld3(s.vy) * sx) - ld3(s.vz) * sy;
  d3 vec = present - old;
  const double d2 = dot(vec, vec);
  if (d2 < 100) vec = vec * (100.0 / d2);
  uint32_t rgb = 0;
  int cls = BH_HIT_NONE, st = 0;
  d3 q = mk3(__builtin_nan(""), __builtin_nan(""), __builtin_nan(""));
  for (int k = 0; k < s.steps; ++k) {
    ++st;
    bool found = false;
    int kind = -1, which = -1;
    double best = 0;
    d3 c;
    for (int i = 0; i < s.n_rects; ++i) {
      if (rect_collide(s.rv[i], old, present, &c)) {
        const d3 d = old - c;
        const double kk = dot(d, d);
        if (!found || kk < best) {
          best = kk;
          found = true;
          kind = 0;
          which = i;
          q = c;
        }
      }
    }
    for (int i = 0; i < s.n_planes; ++i) {
      if (plane_collide(s, i, old, present, &c)) {
        const d3 d = old - c;
        const double kk = dot(d, d);
        if (!found || kk < best) {
          best = kk;
          found = true;
          kind = 1;
          which = i;
          q = c;
        }
      }
    }
    if (found) {
      cls = BH_HIT_OBJECT;
      rgb = kind == 0 ? rect_color(s, which, q) : plane_color(s, which, q);
      break;
    }
    const d3 np = present + vec;
    old = present;
    present = np;
  }
  if (out.rgb) {
    out.rgb[3 * idx + 0] = (uint8_t)(rgb & 0xff);
    out.rgb[3 * idx + 1] = (uint8_t)((rgb >> 8) & 0xff);
    out.rgb[3 * idx + 2] = (uint8_t)((rgb >> 16) & 0xff);
  }
  if (out.hit) out.hit[idx] = (uint8_t)cls;
  if (out.phi) out.phi[idx] = __builtin_nan("");
  if (out.hitpt) {
    out.hitpt[3 * idx + 0] = q.x;
    out.hitpt[3 * idx + 1] = q.y;
    out.hitpt[3 * idx + 2] = q.z;
  }
  if (out.steps) out.steps[idx] = st;
}

}  // namespace

hipError_t launch_flat(const DevFlat& s, const DevOut& out, hipStream_t stream) {
  dim3 grid((s.W + 15) / 16, (s.H + 15) / 16);
  hipLaunchKernelGGL(flat_kernel, grid, dim3(256), 0, stream, s, out);
  return hipGetLastError();
}

}  // namespace bh

// bh_fp64.h — the reference march in FP64, one lane per ray, in the
// reference's operation order (the library is compiled with
// -ffp-contract=off, so every product and sum rounds once, as written).
//
// Device restatement of include/blackhole/blackhole_solution_test.cc:161-298
// and of the objects it calls (blackhole_solution.h:24-88,
// object_manager.h:69-85, vector_object.h:163-179,328-347, camera.h:55-63).
// OpenCV's Vec/Matx arithmetic is written out: normalize = v * (1/|v|),
// Vec / s = Vec * (1/s), 3x3 inverse = adjugate * (1/det), dot products and
// matrix rows summed left to right from zero.
#pragma once

#include "bh_internal.h"

namespace bh {

struct d3 {
  double x, y, z;
};

__device__ __forceinline__ d3 mk3(double a, double b, double c) { return d3{a, b, c}; }
__device__ __forceinline__ d3 ld3(const double* p) { return d3{p[0], p[1], p[2]}; }
__device__ __forceinline__ d3 operator+(d3 a, d3 b) { return d3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ d3 operator-(d3 a, d3 b) { return d3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3 operator*(d3 a, double s) { return d3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double dot(d3 a, d3 b) {
  double s = 0;
  s += a.x * b.x;
  s += a.y * b.y;
  s += a.z * b.z;
  return s;
}
__device__ __forceinline__ d3 cross(d3 a, d3 b) {
  return d3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ d3 normalize(d3 a) {
  double nv = sqrt(dot(a, a));
  return a * (nv != 0.0 ? 1. / nv : 0.);
}

// Camera::PixelVector(x, y, fv) (camera.h:55-59)
__device__ __forceinline__ d3 pixel_point(const DevScene& s, int x, int y) {
  double sx = s.W / 2.0 - x;
  double sy = s.H / 2.0 - y;
  return (ld3(s.fv) - ld3(s.vy) * sx) - ld3(s.vz) * sy;
}

// StaticBlackhole::G (blackhole_solution.h:27-29)
__device__ __forceinline__ double G64(double u, double b, double M) {
  return (u * u * (2.0 * M * u - 1.0)) + (1.0 / (b * b));
}

// StaticBlackhole::SolveG (blackhole_solution.h:35-53); l0 = cbrt(eps_double)
// is computed on the host.  BH_ROOT_EXACT bisects until the bracket stops
// shrinking.  G is strictly decreasing on (0, 1/(3M)), so the decisions are
// those of the reference; 1/(b*b) is hoisted (same value every iteration).
__device__ inline double solve_g64(double b, double M, double l0, int root) {
  double l = 0.0 + l0;
  double r = 1.0 / (3.0 * M);
  const double ib2 = 1.0 / (b * b);
  const int count = root == BH_ROOT_EXACT ? 200 : 20;
  for (int i = 0; i < count; ++i) {
    double mid = (l + r) / 2.0;
    if (root == BH_ROOT_EXACT && (mid == l || mid == r)) break;
    double g = (mid * mid * (2.0 * M * mid - 1.0)) + ib2;
    if (g > 0.0)
      l = mid;
    else
      r = mid;
  }
  return l;
}

// Orbit-plane frame of one pixel (blackhole_solution_test.cc:167-183).
struct Frame64 {
  double R[9];  // [zv yv xv] as columns
  double b, r0, phi0, u0, u_end, du, du_h;
  bool captured, degenerate;
};

__device__ inline void inv3(const double* a, double* b, bool* ok) {
#define A(i, j) a[3 * (i) + (j)]
#define Bm(i, j) b[3 * (i) + (j)]
  double d = A(0, 0) * (A(1, 1) * A(2, 2) - A(2, 1) * A(1, 2)) -
             A(0, 1) * (A(1, 0) * A(2, 2) - A(2, 0) * A(1, 2)) +
             A(0, 2) * (A(1, 0) * A(2, 1) - A(2, 0) * A(1, 1));
  if (d == 0) {
    for (int i = 0; i < 9; ++i) b[i] = 0;
    *ok = false;
    return;
  }
  d = 1 / d;
  Bm(0, 0) = (A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1)) * d;
  Bm(0, 1) = (A(0, 2) * A(2, 1) - A(0, 1) * A(2, 2)) * d;
  Bm(0, 2) = (A(0, 1) * A(1, 2) - A(0, 2) * A(1, 1)) * d;
  Bm(1, 0) = (A(1, 2) * A(2, 0) - A(1, 0) * A(2, 2)) * d;
  Bm(1, 1) = (A(0, 0) * A(2, 2) - A(0, 2) * A(2, 0)) * d;
  Bm(1, 2) = (A(0, 2) * A(1, 0) - A(0, 0) * A(1, 2)) * d;
  Bm(2, 0) = (A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0)) * d;
  Bm(2, 1) = (A(0, 1) * A(2, 0) - A(0, 0) * A(2, 1)) * d;
  Bm(2, 2) = (A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0)) * d;
  *ok = true;
#undef A
#undef Bm
}

__device__ __forceinline__ d3 matrow(const double* m, d3 v) {
  double r[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double s = 0;
    s += m[3 * i + 0] * v.x;
    s += m[3 * i + 1] * v.y;
    s += m[3 * i + 2] * v.z;
    r[i] = s;
  }
  return d3{r[0], r[1], r[2]};
}

// Orbit-plane axes only (yv, xv), for callers that keep their own march.
__device__ inline void orbit_axes64(const DevScene& s, int x, int y, d3* yv_out, d3* xv_out,
                                    d3* zv_out) {
  d3 B = ld3(s.B);
  d3 pv = pixel_point(s, x, y) - B;
  d3 F = ld3(s.C) - B;
  d3 yv = normalize(F - pv);
  d3 zv = normalize(cross(pv, yv));
  if (isnan(zv.x)) zv = mk3(1, 0, 0);
  d3 xv = normalize(cross(yv, zv));
  *yv_out = yv;
  *xv_out = xv;
  if (zv_out) *zv_out = zv;
}

__device__ inline void frame64(const DevScene& s, int x, int y, Frame64& f) {
  d3 yv, xv, zv;
  orbit_axes64(s, x, y, &yv, &xv, &zv);
  d3 F = ld3(s.C) - ld3(s.B);
  f.R[0] = zv.x; f.R[1] = yv.x; f.R[2] = xv.x;
  f.R[3] = zv.y; f.R[4] = yv.y; f.R[5] = xv.y;
  f.R[6] = zv.z; f.R[7] = yv.z; f.R[8] = xv.z;
  double Ri[9];
  bool ok;
  inv3(f.R, Ri, &ok);
  d3 Fp = matrow(Ri, F);
  f.degenerate = !ok || dot(zv, zv) == 0.0;
  const double M = s.M;
  double b = Fp.z;
  double sol = -1;
  if (b >= s.bc) sol = solve_g64(b, M, s.l0, s.root);
  f.captured = sol < 0;
  f.u_end = b < s.bc ? 1.0 / (3 * M) : sol;
  f.b = b;
  f.r0 = sqrt(dot(Fp, Fp));
  f.phi0 = atan(Fp.z / Fp.y);
  f.u0 = 1. / f.r0;
  f.du = (f.u_end - f.u0) * (1.0 / s.nstep);
  f.du_h = f.du / 2.;
}

// Segment vs sphere: StaticBlackhole::Collide (blackhole_solution.h:65-88).
__device__ inline bool seg_sphere64(d3 q1, d3 q2, d3 c, double R, d3* inter) {
  d3 p1 = q1 - c;
  d3 p2 = q2 - c;
  double d1 = sqrt(dot(p1, p1));
  double d2 = sqrt(dot(p2, p2));
  if (d1 < R && d2 < R) return false;
  d3 Q1 = q1 - c;
  d3 Q2 = q2 - c;
  d3 Q = Q2 - Q1;
  double qd = dot(Q, Q1);
  double cc = qd * qd - dot(Q, Q) * (dot(Q1, Q1) - R * R);
  if (cc <= 0) return false;
  double t = (((double)1) / dot(Q, Q)) * (-dot(Q, Q1) - sqrt(cc));
  if (t <= 0 || t >= 1) return false;
  *inter = (Q1 + Q * t) + c;
  return true;
}

// Annulus::Collide (vector_object.h:328-347).
__device__ inline bool annulus64(const DevScene& s, d3 q1, d3 q2, d3* inter) {
  d3 c0 = ld3(s.dv0);
  d3 n = ld3(s.dn);
  d3 p1 = q1 - c0;
  d3 p2 = q2 - c0;
  if (dot(n, p1) * dot(n, p2) >= 0) return false;
  double t1 = dot(n, p1);
  double t2 = dot(n, p2);
  d3 num = p1 * fabs(t2) + p2 * fabs(t1);
  d3 c = num * (1. / (fabs(t1) + fabs(t2)));
  if (sqrt(dot(c, c)) > s.r_out) return false;
  if (sqrt(dot(c, c)) < s.r_in) return false;
  *inter = c + c0;
  return true;
}

// ObjectManager::FindCollision (object_manager.h:69-85): nearest hit by
// |p1 - q|^2, ties keep the object tested first (hole, disc, spheres).
struct Hit64 {
  int cls;
  int index;
  d3 q;
};

// (sphere i is tested when bit i of smask is set; the hole and the disc always)
__device__ inline bool find_collision64(const DevScene& s, d3 p1, d3 p2, Hit64* h,
                                        uint32_t smask = ~0u) {
  bool found = false;
  double best = 0;
  d3 q;
  if (seg_sphere64(p1, p2, ld3(s.B), s.Rh, &q)) {
    d3 d = p1 - q;
    best = dot(d, d);
    found = true;
    h->cls = BH_HIT_HORIZON;
    h->index = -1;
    h->q = q;
  }
  if (s.has_disc && annulus64(s, p1, p2, &q)) {
    d3 d = p1 - q;
    double k = dot(d, d);
    if (!found || k < best) {
      best = k;
      found = true;
      h->cls = BH_HIT_DISC;
      h->index = -1;
      h->q = q;
    }
  }
  for (int i = 0; i < s.n_spheres; ++i) {
    if (!((smask >> i) & 1u)) continue;
    if (seg_sphere64(p1, p2, ld3(s.sc[i]), s.sr[i], &q)) {
      d3 d = p1 - q;
      double k = dot(d, d);
      if (!found || k < best) {
        best = k;
        found = true;
        h->cls = BH_HIT_OBJECT;
        h->index = i;
        h->q = q;
      }
    }
  }
  return found;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// Truncating texel index in [0, n-1]; NaN maps to 0 (the reference's (int)
// cast of an out-of-range value is undefined and reads out of bounds).
__device__ __forceinline__ int tex_index(double v, int n) {
  if (!(v >= 0.0)) return 0;
  if (v >= (double)(n - 1)) return n - 1;
  return (int)v;
}

__device__ __forceinline__ uint32_t texel(const DevTex& t, int col, int row) {
  return t.data[(size_t)row * t.cols + col];
}

// Bilinear sample at continuous texel coordinates (col a, row b), centres at
// +0.5, clamped to the edge, rounded to u8 per channel.
__device__ inline uint32_t bilinear64(const DevTex& t, double a, double b) {
  double x = a - 0.5, y = b - 0.5;
  if (!(x > -2.0)) x = -2.0;
  if (!(y > -2.0)) y = -2.0;
  if (x > (double)t.cols) x = (double)t.cols;
  if (y > (double)t.rows) y = (double)t.rows;
  double fx = floor(x), fy = floor(y);
  double tx = x - fx, ty = y - fy;
  int x0 = (int)fx, y0 = (int)fy;
  int x1 = clampi(x0 + 1, 0, t.cols - 1), y1 = clampi(y0 + 1, 0, t.rows - 1);
  x0 = clampi(x0, 0, t.cols - 1);
  y0 = clampi(y0, 0, t.rows - 1);
  uint32_t t00 = texel(t, x0, y0), t01 = texel(t, x1, y0);
  uint32_t t10 = texel(t, x0, y1), t11 = texel(t, x1, y1);
  uint32_t out = 0;
  for (int ch = 0; ch < 3; ++ch) {
    double v00 = (double)((t00 >> (8 * ch)) & 0xff), v01 = (double)((t01 >> (8 * ch)) & 0xff);
    double v10 = (double)((t10 >> (8 * ch)) & 0xff), v11 = (double)((t11 >> (8 * ch)) & 0xff);
    double top = v00 + (v01 - v00) * tx;
    double bot = v10 + (v11 - v10) * tx;
    double v = top + (bot - top) * ty;
    int iv = clampi((int)floor(v + 0.5), 0, 255);
    out |= (uint32_t)iv << (8 * ch);
  }
  return out;
}

// Rectangle::color (vector_object.h:163-179), indices clamped into the image.
__device__ inline uint32_t disc_color64(const DevScene& s, d3 p) {
  const DevTex& t = s.disc_tex;
  d3 v1 = ld3(s.dv1);
  d3 s1 = ld3(s.dv2) - v1;
  d3 s2 = ld3(s.dv4) - v1;
  d3 v = p - v1;
  double r = sqrt(dot(v, v));
  double theta = acos(dot(s1, v) / (sqrt(dot(s1, s1)) * r));
  theta = isnan(theta) ? 0 : theta;
  double a = (r * cos(theta) / sqrt(dot(s1, s1))) * t.cols;
  double b = (r * sin(theta) / sqrt(dot(s2, s2))) * t.rows;
  if (s.filter == BH_FILTER_BILINEAR) return bilinear64(t, a, b);
  return texel(t, tex_index(a, t.cols), tex_index(b, t.rows));
}

// Equirectangular lookup of a unit direction (extension).
__device__ inline uint32_t equirect64(const DevTex& t, d3 d) {
  const double kPi = 3.141592653589793238462643383279;
  double z = d.z < -1.0 ? -1.0 : (d.z > 1.0 ? 1.0 : d.z);
  double lon = atan2(d.y, d.x);
  double lat = asin(z);
  double c = (lon + kPi) * (1.0 / (2.0 * kPi)) * t.cols;
  double r = (kPi / 2.0 - lat) * (1.0 / kPi) * t.rows;
  return texel(t, tex_index(floor(c), t.cols), tex_index(floor(r), t.rows));
}

__device__ inline uint32_t hit_color64(const DevScene& s, const Hit64& h) {
  if (h.cls == BH_HIT_DISC) {
    if (!s.disc_tex.data) return 0;
    return disc_color64(s, h.q);
  }
  if (h.cls == BH_HIT_OBJECT) {
    const DevTex& t = s.stex[h.index];
    if (!t.data) return 0;
    d3 n = (h.q - ld3(s.sc[h.index])) * (1.0 / s.sr[h.index]);
    return equirect64(t, n);
  }
  return 0;  // StaticBlackhole::color is black (blackhole_solution.h:61-63)
}

struct Pixel64 {
  uint32_t rgb;
  int cls;
  double phi;
  d3 q;
  int steps;
};

// The spheres (a bit each) that a chord of the march in frame f can meet.
// Every point of the march is B + yv r cos(phi) + xv r sin(phi) (matrow with
// the first component 0) or the camera or B, all in the plane through B
// spanned by yv and xv (R's second and third columns) up to rounding of
// ~1e-15 (r + |B|), with r <= reach when du > 0 (u only grows from u0, or
// the escape march stops short of r_esc).  A sphere whose centre lies farther
// from that plane than R + tol misses every chord in exact arithmetic by
// D^2 - R^2 >= 1e-12 L^2, L bounding |q - c| over the march's points q, far
// above the rounding of seg_sphere64's discriminant (~1e-15 |Q|^2 L^2 against
// its exact |Q|^2 (R^2 - D^2)), so its test would return false on every
// chord: skipping it changes no outcome.  With du <= 0 (or not finite) the
// points are not bounded and every sphere is kept.
__device__ inline uint32_t plane_spheres64(const DevScene& s, const Frame64& f) {
  if (!(f.du > 0.0) || !(f.r0 < 1e300)) return ~0u;
  const d3 yv = mk3(f.R[1], f.R[4], f.R[7]), xv = mk3(f.R[2], f.R[5], f.R[8]);
  const d3 n = normalize(cross(yv, xv));
  if (!(dot(n, n) > 0.5)) return ~0u;
  const d3 B = ld3(s.B);
  const double reach = s.escape ? fmax(f.r0, s.r_esc) : f.r0;
  const double bs = sqrt(dot(B, B));
  uint32_t m = 0u;
  for (int i = 0; i < s.n_spheres; ++i) {
    const d3 c = ld3(s.sc[i]) - B;
    const double R = s.sr[i];
    const double tol = 1e-6 * (sqrt(dot(c, c)) + reach + R) + 1e-9 * bs;
    if (!(fabs(dot(n, c)) > R + tol)) m |= 1u << i;
  }
  return m;
}

// The whole march of one pixel (blackhole_solution_test.cc:167-298) with the
// escape extension.  Returns through px.
__device__ inline void trace64(const DevScene& s, int x, int y, Pixel64& px) {
  Frame64 f;
  frame64(s, x, y, f);
  const double nan = __builtin_nan("");
  px.q = mk3(nan, nan, nan);
  if (f.degenerate) {
    px.rgb = 0;
    px.cls = BH_HIT_HORIZON;
    px.phi = nan;
    px.steps = 0;
    return;
  }
  const d3 B = ld3(s.B);
  const double M = s.M;
  double u = f.u0, w_prev = 0, phi = f.phi0;
  d3 P_prev = ld3(s.C), P_prev2 = P_prev;
  int steps = 0;
  Hit64 h;
  bool hit = false;
  const int nstep = s.nstep;
  const uint32_t smask = plane_spheres64(s, f);

  // One step at the current u (the caller advances u).
  auto step = [&]() -> bool {
    double w = 1.0 / sqrt(G64(u, f.b, M));
    phi += (w_prev + w) * f.du_h;
    double r = 1. / u;
    d3 lv = mk3(0, r * cos(phi), r * sin(phi));
    d3 P = matrow(f.R, lv) + B;
    w_prev = w;
    ++steps;
    if (find_collision64(s, P_prev, P, &h, smask)) return true;
    P_prev2 = P_prev;
    P_prev = P;
    return false;
  };

  for (int i = 0; i < nstep - 1 && !hit; ++i) {
    u += f.du;
    hit = step();
  }
  if (!hit) {
    u += f.du * 0.9;
    hit = step();
  }
  if (!hit && f.captured) {
    if (find_collision64(s, P_prev, B, &h, smask)) {
      hit = true;
    } else {
      px.rgb = 0;
      px.cls = BH_HIT_NONE;
      px.phi = phi;
      px.steps = steps;
      return;
    }
  }
  if (!hit) {
    for (int i = 0; i < nstep - 1 && !hit; ++i) {
      u -= f.du;
      hit = step();
    }
  }
  if (!hit && s.escape) {
    while (steps < s.step_cap && !hit) {
      double un = u - f.du;
      if (un * s.r_esc <= 1.0) break;
      u = un;
      hit = step();
    }
  }
  if (hit) {
    px.rgb = hit_color64(s, h);
    px.cls = h.cls;
    px.q = h.q;
  } else {
    px.rgb = 0;
    px.cls = BH_HIT_NONE;
    if (s.escape && s.has_bg) px.rgb = equirect64(s.bg, normalize(P_prev - P_prev2));
  }
  px.phi = phi;
  px.steps = steps;
}

// The reference march of pixel (x, y) replayed for a known outcome: the march
// ends at step n with a disc hit (cls BH_HIT_DISC) or with no hit (escape
// background along the last chord).  The swept angle of steps 1..n follows
// trace64's loop to the bit (the same updates of u, w and phi in the same
// order); only the last chord's two end points are placed, and the colour is
// taken as trace64 takes it, without the per-step positions and object tests.
// Returns false when the last chord does not cross the disc after all (the
// outcome was not the reference's); px is then unset.
__device__ inline bool replay64(const DevScene& s, int x, int y, int n, int cls, Pixel64& px) {
  Frame64 f;
  frame64(s, x, y, f);
  if (f.degenerate || n < 1) return false;
  const d3 B = ld3(s.B);
  const double M = s.M;
  const int nstep = s.nstep;
  double u = f.u0, w_prev = 0, phi = f.phi0;
  double u_a = u, phi_a = phi;  // state at the start of the last chord
  for (int k = 1; k <= n; ++k) {
    u_a = u;
    phi_a = phi;
    if (k < nstep)
      u += f.du;
    else if (k == nstep)
      u += f.du * 0.9;
    else
      u -= f.du;
    double w = 1.0 / sqrt(G64(u, f.b, M));
    phi += (w_prev + w) * f.du_h;
    w_prev = w;
  }
  auto place = [&](double uu, double ph) {
    double r = 1. / uu;
    d3 lv = mk3(0, r * cos(ph), r * sin(ph));
    return matrow(f.R, lv) + B;
  };
  const d3 Pb = place(u, phi);
  const d3 Pa = n == 1 ? ld3(s.C) : place(u_a, phi_a);
  const double nan = __builtin_nan("");
  px.phi = phi;
  px.steps = n;
  if (cls == BH_HIT_DISC) {
    Hit64 h;
    h.cls = BH_HIT_DISC;
    h.index = -1;
    if (!s.has_disc || !annulus64(s, Pa, Pb, &h.q)) return false;
    px.rgb = hit_color64(s, h);
    px.cls = BH_HIT_DISC;
    px.q = h.q;
    return true;
  }
  if (cls != BH_HIT_NONE) return false;
  px.rgb = 0;
  px.cls = BH_HIT_NONE;
  px.q = mk3(nan, nan, nan);
  if (s.escape && s.has_bg) px.rgb = equirect64(s.bg, normalize(Pb - Pa));
  return true;
}

__device__ inline void store_pixel64(const DevOut& o, size_t idx, const Pixel64& p) {
  if (o.rgb) {
    o.rgb[3 * idx + 0] = (uint8_t)(p.rgb & 0xff);
    o.rgb[3 * idx + 1] = (uint8_t)((p.rgb >> 8) & 0xff);
    o.rgb[3 * idx + 2] = (uint8_t)((p.rgb >> 16) & 0xff);
  }
  if (o.hit) o.hit[idx] = (uint8_t)p.cls;
  if (o.phi) o.phi[idx] = p.phi;
  if (o.hitpt) {
    o.hitpt[3 * idx + 0] = p.q.x;
    o.hitpt[3 * idx + 1] = p.q.y;
    o.hitpt[3 * idx + 2] = p.q.z;
  }
  if (o.steps) o.steps[idx] = p.steps;
}

}  // namespace bh

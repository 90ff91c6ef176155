// bh_diff.hip — differentiable render of a hole + disc scene and its adjoint
// with respect to (r_s, inclination, camera distance, disc inner radius).
//
// The march is the reference rule (include/blackhole/blackhole_solution_test.cc:
// 161-298: orbit-plane frame, trapezoid sum with w_prev = 0 at the first step,
// 0.9 du final approach, captured chord to the centre, outbound sweep) with
// three smooth replacements that make the image a differentiable function of
// the parameters: the exact turning point instead of the 20-step bisection,
// bilinear texels instead of nearest, and a soft disc edge
//   alpha = S((rho - r_in)/w) S((r_out - rho)/w),  S(x) = smoothstep(clamp(x + 1/2)),
// composited front to back over the disc crossings of a ray until the horizon
// or full coverage.  One lane per ray.
//
// Adjoint (reverse mode, fused with the forward in one kernel):
//   pass 1   forward march; each disc crossing records its chord's angles
//            (FP32: psi_j alone, in fixed point), the transmittance T_k and
//            four prefix sums in LDS and sets bit j of a chord mask;
//   pass 2   reverse over the crossing records only.  The angle psi_m =
//            h sum_{i <= m} (w_{i-1} + w_i) makes w_i enter every later angle,
//            so the adjoint of w_i is h (A_i + A_{i+1}) with A_i the sum of
//            the angle adjoints injected at m >= i.  Angle adjoints arise only
//            at the two ends of a crossing chord, so A is piecewise constant
//            and sum_i (A_i + A_{i+1}) X_i = sum_m g_m (S(m) + S(m - 1)) for
//            the prefix sums S of each per-step factor X_i of dw_i
//            (w^3 G', kappa w^3 G', w^3 u^3, w^3; w itself gives psi_m / h).
//            Pass 1 carries the four prefix sums and records them at each
//            crossing; S(m - 1), S(m - 2) follow by recomputing X_m, X_{m-1}
//            in closed form.  No step of the march is revisited.  Compositing
//            adjoint with the behind colour B built back to front; crossing
//            point, coverage and bilinear texel adjoints give P_bar, hence
//            (psi_bar, u_bar, e1_bar, e2_bar);
//   set-up   turning point by implicit differentiation of G(l) = 0, then the
//            orbit-plane frame in forward mode (tangents: inclination,
//            distance) contracted with the set-up adjoints.
// The forward visits a chord's geometry only where something can happen on
// it: when the orbit angle passes a crossing of the disc plane (closed form
// per ray) or when a bound on the chord's distance from the centre does not
// exclude the horizon; every other chord provably has neither.
// Per-block partial sums are reduced in a fixed order by a second kernel, so
// the gradient is deterministic.
#include "bh_internal.h"
#include "bh_math.h"

#include <type_traits>

namespace bh {
namespace {

constexpr int kTile = 16;
constexpr int kThreads = kTile * kTile;
constexpr int kMaxRec = BH_DIFF_MAX_CROSSINGS;
constexpr int kPartial = 5;  // r_s, incl, dist, r_in, overflowed rays

// ---------------------------------------------------------------------------
// Forward-mode FP64 number: value and derivatives along inclination, distance
// ---------------------------------------------------------------------------
struct D2 {
  double v, i, d;
};
__device__ __forceinline__ D2 operator+(D2 a, D2 b) { return {a.v + b.v, a.i + b.i, a.d + b.d}; }
__device__ __forceinline__ D2 operator-(D2 a, D2 b) { return {a.v - b.v, a.i - b.i, a.d - b.d}; }
__device__ __forceinline__ D2 operator*(D2 a, D2 b) {
  return {a.v * b.v, a.i * b.v + a.v * b.i, a.d * b.v + a.v * b.d};
}
__device__ __forceinline__ double val(double a) { return a; }
__device__ __forceinline__ double val(D2 a) { return a.v; }
__device__ __forceinline__ double rsqrt_n(double a) { return rsq<double>(a); }
__device__ __forceinline__ D2 rsqrt_n(D2 a) {
  const double y = rsq<double>(a.v);
  const double k = -0.5 * y * y * y;
  return {y, k * a.i, k * a.d};
}
__device__ __forceinline__ double sqrt_n(double a) { return sqrt(a); }
__device__ __forceinline__ D2 sqrt_n(D2 a) {
  const double y = sqrt(a.v);
  const double k = 0.5 / y;
  return {y, k * a.i, k * a.d};
}
__device__ __forceinline__ double rcp_n(double a) { return 1.0 / a; }
__device__ __forceinline__ D2 rcp_n(D2 a) {
  const double y = 1.0 / a.v;
  const double k = -y * y;
  return {y, k * a.i, k * a.d};
}
__device__ __forceinline__ void lift(double v, double ti, double td, double& o) { o = v; }
__device__ __forceinline__ void lift(double v, double ti, double td, D2& o) { o = {v, ti, td}; }

template <typename N> __device__ __forceinline__ N dot3(const N* a, const N* b) {
  N s = a[0] * b[0];
  s = s + a[1] * b[1];
  return s + a[2] * b[2];
}
template <typename N> __device__ __forceinline__ void cross3(const N* a, const N* b, N* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

// Orbit-plane frame of pixel (px, py) (blackhole_solution_test.cc:167-193):
// the ray is the line from the camera F = C through the pixel point
// pv = f axis_x - axis_y (W/2 - px) - axis_z (H/2 - py) (camera.h:55-63, the
// hole at the origin); yv = N(F - pv), zv = N(pv x yv), xv = yv x zv;
// b = xv.F, Y0 = yv.F, r0 = |F|.  The reference puts the orbit point at angle
// phi and radius r at r (cos(phi) yv + sin(phi) xv) starting from
// phi0 = atan(b / Y0); here the angle psi = phi - phi0 is measured from the
// camera direction e1 = F / r0 with e2 = (Y0 xv - b yv) / r0, the same point
// r (cos(psi) e1 + sin(psi) e2) without an arctangent.  (Y0 > 0 for a camera
// looking at the hole: F.(F - pv) = dist^2 + dist f, so atan(b/Y0) is the
// polar angle of F and psi = 0 is the camera.)
template <typename N> struct Frame {
  N e1[3], e2[3];
  N b, r0;
  bool degenerate;
};

template <typename N>
__device__ inline void diff_frame(const DiffScene& s, int px, int py, Frame<N>& f) {
  const double sx = s.W / 2.0 - px, sy = s.H / 2.0 - py;
  N pv[3], F[3], dv[3], yv[3], cv[3], zv[3], xv[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    lift(s.fx[k] - s.ay[k] * sx - s.az[k] * sy, s.fx_i[k] - s.ay_i[k] * sx - s.az_i[k] * sy,
         0.0, pv[k]);
    lift(s.C[k], s.C_i[k], s.C_d[k], F[k]);
    dv[k] = F[k] - pv[k];
  }
  const N ny = rsqrt_n(dot3(dv, dv));
#pragma unroll
  for (int k = 0; k < 3; ++k) yv[k] = dv[k] * ny;
  cross3(pv, yv, cv);
  const N cc = dot3(cv, cv);
  f.degenerate = !(val(cc) > 0.0);
  const N nz = rsqrt_n(f.degenerate ? dot3(dv, dv) : cc);
#pragma unroll
  for (int k = 0; k < 3; ++k) zv[k] = cv[k] * nz;
  cross3(yv, zv, xv);
  f.b = dot3(xv, F);
  const N Y0 = dot3(yv, F);
  f.r0 = sqrt_n(dot3(F, F));
  const N ir0 = rcp_n(f.r0);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    f.e1[k] = F[k] * ir0;
    f.e2[k] = (xv[k] * Y0 - yv[k] * f.b) * ir0;
  }
}

// Inbound turning point of G(u) = 2M u^3 - u^2 + 1/b^2 for b >= b_c.  With
// y = 3 M u and q = 27 M^2 / b^2, G = 0 reads 3y^2 - 2y^3 = q; the root
// starts from s P(s), s = sqrt(q), for q <= 1/2 and from 1 - s P(s),
// s = sqrt(1 - q), above (P: the degree-9 fit of y/s on [0, sqrt(1/2)],
// good to 3e-7 of min(y, 1 - y) in FP32) and is polished by two FP64 Newton
// steps with an FP32 1/G' while G' < 0: within 1e-15 of the root outside
// the photon-sphere band, and as close as the double root's conditioning
// allows within it.
__device__ inline double turning_point(double ib2, double M, double Uc) {
  const double q = fmin(1.0, 27.0 * M * M * ib2);
  const float qf = (float)q, qc = (float)(1.0 - q);
  const float sq = __builtin_amdgcn_sqrtf(fmaxf(fminf(qf, qc), 0.f));
  float p = 0.68455708f;
  p = __builtin_fmaf(p, sq, -1.7155893f);
  p = __builtin_fmaf(p, sq, 1.9122510f);
  p = __builtin_fmaf(p, sq, -1.1384754f);
  p = __builtin_fmaf(p, sq, 0.42832887f);
  p = __builtin_fmaf(p, sq, -0.063991360f);
  p = __builtin_fmaf(p, sq, 0.043269463f);
  p = __builtin_fmaf(p, sq, 0.052839678f);
  p = __builtin_fmaf(p, sq, 0.11112535f);
  p = __builtin_fmaf(p, sq, 0.57735020f);
  const double z = (double)(sq * p);
  double u = Uc * (qf <= qc ? z : 1.0 - z);
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const double g = __builtin_fma(u * u, __builtin_fma(2.0 * M, u, -1.0), ib2);
    const double gp = u * __builtin_fma(6.0 * M, u, -2.0);
    const double ig = (double)__builtin_amdgcn_rcpf((float)gp);
    if (gp < 0.0) u = __builtin_fma(-g, ig, u);
  }
  return u;
}

// Per-ray march constants in the march precision T.
template <typename T> struct DRay {
  T e1[3], e2[3], C[3];
  T u0, du, h;
  T c0, c1, c2, c3;  // G(l + s) = c0 + s (c1 + s (c2 + s c3))
  // disc-plane crossings: z(psi) = 0 at psi = a0 + k pi; an event is due
  // once psi >= (crossing - guard)
  double a0, guard;
  bool captured, degenerate;
};

// FP64 set-up values shared by the march and the adjoint's epilogue.  The
// orbit angle psi is accumulated in FP64 in both precisions.
struct Setup64 {
  Frame<double> f;
  double ib2, l, u0, du;
  bool captured;
};

__device__ inline void diff_setup(const DiffScene& s, int px, int py, Setup64& r) {
  diff_frame(s, px, py, r.f);
  const double b = r.f.b;
  r.ib2 = 1.0 / (b * b);
  r.captured = b < s.bc;
  r.l = r.captured ? s.Uc : turning_point(r.ib2, s.M, s.Uc);
  r.u0 = 1.0 / r.f.r0;
  r.du = (r.l - r.u0) * (1.0 / s.nstep);
}

template <typename T>
__device__ inline void to_ray(const DiffScene& s, const Setup64& r, DRay<T>& d) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    d.e1[k] = (T)r.f.e1[k];
    d.e2[k] = (T)r.f.e2[k];
    d.C[k] = (T)s.C[k];
  }
  d.u0 = (T)r.u0;
  d.du = (T)r.du;
  d.h = d.du * (T)0.5;
  // z(psi) ~ cos(psi) e1_z + sin(psi) e2_z vanishes at atan2(-e1_z, e2_z) + k pi
  const float ez1 = (float)r.f.e1[2], ez2 = (float)r.f.e2[2];
  const float amp2 = ez1 * ez1 + ez2 * ez2;
  if (amp2 > 0.f) {
    double a = (double)atan2_fast(-ez1, ez2);
    if (a < 0.0) a += kPiD;
    d.a0 = a;
    d.guard = 4e-6 + 4e-7 * (double)__builtin_amdgcn_rsqf(amp2);
  } else {  // orbit in the disc plane: z = 0 throughout, never a crossing
    d.a0 = 1e300;
    d.guard = 0.0;
  }
  const double l = r.l, M = s.M;
  d.c0 = (T)__builtin_fma(l * l, __builtin_fma(2.0 * M, l, -1.0), r.ib2);
  d.c1 = (T)(l * __builtin_fma(6.0 * M, l, -2.0));
  d.c2 = (T)__builtin_fma(6.0 * M, l, -1.0);
  d.c3 = (T)(2.0 * M);
  d.captured = r.captured;
  d.degenerate = r.f.degenerate;
}

// Step j (1-based) of the reference rule: u_j = u0 + kappa_j du with kappa
// = j inbound (j < n), n - 0.1 at the final approach (j = n) and n - 0.1 - i
// on the i-th outbound step; sigma_j = n - kappa_j so that u_j = l - sigma_j du.
template <typename T> __device__ __forceinline__ T kappa(int j, int n) {
  return j < n ? (T)j : (T)(2 * n - j) - (T)0.1;
}
template <typename T> __device__ __forceinline__ T sigma(int j, int n) {
  return j < n ? (T)(n - j) : (T)(j - n) + (T)0.1;
}

template <typename T> __device__ __forceinline__ T g_at(const DRay<T>& d, T sj) {
  return fma_t(sj, fma_t(sj, fma_t(sj, d.c3, d.c2), d.c1), d.c0);
}

// sin/cos of an FP64 orbit angle: FP64 reduction, then (T = float) the
// hardware sine and cosine of the angle in revolutions (the revolution's
// FP32 rounding, ~2e-7 rad, and the hardware's 2.7e-7; the adjoint's
// gradients stay within 4e-5 of central differences), or the FP64 series
__device__ __forceinline__ void angle_sincos(double a, float* s, float* c) {
  const double rv = a * (0.5 * kInvPiD);
  const float f = (float)(rv - rint(rv));
  *s = __builtin_amdgcn_sinf(f);
  *c = __builtin_amdgcn_cosf(f);
}
__device__ __forceinline__ void angle_sincos(double a, double* s, double* c) {
  sincos_small64(a, s, c);
}

template <typename T>
__device__ __forceinline__ void orbit_point(const DRay<T>& d, double psi, T r, T* P, T* cph,
                                            T* sph) {
  T sn, cs;
  angle_sincos(psi, &sn, &cs);
#pragma unroll
  for (int k = 0; k < 3; ++k) P[k] = r * fma_t(cs, d.e1[k], sn * d.e2[k]);
  *cph = cs;
  *sph = sn;
}

// The first crossing threshold at or after psi: a0 + k pi - guard with
// a0 + k pi >= psi - guard.
template <typename T> __device__ __forceinline__ double next_crossing(const DRay<T>& d, double psi) {
  const double k = fmax(0.0, ceil((psi - d.guard - d.a0) * kInvPiD));
  return __builtin_fma(k, kPiD, d.a0) - d.guard;
}

// StaticBlackhole::Collide (blackhole_solution.h:65-88) for the horizon at the
// origin: the entry point's squared distance from P1 when the chord enters.
template <typename T>
__device__ inline bool seg_horizon(const T* P1, const T* P2, T R, T* dh) {
  const T R2 = R * R;
  const T d1 = P1[0] * P1[0] + P1[1] * P1[1] + P1[2] * P1[2];
  const T d2 = P2[0] * P2[0] + P2[1] * P2[1] + P2[2] * P2[2];
  if (d1 < R2 && d2 < R2) return false;
  const T Q[3] = {P2[0] - P1[0], P2[1] - P1[1], P2[2] - P1[2]};
  const T qd = Q[0] * P1[0] + Q[1] * P1[1] + Q[2] * P1[2];
  const T QQ = Q[0] * Q[0] + Q[1] * Q[1] + Q[2] * Q[2];
  const T cc = qd * qd - QQ * (d1 - R2);
  if (!(cc > (T)0)) return false;
  const T t = (-qd - sqrt_t(cc)) * rcp(QQ);
  if (t <= (T)0 || t >= (T)1) return false;
  *dh = t * t * QQ;
  return true;
}

template <typename T> __device__ __forceinline__ T sstep(T x) {
  T t = x + (T)0.5;
  t = t < (T)0 ? (T)0 : (t > (T)1 ? (T)1 : t);
  return t * t * ((T)3 - (T)2 * t);
}
template <typename T> __device__ __forceinline__ T sstep_d(T x) {
  const T t = x + (T)0.5;
  return (t > (T)0 && t < (T)1) ? (T)6 * t * ((T)1 - t) : (T)0;
}

// Soft coverage of the disc at radius rho and its derivatives.
template <typename T, bool GRAD>
__device__ inline T coverage(const DiffScene& s, T rho, T* drho, T* drin) {
  const T w = (T)s.edge;
  if (!(w > (T)0)) {
    if (GRAD) *drho = *drin = (T)0;
    return (rho > (T)s.r_out || rho < (T)s.r_in) ? (T)0 : (T)1;
  }
  const T iw = (T)(1.0 / s.edge);
  const T x1 = (rho - (T)s.r_in) * iw, x2 = ((T)s.r_out - rho) * iw;
  const T a1 = sstep(x1), a2 = sstep(x2);
  if (GRAD) {
    const T g1 = sstep_d(x1) * a2 * iw, g2 = a1 * sstep_d(x2) * iw;
    *drho = g1 - g2;
    *drin = -g1;
  }
  return a1 * a2;
}

// Bilinear texel of a disc point q (texel centres at +1/2, clamped indices):
// U = (half - q_x) cols / (2 half), V = |(q_y - half, q_z)| rows / (2 half),
// the smooth form of Rectangle::color (vector_object.h:159-179) on the
// square (+-half, +-half).  With GRAD, dq receives d(colour . cw)/dq.
template <typename T, bool GRAD>
__device__ inline void disc_texel(const DiffScene& s, const T* q, T* c, const T* cw, T* dq) {
  c[0] = c[1] = c[2] = (T)0;
  if (GRAD) dq[0] = dq[1] = dq[2] = (T)0;
  if (!s.tex.data) return;
  const T half = (T)s.half;
  const T U = (half - q[0]) * (T)s.uscale;
  const T vy = q[1] - half, vz = q[2];
  const T rv = sqrt_t(vy * vy + vz * vz);
  const T V = rv * (T)s.vscale;
  const T x = U - (T)0.5, y = V - (T)0.5;
  T fx = floor(x), fy = floor(y);
  const T tx = x - fx, ty = y - fy;
  const int cols = s.tex.cols, rows = s.tex.rows;
  fx = fx > (T)-1 ? (fx < (T)cols ? fx : (T)cols) : (T)-1;  // NaN maps to -1
  fy = fy > (T)-1 ? (fy < (T)rows ? fy : (T)rows) : (T)-1;
  const int ix = (int)fx, iy = (int)fy;
  const int x0 = min(max(ix, 0), cols - 1), x1 = min(max(ix + 1, 0), cols - 1);
  const int y0 = min(max(iy, 0), rows - 1), y1 = min(max(iy + 1, 0), rows - 1);
  const uint32_t* t = s.tex.data;
  const uint32_t t00 = t[(size_t)y0 * cols + x0], t01 = t[(size_t)y0 * cols + x1];
  const uint32_t t10 = t[(size_t)y1 * cols + x0], t11 = t[(size_t)y1 * cols + x1];
  T gU = (T)0, gV = (T)0;
#pragma unroll
  for (int ch = 0; ch < 3; ++ch) {
    const T v00 = (T)((t00 >> (8 * ch)) & 0xff), v01 = (T)((t01 >> (8 * ch)) & 0xff);
    const T v10 = (T)((t10 >> (8 * ch)) & 0xff), v11 = (T)((t11 >> (8 * ch)) & 0xff);
    const T top = v00 + (v01 - v00) * tx;
    const T bot = v10 + (v11 - v10) * tx;
    c[ch] = top + (bot - top) * ty;
    if (GRAD) {
      gU += cw[ch] * ((v01 - v00) * ((T)1 - ty) + (v11 - v10) * ty);
      gV += cw[ch] * (bot - top);
    }
  }
  if (GRAD) {
    dq[0] = -gU * (T)s.uscale;
    if (rv > (T)0) {
      const T k = gV * (T)s.vscale / rv;
      dq[1] = k * vy;
      dq[2] = k * vz;
    }
  }
}

// Disc crossing of the chord P1 -> P2 (Annulus::Collide, vector_object.h:
// 328-347: sign change across z = 0, crossing by linear interpolation on the
// chord) with the soft radius window; returns the coverage (0: no crossing).
template <typename T>
__device__ inline T disc_cross(const DiffScene& s, const T* P1, const T* P2, T* q, T* sfrac,
                               T* dd) {
  const T z1 = P1[2], z2 = P2[2];
  if (!(z1 * z2 < (T)0)) return (T)0;
  const T sf = z1 * rcp(z1 - z2);
  T D[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    D[k] = P2[k] - P1[k];
    q[k] = fma_t(sf, D[k], P1[k]);
  }
  const T rho = sqrt_t(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
  T g0, g1;
  const T al = coverage<T, false>(s, rho, &g0, &g1);
  *sfrac = sf;
  *dd = sf * sf * (D[0] * D[0] + D[1] * D[1] + D[2] * D[2]);
  return al;
}

template <typename T> struct Acc {
  T rgb[3];
  T tr;  // transmittance
  int cls, n;
  bool done, overflow;
  uint64_t mask;
};

// Per-lane crossing records in LDS, strided by the block size so that a
// wave's lanes hit consecutive words.  FP64 keeps psi_{j-1} and psi_j of
// crossing k at psi[2k kThreads] and psi[(2k + 1) kThreads].  FP32 keeps
// psi_j alone at psi[k kThreads], as a 32-bit fixed-point angle in units of
// 2^-26 rad (error <= 2^-27 rad for psi_j < 64 rad; a crossing at a larger
// angle is not recorded and counts as an overflow), and the adjoint
// recovers psi_{j-1} = psi_j - (w_{j-1} + w_j) h from the step's closed
// form.  That is 24 bytes a record, so three 256-lane blocks fit in a CU's
// LDS.
constexpr double kPsiUnit = 0x1p26;  // fixed-point units per radian
constexpr double kPsiMax = 64.0;     // 2^32 units
template <typename T> struct Rec {
  using Psi = std::conditional_t<sizeof(T) == 4, uint32_t, double>;
  static constexpr int kPsi = sizeof(T) == 4 ? 1 : 2;  // angles per record
  Psi* psi;
  T* tr;  // T_k at tr[k kThreads]
  T* ps;  // prefix sums through step j at ps[(4k + q) kThreads], q < 4
};

// Per-step factors of dw_j (w = G^(-1/2), dw = -w^3 dG / 2 with
// dG = G'(u) du + 2 u^3 dM + d(1/b^2) and du_j = du0 + kappa_j d(du)):
// w^3 G', kappa w^3 G', w^3 u^3 and w^3, in the order of Rec::ps.
template <typename T> struct StepX {
  T w, x[4];
};
template <typename T>
__device__ __forceinline__ StepX<T> step_x(const DRay<T>& d, int j, int n, T c3x3, T c2x2) {
  const T sj = -sigma<T>(j, n) * d.du;
  const T w = rsq(g_at(d, sj));
  const T kj = kappa<T>(j, n);
  const T u = fma_t(kj, d.du, d.u0);
  const T w3 = w * w * w;
  const T t = w3 * fma_t(sj, fma_t(c3x3, sj, c2x2), d.c1);
  return {w, {t, kj * t, w3 * (u * u * u), w3}};
}

// Pass 1: the forward march.  With RECORD, crossing k stores
// (psi_{j-1}, psi_j, T_k) in rec.  Chord j is examined only when psi_j has
// reached a disc-plane crossing threshold or when the bound
//   dist(chord, centre) >= min(r_{j-1}, r_j) cos(dpsi/2)
//                       >= (1 - dpsi^2/8) / max(u_{j-1}, u_j)
// does not clear the horizon radius; other chords have neither a sign change
// of z nor a horizon entry, so skipping them changes nothing.
template <typename T, bool RECORD>
__device__ inline void diff_march(const DiffScene& s, const DRay<T>& d, Rec<T> rec, Acc<T>& a) {
  a.rgb[0] = a.rgb[1] = a.rgb[2] = (T)0;
  a.tr = (T)1;
  a.cls = BH_HIT_NONE;
  a.n = 0;
  a.done = false;
  a.overflow = false;
  a.mask = 0;
  if (d.degenerate) {  // pixel line through the centre: captured, black
    a.cls = BH_HIT_HORIZON;
    a.tr = (T)0;
    return;
  }
  const int n = s.nstep;
  const T Rh = (T)s.Rh;
  const T Rh_hi = (T)(s.Rh * (1.0 + 1e-5));
  T P1[3] = {d.C[0], d.C[1], d.C[2]};
  bool have1 = true;  // P1 holds the point of step j - 1
  double psi = 0.0;
  double a_ev = d.a0 - d.guard;
  T wp = (T)0, u_prev = d.u0;
  T sx[4] = {(T)0, (T)0, (T)0, (T)0};  // RECORD: prefix sums of the StepX factors
  const T c3x3 = (T)3 * d.c3, c2x2 = (T)2 * d.c2;
  for (int j = 1; j <= 2 * n - 1; ++j) {
    if (j > n && d.captured) {
      // chord to the centre: the horizon is entered, the disc plane is not
      // crossed (n.(B - B) = 0)
      if (a.cls == BH_HIT_NONE) a.cls = BH_HIT_HORIZON;
      a.tr = (T)0;
      break;
    }
    const T sj = -sigma<T>(j, n) * d.du;
    const T w = rsq(g_at(d, sj));
    const T dpsi = (wp + w) * d.h;
    const double psi_prev = psi;
    psi += (double)dpsi;
    wp = w;
    const T kj = kappa<T>(j, n);
    const T u = fma_t(kj, d.du, d.u0);
    if (RECORD) {
      const T w3 = w * w * w;
      const T t = w3 * fma_t(sj, fma_t(c3x3, sj, c2x2), d.c1);
      sx[0] += t;
      sx[1] = fma_t(kj, t, sx[1]);
      sx[2] = fma_t(w3, u * u * u, sx[2]);
      sx[3] += w3;
    }
    const T umax = u > u_prev ? u : u_prev;
    const bool need_h = !(dpsi < (T)3) || fma_t(-dpsi * dpsi, (T)0.125, (T)1) <= Rh_hi * umax;
    const bool need_d = psi >= a_ev;
    if (need_h || need_d) {
      T cs, sn;
      if (!have1) orbit_point(d, psi_prev, rcp(u_prev), P1, &cs, &sn);
      T P2[3];
      orbit_point(d, psi, rcp(u), P2, &cs, &sn);
      T dh = (T)0, q[3], sf, dd;
      const bool hh = need_h && seg_horizon(P1, P2, Rh, &dh);
      const T al = disc_cross(s, P1, P2, q, &sf, &dd);
      if (al > (T)0 && (!hh || dd < dh)) {
        T c[3];
        disc_texel<T, false>(s, q, c, nullptr, nullptr);
        if (RECORD) {
          // psi >= 0; the negated test also catches a NaN angle
          if (a.n < kMaxRec && (Rec<T>::kPsi == 2 || psi < kPsiMax)) {
            if constexpr (Rec<T>::kPsi == 1) {
              rec.psi[a.n * kThreads] = (uint32_t)rint(psi * kPsiUnit);
            } else {
              rec.psi[(2 * a.n + 0) * kThreads] = psi_prev;
              rec.psi[(2 * a.n + 1) * kThreads] = psi;
            }
            rec.tr[a.n * kThreads] = a.tr;
#pragma unroll
            for (int q = 0; q < 4; ++q) rec.ps[(4 * a.n + q) * kThreads] = sx[q];
            a.mask |= 1ull << j;
          } else {
            a.overflow = true;
          }
        }
        const T ta = a.tr * al;
#pragma unroll
        for (int k = 0; k < 3; ++k) a.rgb[k] = fma_t(ta, c[k], a.rgb[k]);
        a.tr = a.tr * ((T)1 - al);
        if (a.cls == BH_HIT_NONE) a.cls = BH_HIT_DISC;
        ++a.n;
        if (!(a.tr > (T)0)) break;
      }
      if (hh) {
        if (a.cls == BH_HIT_NONE) a.cls = BH_HIT_HORIZON;
        a.tr = (T)0;
        break;
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) P1[k] = P2[k];
      have1 = true;
      if (need_d) a_ev = next_crossing(d, psi);
    } else {
      have1 = false;
    }
    u_prev = u;
  }
}

// Adjoint of P = r (cos(psi) e1 + sin(psi) e2), r = 1/u.
template <typename T>
__device__ __forceinline__ void point_adjoint(const DRay<T>& d, const T* Pb, T cs, T sn, T r,
                                              T* e1b, T* e2b, T* psib, T* ub) {
  T p1 = (T)0, p2 = (T)0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    e1b[k] = fma_t(Pb[k], r * cs, e1b[k]);
    e2b[k] = fma_t(Pb[k], r * sn, e2b[k]);
    p1 = fma_t(Pb[k], d.e1[k], p1);
    p2 = fma_t(Pb[k], d.e2[k], p2);
  }
  *psib = r * (cs * p2 - sn * p1);
  *ub = -(cs * p1 + sn * p2) * r * r;
}

template <typename T> __device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

template <typename T, bool ADJ>
__global__ __launch_bounds__(kThreads, (ADJ && sizeof(T) == 4) ? 3 : 1) void diff_kernel(DiffScene s, const float* __restrict__ grad_rgb,
                                                        float* __restrict__ rgb_out,
                                                        uint8_t* __restrict__ hit_out,
                                                        double* __restrict__ partial) {
  __shared__ typename Rec<T>::Psi rec_psi[ADJ ? Rec<T>::kPsi * kMaxRec * kThreads : 1];
  __shared__ T rec_tr[ADJ ? kMaxRec * kThreads : 1];
  __shared__ T rec_ps[ADJ ? 4 * kMaxRec * kThreads : 1];
  __shared__ double red[ADJ ? kPartial * (kThreads / 64) : 1];
  const int tid = threadIdx.x;
  const int px = blockIdx.x * kTile + (tid & (kTile - 1));
  const int py = blockIdx.y * kTile + (tid / kTile);
  const bool inside = px < s.W && py < s.H;
  double g_rs = 0, g_in = 0, g_di = 0, g_ri = 0, g_ov = 0;
  if (inside) {
    const size_t idx = (size_t)py * s.W + px;
    Setup64 su;
    diff_setup(s, px, py, su);
    DRay<T> d;
    to_ray(s, su, d);
    Acc<T> a;
    const Rec<T> rec{rec_psi + tid, rec_tr + tid, rec_ps + tid};
    diff_march<T, ADJ>(s, d, rec, a);
    if (rgb_out) {
      rgb_out[3 * idx + 0] = (float)a.rgb[0];
      rgb_out[3 * idx + 1] = (float)a.rgb[1];
      rgb_out[3 * idx + 2] = (float)a.rgb[2];
    }
    if (hit_out) hit_out[idx] = (uint8_t)a.cls;
    if (ADJ) g_ov = a.overflow ? 1.0 : 0.0;
    if (ADJ && a.mask) {
      const T gb[3] = {(T)grad_rgb[3 * idx], (T)grad_rgb[3 * idx + 1], (T)grad_rgb[3 * idx + 2]};
      const int n = s.nstep;
      T B[3] = {(T)0, (T)0, (T)0};
      T e1b[3] = {(T)0, (T)0, (T)0}, e2b[3] = {(T)0, (T)0, (T)0}, Cb[3] = {(T)0, (T)0, (T)0};
      T hb = (T)0, u0b = (T)0, dub = (T)0, Mb = (T)0, ib2b = (T)0, rinb = (T)0;
      // sum over angle injections m of g_m (S(m) + S(m - 1)), per factor
      T wx[4] = {(T)0, (T)0, (T)0, (T)0};
      const T c3x3 = (T)3 * d.c3, c2x2 = (T)2 * d.c2;
      const T ih = (T)1 / d.h;
      int k = __popcll(a.mask) - 1;  // records 0 .. k, one per mask bit
      uint64_t mask = a.mask;
      while (mask) {
        const int j = 63 - __builtin_clzll(mask);
        mask &= ~(1ull << j);
        const StepX<T> xj = step_x(d, j, n, c3x3, c2x2);
        const StepX<T> xi = step_x(d, j > 1 ? j - 1 : 1, n, c3x3, c2x2);  // used for j > 1
        double ps_a, ps_b;
        if constexpr (Rec<T>::kPsi == 1) {
          ps_b = (double)rec.psi[k * kThreads] * (1.0 / kPsiUnit);
          ps_a = ps_b - (double)((xi.w + xj.w) * d.h);
        } else {
          ps_a = rec.psi[(2 * k + 0) * kThreads];
          ps_b = rec.psi[(2 * k + 1) * kThreads];
        }
        const T Tk = rec.tr[k * kThreads];
        T S0[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) S0[q] = rec.ps[(4 * k + q) * kThreads];
        --k;
        const T uj = fma_t(kappa<T>(j, n), d.du, d.u0);
        T P1[3], P2[3], ca = (T)0, sa = (T)0, ra = (T)0, cb, sb;
        if (j == 1) {
#pragma unroll
          for (int m = 0; m < 3; ++m) P1[m] = d.C[m];
        } else {
          ra = rcp(fma_t(kappa<T>(j - 1, n), d.du, d.u0));
          orbit_point(d, ps_a, ra, P1, &ca, &sa);
        }
        const T rb = rcp(uj);
        orbit_point(d, ps_b, rb, P2, &cb, &sb);
        T q[3], sf, dd;
        const T z1 = P1[2], z2 = P2[2];
        disc_cross(s, P1, P2, q, &sf, &dd);
        const T rho = sqrt_t(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        T da_rho, da_rin;
        const T al = coverage<T, true>(s, rho, &da_rho, &da_rin);
        T c[3], cw[3], qb[3];
#pragma unroll
        for (int m = 0; m < 3; ++m) cw[m] = gb[m] * Tk * al;
        disc_texel<T, true>(s, q, c, cw, qb);
        // compositing: rgb = F + T_k al c + T_k (1 - al) B
        T alb = (T)0;
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          alb = fma_t(gb[m], Tk * (c[m] - B[m]), alb);
          B[m] = fma_t(al, c[m] - B[m], B[m]);
        }
        rinb = fma_t(alb, da_rin, rinb);
        const T rhob = alb * da_rho;
        if (rho > (T)0) {
          const T kr = rhob / rho;
#pragma unroll
          for (int m = 0; m < 3; ++m) qb[m] = fma_t(kr, q[m], qb[m]);
        }
        // q = P1 + s (P2 - P1), s = z1 / (z1 - z2)
        T sb_ = (T)0, P1b[3], P2b[3];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          sb_ = fma_t(qb[m], P2[m] - P1[m], sb_);
          P1b[m] = ((T)1 - sf) * qb[m];
          P2b[m] = sf * qb[m];
        }
        const T den = z1 - z2;
        const T kd = sb_ / (den * den);
        P1b[2] = fma_t(-kd, z2, P1b[2]);
        P2b[2] = fma_t(kd, z1, P2b[2]);
        // the chord's end point: angle psi_j (injection m = j), radius 1/u_j
        T phb, ub;
        point_adjoint(d, P2b, cb, sb, rb, e1b, e2b, &phb, &ub);
        u0b += ub;
        dub = fma_t(kappa<T>(j, n), ub, dub);
        T S1[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          S1[m] = S0[m] - xj.x[m];
          wx[m] = fma_t(phb, S0[m] + S1[m], wx[m]);
        }
        hb = fma_t(phb, (T)ps_b * ih, hb);
        if (j == 1) {
#pragma unroll
          for (int m = 0; m < 3; ++m) Cb[m] += P1b[m];
        } else {
          // the start point: angle psi_{j-1} (injection m = j - 1)
          point_adjoint(d, P1b, ca, sa, ra, e1b, e2b, &phb, &ub);
          u0b += ub;
          dub = fma_t(kappa<T>(j - 1, n), ub, dub);
#pragma unroll
          for (int m = 0; m < 4; ++m) wx[m] = fma_t(phb, (T)2 * S1[m] - xi.x[m], wx[m]);
          hb = fma_t(phb, (T)ps_a * ih, hb);
        }
      }
      // G_bar_i = -w_i^3 h (A_i + A_{i+1}) / 2 times the factors of dG
      const T gh = (T)-0.5 * d.h;
      u0b = fma_t(gh, wx[0], u0b);
      dub = fma_t(gh, wx[1], dub);
      Mb = (T)2 * gh * wx[2];
      ib2b = gh * wx[3];
      // epilogue in FP64: h = du/2, du = (l - u0)/n, l(M, 1/b^2), set-up
      // (psi starts at 0 whatever the parameters, so A_1 goes nowhere)
      double dub64 = (double)dub + 0.5 * (double)hb;
      const double lb = dub64 / n;
      double u0b64 = (double)u0b - lb;
      double Mb64 = (double)Mb, ib2b64 = (double)ib2b;
      const double M = s.M, l = su.l;
      if (su.captured) {
        Mb64 += lb * (-1.0 / (3.0 * M * M));
      } else {
        const double Gu = l * __builtin_fma(6.0 * M, l, -2.0);
        Mb64 += lb * (-2.0 * l * l * l / Gu);
        ib2b64 += lb * (-1.0 / Gu);
      }
      const double b = su.f.b, r0 = su.f.r0;
      const double bb = ib2b64 * (-2.0 / (b * b * b));
      const double r0b = -u0b64 / (r0 * r0);
      Frame<D2> fd;
      diff_frame(s, px, py, fd);
      double gi = bb * fd.b.i + r0b * fd.r0.i;
      double gd = bb * fd.b.d + r0b * fd.r0.d;
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        gi += (double)e1b[m] * fd.e1[m].i + (double)e2b[m] * fd.e2[m].i + (double)Cb[m] * s.C_i[m];
        gd += (double)e1b[m] * fd.e1[m].d + (double)e2b[m] * fd.e2[m].d + (double)Cb[m] * s.C_d[m];
      }
      g_rs = 0.5 * Mb64;  // M = r_s / 2
      g_in = gi;
      g_di = gd;
      g_ri = (double)rinb;
    }
  }
  if (ADJ) {
    double v[kPartial] = {g_rs, g_in, g_di, g_ri, g_ov};
    const int wave = tid >> 6, lane = tid & 63;
#pragma unroll
    for (int m = 0; m < kPartial; ++m) {
      v[m] = wave_sum(v[m]);
      if (lane == 0) red[m * (kThreads / 64) + wave] = v[m];
    }
    __syncthreads();
    if (tid < kPartial) {
      double t = 0;
      for (int w = 0; w < kThreads / 64; ++w) t += red[tid * (kThreads / 64) + w];
      partial[((size_t)blockIdx.y * gridDim.x + blockIdx.x) * kPartial + tid] = t;
    }
  }
}

// Fixed-order sum of the per-block partials: thread t adds blocks t, t + 256,
// ... then a tree over the 256 threads.
__global__ __launch_bounds__(256) void diff_reduce_kernel(const double* __restrict__ partial,
                                                          int nblocks, double* __restrict__ grad,
                                                          uint32_t* __restrict__ overflow) {
  __shared__ double sm[kPartial][256];
  const int t = threadIdx.x;
  double acc[kPartial] = {0, 0, 0, 0, 0};
  for (int b = t; b < nblocks; b += 256)
#pragma unroll
    for (int m = 0; m < kPartial; ++m) acc[m] += partial[(size_t)b * kPartial + m];
#pragma unroll
  for (int m = 0; m < kPartial; ++m) sm[m][t] = acc[m];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w)
#pragma unroll
      for (int m = 0; m < kPartial; ++m) sm[m][t] += sm[m][t + w];
    __syncthreads();
  }
  if (t < 4) grad[t] = sm[t][0];
  if (t == 4 && overflow) overflow[0] = (uint32_t)sm[4][0];
}

}  // namespace

int diff_partial_blocks(int W, int H) {
  return ((W + kTile - 1) / kTile) * ((H + kTile - 1) / kTile);
}

hipError_t launch_diff_forward(const DiffScene& s, bool fp64, float* rgb, uint8_t* hit,
                               hipStream_t stream) {
  const dim3 grid((s.W + kTile - 1) / kTile, (s.H + kTile - 1) / kTile);
  if (fp64)
    hipLaunchKernelGGL((diff_kernel<double, false>), grid, dim3(kThreads), 0, stream, s, nullptr,
                       rgb, hit, nullptr);
  else
    hipLaunchKernelGGL((diff_kernel<float, false>), grid, dim3(kThreads), 0, stream, s, nullptr,
                       rgb, hit, nullptr);
  return hipGetLastError();
}

hipError_t launch_diff_adjoint(const DiffScene& s, bool fp64, const float* grad_rgb, float* rgb,
                               uint8_t* hit, double* partial, double* grad_out,
                               uint32_t* overflow_out, hipStream_t stream) {
  const dim3 grid((s.W + kTile - 1) / kTile, (s.H + kTile - 1) / kTile);
  if (fp64)
    hipLaunchKernelGGL((diff_kernel<double, true>), grid, dim3(kThreads), 0, stream, s, grad_rgb,
                       rgb, hit, partial);
  else
    hipLaunchKernelGGL((diff_kernel<float, true>), grid, dim3(kThreads), 0, stream, s, grad_rgb,
                       rgb, hit, partial);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(diff_reduce_kernel, dim3(1), dim3(256), 0, stream, partial,
                     (int)(grid.x * grid.y), grad_out, overflow_out);
  return hipGetLastError();
}

}  // namespace bh

// bh_march.h — device code of the FP32/FP64 event march shared by the
// per-frame kernels (bh_march.hip) and the persistent batch kernel
// (bh_batch.hip): per-ray set-up, the step, the exact chord tests and the
// output of one ray.  Both kernels run these functions on the same per-ray
// inputs, so a ray's result does not depend on which kernel marched it.
#pragma once

#include <type_traits>

#include "bh_math.h"

namespace bh {

// ---------------------------------------------------------------------------
// FP32 event march
// ---------------------------------------------------------------------------
//
// Orbit-plane reduction.  The reference places step k at
//   P_k = B + r_k (cos phi_k yv + sin phi_k xv),  r_k = 1/u_k,
// so every chord lies in the plane through B spanned by (yv, xv).  With
// e1 = cos phi0 yv + sin phi0 xv = sign(Y0) F/|F| and e2 = e1 rotated by +90
// degrees in that plane, P_k = B + r_k (cos a_k e1 + sin a_k e2) where
// a_k = phi_k - phi0 is the angle swept since phi0.
//
// G on the march.  With l the turning point and s = l - u,
//   G(u) = G(l) + s [2 l d + s (1 - 2 d) - 2 M s^2],  d = 1 - 3 M l,
// a cubic in s whose terms do not cancel away from the photon sphere; the
// coefficients come from FP64.  The reference's u_k = u0 + kappa_k du gives
// s_k = sigma_k du with sigma_k = nstep - kappa_k exact, so a step is three
// FMAs, one rsqrt and the trapezoid update of a (FP64 sum); in FP32 the
// cubic carries the factor 1/h^2 of the trapezoid, so the update is one add.
//
// Events.  A chord can only reach the disc plane if sign(n . (P - B)) =
// sign(alpha cos a + beta sin a) changes, i.e. a passes a crossing angle
// chi_0 + j pi; the step is tested exactly when a passes the next one (minus
// a guard for the FP32 angle).  A chord between radii >= r_min subtending Da
// stays at distance >= r_min cos(Da/2) from B, so it can only reach the
// horizon when Da > 2 acos(R_h / r_min) >= 2 sqrt(2t) (1 + t/12) with
// t = 1 - R_h / r_min (acos(1 - t) = 2 asin(sqrt(t/2)), whose series in t has
// positive terms); that step is tested exactly too.  Every other step
// provably hits nothing.
//
// Spheres.  A sphere meets the orbit plane in a circle (or not at all), and a
// chord spanning a swept angle Da < pi stays inside the wedge of polar angles
// it spans, while a circle of radius rho whose centre lies at distance D > rho
// from B stays inside the wedge theta_c +- asin(rho / D).  A step is tested
// against a sphere's circle only when its wedge meets that window (plus a
// margin far above the FP32 angle error), once per revolution of the ray.
// An FP32 lane keeps two spheres (96 VGPRs: 5 waves per SIMD; a ray with more
// in its plane goes to the FP64 pass), an FP64 lane four.
template <typename T> constexpr int kSphSlots = sizeof(T) == 4 ? 2 : 4;

template <typename T>
struct Lane {
  // G at the steps as a cubic c0 + x (c1 + x (c2 + x c3)) in a uniform x,
  // with s = l - u = sigma du:
  //   inbound  x = sigma = nstep - n
  //   approach G = c0 after to_outbound                  (sigma = 0.1)
  //   outbound x = j = n - nstep                         (sigma = 0.1 + j)
  // to_outbound re-expands the cubic about sigma = 0.1 once the inbound steps
  // are done, so a step needs no per-lane s.  FP32 holds G / h^2.
  T c0, c1, c2, c3;
  T du, h;            // du and du/2
  T u0;               // 1/|C - B|
  float phi_safe;     // horizon guard angle
  float guard;        // disc crossing-angle guard
  // a disc crossing angle, chi0 (crossings at chi0 + j pi): FP64 lanes hold
  // it, FP32 lanes the FP32 atan2 it is formed from (chi0 = that + pi/2 in
  // FP64, chi0_of), one VGPR fewer live through the march
  std::conditional_t<sizeof(T) == 4, float, double> chi0s;
  double a_ev;        // next disc event angle (crossing - guard)
  // Scenes with spheres only (setup_spheres): the spheres that cut the orbit
  // plane, at most kSphSlots<T>, in index order.  Slot k holds sphere
  // (sids >> 8k) & 0xff as the circle it cuts from the plane (centre
  // (sx, sy) in the (e1, e2) frame, radius^2 srho2) and the window of swept
  // angle [slo, slo + sw] of the ray's next pass by that circle; a_sph is the
  // earliest window start (1e308 once the ray has ended).
  T sx[kSphSlots<T>], sy[kSphSlots<T>], srho2[kSphSlots<T>];
  float slo[kSphSlots<T>], sw[kSphSlots<T>];
  uint32_t sids;
  int ns;
  double a_sph;
  // FP32: a chord passed a sphere's surface within the FP32 position error,
  // or more than kSphSlots<T> spheres cut the plane (the ray is settled in FP64)
  bool sredo;
};

template <typename T> __device__ __forceinline__ double chi0_of(const Lane<T>& L) {
  if constexpr (sizeof(T) == 4)
    return (double)L.chi0s + kPiD / 2.0;
  else
    return L.chi0s;
}

template <typename T>
struct HitRec {
  int cls;       // BH_HIT_*
  int obj;       // sphere index of an object hit
  double ang;    // rotation of the frame the hit point is expressed in
                 // (FP32 disc hits: the frame of chi0, ang not set)
  T hx, hy;      // hit point relative to B in that frame
};

// Exact tests of the chord from (kappa_a, a_a) to (kappa_b, a_b), where
// u = u0 + kappa du and a is the swept angle; da = a_b - a_a.
// S is DevScene or any type with its fields Rh, r_in2 and r_out2.
template <typename T, typename S>
__device__ inline bool event_chord(const S& s, const Lane<T>& L, T ka, double aa, T kb,
                                   double ab, T da, bool th, bool td, HitRec<T>& hit,
                                   T* d2out) {
  const T ra = rcp<T>(fma_t<T>(ka, L.du, L.u0));
  const T rb = rcp<T>(fma_t<T>(kb, L.du, L.u0));
  bool found = false;
  T best = 0;
  if (th) {
    // frame rotated by aa: P_a = (ra, 0), P_b = rb (cos da, sin da);
    // StaticBlackhole::Collide (blackhole_solution.h:65-88) in the plane
    const T R = (T)s.Rh;
    T sd, cd;
    sincos_chord<T>(da, &sd, &cd);
    const T qx = rb * cd - ra, qy = rb * sd;
    const T qd = qx * ra;
    const T qq = qx * qx + qy * qy;
    if (!(ra < R && rb < R)) {
      const T cc = qd * qd - qq * (ra * ra - R * R);
      if (cc > (T)0) {
        const T t = (-qd - sqrt_t<T>(cc)) * rcp<T>(qq);
        if (t > (T)0 && t < (T)1) {
          found = true;
          best = t * t * qq;
          hit.cls = BH_HIT_HORIZON;
          hit.ang = aa;
          hit.hx = ra + t * qx;
          hit.hy = t * qy;
        }
      }
    }
  }
  if (td) {
    // Annulus::Collide (vector_object.h:328-347): n.(P - B) ~ r sin(a - chi)
    // changes sign an odd number of times; crossing by linear interpolation
    // on the chord with weights |t_b|, |t_a|
    const double chi0 = chi0_of(L);
    const double fa = floor((aa - chi0) * kInvPiD);
    const double fb = floor((ab - chi0) * kInvPiD);
    const int ja = (int)fa, jb = (int)fb;
    if ((jb - ja) & 1) {
      const double chik = chi0 + fmax(fa, fb) * kPiD;
      T sa, ca, sb, cb;
      // aa <= chik <= ab: both angles lie within the chord's span
      sincos_chord<T>((T)(aa - chik), &sa, &ca);
      sincos_chord<T>((T)(ab - chik), &sb, &cb);
      const T wa = fabs(ra * sa), wb = fabs(rb * sb);
      const T inv = rcp<T>(wa + wb);
      const T px = (wb * ra * ca + wa * rb * cb) * inv;
      // sa <= 0 <= sb, so the weighted heights cancel exactly: the point
      // lies on the crossing line (FP32 lanes take py = 0, FP64 lanes keep
      // the reference's interpolation and its rounding)
      const T py = sizeof(T) == 4 ? (T)0 : (wb * ra * sa + wa * rb * sb) * inv;
      const T rho2 = px * px + py * py;
      if (!(rho2 > (T)s.r_out2) && !(rho2 < (T)s.r_in2)) {
        const T dx = ra * ca - px, dy = ra * sa - py;
        const T d2 = dx * dx + dy * dy;
        if (!found || d2 < best) {
          found = true;
          best = d2;
          hit.cls = BH_HIT_DISC;
          if constexpr (sizeof(T) == 4) {
            // FP32 lanes keep the point in the frame of chi0 (finish_ray
            // forms its direction without trig): (-1)^j folded into it
            const bool odd = max(ja, jb) & 1;
            hit.hx = odd ? -px : px;
            hit.hy = odd ? -py : py;
          } else {
            hit.ang = chik;
            hit.hx = px;
            hit.hy = py;
          }
        }
      }
    }
  }
  *d2out = best;
  return found;
}

// sin and cos of a swept angle in precision T (FP64 reduction for both).
template <typename T> __device__ __forceinline__ void sincos_swept(double a, T* s, T* c);
template <> __device__ __forceinline__ void sincos_swept<float>(double a, float* s, float* c) {
  sincos_angle(a, s, c);
}
template <> __device__ __forceinline__ void sincos_swept<double>(double a, double* s, double* c) {
  sincos_small64(a, s, c);
}

constexpr float kTwoPiF = 6.28318531f;

// Spheres (the slots of L).  A sphere of radius R whose centre lies at
// distance z from the plane cuts it in a circle of radius sqrt(R^2 - z^2), and
// |P - c|^2 = |P - c'|^2 + z^2 for a point P of the plane and the centre's
// projection c', so the reference's segment/sphere test
// (blackhole_solution.h:65-88, the same test as the horizon's) of an in-plane
// chord is exactly this circle test in the plane.  The chord from
// P_a = (ra, 0) to P_b = (bx, by), in the frame rotated by the swept angle aa
// (sa, ca its sine and cosine), is tested against the slots whose window its
// wedge meets (every slot when all); a hit nearer to the chord start than
// best (found) replaces hit.  ab is the swept angle at the chord's end: a
// window that ends before it moves on to the ray's next revolution.
template <typename T>
__device__ inline bool sphere_slots(const DevScene& s, Lane<T>& L, T ra, T sa, T ca, double aa,
                                    T bx, T by, float ab, bool all, bool found, T best,
                                    HitRec<T>& hit) {
  const T qx = bx - ra, qy = by;
  const T qq = qx * qx + qy * qy;
  const T iqq = rcp<T>(qq);
  double nxt = 1e308;
#pragma unroll
  for (int k = 0; k < kSphSlots<T>; ++k) {
    if (k >= L.ns) continue;
    if (all || ab >= L.slo[k]) {
      const T X = L.sx[k], Y = L.sy[k], rho2 = L.srho2[k];
      // centre in the rotated frame, then P_a - c' and P_b - c'
      const T xc = X * ca + Y * sa, yc = Y * ca - X * sa;
      const T ax = ra - xc, ay = -yc;
      const T q1 = ax * ax + ay * ay;
      const T q2 = (bx - xc) * (bx - xc) + (by - yc) * (by - yc);
      const T qd = qx * ax + qy * ay;
      const T cc = qd * qd - qq * (q1 - rho2);
      if constexpr (sizeof(T) == 4) {
        // grazing: the line's nearest approach to c' lies on the chord and
        // cc = |Q|^2 (rho^2 - dist^2) is within the FP32 position error
        // (relative pos_guard of the scale |c'| + rho) of zero
        const float rho = __builtin_sqrtf(fmaxf(rho2, 0.f));
        const float dc = __builtin_sqrtf(X * X + Y * Y);
        const float t0 = -qd * iqq;
        if (t0 >= 0.f && t0 <= 1.f &&
            fabsf(cc) < qq * (4.f * s.pos_guard) * rho * (dc + rho))
          L.sredo = true;
      }
      if (!(q1 < rho2 && q2 < rho2) && cc > (T)0) {
        const T t = (-qd - sqrt_t<T>(cc)) * iqq;
        if (t > (T)0 && t < (T)1) {
          const T d2 = t * t * qq;
          if (!found || d2 < best) {
            found = true;
            best = d2;
            hit.cls = BH_HIT_OBJECT;
            hit.obj = (int)((L.sids >> (8 * k)) & 0xffu);
            hit.ang = aa;
            hit.hx = ra + t * qx;
            hit.hy = t * qy;
          }
        }
      }
    }
    const float hi = L.slo[k] + L.sw[k];
    if (ab > hi) L.slo[k] += kTwoPiF * (__builtin_floorf((ab - hi) * (1.f / kTwoPiF)) + 1.f);
    nxt = fmin(nxt, (double)L.slo[k]);
  }
  L.a_sph = nxt;
  return found;
}

// The step chord from (ka, aa) to (kb, aa + da) against the spheres; ab and
// all as for sphere_slots.
template <typename T>
__device__ inline bool sphere_chord(const DevScene& s, Lane<T>& L, T ka, double aa, T kb, T da,
                                    float ab, bool all, bool found, T best, HitRec<T>& hit) {
  const T ra = rcp<T>(fma_t<T>(ka, L.du, L.u0));
  const T rb = rcp<T>(fma_t<T>(kb, L.du, L.u0));
  T sd, cd, sa, ca;
  sincos_t<T>(da, &sd, &cd);
  sincos_swept<T>(aa, &sa, &ca);
  return sphere_slots<T>(s, L, ra, sa, ca, aa, rb * cd, rb * sd, ab, all, found, best, hit);
}

// A captured ray's last chord, from its last point (kappa = nstep - 0.1, at
// the swept angle a) straight to B (blackhole_solution_test.cc:264-272): it
// enters the horizon at distance r - R_h, unless a sphere is met before.
template <typename T>
__device__ inline void captured_chord(const DevScene& s, Lane<T>& L, double a, HitRec<T>& hit) {
  hit.cls = BH_HIT_HORIZON;
  hit.ang = a;
  hit.hx = (T)s.Rh;
  hit.hy = (T)0;
  if (L.ns > 0) {
    const T ra = rcp<T>(fma_t<T>((T)s.nstep - (T)0.1, L.du, L.u0));
    const T dh = ra - (T)s.Rh;
    T sa, ca;
    sincos_swept<T>(a, &sa, &ca);
    sphere_slots<T>(s, L, ra, sa, ca, a, (T)0, (T)0, (float)a, true, true, dh * dh, hit);
  }
}

__device__ __forceinline__ double next_event(double a, double chi0, float guard) {
  return chi0 + (floor((a - (double)guard - chi0) * kInvPiD) + 1.0) * kPiD - (double)guard;
}

template <typename T>
struct MarchOut {
  HitRec<T> hit;
  int steps;     // steps taken
  double a;      // swept angle after the last step taken
  double du64;   // du of the ray
  T d;           // last increment taken (for the escape direction)
};

// Step n of the reference rule has u = u0 + kappa_n du and s = l - u =
// sigma_n du with
//   inbound   n < nstep:  kappa = n,                 sigma = nstep - n
//   approach  n = nstep:  kappa = nstep - 0.1,       sigma = 0.1
//   outbound  n > nstep:  kappa = 2 nstep - 0.1 - n, sigma = 0.1 + (n - nstep)
// The coeffs_* helpers give G at step n and the kappa of the chord's two
// ends for each phase (the approach and outbound ones once to_outbound has
// re-expanded the cubic); step_coeffs picks the phase from n.
// x = (T)k for the wave-uniform step index k (one conversion per step;
// assembling the float from k with scalar integer operations was measured
// slower: a dependent chain of 8 scalar instructions per step).
template <typename T> __device__ __forceinline__ T step_x(int k) { return (T)k; }

template <typename T>
__device__ __forceinline__ T step_g(const Lane<T>& L, T x) {
  return fma_t<T>(fma_t<T>(fma_t<T>(L.c3, x, L.c2), x, L.c1), x, L.c0);
}
// The step cubic in sigma re-expanded about sigma = 0.1 in powers of
// j = sigma - 0.1 (the approach and outbound steps).
template <typename T> __device__ __forceinline__ void to_outbound(Lane<T>& L) {
  const T a = (T)0.1, a3 = (T)0.3;
  const T c0 = fma_t<T>(a, fma_t<T>(a, fma_t<T>(a, L.c3, L.c2), L.c1), L.c0);
  const T c1 = fma_t<T>(a, fma_t<T>(a3, L.c3, (T)2 * L.c2), L.c1);
  L.c2 = fma_t<T>(a3, L.c3, L.c2);
  L.c1 = c1;
  L.c0 = c0;
}
template <typename T>
__device__ __forceinline__ void coeffs_in(int n, int nstep, const Lane<T>& L, T* G, T* ka,
                                          T* kb) {
  *G = step_g<T>(L, step_x<T>(nstep - n));
  const T kn = (T)n;
  *ka = kn - (T)1;
  *kb = kn;
}
// (after to_outbound)
template <typename T>
__device__ __forceinline__ void coeffs_final(int nstep, const Lane<T>& L, T* G, T* ka, T* kb) {
  const T kt = (T)nstep - (T)0.1;
  *G = L.c0;
  *ka = kt - (T)0.9;
  *kb = kt;
}
template <typename T>
__device__ __forceinline__ void coeffs_out(int n, int nstep, const Lane<T>& L, T* G, T* ka,
                                           T* kb) {
  const T kt = (T)nstep - (T)0.1;
  const T j = (T)(n - nstep);
  *G = step_g<T>(L, step_x<T>(n - nstep));
  *ka = kt - j + (T)1;
  *kb = kt - j;
}
template <typename T>
__device__ __forceinline__ void step_coeffs(int n, int nstep, const Lane<T>& L, T* G, T* ka,
                                            T* kb) {
  if (n < nstep)
    coeffs_in<T>(n, nstep, L, G, ka, kb);
  else if (n == nstep)
    coeffs_final<T>(nstep, L, G, ka, kb);
  else
    coeffs_out<T>(n, nstep, L, G, ka, kb);
}

// Running state of one ray's march.
template <typename T> struct MarchState {
  T w;       // w of the previous step (FP32: w h)
  double a;  // swept angle
  T d;       // last increment
};

// The trapezoid update of the swept angle by one step (G = G(u) at the step).
template <typename T>
__device__ __forceinline__ void advance(const Lane<T>& L, MarchState<T>& st, T G) {
  const T wn = rsq<T>(G);  // FP32: w h (setup_lane scales the cubic)
  if constexpr (sizeof(T) == 4)
    st.d = st.w + wn;
  else
    st.d = (st.w + wn) * L.h;
  st.w = wn;
  st.a += (double)st.d;
}

// The exact chord tests of the step just advanced (from angle a_prev to st.a,
// increment st.d), if it is an event (a crossing angle passed, the horizon
// guard exceeded, or force).  Returns true when the chord hits (hit is
// filled).  With SPH the chord is also tested against the spheres whose
// window it meets.
template <typename T, typename S, bool SPH = false>
__device__ __forceinline__ bool step_tests(const S& s, Lane<T>& L, const MarchState<T>& st,
                                           double a_prev, double first_aa, T ka, T kb,
                                           bool force, HitRec<T>& hit) {
  const bool eh = (float)st.d > L.phi_safe;
  const double aa = force ? first_aa : a_prev;
  const T da = force ? (T)(st.a - first_aa) : st.d;
  bool found = false;
  T best = (T)0;
  if (eh || st.a >= L.a_ev || force) {
    found = event_chord<T, S>(s, L, ka, aa, kb, st.a, da, eh || force, s.has_disc != 0, hit,
                              &best);
    if (!found && st.a >= L.a_ev) L.a_ev = next_event(st.a, chi0_of(L), L.guard);
  }
  if constexpr (SPH) {
    // every slot on a chord of Da >= 1 (the wedge argument needs Da < pi;
    // phi_safe < pi on an armed lane, +inf on an ended one)
    const bool all = force || (float)st.d > fmaxf(1.f, L.phi_safe);
    if (all || st.a >= L.a_sph)
      found = sphere_chord<T>(s, L, ka, aa, kb, da, (float)st.a, all, found, best, hit);
  }
  return found;
}

// One step: advance, then its tests.
template <typename T, typename S, bool SPH = false>
__device__ __forceinline__ bool march_step(const S& s, Lane<T>& L, MarchState<T>& st,
                                           double first_aa, T G, T ka, T kb, bool force,
                                           HitRec<T>& hit) {
  advance<T>(L, st, G);
  return step_tests<T, S, SPH>(s, L, st, st.a - (double)st.d, first_aa, ka, kb, force, hit);
}

// The march of one ray in precision T.  first_aa is the angle of the camera
// in the (e1, e2) frame: 0 when Y0 > 0, pi when Y0 < 0 (the reference's
// phi0 = atan(b/Y0) then points away from the camera, and only the first
// chord starts at the camera itself).
//
// Every lane of a wave is at the same step, so the loops run wave-uniform
// trip counts and keep no per-lane flag: a ray that has ended is disarmed by
// setting its event thresholds (and sphere windows) to +inf (its results
// were recorded when it ended; it keeps stepping harmlessly), and after each
// block of four (or six) steps the wave leaves once no lane is armed.  The
// steps of a block take one event check between them, which keeps the
// exec-mask bookkeeping of the hot loop to a handful of scalar instructions
// per block.
// The step count of the reference's step rule (blackhole_solution_test.cc
// and the shipped scene): the FP32 march has an instance compiled for it.
constexpr int kRefSteps = 20;

template <typename T, bool ESCAPE, bool SPH = false, int NS = 0>
__device__ inline void march(const DevScene& s, Lane<T>& L, bool captured, double first_aa,
                             double du64, MarchOut<T>& o) {
  // NS > 0: the scene's nstep, known at compile time: the blocks unroll, and
  // their step values and chord ends become literal operands (no per-step
  // integer-to-float conversion, no loop bookkeeping)
  const int nstep = NS > 0 ? NS : s.nstep;
  // block length: six steps in the compiled instance of an FP32 lane
  // without sphere slots (fewer checks and single tail steps: 20 steps leave
  // none inbound and one outbound; nine-step blocks measured no faster and
  // spill), four otherwise (six would cost the sphere lanes occupancy)
  constexpr int kB = (NS > 0 && sizeof(T) == 4 && !SPH) ? 6 : 4;
  MarchState<T> st{(T)0, 0.0, (T)0};
  o.hit.cls = BH_HIT_NONE;
  o.hit.obj = 0;
  o.hit.ang = 0.0;
  o.hit.hx = o.hit.hy = (T)0;
  o.du64 = du64;
  o.steps = 0;
  o.a = 0.0;
  o.d = (T)0;
  const float kInf = __builtin_inff();
  auto end_ray = [&](int n) {
    o.steps = n;
    o.a = st.a;
    o.d = st.d;
    L.phi_safe = kInf;
    L.a_ev = 1e308;
    if constexpr (SPH) L.a_sph = 1e308;  // no sphere tests either once ended
  };
  auto armed = [&]() { return L.phi_safe < kInf; };
  auto step = [&](int n, T G, T ka, T kb, bool force) {
    if (march_step<T, DevScene, SPH>(s, L, st, first_aa, G, ka, kb, force, o.hit)) end_ray(n);
  };
  auto step_in = [&](int n, bool force) {
    T G, ka, kb;
    coeffs_in<T>(n, nstep, L, &G, &ka, &kb);
    step(n, G, ka, kb, force);
  };
  auto step_out = [&](int n) {
    T G, ka, kb;
    coeffs_out<T>(n, nstep, L, &G, &ka, &kb);
    step(n, G, ka, kb, false);
  };
  // K steps n0..n0+K-1 (coeffs: coeffs_in or coeffs_out; len carries K,
  // four or six): the increments first, then one check of the block.  The
  // swept angle only grows, so a crossing threshold or sphere window is
  // passed in the block iff the block's last angle passes it, and the horizon
  // guard iff the largest increment exceeds it; only a lane whose block holds
  // such an event replays the block's step tests in order, on angles that
  // rise monotonely to the block's last one, so every step is tested as one
  // at a time.
  // The increments grow along the inbound steps and shrink along the
  // outbound ones (G falls as u rises below 1/(3M)), so the largest of a
  // block is its last inbound or its first outbound one (the 1e-4 margin of
  // phi_safe covers rounding).
  //
  // FP32 lanes add the block's trapezoid increments in FP32 before the
  // FP64 sum: with w_k the half-steps w h, the increments are w_{k-1} + w_k
  // and their sum telescopes to (w_0 + w_K) + 2 (w_1 + ... + w_{K-1}), K
  // FP32 operations and one FP64 addition in place of K increments and K
  // FP64 additions.  The replay forms its intermediate angles as a0
  // plus the FP32 partial sums of the increments (monotone, below the
  // block's sum by at least the last increment) and ends on the block's own
  // angle; each chord starts where the previous one ended.  FP64 lanes sum
  // every increment in FP64.
  auto block = [&](int n0, auto coeffs, bool inbound, auto len) {
    constexpr int K = decltype(len)::value;
    const double a0 = st.a;
    if constexpr (sizeof(T) == 4) {
      const T w0 = st.w;
      T wk[K];
#pragma unroll
      for (int k = 0; k < K; ++k) {
        T G, ka, kb;
        coeffs(n0 + k, &G, &ka, &kb);
        wk[k] = rsq<T>(G);
      }
      T mid = wk[0];
#pragma unroll
      for (int k = 1; k < K - 1; ++k) mid += wk[k];
      const T sum = __builtin_fmaf(2.f, mid, w0 + wk[K - 1]);
      st.w = wk[K - 1];
      st.a = a0 + (double)sum;
      const T dlast = wk[K - 2] + wk[K - 1];
      // the last increment is read after the block only by end_ray, which
      // an outbound block's successor may call before stepping
      if (inbound || n0 + 2 * K > 2 * nstep) st.d = dlast;
      const float dmax = inbound ? dlast : w0 + wk[0];
      bool ev = dmax > L.phi_safe || st.a >= L.a_ev;
      if constexpr (SPH) ev = ev || dmax > fmaxf(1.f, L.phi_safe) || st.a >= L.a_sph;
      if (ev) {
        // the block's angles are recomputed from a0 rather than kept (more
        // live VGPRs in the fast path would cost occupancy): an empty asm
        // statement hides from the compiler that the copy equals a0
        double ar = a0;
        __asm__ volatile("" : "+v"(ar));
        const double a_end = st.a;
        double ap = ar;
        T part = (T)0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
          T G, ka, kb;
          coeffs(n0 + k, &G, &ka, &kb);
          st.d = (k == 0 ? w0 : wk[k - 1]) + wk[k];
          part = k == 0 ? st.d : part + st.d;
          st.a = k == K - 1 ? a_end : ar + (double)part;
          if (step_tests<T, DevScene, SPH>(s, L, st, ap, first_aa, ka, kb, false, o.hit))
            end_ray(n0 + k);
          ap = st.a;
        }
      }
    } else {
      static_assert(K == 4, "FP64 lanes march four-step blocks");
      T dk[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        T G, ka, kb;
        coeffs(n0 + k, &G, &ka, &kb);
        advance<T>(L, st, G);
        dk[k] = st.d;
      }
      const float dmax = (float)(inbound ? dk[3] : dk[0]);
      bool ev = dmax > L.phi_safe || st.a >= L.a_ev;
      if constexpr (SPH) ev = ev || dmax > fmaxf(1.f, L.phi_safe) || st.a >= L.a_sph;
      if (ev) {
        // the replay recomputes the four sums from a0 (the same FP64
        // additions) rather than keeping them live
        double ar = a0;
        __asm__ volatile("" : "+v"(ar));
        st.a = ar;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          T G, ka, kb;
          coeffs(n0 + k, &G, &ka, &kb);
          st.d = dk[k];
          st.a += (double)dk[k];
          if (step_tests<T, DevScene, SPH>(s, L, st, st.a - (double)st.d, first_aa, ka, kb,
                                           false, o.hit))
            end_ray(n0 + k);
        }
      }
    }
  };
  auto in_coeffs = [&](int n, T* G, T* ka, T* kb) { coeffs_in<T>(n, nstep, L, G, ka, kb); };
  auto out_coeffs = [&](int n, T* G, T* ka, T* kb) { coeffs_out<T>(n, nstep, L, G, ka, kb); };
  // inbound (blackhole_solution_test.cc:217-239)
  step_in(1, first_aa != 0.0);
  int n = 2;
  if constexpr (NS > 0) {
#pragma unroll
    for (int b = 0; b < (NS - 2) / kB; ++b) {
      block(2 + kB * b, in_coeffs, true, std::integral_constant<int, kB>{});
      if (!__any(armed())) goto done;
    }
    n = 2 + kB * ((NS - 2) / kB);
  } else {
    for (; n + 3 < nstep; n += 4) {
      block(n, in_coeffs, true, std::integral_constant<int, 4>{});
      if (!__any(armed())) goto done;
    }
  }
  for (; n < nstep; ++n) step_in(n, false);
  {
    // final approach: u += 0.9 du, weighted du/2 (:240-262)
    {
      T G, ka, kb;
      to_outbound<T>(L);
      coeffs_final<T>(nstep, L, &G, &ka, &kb);
      step(nstep, G, ka, kb, false);
    }
    // captured: chord to the centre; the horizon is always entered
    // (r > 3M > 2M) and the disc test sees n.(B - B) = 0 (:264-272)
    if (captured && armed()) {
      if constexpr (SPH) {
        captured_chord<T>(s, L, st.a, o.hit);
      } else {
        o.hit.cls = BH_HIT_HORIZON;
        o.hit.ang = st.a;
        o.hit.hx = (T)s.Rh;
        o.hit.hy = (T)0;
      }
      end_ray(nstep);
    }
    if (!__any(armed())) goto done;
    // outbound: u -= du (:274-296)
    n = nstep + 1;
    if constexpr (NS > 0) {
#pragma unroll
      for (int b = 0; b < (NS - 1) / kB; ++b) {
        block(NS + 1 + kB * b, out_coeffs, false, std::integral_constant<int, kB>{});
        if (!__any(armed())) goto done;
      }
      n = NS + 1 + kB * ((NS - 1) / kB);
    } else {
      for (; n + 3 < 2 * nstep; n += 4) {
        block(n, out_coeffs, false, std::integral_constant<int, 4>{});
        if (!__any(armed())) goto done;
      }
    }
    for (; n < 2 * nstep; ++n) step_out(n);
    if (ESCAPE) {
      // extension: keep stepping outward until the next point would lie at
      // r >= r_escape (u <= 1/r_escape), or the step cap is reached
      const T kt = (T)nstep - (T)0.1;
      for (; n <= s.step_cap; ++n) {
        if (!__any(armed())) goto done;
        const T j = (T)(n - nstep);
        const double un = __builtin_fma((double)kt - (double)j, du64, s.u0);
        if (armed() && un * s.r_esc <= 1.0) end_ray(n - 1);
        step_out(n);
      }
    }
    if (armed()) end_ray(n - 1);
  }
done:;
}

// tex_index for an FP32 coordinate, without the FP64 round trip: the same
// index for every float v (NaN -> 0; n - 1 < 2^24 is exact in FP32).
__device__ __forceinline__ int tex_index32(float v, int n) {
  return (int)fminf(fmaxf(v, 0.f), (float)(n - 1));  // max(NaN, 0) = 0
}

// True when v lies within tol of an integer, a texel edge: the nearest-texel
// index of a point within tol of v may then differ from v's (false for NaN).
__device__ __forceinline__ bool near_edge(float v, float tol) {
  const float f = v - __builtin_floorf(v);
  return f < tol || f > 1.f - tol;
}

// Disc texel of a world point q (Rectangle::color, vector_object.h:163-179)
// with r cos(theta) = s1.v/|s1| and r sin(theta) = |s1 x v|/|s1|.  *redo:
// the texel coordinates lie within the scene's guard of a texel edge.
__device__ inline uint32_t disc_color32(const DevScene& s, float qx, float qy, float qz,
                                        bool* redo) {
  const DevTex& t = s.disc_tex;
  const float vx = qx - s.tv1[0], vy = qy - s.tv1[1], vz = qz - s.tv1[2];
  const float a = (s.ts1[0] * vx + s.ts1[1] * vy + s.ts1[2] * vz) * s.tka;
  const float cx = s.ts1[1] * vz - s.ts1[2] * vy;
  const float cy = s.ts1[2] * vx - s.ts1[0] * vz;
  const float cz = s.ts1[0] * vy - s.ts1[1] * vx;
  const float b = __builtin_amdgcn_sqrtf(cx * cx + cy * cy + cz * cz) * s.tkb;
  *redo = near_edge(a, s.edge_a) || near_edge(b, s.edge_b);
  return texel(t, tex_index32(a, t.cols), tex_index32(b, t.rows));
}

// Equirectangular texel of a unit direction.  *redo: a direction within eps
// radians may fall in another texel (a longitude error is the angle over
// cos(latitude); asin's conditioning near the poles is covered by the same
// factor on the rows).
__device__ inline uint32_t equirect32(const DevTex& t, float dx, float dy, float dz, float eps,
                                      bool* redo) {
  const float kPi = 3.14159265358979323846f;
  const float z = fminf(1.f, fmaxf(-1.f, dz));
  const float lon = atan2f(dy, dx);
  const float lat = asinf(z);
  const float c = (lon + kPi) * (1.0f / (2.0f * kPi)) * (float)t.cols;
  const float r = (kPi / 2.0f - lat) * (1.0f / kPi) * (float)t.rows;
  const float e = eps * __builtin_amdgcn_rsqf(fmaxf(__builtin_fmaf(-z, z, 1.f), 1e-12f));
  *redo = near_edge(c, e * (float)t.cols * (1.0f / (2.0f * kPi))) ||
          near_edge(r, e * (float)t.rows * (1.0f / kPi));
  return texel(t, tex_index32(floorf(c), t.cols), tex_index32(floorf(r), t.rows));
}

// Lane set-up in precision T from the FP64 quantities of the ray.
template <typename T>
__device__ inline void setup_lane(const DevScene& s, double l, double ib2, double du, float e2x,
                                  float e2y, float e2z, float sgn, Lane<T>& L) {
  // G(l - s) = c0 + s (c1 + s (c2 + s c3)), no cancellation near the
  // photon sphere; the step cubics of Lane in units of du
  const double dl = __builtin_fma(-s.M3, l, 1.0);
  const double c0 = __builtin_fma(l * l, __builtin_fma(s.M2, l, -1.0), ib2);
  const double c1 = 2.0 * l * dl, c2 = __builtin_fma(-2.0, dl, 1.0), c3 = -s.M2;
  const double du2 = du * du;
  if constexpr (sizeof(T) == 4) {
    // FP32: the cubic is scaled by 1/h^2 = 4/du^2, so that its rsqrt is
    // w h, the trapezoid half-step, and a step adds two of them with no
    // multiply.  du > 0 on every FP32 ray (the camera is outside the photon
    // sphere) save a ray leaving the camera at a right angle to the hole,
    // whose turning point can round onto u0: it marches nowhere (G = +inf,
    // every increment 0), as it does unscaled with du = 0 (and, to within
    // 1e-30 w per step, with du below 1e-30).
    if (du > 1e-30) {
      // 1/du: FP32 reciprocal and one FP64 Newton step (relative error
      // ~1e-14; an FP64 division here costs more than the multiply saved)
      double r = (double)__builtin_amdgcn_rcpf((float)du);
      r = r * __builtin_fma(-du, r, 2.0);
      const double r4 = 4.0 * r;
      L.c0 = (T)(c0 * (r4 * r));
      L.c1 = (T)(c1 * r4);
      L.c2 = (T)(c2 * 4.0);
      L.c3 = (T)(c3 * (4.0 * du));
    } else {
      L.c0 = (T)__builtin_inf();
      L.c1 = L.c2 = L.c3 = (T)0;
    }
  } else {
    L.c0 = (T)c0;
    L.c1 = (T)(c1 * du);
    L.c2 = (T)(c2 * du2);
    L.c3 = (T)(c3 * du2 * du);
  }
  L.du = (T)du;
  L.h = (T)(0.5 * du);
  if constexpr (sizeof(T) == 4)
    L.u0 = s.u0f;  // (float)u0, set with the scene
  else
    L.u0 = (T)s.u0;
  // smallest radius on the march: u = u0 + (nstep - 0.1) du; the guard is
  // the lower bound of 2 acos(x) above, x = R_h / r_min (FP32: a 1e-4
  // relative margin; at most 2 sqrt(2) 13/12 < pi)
  const float x = s.Rhf * __builtin_fmaf(s.ktf, (float)du, s.u0f);
  const float t = 1.f - x;
  L.phi_safe = x < 1.f ? 2.f * __builtin_amdgcn_sqrtf(2.f * t) *
                             __builtin_fmaf(t, 1.f / 12.f, 1.f) * 0.9999f
                       : 0.f;
  L.chi0s = 0;
  L.a_ev = 1e300;
  L.guard = 0.f;
  if (s.has_disc) {
    const float al = sgn * s.alpha_f;
    const float be = s.nf[0] * e2x + s.nf[1] * e2y + s.nf[2] * e2z;
    const float amp2 = al * al + be * be;
    if (amp2 > 0.f) {
      // al cos a + be sin a vanishes at a = atan2(be, al) + pi/2 (mod pi)
      const float at = atan2_fast(be, al);
      const float chi = at + 1.57079633f;
      if constexpr (sizeof(T) == 4)
        L.chi0s = at;
      else
        L.chi0s = (double)at + kPiD / 2.0;
      L.guard = 4e-6f + 4e-7f * __builtin_amdgcn_rsqf(amp2);
      // first crossing threshold at or after 0 (FP32: far inside the guard)
      L.a_ev = (double)(chi + (__builtin_floorf((-L.guard - chi) * 0.318309886f) + 1.f) *
                                  3.14159265f - L.guard);
    }
  }
  // no crossings: parity tests stay even
  if (L.a_ev == 1e300) L.chi0s = sizeof(T) == 4 ? (T)1e30f : (T)1e300;
}

// The spheres that cut the orbit plane spanned by the unit vectors e1, e2
// (precision T) into the slots of L, with their first windows: the pass of
// the circle's wedge that ends at or after swept angle 0.  A circle around B
// (D <= rho) has a window of all angles.  FP32 keeps a sphere within 1e-5 R
// of the plane (the FP32 normal's error) and widens each window by 1e-4 of
// its half-width plus 1e-4 rad (FP32 angles are good to ~1e-6 rad here).
// Returns false when more than kSphSlots<T> spheres cut the plane: the first
// kSphSlots<T> are kept and L.sredo is set.
template <typename T>
__device__ inline bool setup_spheres(const DevScene& s, const T e1[3], const T e2[3],
                                     Lane<T>& L) {
  const T n0 = e1[1] * e2[2] - e1[2] * e2[1];
  const T n1 = e1[2] * e2[0] - e1[0] * e2[2];
  const T n2 = e1[0] * e2[1] - e1[1] * e2[0];
  const T margin = sizeof(T) == 4 ? (T)1.00001f : (T)(1.0 + 1e-9);
  L.ns = 0;
  L.sids = 0u;
  L.a_sph = 1e308;
  L.sredo = false;
  bool ok = true;
  for (int i = 0; i < s.n_spheres; ++i) {
    T c0, c1, c2, R;
    if constexpr (sizeof(T) == 4) {
      c0 = s.sphf[i][0];
      c1 = s.sphf[i][1];
      c2 = s.sphf[i][2];
      R = s.sphf[i][3];
    } else {
      c0 = s.sc[i][0] - s.B[0];
      c1 = s.sc[i][1] - s.B[1];
      c2 = s.sc[i][2] - s.B[2];
      R = s.sr[i];
    }
    const T Z = c0 * n0 + c1 * n1 + c2 * n2;
    const T Rm = R * margin;
    if (!(Z * Z < Rm * Rm)) continue;
    if (L.ns == kSphSlots<T>) {
      ok = false;
      continue;
    }
    const T X = c0 * e1[0] + c1 * e1[1] + c2 * e1[2];
    const T Y = c0 * e2[0] + c1 * e2[1] + c2 * e2[2];
    const T rho2 = R * R - Z * Z;
    const float dc = __builtin_sqrtf((float)(X * X + Y * Y));
    const float rho = __builtin_sqrtf(fmaxf((float)rho2, 0.f));
    float lo = -__builtin_inff(), w2 = __builtin_inff();
    if (dc > rho * 1.0001f + 1e-6f) {
      const float w = asinf(rho / dc) * 1.0001f + 1e-4f;
      lo = atan2_fast((float)Y, (float)X) - w;
      w2 = 2.f * w;
      if (lo + w2 < 0.f) lo += kTwoPiF;
    }
#pragma unroll
    for (int k = 0; k < kSphSlots<T>; ++k) {
      if (k == L.ns) {
        L.sx[k] = X;
        L.sy[k] = Y;
        L.srho2[k] = rho2;
        L.slo[k] = lo;
        L.sw[k] = w2;
      }
    }
    L.sids |= (uint32_t)i << (8 * L.ns);
    L.a_sph = fmin(L.a_sph, (double)lo);
    ++L.ns;
  }
  if (!ok) L.sredo = true;
  return ok;
}

// Per-ray set-up shared by the FP32 march and the FP64 band kernel.
struct RaySetup {
  double cc, dd, dF, ib2, l, first_aa;
  float e2x, e2y, e2z, sgn;
  bool degenerate, captured, band;
};

__device__ inline void setup_ray(const DevScene& s, int x, int y, bool force_bisect,
                                 RaySetup& r) {
  // the pixel point pv = fvB - sx vy - sy vz (relative to B) enters only
  // through dv = F - pv (along yv), cv = pv x F (|cv| = b |dv|) and dv.F,
  // each linear in (sx, sy) with coefficients formed once per frame
  const double sx = s.hW - (double)x, sy = s.hH - (double)y;
  const double Fx = s.F[0], Fy = s.F[1], Fz = s.F[2];
  // (a vector instruction reads one scalar operand: the pixel terms first,
  // the frame's constant term last, so no constant is copied to a register)
  const double dvx = __builtin_fma(s.vz[0], sy, s.vy[0] * sx) + s.pxD[0];
  const double dvy = __builtin_fma(s.vz[1], sy, s.vy[1] * sx) + s.pxD[1];
  const double dvz = __builtin_fma(s.vz[2], sy, s.vy[2] * sx) + s.pxD[2];
  const double cvx = s.pxC[0] - __builtin_fma(s.pxCz[0], sy, s.pxCy[0] * sx);
  const double cvy = s.pxC[1] - __builtin_fma(s.pxCz[1], sy, s.pxCy[1] * sx);
  const double cvz = s.pxC[2] - __builtin_fma(s.pxCz[2], sy, s.pxCy[2] * sx);
  const double cc = __builtin_fma(cvz, cvz, __builtin_fma(cvy, cvy, cvx * cvx));
  const double dd = __builtin_fma(dvz, dvz, __builtin_fma(dvy, dvy, dvx * dvx));
  const double dF = __builtin_fma(s.pxF[2], sy, s.pxF[1] * sx) + s.pxF[0];
  r.cc = cc;
  r.dd = dd;
  r.dF = dF;
  r.degenerate = !(cc > 0.0) || !(dd > 0.0);
  r.sgn = dF > 0.0 ? 1.f : -1.f;
  r.first_aa = dF > 0.0 ? 0.0 : kPiD;
  r.captured = cc < dd * s.bc2;  // b < b_c
  r.band = cc > dd * s.lo2 && cc < dd * s.hi2;
  r.e2x = r.e2y = r.e2z = 0.f;
  r.ib2 = 0.0;
  r.l = s.Uc;
  if (r.degenerate) return;
  // 1 / b^2 = dd / cc: hardware reciprocal and two Newton steps (<= 2 ulp)
  double rc = __builtin_amdgcn_rcp(cc);
  rc = rc * __builtin_fma(-cc, rc, 2.0);
  rc = rc * __builtin_fma(-cc, rc, 2.0);
  r.ib2 = dd * rc;
  // e2 = sign(Y0) ((dv.F) F - r0^2 dv) / (r0 |cv|), kept in FP32: the
  // numerator in FP64 (it cancels), its scale in FP32, which sets the length
  // of e2 to 1 + O(2e-7)
  const float k = r.sgn * __builtin_amdgcn_rsqf((float)cc) * s.u0f;
  r.e2x = (float)__builtin_fma(dF, Fx, -s.r02 * dvx) * k;
  r.e2y = (float)__builtin_fma(dF, Fy, -s.r02 * dvy) * k;
  r.e2z = (float)__builtin_fma(dF, Fz, -s.r02 * dvz) * k;
  if (r.captured) return;
  // Turning point: StaticBlackhole::SolveG's lower bisection end.  With
  // y = 3 M u and q = (b_c/b)^2, G(u) = 0 reads 3y^2 - 2y^3 = q, a cubic
  // symmetric under y -> 1 - y, q -> 1 - q.  Its root y* in [0, 1) starts
  // from s P(s), s = sqrt(q), for q <= 1/2 and from 1 - s P(s),
  // s = sqrt(1 - q), above, where P is a least-squares degree-9 fit to y/s on
  // [0, sqrt(1/2)] (relative error 9e-8): the start, evaluated in FP32 with
  // 1 - s P(s) formed in FP64, is good to 3e-7 of min(y*, 1 - y*), which
  // cancels the Newton constant's growth as y* -> 1, and one Newton step in
  // FP64 brings u* within 1e-13 of the root outside the photon-sphere band
  // (6e-12 at q = 1 - 1e-7, where two steps from a degree-5 start left
  // 2e-11).  The bisection cell containing it is then exact unless u* sits
  // within 1e-5 cells of a cell edge, where (as in the band) the reference's
  // bisection is rerun.
  bool bisect = force_bisect || r.band || s.root == BH_ROOT_EXACT;
  if (!bisect) {
    const double q = s.bc2 * r.ib2;
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
    double u = s.Uc * (qf <= qc ? z : 1.0 - z);
    // (the FP32 1/G' leaves the quadratic convergence a 1e-7 linear term)
    {
      const double g = __builtin_fma(u * u, __builtin_fma(s.M2, u, -1.0), r.ib2);
      const double gp = u * __builtin_fma(s.M6, u, -2.0);
      const double ig = (double)__builtin_amdgcn_rcpf((float)gp);
      u = __builtin_fma(-g, ig, u);
    }
    const double t = (u - s.l0) * s.inv_cell;
    const double kc = floor(t);
    const double fr = t - kc;
    if (fr > 1e-5 && fr < 1.0 - 1e-5 && kc >= 0.0 && kc < 1048576.0)
      r.l = __builtin_fma(kc, s.cell, s.l0);
    else
      bisect = true;
  }
  if (bisect) r.l = solve_g64(sqrt(cc / dd), s.M, s.l0, s.root);
}

// The sphere slots of an FP32 lane, from its FP32 basis (e1 = sgn F/|F|).
__device__ inline void setup_spheres32(const DevScene& s, const RaySetup& r, Lane<float>& L) {
  const float e1[3] = {r.sgn * s.e1f[0], r.sgn * s.e1f[1], r.sgn * s.e1f[2]};
  const float e2[3] = {r.e2x, r.e2y, r.e2z};
  setup_spheres<float>(s, e1, e2, L);
}

// The reference's phi0 = atan(b / Y0) of the ray in FP32, for the optional
// phi plane (the FP32 march's angles are good to ~1e-6 rad), formed before
// the march so that cc and dF need not stay live through it.
__device__ __forceinline__ float phi0_of(const DevOut& out, const RaySetup& r) {
  return (out.phi && !r.degenerate) ? (float)atan(sqrt(r.cc) / r.dF) : 0.f;
}

// Hit point, colour and output planes of one ray.  Returns true when the
// ray's texel coordinate lies within the scene's edge guard of a texel edge:
// its texel is then settled by the FP64 pass.
template <bool ESCAPE>
__device__ inline bool finish_ray(const DevScene& s, const DevOut& out, size_t idx,
                                  const RaySetup& r, const MarchOut<float>& mo, float phi0) {
  const float e1x = r.sgn * s.e1f[0], e1y = r.sgn * s.e1f[1], e1z = r.sgn * s.e1f[2];
  uint32_t rgb = 0;
  bool redo = false;
  float qx = __builtin_nanf(""), qy = qx, qz = qx;
  if (mo.hit.cls != BH_HIT_NONE) {
    // a horizon hit is black: its point is formed only for the hit-point plane
    if (!r.degenerate && (mo.hit.cls != BH_HIT_HORIZON || out.hitpt)) {
      float sa, ca;
      if (mo.hit.cls == BH_HIT_DISC) {
        // a disc hit's point is given in the frame of chi0, where the orbit
        // plane meets the disc plane: al cos a + be sin a = 0 (setup_lane's
        // al, be), so (cos, sin)(chi0) = (-be, al) / |(al, be)|, without trig
        const float al = r.sgn * s.alpha_f;
        const float be = s.nf[0] * r.e2x + s.nf[1] * r.e2y + s.nf[2] * r.e2z;
        const float ia = __builtin_amdgcn_rsqf(al * al + be * be);
        ca = -be * ia;
        sa = al * ia;
      } else {
        sincos_angle(mo.hit.ang, &sa, &ca);
      }
      const float X = mo.hit.hx * ca - mo.hit.hy * sa;
      const float Y = mo.hit.hx * sa + mo.hit.hy * ca;
      qx = __builtin_fmaf(Y, r.e2x, __builtin_fmaf(X, e1x, s.Bf[0]));
      qy = __builtin_fmaf(Y, r.e2y, __builtin_fmaf(X, e1y, s.Bf[1]));
      qz = __builtin_fmaf(Y, r.e2z, __builtin_fmaf(X, e1z, s.Bf[2]));
      if (mo.hit.cls == BH_HIT_DISC) {
        rgb = disc_color32(s, qx, qy, qz, &redo);
      } else if (mo.hit.cls == BH_HIT_OBJECT) {
        // sphere texel: equirectangular lookup of the outward normal
        const float* c = s.sphf[mo.hit.obj];
        const float iR = 1.0f / c[3];
        rgb = equirect32(s.stex[mo.hit.obj], (qx - s.Bf[0] - c[0]) * iR,
                         (qy - s.Bf[1] - c[1]) * iR, (qz - s.Bf[2] - c[2]) * iR,
                         s.sph_edge[mo.hit.obj], &redo);
      }
    }
  } else if (ESCAPE && s.has_bg) {
    // direction of the last chord; outbound points are one du apart, the
    // last one at sigma = 0.1 + steps - nstep
    const double kb = (double)s.nstep - (0.1 + (double)(mo.steps - s.nstep));
    const float rb = 1.0f / (float)__builtin_fma(kb, mo.du64, s.u0);
    const float ra = 1.0f / (float)__builtin_fma(kb + 1.0, mo.du64, s.u0);
    float sb, cb, sa, ca;
    sincos_angle(mo.a, &sb, &cb);
    sincos_angle(mo.a - (double)mo.d, &sa, &ca);
    const float X = rb * cb - ra * ca, Y = rb * sb - ra * sa;
    const float dx = X * e1x + Y * r.e2x, dy = X * e1y + Y * r.e2y, dz = X * e1z + Y * r.e2z;
    const float inv = __builtin_amdgcn_rsqf(dx * dx + dy * dy + dz * dz);
    rgb = equirect32(s.bg, dx * inv, dy * inv, dz * inv, s.edge_bg, &redo);
  }
  if (out.rgb) {
    out.rgb[3 * idx + 0] = (uint8_t)(rgb & 0xff);
    out.rgb[3 * idx + 1] = (uint8_t)((rgb >> 8) & 0xff);
    out.rgb[3 * idx + 2] = (uint8_t)((rgb >> 16) & 0xff);
  }
  if (out.hit) out.hit[idx] = (uint8_t)mo.hit.cls;
  if (out.steps) out.steps[idx] = mo.steps;
  if (out.hitpt) {
    out.hitpt[3 * idx + 0] = (double)qx;
    out.hitpt[3 * idx + 1] = (double)qy;
    out.hitpt[3 * idx + 2] = (double)qz;
  }
  if (out.phi) out.phi[idx] = r.degenerate ? mo.a : (double)phi0 + mo.a;
  return redo;
}

template <typename T>
__device__ inline void degenerate_out(MarchOut<T>& mo) {
  // the pixel line passes through B: the reference marches NaNs; defined
  // here as captured / black
  mo.hit.cls = BH_HIT_HORIZON;
  mo.hit.obj = 0;
  mo.hit.ang = 0.0;
  mo.hit.hx = mo.hit.hy = (T)0;
  mo.steps = 0;
  mo.a = __builtin_nan("");
  mo.du64 = 0.0;
  mo.d = (T)0;
}

// ---------------------------------------------------------------------------
// FP64 finish of the FP64 event march (the FP64 pass over queued rays)
// ---------------------------------------------------------------------------
// Texel coordinates within this of a texel edge count as ties: the FP64 event
// march keeps a hit point to ~1e-12 relative, far inside it, so only hit
// points whose exact position lies on a texel edge (under a mirror symmetry
// of the camera and the texture) are ties, and their texel is decided by the
// rounding of the reference's own operation order.
constexpr double kTieTexel = 1e-8;

__device__ __forceinline__ bool near_edge64(double v, double tol) {
  const double f = v - floor(v);
  return f < tol || f > 1.0 - tol;
}

// Equirectangular texel of a unit direction in FP64 (equirect64's mapping);
// *tie: a coordinate lies within kTieTexel of a texel edge.
__device__ inline uint32_t equirect64_tie(const DevTex& t, double dx, double dy, double dz,
                                          bool* tie) {
  const double z = dz < -1.0 ? -1.0 : (dz > 1.0 ? 1.0 : dz);
  const double c = (atan2(dy, dx) + kPiD) * (0.5 * kInvPiD) * t.cols;
  const double r = (kPiD / 2.0 - asin(z)) * kInvPiD * t.rows;
  *tie = near_edge64(c, kTieTexel) || near_edge64(r, kTieTexel);
  return texel(t, tex_index(floor(c), t.cols), tex_index(floor(r), t.rows));
}

// Orbit-plane basis of pixel (x, y) in FP64 (setup_ray keeps e2 in FP32):
// e1 = sign(Y0) F/|F|, e2 = sign(Y0) ((dv.F) F - r0^2 dv) / (r0 |pv x F|).
__device__ inline void basis64(const DevScene& s, int x, int y, const RaySetup& r, double e1[3],
                               double e2[3]) {
  const double sx = s.W / 2.0 - x, sy = s.H / 2.0 - y;
  const double pv[3] = {__builtin_fma(-s.vz[0], sy, __builtin_fma(-s.vy[0], sx, s.fvB[0])),
                        __builtin_fma(-s.vz[1], sy, __builtin_fma(-s.vy[1], sx, s.fvB[1])),
                        __builtin_fma(-s.vz[2], sy, __builtin_fma(-s.vy[2], sx, s.fvB[2]))};
  const double sg = (double)r.sgn;
  const double ke = sg * s.u0 / sqrt(r.cc);
  for (int i = 0; i < 3; ++i) {
    e1[i] = sg * s.F[i] * s.u0;
    e2[i] = __builtin_fma(r.dF, s.F[i], -s.r02 * (s.F[i] - pv[i])) * ke;
  }
}

// The disc crossing angles of an FP64 lane from the FP64 basis (setup_lane
// finds them with an FP32 atan2, which would limit the crossing point to
// FP32 accuracy): n.(P - B) ~ al cos a + be sin a vanishes at
// a = atan2(be, al) + pi/2 (mod pi).
__device__ inline void disc_events64(const DevScene& s, const double e1[3], const double e2[3],
                                     Lane<double>& L) {
  L.chi0s = 1e300;
  L.a_ev = 1e300;
  if (!s.has_disc) return;
  const double al = s.dn[0] * e1[0] + s.dn[1] * e1[1] + s.dn[2] * e1[2];
  const double be = s.dn[0] * e2[0] + s.dn[1] * e2[1] + s.dn[2] * e2[2];
  if (!(al * al + be * be > 0.0)) return;
  L.chi0s = atan2(be, al) + kPiD / 2.0;
  L.a_ev = next_event(0.0, L.chi0s, L.guard);
}

// Hit point, colour and output planes of pixel idx marched in FP64, with the
// ray's FP64 basis.  Returns true for a tie (a texel coordinate on a texel
// edge to FP64 accuracy): the caller then renders the pixel with the
// reference march.
template <bool ESCAPE>
__device__ inline bool finish_ray64(const DevScene& s, const DevOut& out, size_t idx,
                                    const double e1[3], const double e2[3], const RaySetup& r,
                                    const MarchOut<double>& mo) {
  uint32_t rgb = 0;
  bool tie = false;
  const double nan = __builtin_nan("");
  double q[3] = {nan, nan, nan};
  if (mo.hit.cls != BH_HIT_NONE) {
    double sa, ca;
    sincos_small64(mo.hit.ang, &sa, &ca);
    const double X = mo.hit.hx * ca - mo.hit.hy * sa;
    const double Y = mo.hit.hx * sa + mo.hit.hy * ca;
    for (int i = 0; i < 3; ++i) q[i] = __builtin_fma(Y, e2[i], __builtin_fma(X, e1[i], s.B[i]));
    if (mo.hit.cls == BH_HIT_DISC) {
      // Rectangle::color (vector_object.h:163-179) in closed form
      const DevTex& t = s.disc_tex;
      double s1[3], s2[3], v[3];
      for (int i = 0; i < 3; ++i) {
        s1[i] = s.dv2[i] - s.dv1[i];
        s2[i] = s.dv4[i] - s.dv1[i];
        v[i] = q[i] - s.dv1[i];
      }
      const double l11 = s1[0] * s1[0] + s1[1] * s1[1] + s1[2] * s1[2];
      const double l22 = s2[0] * s2[0] + s2[1] * s2[1] + s2[2] * s2[2];
      const double cx = s1[1] * v[2] - s1[2] * v[1];
      const double cy = s1[2] * v[0] - s1[0] * v[2];
      const double cz = s1[0] * v[1] - s1[1] * v[0];
      const double a = (s1[0] * v[0] + s1[1] * v[1] + s1[2] * v[2]) / l11 * t.cols;
      const double b = sqrt((cx * cx + cy * cy + cz * cz) / (l11 * l22)) * t.rows;
      tie = near_edge64(a, kTieTexel) || near_edge64(b, kTieTexel);
      rgb = texel(t, tex_index(a, t.cols), tex_index(b, t.rows));
    } else if (mo.hit.cls == BH_HIT_OBJECT) {
      const double iR = 1.0 / s.sr[mo.hit.obj];
      const double* c = s.sc[mo.hit.obj];
      rgb = equirect64_tie(s.stex[mo.hit.obj], (q[0] - c[0]) * iR, (q[1] - c[1]) * iR,
                           (q[2] - c[2]) * iR, &tie);
    }
  } else if (ESCAPE && s.has_bg) {
    // direction of the last chord; outbound points are one du apart, the
    // last one at sigma = 0.1 + steps - nstep
    const double kb = (double)s.nstep - (0.1 + (double)(mo.steps - s.nstep));
    const double rb = 1.0 / __builtin_fma(kb, mo.du64, s.u0);
    const double ra = 1.0 / __builtin_fma(kb + 1.0, mo.du64, s.u0);
    double sb, cb, sa, ca;
    sincos_small64(mo.a, &sb, &cb);
    sincos_small64(mo.a - mo.d, &sa, &ca);
    const double X = rb * cb - ra * ca, Y = rb * sb - ra * sa;
    double d[3];
    for (int i = 0; i < 3; ++i) d[i] = X * e1[i] + Y * e2[i];
    const double inv = 1.0 / sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    rgb = equirect64_tie(s.bg, d[0] * inv, d[1] * inv, d[2] * inv, &tie);
  }
  if (out.rgb) {
    out.rgb[3 * idx + 0] = (uint8_t)(rgb & 0xff);
    out.rgb[3 * idx + 1] = (uint8_t)((rgb >> 8) & 0xff);
    out.rgb[3 * idx + 2] = (uint8_t)((rgb >> 16) & 0xff);
  }
  if (out.hit) out.hit[idx] = (uint8_t)mo.hit.cls;
  if (out.steps) out.steps[idx] = mo.steps;
  if (out.hitpt) {
    out.hitpt[3 * idx + 0] = q[0];
    out.hitpt[3 * idx + 1] = q[1];
    out.hitpt[3 * idx + 2] = q[2];
  }
  if (out.phi) out.phi[idx] = atan(sqrt(r.cc) / r.dF) + mo.a;
  return tie;
}

// FP64 refinement of a texel-edge ray whose FP32 march ended at step n with
// class cls (a disc hit, or no hit in an escape scene): the swept angle of
// steps 1..n in FP64 with no object tests (the same work on every lane, so
// the rays of a wave, taken from all over the frame, do not diverge), then
// the last chord's disc test in FP64.  Returns false when that outcome does
// not hold in FP64 (the caller then runs the FP64 event march).  The earlier
// chords are taken to hit nothing, as the FP32 march found: an FP32 miss of
// an earlier hit needs a crossing within FP32 rounding of the disc's rim,
// independent of the texel-edge condition that queued the ray.
template <bool ESCAPE>
__device__ inline bool refine64(const DevScene& s, Lane<double> L, const RaySetup& r, double du,
                                int cls, int n, MarchOut<double>& o) {
  const int nstep = s.nstep;
  if (n < 1 || (r.captured && n > nstep)) return false;
  if (cls != BH_HIT_DISC && !(ESCAPE && cls == BH_HIT_NONE)) return false;
  double w = 0.0, a = 0.0, a_prev = 0.0, d = 0.0, ka = 0.0, kb = 0.0;
  // the steps in the order of the march, one phase (one step cubic) at a
  // time; within a phase two steps' G and w are formed together, so that
  // their dependent FP64 chains overlap (the sums keep their order)
  auto add = [&](double wn, double ka_n, double kb_n) {
    d = (w + wn) * L.h;
    w = wn;
    a_prev = a;
    a += d;
    ka = ka_n;
    kb = kb_n;
  };
  auto run = [&](int k0, int k1, auto coeffs) {
    int k = k0;
    for (; k + 1 <= k1; k += 2) {
      double G1, ka1, kb1, G2, ka2, kb2;
      coeffs(k, &G1, &ka1, &kb1);
      coeffs(k + 1, &G2, &ka2, &kb2);
      const double w1 = rsq<double>(G1), w2 = rsq<double>(G2);
      add(w1, ka1, kb1);
      add(w2, ka2, kb2);
    }
    if (k <= k1) {
      double G1, ka1, kb1;
      coeffs(k, &G1, &ka1, &kb1);
      add(rsq<double>(G1), ka1, kb1);
    }
  };
  run(1, min(n, nstep - 1), [&](int k, double* G, double* ka_n, double* kb_n) {
    coeffs_in<double>(k, nstep, L, G, ka_n, kb_n);
  });
  if (n >= nstep) {
    to_outbound<double>(L);
    double G1, ka1, kb1;
    coeffs_final<double>(nstep, L, &G1, &ka1, &kb1);
    add(rsq<double>(G1), ka1, kb1);
  }
  run(nstep + 1, n, [&](int k, double* G, double* ka_n, double* kb_n) {
    coeffs_out<double>(k, nstep, L, G, ka_n, kb_n);
  });
  o.steps = n;
  o.a = a;
  o.d = d;
  o.du64 = du;
  o.hit.cls = BH_HIT_NONE;
  o.hit.obj = 0;
  o.hit.ang = 0.0;
  o.hit.hx = o.hit.hy = 0.0;
  if (cls == BH_HIT_NONE) return true;
  // the first chord of a ray with Y0 < 0 starts at the camera (angle pi)
  const double aa = n == 1 ? r.first_aa : a_prev;
  double d2;
  return event_chord<double, DevScene>(s, L, ka, aa, kb, a, a - aa, false, true, o.hit, &d2) &&
         o.hit.cls == BH_HIT_DISC;
}

// FP64 march of one queued ray: a texel-edge ray's FP32 outcome refined in
// FP64 (refine64), or, for a band ray and when the refinement fails, the event
// march in FP64; both with an FP64 basis, crossing angles and finish.  A texel
// tie, whose texel the reference's rounding order decides, is replayed in
// that order (replay64) for the outcome found.  The reference march itself
// (march64_kernel's per-pixel code) serves a degenerate ray, a replay that
// disagrees, and scenes with spheres (the event march tests spheres with an
// FP32 table).
template <bool ESCAPE, bool SPH>
__device__ inline void fp64_ray(const DevScene& s, const DevOut& out, size_t idx, int x, int y,
                                uint32_t outcome) {
  Pixel64 px;
  RaySetup r;
  setup_ray(s, x, y, false, r);
  if (!r.degenerate) {
    double e1[3], e2[3];
    basis64(s, x, y, r, e1, e2);
    Lane<double> L;
    const double du = (r.l - s.u0) * s.inv_nstep;
    setup_lane<double>(s, r.l, r.ib2, du, r.e2x, r.e2y, r.e2z, r.sgn, L);
    disc_events64(s, e1, e2, L);
    // more spheres in the orbit plane than the slots hold: reference march
    bool ok = true;
    if constexpr (SPH) ok = setup_spheres<double>(s, e1, e2, L);
    if (ok) {
      MarchOut<double> mo;
      // spheres: the FP32 outcome is not refined (a sphere may have been
      // grazed before the step it names)
      if (SPH || !(outcome != 0u && refine64<ESCAPE>(s, L, r, du, (int)(outcome >> 28),
                                                     (int)(outcome & 0x0fffffffu), mo)))
        march<double, ESCAPE, SPH>(s, L, r.captured, r.first_aa, du, mo);
      if (!finish_ray64<ESCAPE>(s, out, idx, e1, e2, r, mo)) return;
      if (replay64(s, x, y, mo.steps, mo.hit.cls, px)) {
        store_pixel64(out, idx, px);
        return;
      }
    }
  }
  trace64(s, x, y, px);
  store_pixel64(out, idx, px);
}

}  // namespace bh

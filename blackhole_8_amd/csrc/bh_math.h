// bh_math.h — small device math shared by the march kernels: precision-
// generic wrappers of the hardware reciprocal / rsqrt / sqrt, FP32 and FP64
// sin/cos of bounded arguments, an FP32 atan2, and forward-mode dual numbers.
#pragma once

#include "bh_fp64.h"

namespace bh {

constexpr double kPiD = 3.141592653589793238462643383279;
constexpr double kInvPiD = 0.318309886183790671537767526745;

// ---------------------------------------------------------------------------
// small math
// ---------------------------------------------------------------------------
template <typename T> __device__ __forceinline__ T rsq(T x);
template <> __device__ __forceinline__ float rsq<float>(float x) { return __builtin_amdgcn_rsqf(x); }
// FP64 1/sqrt: hardware estimate refined by two Newton steps y (3 - x y^2)/2
template <> __device__ __forceinline__ double rsq<double>(double x) {
  double y = __builtin_amdgcn_rsq(x);
  y = y * __builtin_fma(-0.5 * x, y * y, 1.5);
  y = y * __builtin_fma(-0.5 * x, y * y, 1.5);
  return y;
}
template <typename T> __device__ __forceinline__ T rcp(T x);
template <> __device__ __forceinline__ float rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <> __device__ __forceinline__ double rcp<double>(double x) { return 1.0 / x; }
template <typename T> __device__ __forceinline__ T sqrt_t(T x);
template <> __device__ __forceinline__ float sqrt_t<float>(float x) { return __builtin_amdgcn_sqrtf(x); }
template <> __device__ __forceinline__ double sqrt_t<double>(double x) { return sqrt(x); }
template <typename T> __device__ __forceinline__ T fma_t(T a, T b, T c);
template <> __device__ __forceinline__ float fma_t<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <> __device__ __forceinline__ double fma_t<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }

// sin and cos of r, |r| <= pi/4 + a little: Taylor series to r^9 / r^10
// (truncation below 2e-9 on that interval), evaluated by Horner.
__device__ __forceinline__ void sincos_core(float r, float* s, float* c) {
  const float r2 = r * r;
  float ps = __builtin_fmaf(r2, 2.7557319e-6f, -1.9841270e-4f);  // 1/9!, -1/7!
  ps = __builtin_fmaf(ps, r2, 8.3333333e-3f);                      // 1/5!
  ps = __builtin_fmaf(ps, r2, -1.6666667e-1f);                     // -1/3!
  *s = __builtin_fmaf(ps * r2, r, r);
  float pc = __builtin_fmaf(r2, -2.7557319e-7f, 2.4801587e-5f);   // -1/10!, 1/8!
  pc = __builtin_fmaf(pc, r2, -1.3888889e-3f);                     // -1/6!
  pc = __builtin_fmaf(pc, r2, 4.1666667e-2f);                      // 1/4!
  pc = __builtin_fmaf(pc, r2, -0.5f);
  *c = __builtin_fmaf(pc, r2, 1.0f);
}

__device__ __forceinline__ void quadrant(int q, float sr, float cr, float* s, float* c) {
  const float ss = (q & 1) ? cr : sr;
  const float cc = (q & 1) ? sr : cr;
  *s = (q & 2) ? -ss : ss;
  *c = ((q + 1) & 2) ? -cc : cc;
}

// FP32 sin/cos for |x| up to a few times 2 pi: quadrant reduction with a
// two-part pi/2 (Cody and Waite), then the core series.
__device__ __forceinline__ void sincos_small(float x, float* s, float* c) {
  const float k = __builtin_rintf(x * 0.63661977f);
  float r = __builtin_fmaf(-k, 1.5707964f, x);
  r = __builtin_fmaf(-k, -4.3711388e-8f, r);
  float sr, cr;
  sincos_core(r, &sr, &cr);
  quadrant((int)k & 3, sr, cr, s, c);
}

// sin/cos of an FP64 angle of any size: reduce in FP64, evaluate in FP32.
__device__ __forceinline__ void sincos_angle(double a, float* s, float* c) {
  const double k = rint(a * (2.0 * kInvPiD));
  const float r = (float)__builtin_fma(-k, kPiD / 2.0, a);
  float sr, cr;
  sincos_core(r, &sr, &cr);
  quadrant((int)(int64_t)k & 3, sr, cr, s, c);
}

// FP64 sin/cos for |x| up to a few times 2 pi: three-part pi/2 reduction
// (Cody and Waite), Taylor series to r^21 / r^22 on |r| <= pi/4.
__device__ inline void sincos_small64(double x, double* s, double* c) {
  const double k = rint(x * (2.0 * kInvPiD));
  double r = __builtin_fma(-k, 1.5707963267341256, x);
  r = __builtin_fma(-k, 6.077100506303966e-11, r);
  r = __builtin_fma(-k, 2.0222662487959506e-21, r);
  const double r2 = r * r;
  double ps = 1.9572941063391263e-20;   // 1/21!
  ps = __builtin_fma(ps, r2, -8.22063524662433e-18);
  ps = __builtin_fma(ps, r2, 2.8114572543455206e-15);
  ps = __builtin_fma(ps, r2, -7.647163731819816e-13);
  ps = __builtin_fma(ps, r2, 1.6059043836821613e-10);
  ps = __builtin_fma(ps, r2, -2.505210838544172e-08);
  ps = __builtin_fma(ps, r2, 2.7557319223985893e-06);
  ps = __builtin_fma(ps, r2, -0.0001984126984126984);
  ps = __builtin_fma(ps, r2, 0.008333333333333333);
  ps = __builtin_fma(ps, r2, -0.16666666666666666);
  const double sr = __builtin_fma(ps * r2, r, r);
  double pc = -8.896791392450574e-22;  // -1/22!
  pc = __builtin_fma(pc, r2, 4.110317623312165e-19);
  pc = __builtin_fma(pc, r2, -1.5619206968586225e-16);
  pc = __builtin_fma(pc, r2, 4.779477332387385e-14);
  pc = __builtin_fma(pc, r2, -1.1470745597729725e-11);
  pc = __builtin_fma(pc, r2, 2.08767569878681e-09);
  pc = __builtin_fma(pc, r2, -2.755731922398589e-07);
  pc = __builtin_fma(pc, r2, 2.48015873015873e-05);
  pc = __builtin_fma(pc, r2, -0.001388888888888889);
  pc = __builtin_fma(pc, r2, 0.041666666666666664);
  pc = __builtin_fma(pc, r2, -0.5);
  const double cr = __builtin_fma(pc, r2, 1.0);
  const int q = (int)(int64_t)k & 3;
  const double ss = (q & 1) ? cr : sr;
  const double cs = (q & 1) ? sr : cr;
  *s = (q & 2) ? -ss : ss;
  *c = ((q + 1) & 2) ? -cs : cs;
}

template <typename T> __device__ __forceinline__ void sincos_t(T x, T* s, T* c);
template <> __device__ __forceinline__ void sincos_t<float>(float x, float* s, float* c) { sincos_small(x, s, c); }
template <> __device__ __forceinline__ void sincos_t<double>(double x, double* s, double* c) { sincos_small64(x, s, c); }
// sin/cos of a chord's span and the angles within it (|x| < pi).  FP32
// lanes use the hardware sine and cosine (argument in revolutions): within
// 1.4e-7 of FP64 for |x| <= 1 and 2.7e-7 up to pi, the sine within 1e-6 of
// itself at small angles (measured with bh_chord_sincos; the GPU test holds
// these bounds).  Both run on the transcendental unit, beside the VALU work.
template <typename T> __device__ __forceinline__ void sincos_chord(T x, T* s, T* c);
template <> __device__ __forceinline__ void sincos_chord<float>(float x, float* s, float* c) {
  const float r = x * 0.15915494309189535f;  // 1 / (2 pi)
  *s = __builtin_amdgcn_sinf(r);
  *c = __builtin_amdgcn_cosf(r);
}
template <> __device__ __forceinline__ void sincos_chord<double>(double x, double* s, double* c) { sincos_small64(x, s, c); }

// atan2 in FP32 to ~1e-7 rad: octant reduction and an odd polynomial for
// atan on [0, 1] (coefficients from a least-squares fit at Chebyshev nodes;
// fit error 3e-8).
__device__ inline float atan2_fast(float y, float x) {
  const float ax = fabsf(x), ay = fabsf(y);
  const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
  const float t = mx > 0.f ? mn * __builtin_amdgcn_rcpf(mx) : 0.f;
  const float t2 = t * t;
  float p = 0.0024682467f;
  p = __builtin_fmaf(p, t2, -0.014458697f);
  p = __builtin_fmaf(p, t2, 0.039899562f);
  p = __builtin_fmaf(p, t2, -0.072479509f);
  p = __builtin_fmaf(p, t2, 0.10507320f);
  p = __builtin_fmaf(p, t2, -0.14164333f);
  p = __builtin_fmaf(p, t2, 0.19986537f);
  p = __builtin_fmaf(p, t2, -0.33332658f);
  p = __builtin_fmaf(p, t2, 0.99999988f);
  float r = p * t;
  if (ay > ax) r = 1.57079633f - r;
  if (x < 0.f) r = 3.14159265f - r;
  return y < 0.f ? -r : r;
}

}  // namespace bh

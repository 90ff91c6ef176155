// bh_march.hip — CDNA4 (gfx950) kernels for the per-pixel photon march.
//
//  march64_kernel  FP64, reference operation order, one lane per pixel
//                  (include/blackhole/blackhole_solution_test.cc:161-298).
//  march32_kernel  FP32 event march: FP64 set-up, a 3-FMA + 1-rsqrt step in
//                  the orbit plane, exact object tests only on the steps
//                  where an object can be reached, FP64 band queue.
//  band64_kernel   drains the band queue with the FP64 march.
#include "bh_fp64.h"

namespace bh {

constexpr double kPiD = 3.141592653589793238462643383279;

// ---------------------------------------------------------------------------
// FP64 reference march
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void march64_kernel(const DevScene* __restrict__ sp,
                                                      DevOut out) {
  const DevScene& s = *sp;
  const int x = blockIdx.x * 16 + (threadIdx.x & 15);
  const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= s.W || y >= s.H) return;
  Pixel64 px;
  trace64(s, x, y, px);
  store_pixel64(out, (size_t)y * s.W + x, px);
}

__global__ __launch_bounds__(64) void band64_kernel(const DevScene* __restrict__ sp, DevOut out,
                                                    const uint32_t* __restrict__ items,
                                                    const uint32_t* __restrict__ count,
                                                    uint32_t capacity) {
  const DevScene& s = *sp;
  const uint32_t n = min(*count, capacity);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t idx = items[i];
    const int x = (int)(idx % (uint32_t)s.W), y = (int)(idx / (uint32_t)s.W);
    Pixel64 px;
    trace64(s, x, y, px);
    store_pixel64(out, idx, px);
  }
}

__global__ void zero_u32_kernel(uint32_t* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
// FP32 event march
// ---------------------------------------------------------------------------
//
// Orbit-plane reduction.  The reference places step k at
//   P_k = B + r_k (cos phi_k * yv + sin phi_k * xv),  r_k = 1/u_k,
// so every chord lies in the plane through B spanned by (yv, xv).  With
// e1 = cos phi0 yv + sin phi0 xv = sign(Y0) F/|F| and e2 = e1 rotated by +90
// degrees in that plane, P_k = B + r_k (cos a_k e1 + sin a_k e2) where
// a_k = phi_k - phi0 is the angle swept since the camera.
//
// G on the march.  With l the turning point and s = l - u,
//   G(u) = G(l) + s [2 l d + s (1 - 2 d) - 2 M s^2],  d = 1 - 3 M l,
// a cubic in s whose terms do not cancel away from the photon sphere; the
// coefficients come from FP64.  The reference's u_k = u0 + kappa_k du gives
// s_k = (nstep - kappa_k) du exactly, so each step is s = sigma * du, three
// FMAs, one rsqrt, and the trapezoid update of a (accumulated in FP64).
//
// Events.  A chord can only reach the disc plane if sign(n . (P - B)) =
// sign(cos(a - psi)) changes, i.e. a crosses chi_j = psi + pi/2 + j pi; the
// step is tested exactly when a passes the next chi_j (minus a guard for the
// FP32 angle).  A chord between radii r_a, r_b >= r_min subtending Da stays at
// distance >= r_min cos(Da/2) from B, so it can only reach the horizon when
// Da > 2 acos(R_h / r_min); that step is tested exactly too.  Every other
// step provably hits nothing.
struct Lane32 {
  float c0, c1, c2, c3;  // G(s) cubic
  float duf, h;          // du and du/2 in FP32
  float phi_safe;        // horizon guard angle
  float guard;           // disc crossing-angle guard
  double du64;           // du in FP64 (event path positions)
  double chi0;           // first disc crossing angle (chi_0 = psi + pi/2)
  double a_ev;           // next disc event angle (crossing - guard)
};

struct Hit32 {
  int cls;      // BH_HIT_*
  double ang;   // swept angle of the hit point
  float rho;    // distance of the hit point from B
};

// Swept angle of the crossing chi_j nearest below or at a, and the count of
// crossings in (a_lo, a_hi].
__device__ __forceinline__ double crossing_floor(double a, double chi0) {
  return floor((a - chi0) * (1.0 / kPiD));
}

// Exact tests of the chord between step (sa_, aa) and step (sb_, ab), where
// sigma = nstep - kappa is the distance to the turning point in units of du
// (u = u0 + kappa du) and a the swept angle.  Returns true on a hit.
__device__ inline bool event32(const DevScene& s, const Lane32& L, float sa_, double aa,
                               float sb_, double ab, float da, bool test_h, bool test_d,
                               Hit32& hit) {
  const double ua = s.u0 + ((double)s.nstep - (double)sa_) * L.du64;
  const double ub = s.u0 + ((double)s.nstep - (double)sb_) * L.du64;
  const float ra = 1.0f / (float)ua;
  const float rb = 1.0f / (float)ub;
  bool found = false;
  float best = 0.f;
  const float R = (float)s.Rh;
  if (test_h) {
    // frame rotated by aa: P_a = (ra, 0), P_b = rb (cos da, sin da)
    float sd, cd;
    sincosf(da, &sd, &cd);
    const float qx = rb * cd - ra, qy = rb * sd;
    const float qd = qx * ra;
    const float qq = qx * qx + qy * qy;
    if (!(ra < R && rb < R)) {
      const float cc = qd * qd - qq * (ra * ra - R * R);
      if (cc > 0.f) {
        const float t = (-qd - sqrtf(cc)) / qq;
        if (t > 0.f && t < 1.f) {
          const float hx = ra + t * qx, hy = t * qy;
          found = true;
          best = t * t * qq;
          hit.cls = BH_HIT_HORIZON;
          hit.ang = aa + (double)atan2f(hy, hx);
          hit.rho = sqrtf(hx * hx + hy * hy);
        }
      }
    }
  }
  if (test_d) {
    const double fa = crossing_floor(aa, L.chi0);
    const double fb = crossing_floor(ab, L.chi0);
    const double n = fb - fa;
    if (fmod(n, 2.0) != 0.0) {
      // crossing chi_k inside (aa, ab]; reference interpolation on the chord
      // with weights |t_b|, |t_a|, t = n.(P - B) ~ r sin(a - chi_k)
      const double chik = L.chi0 + fb * kPiD;
      const float xa = (float)(aa - chik), xb = (float)(ab - chik);
      float sa, ca, sb, cb;
      sincosf(xa, &sa, &ca);
      sincosf(xb, &sb, &cb);
      const float wa = fabsf(ra * sa), wb = fabsf(rb * sb);  // |t_a|, |t_b|
      const float inv = 1.0f / (wa + wb);
      // crossing point in the frame whose x axis points along chi_k
      const float px = (wb * ra * ca + wa * rb * cb) * inv;
      const float py = (wb * ra * sa + wa * rb * sb) * inv;
      const float rho = sqrtf(px * px + py * py);
      if (!(rho > (float)s.r_out) && !(rho < (float)s.r_in)) {
        const float dx = ra * ca - px, dy = ra * sa - py;
        const float d2 = dx * dx + dy * dy;
        if (!found || d2 < best) {
          found = true;
          best = d2;
          hit.cls = BH_HIT_DISC;
          hit.ang = chik + (double)atan2f(py, px);
          hit.rho = rho;
        }
      }
    }
  }
  return found;
}

// Disc texel of a world point (Rectangle::color, vector_object.h:163-179), with
// r cos(theta) = s1.v/|s1| and r sin(theta) = |s1 x v|/|s1| in place of the
// acos round trip.
__device__ inline uint32_t disc_color32(const DevScene& s, float qx, float qy, float qz) {
  const DevTex& t = s.disc_tex;
  if (!t.data) return 0;
  const float v1x = (float)s.dv1[0], v1y = (float)s.dv1[1], v1z = (float)s.dv1[2];
  const float s1x = (float)(s.dv2[0] - s.dv1[0]), s1y = (float)(s.dv2[1] - s.dv1[1]),
              s1z = (float)(s.dv2[2] - s.dv1[2]);
  const float s2x = (float)(s.dv4[0] - s.dv1[0]), s2y = (float)(s.dv4[1] - s.dv1[1]),
              s2z = (float)(s.dv4[2] - s.dv1[2]);
  const float vx = qx - v1x, vy = qy - v1y, vz = qz - v1z;
  const float l1 = sqrtf(s1x * s1x + s1y * s1y + s1z * s1z);
  const float l2 = sqrtf(s2x * s2x + s2y * s2y + s2z * s2z);
  const float proj = (s1x * vx + s1y * vy + s1z * vz) / l1;
  const float cx = s1y * vz - s1z * vy, cy = s1z * vx - s1x * vz, cz = s1x * vy - s1y * vx;
  const float perp = sqrtf(cx * cx + cy * cy + cz * cz) / l1;
  const float a = (proj / l1) * (float)t.cols;
  const float b = (perp / l2) * (float)t.rows;
  if (s.filter == BH_FILTER_BILINEAR) return bilinear64(t, (double)a, (double)b);
  return texel(t, tex_index((double)a, t.cols), tex_index((double)b, t.rows));
}

__device__ inline uint32_t equirect32(const DevTex& t, float dx, float dy, float dz) {
  const float kPi = 3.14159265358979323846f;
  const float z = fminf(1.f, fmaxf(-1.f, dz));
  const float lon = atan2f(dy, dx);
  const float lat = asinf(z);
  const float c = (lon + kPi) * (1.0f / (2.0f * kPi)) * (float)t.cols;
  const float r = (kPi / 2.0f - lat) * (1.0f / kPi) * (float)t.rows;
  return texel(t, tex_index((double)floorf(c), t.cols), tex_index((double)floorf(r), t.rows));
}

template <bool ESCAPE>
__global__ __launch_bounds__(256) void march32_kernel(const DevScene* __restrict__ sp,
                                                      DevOut out, DevQueue q) {
  const DevScene& s = *sp;
  const int x = blockIdx.x * 16 + (threadIdx.x & 15);
  const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= s.W || y >= s.H) return;
  const size_t idx = (size_t)y * s.W + x;

  // ---- FP64 set-up ------------------------------------------------------
  const d3 B = ld3(s.B);
  const d3 F = ld3(s.F);
  const d3 pv = pixel_point(s, x, y) - B;
  const d3 dv = F - pv;                  // along yv
  const d3 cv = cross(pv, F);            // |cv| = b |dv|
  const double cc = dot(cv, cv);
  const double dd = dot(dv, dv);
  const double b = sqrt(cc / dd);
  const double M = s.M;
  const bool captured = b < s.bc;
  // photon-sphere band and degenerate rays go to the FP64 queue
  const double dF = dot(dv, F);
  // rays in the photon-sphere band, degenerate rays, and rays whose start
  // angle atan(b/Y0) is not the camera's own direction (Y0 <= 0, quirk Q2)
  // go to the FP64 queue
  if (fabs(b / s.bc - 1.0) < s.band || cc == 0.0 || dd == 0.0 || !(dF > 0.0)) {
    const uint32_t slot = atomicAdd(q.count, 1u);
    if (slot < q.capacity) q.items[slot] = (uint32_t)idx;
    return;
  }
  const double l = captured ? 1.0 / (3 * M) : solve_g64(b, M, s.l0, s.root);
  const double du = (l - s.u0) * (1.0 / s.nstep);
  const double dl = 1.0 - 3.0 * M * l;
  Lane32 L;
  L.c0 = (float)G64(l, b, M);
  L.c1 = (float)(2.0 * l * dl);
  L.c2 = (float)(1.0 - 2.0 * dl);
  L.c3 = (float)(-2.0 * M);
  L.duf = (float)du;
  L.h = (float)(du / 2.);
  L.du64 = du;
  const double r0 = 1.0 / s.u0;
  // Y0 > 0 here, so e1 = F / r0 and e2 = ((dv.F) F - r0^2 dv) / (r0 |cv|)
  const double ic = 1.0 / (r0 * sqrt(cc));
  const float e1x = (float)(F.x / r0), e1y = (float)(F.y / r0), e1z = (float)(F.z / r0);
  const float e2x = (float)((dF * F.x - r0 * r0 * dv.x) * ic);
  const float e2y = (float)((dF * F.y - r0 * r0 * dv.y) * ic);
  const float e2z = (float)((dF * F.z - r0 * r0 * dv.z) * ic);
  // horizon guard from the smallest radius on the march, u = u0 + (n - 0.1) du
  {
    const double umax = s.u0 + ((double)s.nstep - 0.1) * du;
    const float cr = (float)(s.Rh * umax);
    L.phi_safe = cr < 1.f ? 2.0f * acosf(fminf(1.f, cr * 1.0001f)) : 0.f;
  }
  // disc crossing angles: n.(P-B) = r (al cos a + be sin a)
  L.chi0 = 0.0;
  L.a_ev = 1e300;
  L.guard = 0.f;
  if (s.has_disc) {
    const float nx = (float)s.dn[0], ny = (float)s.dn[1], nz = (float)s.dn[2];
    const float al = nx * e1x + ny * e1y + nz * e1z;
    const float be = nx * e2x + ny * e2y + nz * e2z;
    const float amp = sqrtf(al * al + be * be);
    if (amp > 0.f) {
      const double psi = (double)atan2f(be, al);
      L.chi0 = psi + kPiD / 2.0;
      L.guard = 4e-6f + 2e-7f / amp;
      const double j = floor((0.0 - (double)L.guard - L.chi0) * (1.0 / kPiD)) + 1.0;
      L.a_ev = L.chi0 + j * kPiD - (double)L.guard;
    }
  }

  // ---- FP32 march ---------------------------------------------------------
  // Step argument sigma: s = l - u = sigma du.  Inbound sigma = nstep - i,
  // final approach 0.1, outbound 0.1 + j.
  const int nstep = s.nstep;
  float wprev = 0.f;
  double a = 0.0;          // swept angle
  float da = 0.f;          // last increment
  float sprev = (float)nstep;  // the camera sits at sigma = nstep
  int steps = 0;
  bool alive = true;
  Hit32 hit;
  hit.cls = BH_HIT_NONE;
  hit.ang = 0.0;
  hit.rho = 0.f;

  auto step = [&](float sg) {
    const float sv = sg * L.duf;
    const float G = fmaf(fmaf(fmaf(L.c3, sv, L.c2), sv, L.c1), sv, L.c0);
    const float w = __builtin_amdgcn_rsqf(G);
    const float d = (wprev + w) * L.h;
    const double an = a + (double)d;
    ++steps;
    const bool eh = d > L.phi_safe;
    const bool ed = an >= L.a_ev;
    if (eh || ed) {
      if (event32(s, L, sprev, a, sg, an, d, eh, ed, hit)) alive = false;
      if (ed && alive) {
        const double j = floor((an - (double)L.guard - L.chi0) * (1.0 / kPiD)) + 1.0;
        L.a_ev = L.chi0 + j * kPiD - (double)L.guard;
      }
    }
    a = an;
    da = d;
    wprev = w;
    sprev = sg;
  };

  // inbound: u = u0 + i du, i = 1 .. nstep-1 (blackhole_solution_test.cc:217-239)
  for (int i = 1; i < nstep && alive; ++i) step((float)(nstep - i));
  // final approach: u += 0.9 du, still weighted du/2 (:240-262)
  if (alive) step(0.1f);
  if (alive && captured) {
    // chord to the centre: the horizon is always entered (r > 3M > 2M) and
    // the disc test sees n.(B - B) = 0 (:264-272)
    hit.cls = BH_HIT_HORIZON;
    hit.ang = a;
    hit.rho = (float)s.Rh;
    alive = false;
  }
  // outbound: u -= du (:274-296)
  for (int j = 1; j < nstep && alive; ++j) step(0.1f + (float)j);
  if (ESCAPE) {
    // keep stepping outward until the next point would lie at r >= r_escape
    for (int j = nstep; alive && steps < s.step_cap; ++j) {
      const float sg = 0.1f + (float)j;
      const double un = s.u0 + ((double)nstep - (double)sg) * L.du64;
      if (un * s.r_esc <= 1.0) break;
      step(sg);
    }
  }

  // ---- finalisation -------------------------------------------------------
  uint32_t rgb = 0;
  float qx = __builtin_nanf(""), qy = qx, qz = qx;
  if (hit.cls != BH_HIT_NONE) {
    float sh, ch;
    sincosf((float)hit.ang, &sh, &ch);
    qx = (float)s.B[0] + hit.rho * (ch * e1x + sh * e2x);
    qy = (float)s.B[1] + hit.rho * (ch * e1y + sh * e2y);
    qz = (float)s.B[2] + hit.rho * (ch * e1z + sh * e2z);
    if (hit.cls == BH_HIT_DISC) rgb = disc_color32(s, qx, qy, qz);
  } else if (ESCAPE && s.has_bg) {
    // direction of the last chord
    // outbound steps are one du apart: the previous point sits at sigma - 1
    const float rb = 1.0f / (float)(s.u0 + ((double)nstep - (double)sprev) * L.du64);
    const float ra = 1.0f / (float)(s.u0 + ((double)nstep - (double)sprev + 1.0) * L.du64);
    float sb, cb, sa, ca;
    sincosf((float)a, &sb, &cb);
    sincosf((float)(a - (double)da), &sa, &ca);
    const float px = rb * cb - ra * ca, py = rb * sb - ra * sa;
    float dx = px * e1x + py * e2x, dy = px * e1y + py * e2y, dz = px * e1z + py * e2z;
    const float inv = rsqrtf(dx * dx + dy * dy + dz * dz);
    rgb = equirect32(s.bg, dx * inv, dy * inv, dz * inv);
  }

  if (out.rgb) {
    out.rgb[3 * idx + 0] = (uint8_t)(rgb & 0xff);
    out.rgb[3 * idx + 1] = (uint8_t)((rgb >> 8) & 0xff);
    out.rgb[3 * idx + 2] = (uint8_t)((rgb >> 16) & 0xff);
  }
  if (out.hit) out.hit[idx] = (uint8_t)hit.cls;
  if (out.steps) out.steps[idx] = steps;
  if (out.hitpt) {
    out.hitpt[3 * idx + 0] = (double)qx;
    out.hitpt[3 * idx + 1] = (double)qy;
    out.hitpt[3 * idx + 2] = (double)qz;
  }
  if (out.phi) {
    // phi0 = atan(b / Y0) (quirk Q2); b / Y0 = |cv| / (dv . F)
    out.phi[idx] = atan(sqrt(cc) / dF) + a;
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_march_fp64(const DevScene* scene, const DevOut& out, int W, int H,
                             hipStream_t stream) {
  dim3 grid((W + 15) / 16, (H + 15) / 16);
  hipLaunchKernelGGL(march64_kernel, grid, dim3(256), 0, stream, scene, out);
  return hipGetLastError();
}

hipError_t launch_march_fp32(const DevScene* scene, const DevOut& out, int W, int H,
                             const DevQueue& q, hipStream_t stream, bool escape) {
  dim3 grid((W + 15) / 16, (H + 15) / 16);
  if (escape)
    hipLaunchKernelGGL(march32_kernel<true>, grid, dim3(256), 0, stream, scene, out, q);
  else
    hipLaunchKernelGGL(march32_kernel<false>, grid, dim3(256), 0, stream, scene, out, q);
  return hipGetLastError();
}

hipError_t launch_band_fp64(const DevScene* scene, const DevOut& out, const DevQueue& q,
                            hipStream_t stream) {
  hipLaunchKernelGGL(band64_kernel, dim3(512), dim3(64), 0, stream, scene, out, q.items,
                     q.count, q.capacity);
  return hipGetLastError();
}

hipError_t launch_zero_u32(uint32_t* p, int n, hipStream_t stream) {
  hipLaunchKernelGGL(zero_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, p, n);
  return hipGetLastError();
}

}  // namespace bh

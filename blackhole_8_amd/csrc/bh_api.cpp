// bh_api.cpp — C ABI implementation: contexts, device textures, scene
// validation and upload, kernel launches, camera-path graphs, timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <new>
#include <condition_variable>
#include <mutex>
#include <string>
#include <vector>

#include "bh_api.h"
#include "bh_internal.h"

using bh::DevOut;
using bh::DevQueue;
using bh::DevScene;
using bh::DevTex;

struct DeviceTexture {
  uint32_t* d = nullptr;
  int rows = 0, cols = 0;
};

// Grow-only device buffer.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct bh_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;  // band kernels of captured paths
  bool own_stream = false;
  std::string err;
  DeviceTexture tex[BH_MAX_TEXTURES];
  DevBuf scene, rgb, hit, phi, hitpt, steps, qitems, qcount;
  DevBuf dpart, dgrad, drgb, dhit, dgin;  // differentiable render
  DevBuf bscenes, bouts, boff, bctl, bband, bbandc, bend, btile, btscene;  // batches
  DevBuf probe;   // issue-rate probe; turning points of bh_turning_points
  int wall_khz = 0;
  void* noalloc = nullptr;  // this device's allocation-free scratch array
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
  char name[256] = {0};
  int n_cu = 0, clock_khz = 0;
};

struct bh_path {
  bh_ctx* ctx = nullptr;
  int n_frames = 0, W = 0, H = 0;
  DevScene* d_scenes = nullptr;
  uint8_t* d_rgb = nullptr;
  uint8_t* d_hit = nullptr;
  uint64_t* d_qitems = nullptr;
  uint32_t* d_qcount = nullptr;
  uint32_t sub_cap = 0;              // FP64 queue of each frame: kQueueSubs x sub_cap items
  bool fused = false;                // all frames FP32: one march + one band launch
  bool esc = false, sph = false;     // scene class of the fused launches
  int nstep = 0;                     // the frames' step count (one options record)
  std::vector<hipEvent_t> events;
  // fused paths: the march over all frames and the FP64 pass over all frames
  // are two graphs, launched back to back with timing events around each
  // (HIP does not time events recorded inside a graph)
  hipEvent_t k0 = nullptr, k1 = nullptr, k2 = nullptr;
  hipGraph_t graph = nullptr, graph_band = nullptr;
  hipGraphExec_t exec = nullptr, exec_band = nullptr;
};

namespace {

const char* g_static_err = "";
thread_local std::string g_tls_err;

int fail(bh_ctx* ctx, int code, const std::string& msg) {
  if (ctx)
    ctx->err = msg;
  else
    g_tls_err = msg;
  return code;
}

// Records an error without letting a second exception escape.
int fail_nothrow(bh_ctx* ctx, int code, const char* msg) noexcept {
  try {
    fail(ctx, code, msg);
  } catch (...) {
  }
  return code;
}

// Body of a C entry point: no exception (host allocation, thread or mutex
// failure) crosses the C boundary.
template <class F>
int guarded(bh_ctx* ctx, F&& body) noexcept {
  try {
    return body();
  } catch (const std::bad_alloc&) {
    return fail_nothrow(ctx, BH_ERR_OUT_OF_MEMORY, "out of host memory");
  } catch (const std::exception& e) {
    return fail_nothrow(ctx, BH_ERR_INVALID_ARG, e.what());
  } catch (...) {
    return fail_nothrow(ctx, BH_ERR_INVALID_ARG, "unknown host error");
  }
}

int hip_fail(bh_ctx* ctx, hipError_t e, const char* what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  return fail(ctx, e == hipErrorOutOfMemory ? BH_ERR_OUT_OF_MEMORY : BH_ERR_HIP, m);
}

#define HIP_TRY(ctx, expr)                          \
  do {                                              \
    hipError_t e_ = (expr);                         \
    if (e_ != hipSuccess) return hip_fail(ctx, e_, #expr); \
  } while (0)

int ensure(bh_ctx* ctx, DevBuf& b, size_t bytes) {
  if (b.bytes >= bytes) return BH_OK;
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
  b.bytes = bytes;
  return BH_OK;
}

bool finite3(const double* v) {
  return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]);
}

DevTex devtex(const bh_ctx* ctx, int slot) {
  DevTex t{nullptr, 0, 0};
  if (slot >= 0 && slot < BH_MAX_TEXTURES && ctx->tex[slot].d) {
    t.data = ctx->tex[slot].d;
    t.rows = ctx->tex[slot].rows;
    t.cols = ctx->tex[slot].cols;
  }
  return t;
}

void copy3(double* dst, const double* src) { std::memcpy(dst, src, 3 * sizeof(double)); }

// Options with zeros replaced by the reference defaults.
bh_options normalized(const bh_options* o) {
  bh_options r;
  std::memset(&r, 0, sizeof r);
  if (o) r = *o;
  if (r.nstep == 0) r.nstep = 20;
  if (r.escape && r.step_cap == 0) r.step_cap = 10000;
  return r;
}

// Validate and resolve one scene into the device layout.
int build_scene(bh_ctx* ctx, const bh_scene* sc, const bh_options* opt_in, DevScene* d) {
  if (!sc) return fail(ctx, BH_ERR_INVALID_ARG, "scene is NULL");
  const bh_options o = normalized(opt_in);
  const bh_camera& c = sc->camera;
  if (c.width <= 0 || c.height <= 0 || c.width > 32768 || c.height > 32768)
    return fail(ctx, BH_ERR_INVALID_ARG, "camera width/height must be in 1..32768");
  if ((int64_t)c.width * c.height > (int64_t)1 << 30)
    return fail(ctx, BH_ERR_INVALID_ARG, "frame larger than 2^30 pixels");
  if (!finite3(c.position) || !finite3(c.axis_x) || !finite3(c.axis_y) || !finite3(c.axis_z) ||
      !std::isfinite(c.fov))
    return fail(ctx, BH_ERR_INVALID_ARG, "camera pose/fov must be finite");
  if (!(sc->hole.mass > 0.0) || !std::isfinite(sc->hole.mass) || !finite3(sc->hole.position))
    return fail(ctx, BH_ERR_INVALID_ARG, "hole mass must be positive and finite");
  if (o.precision != BH_PREC_FP64 && o.precision != BH_PREC_FP32)
    return fail(ctx, BH_ERR_INVALID_ARG, "unknown precision");
  if (o.nstep < 2 || o.nstep > 100000) return fail(ctx, BH_ERR_INVALID_ARG, "nstep must be in 2..100000");
  if (o.filter != BH_FILTER_NEAREST && o.filter != BH_FILTER_BILINEAR)
    return fail(ctx, BH_ERR_INVALID_ARG, "unknown filter");
  if (o.root != BH_ROOT_BISECT20 && o.root != BH_ROOT_EXACT)
    return fail(ctx, BH_ERR_INVALID_ARG, "unknown root mode");
  if (o.escape && (!(o.r_escape > 0.0) || o.step_cap <= 0 || o.step_cap > 1000000))
    return fail(ctx, BH_ERR_INVALID_ARG, "escape needs r_escape > 0 and step_cap in 1..1e6");
  if (!(o.band >= 0.0)) return fail(ctx, BH_ERR_INVALID_ARG, "band must be >= 0");
  if (!std::isfinite(o.edge_guard) || o.edge_guard > 0.01)
    return fail(ctx, BH_ERR_INVALID_ARG, "edge_guard must be finite and <= 0.01");
  if (sc->n_spheres < 0 || sc->n_spheres > BH_MAX_SPHERES)
    return fail(ctx, BH_ERR_INVALID_ARG, "n_spheres out of range");

  std::memset(d, 0, sizeof *d);
  copy3(d->C, c.position);
  copy3(d->vy, c.axis_y);
  copy3(d->vz, c.axis_z);
  const double f = bh_focal_length(c.width, c.fov);
  for (int i = 0; i < 3; ++i) d->fv[i] = c.axis_x[i] * f;  // Camera::focus_vector
  d->W = c.width;
  d->H = c.height;
  copy3(d->B, sc->hole.position);
  for (int i = 0; i < 3; ++i) d->F[i] = c.position[i] - sc->hole.position[i];
  d->M = sc->hole.mass;
  d->bc = 3.0 * std::sqrt(3.0) * d->M;  // b_c (blackhole_solution.h:25)
  d->Rh = 2 * d->M;                     // radius() (blackhole_solution.h:57)
  d->l0 = std::cbrt(DBL_EPSILON);       // epsilon<double>() (utility.h:17-21)
  const double r0 = std::sqrt(d->F[0] * d->F[0] + d->F[1] * d->F[1] + d->F[2] * d->F[2]);
  if (!(r0 > 0.0)) return fail(ctx, BH_ERR_INVALID_ARG, "camera coincides with the hole");
  d->u0 = 1.0 / r0;

  d->has_disc = sc->has_disc ? 1 : 0;
  if (d->has_disc) {
    const bh_annulus& a = sc->disc;
    for (int i = 0; i < 5; ++i)
      if (!finite3(a.vertex[i])) return fail(ctx, BH_ERR_INVALID_ARG, "disc vertices must be finite");
    copy3(d->dv0, a.vertex[0]);
    copy3(d->dv1, a.vertex[1]);
    copy3(d->dv2, a.vertex[2]);
    copy3(d->dv4, a.vertex[4]);
    copy3(d->dn, a.normal);
    d->r_in = a.r_inner;
    d->r_out = a.r_outer;
    d->disc_tex = devtex(ctx, a.texture);
    if (!d->disc_tex.data)
      return fail(ctx, BH_ERR_EMPTY_TEXTURE, "disc texture slot is empty (Material::SetTexture)");
  }
  d->n_spheres = sc->n_spheres;
  for (int i = 0; i < sc->n_spheres; ++i) {
    const bh_sphere& s = sc->spheres[i];
    if (!finite3(s.center) || !(s.radius > 0.0))
      return fail(ctx, BH_ERR_INVALID_ARG, "sphere centre/radius invalid");
    copy3(d->sc[i], s.center);
    d->sr[i] = s.radius;
    d->stex[i] = devtex(ctx, s.texture);
    if (!d->stex[i].data)
      return fail(ctx, BH_ERR_EMPTY_TEXTURE, "sphere texture slot is empty");
  }
  d->has_bg = 0;
  if (sc->background >= 0) {
    d->bg = devtex(ctx, sc->background);
    if (!d->bg.data) return fail(ctx, BH_ERR_EMPTY_TEXTURE, "background texture slot is empty");
    d->has_bg = 1;
  }
  d->nstep = o.nstep;
  d->escape = o.escape ? 1 : 0;
  d->step_cap = o.step_cap;
  d->filter = o.filter;
  d->root = o.root;
  d->precision = o.precision;
  d->r_esc = o.r_escape;
  d->band = o.band;
  d->edge = o.edge_width;

  // per-frame constants of the FP32 march
  for (int i = 0; i < 3; ++i) {
    d->fvB[i] = d->fv[i] - d->B[i];
    d->hW = d->W / 2.0;
    d->hH = d->H / 2.0;
    d->e1f[i] = (float)(d->F[i] / r0);
    d->nf[i] = (float)d->dn[i];
    d->Bf[i] = (float)d->B[i];
  }
  {
    auto cross = [](const double a[3], const double b[3], double r[3]) {
      r[0] = a[1] * b[2] - a[2] * b[1];
      r[1] = a[2] * b[0] - a[0] * b[2];
      r[2] = a[0] * b[1] - a[1] * b[0];
    };
    auto dot = [](const double a[3], const double b[3]) {
      return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    };
    for (int i = 0; i < 3; ++i) d->pxD[i] = d->F[i] - d->fvB[i];
    cross(d->fvB, d->F, d->pxC);
    cross(d->vy, d->F, d->pxCy);
    cross(d->vz, d->F, d->pxCz);
    d->pxF[0] = dot(d->pxD, d->F);
    d->pxF[1] = dot(d->vy, d->F);
    d->pxF[2] = dot(d->vz, d->F);
  }
  d->bc2 = d->bc * d->bc;
  const double blo = d->bc * (1.0 - o.band), bhi = d->bc * (1.0 + o.band);
  d->lo2 = o.band > 0.0 ? blo * blo : d->bc2;
  d->hi2 = o.band > 0.0 ? bhi * bhi : d->bc2;
  d->Uc = 1.0 / (3.0 * d->M);
  d->r02 = r0 * r0;
  d->cell = (d->Uc - d->l0) * (1.0 / 1048576.0);
  d->inv_cell = 1048576.0 / (d->Uc - d->l0);
  d->inv_nstep = 1.0 / d->nstep;
  d->M2 = 2.0 * d->M;
  d->M3 = 3.0 * d->M;
  d->M6 = 6.0 * d->M;
  d->Rhf = (float)d->Rh;
  d->u0f = (float)d->u0;
  d->ktf = (float)d->nstep - 0.1f;
  d->alpha_f = d->nf[0] * d->e1f[0] + d->nf[1] * d->e1f[1] + d->nf[2] * d->e1f[2];
  d->r_in2 = d->r_in > 0.0 ? (float)(d->r_in * d->r_in) : -1.0f;
  d->r_out2 = d->r_out >= 0.0 ? (float)(d->r_out * d->r_out) : -1.0f;
  if (d->has_disc) {
    double s1[3], s2[3];
    for (int i = 0; i < 3; ++i) {
      s1[i] = d->dv2[i] - d->dv1[i];
      s2[i] = d->dv4[i] - d->dv1[i];
      d->tv1[i] = (float)d->dv1[i];
      d->ts1[i] = (float)s1[i];
    }
    const double l1 = std::sqrt(s1[0] * s1[0] + s1[1] * s1[1] + s1[2] * s1[2]);
    const double l2 = std::sqrt(s2[0] * s2[0] + s2[1] * s2[1] + s2[2] * s2[2]);
    d->tka = (float)(d->disc_tex.cols / (l1 * l1));
    d->tkb = (float)(d->disc_tex.rows / (l1 * l2));
  }
  for (int i = 0; i < d->n_spheres; ++i) {
    for (int k = 0; k < 3; ++k) d->sphf[i][k] = (float)(d->sc[i][k] - d->B[k]);
    d->sphf[i][3] = (float)d->sr[i];
  }
  // Texel-edge guards from the relative error bound g of an FP32 hit point
  // (the largest measured: 1.51e-6 of the disc's outer radius, 1080p frames
  // of the disc orbit) and of an FP32 escape direction: the disc's in texel
  // units, the spheres' and the background's as angles of the direction that
  // picks the texel.
  const double g = o.edge_guard == 0.0 ? 2e-6 : (o.edge_guard > 0.0 ? o.edge_guard : 0.0);
  if (d->has_disc) {
    double l1 = 0, l2 = 0;
    for (int i = 0; i < 3; ++i) {
      l1 += (d->dv2[i] - d->dv1[i]) * (d->dv2[i] - d->dv1[i]);
      l2 += (d->dv4[i] - d->dv1[i]) * (d->dv4[i] - d->dv1[i]);
    }
    const double E = g * std::fabs(d->r_out);
    d->edge_a = (float)(E * d->disc_tex.cols / std::sqrt(l1));
    d->edge_b = (float)(E * d->disc_tex.rows / std::sqrt(l2));
  }
  d->edge_bg = (float)g;
  d->pos_guard = (float)g;
  for (int i = 0; i < d->n_spheres; ++i) {
    double c2 = 0;
    for (int k = 0; k < 3; ++k) c2 += (d->sc[i][k] - d->B[k]) * (d->sc[i][k] - d->B[k]);
    d->sph_edge[i] = (float)(g * (std::sqrt(c2) + d->sr[i]) / d->sr[i]);
  }
  d->fast_ok = d->precision == BH_PREC_FP32 && d->filter == BH_FILTER_NEAREST &&
               d->nstep <= 4096 && d->u0 < d->Uc &&
               (!d->has_disc || (d->dv0[0] == d->B[0] && d->dv0[1] == d->B[1] &&
                                 d->dv0[2] == d->B[2]));
  return BH_OK;
}

// The FP32 event march assumes the disc is centred on the hole (its plane
// passes through B, so every crossing is an angle in the orbit plane) and a
// camera outside the photon sphere (r0 > 3M: then every ray's march runs
// inward, du > 0, and its swept angle grows, which the event tests rely on);
// other scenes take the FP64 march.
bool fp32_eligible(const DevScene& d) { return d.fast_ok != 0; }

// FP32 march then the FP64 band kernel, or the FP64 reference march.
int launch_frame(bh_ctx* ctx, const DevScene& host_scene, DevScene* d_scene, const DevOut& out,
                 const DevQueue& q, hipStream_t st) {
  const int W = host_scene.W, H = host_scene.H;
  if (fp32_eligible(host_scene)) {
    // the FP32 march appends its band rays to q
    if (!q.items || !q.count || q.sub_cap == 0)
      return fail(ctx, BH_ERR_INVALID_ARG, "FP32 march launched without a band queue");
    const bool esc = host_scene.escape != 0;
    const bool sph = host_scene.n_spheres > 0;
    HIP_TRY(ctx, bh::launch_march_fp32(d_scene, out, W, H, q, st, esc, sph, host_scene.nstep));
    HIP_TRY(ctx, bh::launch_band(d_scene, out, q, st, esc, sph));
  } else {
    HIP_TRY(ctx, bh::launch_march_fp64(d_scene, out, W, H, st));
  }
  return BH_OK;
}

// The context's FP64 queue for frames up to W x H, with zeroed count pairs
// (the FP64 pass keeps them zeroed between frames).
int queue_buffers(bh_ctx* ctx, int W, int H, DevQueue* q) {
  const bool fresh = ctx->qcount.p == nullptr;
  int rc = ensure(ctx, ctx->qitems, bh::queue_items(W, H) * sizeof(uint64_t));
  if (rc) return rc;
  if ((rc = ensure(ctx, ctx->qcount, bh::kQueueCountWords * sizeof(uint32_t)))) return rc;
  if (fresh) {
    HIP_TRY(ctx, bh::launch_zero_u32((uint32_t*)ctx->qcount.p, (int)bh::kQueueCountWords,
                                     ctx->stream));
  }
  *q = DevQueue{(uint64_t*)ctx->qitems.p, (uint32_t*)ctx->qcount.p, bh::queue_sub_cap(W, H)};
  return BH_OK;
}


// ---- batches -----------------------------------------------------------------
struct BatchPlan {
  std::vector<DevScene> scenes;
  std::vector<bh::DevOut> outs;
  std::vector<uint32_t> off;   // ray offsets of the FP32 scenes (n_fast + 1)
  std::vector<int> fast;       // indices of the scenes in the persistent march, by class
  std::vector<int> solo;       // indices rendered one by one
  uint32_t cls_tile[5] = {0, 0, 0, 0, 0};  // tiles of class c: [cls_tile[c], cls_tile[c+1])
  int cls_nstep[4] = {0, 0, 0, 0};         // step count shared by class c's scenes (else 0)
  uint32_t n_rays = 0;
};

// Scene class of the tile-mode kernel instances: escape + 2 spheres.
int scene_class(const DevScene& d) { return (d.escape ? 1 : 0) + (d.n_spheres > 0 ? 2 : 0); }

int plan_batch(bh_ctx* ctx, int n, const bh_scene* scenes, const bh_options* options,
               const bh_outputs* outs, bool lane_mode, BatchPlan* p) {
  if (n <= 0 || !scenes || !options || !outs)
    return fail(ctx, BH_ERR_INVALID_ARG, "batch needs n > 0 and scene, option and output arrays");
  p->scenes.resize(n);
  p->outs.resize(n);
  std::vector<int> by_cls[4];
  for (int i = 0; i < n; ++i) {
    int rc = build_scene(ctx, &scenes[i], &options[i], &p->scenes[i]);
    if (rc) {
      ctx->err = "scene " + std::to_string(i) + ": " + ctx->err;
      return rc;
    }
    const bh_outputs& o = outs[i];
    p->outs[i] = DevOut{o.rgb, o.hit, o.phi_end, o.hit_point, o.steps};
    // the lane-refill mode keeps no sphere table per lane: such scenes go solo
    if (fp32_eligible(p->scenes[i]) && !(lane_mode && p->scenes[i].n_spheres > 0))
      by_cls[scene_class(p->scenes[i])].push_back(i);
    else
      p->solo.push_back(i);
  }
  // scenes grouped by class, in the given order within a class
  uint64_t rays = 0;
  p->off.push_back(0);
  for (int c = 0; c < 4; ++c) {
    p->cls_tile[c] = (uint32_t)(rays >> 6);
    p->cls_nstep[c] = by_cls[c].empty() ? 0 : p->scenes[by_cls[c][0]].nstep;
    for (int i : by_cls[c]) {
      if (p->scenes[i].nstep != p->cls_nstep[c]) p->cls_nstep[c] = 0;
      const DevScene& d = p->scenes[i];
      rays += (uint64_t)((d.W + 7) / 8) * (uint64_t)((d.H + 7) / 8) * 64u;
      if (rays >= ((uint64_t)1 << 31))
        return fail(ctx, BH_ERR_INVALID_ARG, "batch larger than 2^31 rays");
      p->fast.push_back(i);
      p->off.push_back((uint32_t)rays);
    }
  }
  p->cls_tile[4] = (uint32_t)(rays >> 6);
  p->n_rays = (uint32_t)rays;
  return BH_OK;
}

bh_batch_tuning normalized_tuning(const bh_batch_tuning* t) {
  bh_batch_tuning r{BH_BATCH_GRID, 256, 4, 4, 16, 4};
  if (t) {
    if (t->mode) r.mode = t->mode;
    if (t->block) r.block = t->block;
    if (t->waves_per_simd) r.waves_per_simd = t->waves_per_simd;
    if (t->refill_steps) r.refill_steps = t->refill_steps;
    if (t->refill_idle) r.refill_idle = t->refill_idle;
    if (t->chunk) r.chunk = t->chunk;
  }
  return r;
}

int check_tuning(bh_ctx* ctx, const bh_batch_tuning& t) {
  if (t.mode != BH_BATCH_GRID && t.mode != BH_BATCH_TILES && t.mode != BH_BATCH_LANES)
    return fail(ctx, BH_ERR_INVALID_ARG, "tuning.mode must be BH_BATCH_GRID, _TILES or _LANES");
  if (t.block != 64 && t.block != 128 && t.block != 256)
    return fail(ctx, BH_ERR_INVALID_ARG, "tuning.block must be 64, 128 or 256");
  if (t.waves_per_simd < 1 || t.waves_per_simd > 8)
    return fail(ctx, BH_ERR_INVALID_ARG, "tuning.waves_per_simd must be in 1..8");
  if (t.refill_steps < 1 || t.refill_steps > 64)
    return fail(ctx, BH_ERR_INVALID_ARG, "tuning.refill_steps must be in 1..64");
  if (t.refill_idle < 1 || t.refill_idle > 64)
    return fail(ctx, BH_ERR_INVALID_ARG, "tuning.refill_idle must be in 1..64");
  if (t.chunk < 1 || t.chunk > 1024)
    return fail(ctx, BH_ERR_INVALID_ARG, "tuning.chunk must be in 1..1024");
  return BH_OK;
}

// Upload the plan and launch clear + persistent march + band pass (+ solo
// scenes).  Events: ev0 before the march, ev1 after it, ev2 after the band
// pass.  The tables go to the device with one synchronous copy each.
int run_batch(bh_ctx* ctx, const BatchPlan& p, const bh_batch_tuning& t, bh::BatchDev* bd_out) {
  const size_t n = p.scenes.size();
  int rc;
  if ((rc = ensure(ctx, ctx->bscenes, n * sizeof(DevScene)))) return rc;
  if ((rc = ensure(ctx, ctx->bouts, n * sizeof(DevOut)))) return rc;
  if ((rc = ensure(ctx, ctx->boff, (p.off.size() + 1) * sizeof(uint32_t)))) return rc;
  if ((rc = ensure(ctx, ctx->bctl, 64))) return rc;
  if ((rc = ensure(ctx, ctx->bend, (n + 1) * sizeof(unsigned long long)))) return rc;
  if (t.mode == BH_BATCH_GRID &&
      ((rc = ensure(ctx, ctx->btile, ((size_t)(p.n_rays >> 6) + 1) * sizeof(unsigned long long))) ||
       (rc = ensure(ctx, ctx->btscene, ((size_t)(p.n_rays >> 6) + 1) * 2 * sizeof(uint32_t)))))
    return rc;
  // FP64 queue: every ray id is queued at most once (band or texel edge), in
  // the sub-queue of its 64-ray tile
  const uint32_t n_tiles = p.n_rays >> 6;
  const uint32_t sub_cap = std::max<uint32_t>((n_tiles + bh::kQueueSubs - 1) / bh::kQueueSubs, 1u) * 64u;
  if ((rc = ensure(ctx, ctx->bband, (size_t)sub_cap * bh::kQueueSubs * sizeof(uint64_t)))) return rc;
  if ((rc = ensure(ctx, ctx->bbandc, bh::kQueueCountWords * sizeof(uint32_t)))) return rc;
  hipStream_t st = ctx->stream;
  // the persistent march reads the FP32 scenes by position in the plan
  std::vector<DevScene> fs;
  std::vector<DevOut> fo;
  for (int i : p.fast) {
    fs.push_back(p.scenes[i]);
    fo.push_back(p.outs[i]);
  }
  DevScene* d_all = (DevScene*)ctx->bscenes.p;
  DevScene* d_fast = d_all;
  // FP32 scenes first, then the solo ones
  std::vector<DevScene> order(fs);
  for (int i : p.solo) order.push_back(p.scenes[i]);
  HIP_TRY(ctx, hipMemcpyAsync(d_all, order.data(), n * sizeof(DevScene), hipMemcpyHostToDevice, st));
  if (!fo.empty())
    HIP_TRY(ctx, hipMemcpyAsync(ctx->bouts.p, fo.data(), fo.size() * sizeof(DevOut),
                                hipMemcpyHostToDevice, st));
  HIP_TRY(ctx, hipMemcpyAsync(ctx->boff.p, p.off.data(), p.off.size() * sizeof(uint32_t),
                              hipMemcpyHostToDevice, st));
  bh::BatchDev b;
  b.scenes = d_fast;
  b.outs = (const DevOut*)ctx->bouts.p;
  b.ray_off = (const uint32_t*)ctx->boff.p;
  b.n_scenes = (uint32_t)p.fast.size();
  b.n_rays = p.n_rays;
  // control block: four queue heads, the start time
  b.head = (uint32_t*)ctx->bctl.p;
  b.t_start = (unsigned long long*)((char*)ctx->bctl.p + 24);
  b.tile_lo = 0;
  b.tile_hi = p.n_rays >> 6;
  b.nstep = 0;
  b.band = DevQueue{(uint64_t*)ctx->bband.p, (uint32_t*)ctx->bbandc.p, sub_cap};
  b.refill_steps = t.refill_steps;
  b.refill_idle = t.refill_idle;
  b.mode = t.mode;
  b.tile_end = (unsigned long long*)ctx->btile.p;
  b.tile_info = t.mode == BH_BATCH_GRID ? (uint2*)ctx->btscene.p : nullptr;
  b.chunk = t.chunk;
  b.scene_end = (unsigned long long*)ctx->bend.p;
  HIP_TRY(ctx, bh::launch_batch_clear(b, st));
  HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
  if (b.n_scenes > 0) {
    const int waves = std::max(1, ctx->n_cu) * 4 * t.waves_per_simd;
    const int grid = std::max(1, waves / (t.block / 64));
    if (t.mode == BH_BATCH_LANES) {
      HIP_TRY(ctx, bh::launch_batch(b, 0, t.block, grid, st));
    } else {
      // one persistent launch per scene class present, each with its own head
      bh::BatchDev bc = b;
      for (int c = 0; c < 4; ++c) {
        if (p.cls_tile[c + 1] == p.cls_tile[c]) continue;
        bc.tile_lo = p.cls_tile[c];
        bc.tile_hi = p.cls_tile[c + 1];
        bc.nstep = p.cls_nstep[c];
        bc.head = b.head + c;
        HIP_TRY(ctx, bh::launch_batch(bc, c, t.block, grid, st));
      }
    }
  }
  if (b.n_scenes > 0 && t.mode == BH_BATCH_GRID) HIP_TRY(ctx, bh::launch_batch_grid_finish(b, st));
  HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
  if (b.n_scenes > 0) HIP_TRY(ctx, bh::launch_batch_band(b, st));
  // solo scenes: the per-frame kernels, with the context's band queue (the
  // band kernel clears it after each scene)
  if (!p.solo.empty()) {
    int maxw = 0, maxh = 0;
    for (int i : p.solo) {
      maxw = std::max(maxw, p.scenes[i].W);
      maxh = std::max(maxh, p.scenes[i].H);
    }
    DevQueue q;  // large enough for every solo scene: sub_cap grows with W and H
    if ((rc = queue_buffers(ctx, maxw, maxh, &q))) return rc;
    for (size_t k = 0; k < p.solo.size(); ++k) {
      const int i = p.solo[k];
      const DevScene& d = p.scenes[i];
      q.sub_cap = bh::queue_sub_cap(d.W, d.H);
      if ((rc = launch_frame(ctx, d, d_all + p.fast.size() + k, p.outs[i], q, st))) return rc;
    }
  }
  HIP_TRY(ctx, hipEventRecord(ctx->ev2, st));
  if (bd_out) *bd_out = b;
  return BH_OK;
}

// ---- differentiable scene ---------------------------------------------------
// FP64 number carrying d/d(inclination), for the camera frame.
struct HD {
  double v, t;
};
HD operator+(HD a, HD b) { return {a.v + b.v, a.t + b.t}; }
HD operator-(HD a, HD b) { return {a.v - b.v, a.t - b.t}; }
HD operator*(HD a, HD b) { return {a.v * b.v, a.t * b.v + a.v * b.t}; }
struct HV {
  HD x[3];
};
HD hdot(const HV& a, const HV& b) {
  HD s{0.0, 0.0};
  for (int i = 0; i < 3; ++i) s = s + a.x[i] * b.x[i];
  return s;
}
HV hscale(const HV& a, HD k) { return {{a.x[0] * k, a.x[1] * k, a.x[2] * k}}; }
HV hsub(const HV& a, const HV& b) { return {{a.x[0] - b.x[0], a.x[1] - b.x[1], a.x[2] - b.x[2]}}; }
HV hcross(const HV& a, const HV& b) {
  return {{a.x[1] * b.x[2] - a.x[2] * b.x[1], a.x[2] * b.x[0] - a.x[0] * b.x[2],
           a.x[0] * b.x[1] - a.x[1] * b.x[0]}};
}
// v * (1/|v|), the OpenCV normalisation the camera frame uses
HV hnormalize(const HV& a) {
  const HD n2 = hdot(a, a);
  const double nv = std::sqrt(n2.v);
  const double k = 1.0 / nv;
  const HD kk{k, -0.5 * k * k * k * n2.t};
  return hscale(a, kk);
}

// Resolve bh_diff_params: camera at dist (sin i cos az, sin i sin az, cos i)
// looking at the hole with +z up (x toward the hole, z = -(up orthogonal to
// x), y = z x x), hole at the origin, disc in z = 0 with the texture square
// (+-r_out, +-r_out).
int build_diff(bh_ctx* ctx, const bh_diff_params* p, int precision, bh::DiffScene* d) {
  if (!p) return fail(ctx, BH_ERR_INVALID_ARG, "params is NULL");
  if (precision != BH_PREC_FP32 && precision != BH_PREC_FP64)
    return fail(ctx, BH_ERR_INVALID_ARG, "unknown precision");
  if (p->width <= 0 || p->height <= 0 || p->width > 16384 || p->height > 16384)
    return fail(ctx, BH_ERR_INVALID_ARG, "width/height must be in 1..16384");
  const double vals[] = {p->r_s, p->incl, p->dist, p->r_inner, p->r_outer, p->azim, p->fov,
                         p->edge_width};
  for (double v : vals)
    if (!std::isfinite(v)) return fail(ctx, BH_ERR_INVALID_ARG, "parameters must be finite");
  if (!(p->r_s > 0.0)) return fail(ctx, BH_ERR_INVALID_ARG, "r_s must be positive");
  if (!(p->dist > p->r_s)) return fail(ctx, BH_ERR_INVALID_ARG, "camera must be outside the horizon");
  if (!(p->r_outer > 0.0) || !(p->r_inner >= 0.0) || !(p->r_inner < p->r_outer))
    return fail(ctx, BH_ERR_INVALID_ARG, "need 0 <= r_inner < r_outer");
  if (!(p->edge_width >= 0.0)) return fail(ctx, BH_ERR_INVALID_ARG, "edge_width must be >= 0");
  std::memset(d, 0, sizeof *d);
  if (p->texture >= 0) {
    d->tex = devtex(ctx, p->texture);
    if (!d->tex.data) return fail(ctx, BH_ERR_EMPTY_TEXTURE, "disc texture slot is empty");
  }
  const double si = std::sin(p->incl), ci = std::cos(p->incl);
  const double ca = std::cos(p->azim), sa = std::sin(p->azim);
  const double dir[3] = {si * ca, si * sa, ci}, dir_i[3] = {ci * ca, ci * sa, -si};
  HV C;
  for (int k = 0; k < 3; ++k) {
    C.x[k] = {p->dist * dir[k], p->dist * dir_i[k]};
    d->C[k] = C.x[k].v;
    d->C_i[k] = C.x[k].t;
    d->C_d[k] = dir[k];
  }
  const HD m1{-1.0, 0.0};
  const HV x = hnormalize(hscale(C, m1));
  const HV up{{{0.0, 0.0}, {0.0, 0.0}, {1.0, 0.0}}};
  const HD k = hdot(up, x);
  const HV z = hnormalize(hscale(hsub(up, hscale(x, k)), m1));
  const HV y = hnormalize(hcross(z, x));
  const double f = bh_focal_length(p->width, p->fov);
  for (int i = 0; i < 3; ++i) {
    d->fx[i] = x.x[i].v * f;
    d->fx_i[i] = x.x[i].t * f;
    d->ay[i] = y.x[i].v;
    d->ay_i[i] = y.x[i].t;
    d->az[i] = z.x[i].v;
    d->az_i[i] = z.x[i].t;
  }
  d->M = p->r_s / 2.0;
  d->bc = 3.0 * std::sqrt(3.0) * d->M;
  d->Rh = 2 * d->M;
  d->Uc = 1.0 / (3 * d->M);
  d->r_in = p->r_inner;
  d->r_out = p->r_outer;
  d->edge = p->edge_width;
  d->half = p->r_outer;
  d->uscale = d->tex.cols / (2.0 * p->r_outer);
  d->vscale = d->tex.rows / (2.0 * p->r_outer);
  d->W = p->width;
  d->H = p->height;
  d->nstep = 20;
  return BH_OK;
}

// Fused path launches: every frame's march in one launch (up to 1024 frames
// each), then every frame's FP64 band pass in one launch.
constexpr int kFramesPerLaunch = 1024;

hipError_t path_march(const bh_path* p, hipStream_t st) {
  const DevOut o{p->d_rgb, p->d_hit, nullptr, nullptr, nullptr};
  const DevQueue q{p->d_qitems, p->d_qcount, p->sub_cap};
  for (int f0 = 0; f0 < p->n_frames; f0 += kFramesPerLaunch) {
    const int nf = std::min(kFramesPerLaunch, p->n_frames - f0);
    const hipError_t e = bh::launch_march_frames(p->d_scenes, o, p->W, p->H, f0, nf, q, st,
                                                 p->esc, p->sph, p->nstep);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t path_band(const bh_path* p, hipStream_t st) {
  const DevOut o{p->d_rgb, p->d_hit, nullptr, nullptr, nullptr};
  const DevQueue q{p->d_qitems, p->d_qcount, p->sub_cap};
  for (int f0 = 0; f0 < p->n_frames; f0 += kFramesPerLaunch) {
    const int nf = std::min(kFramesPerLaunch, p->n_frames - f0);
    const hipError_t e = bh::launch_band_frames(p->d_scenes, o, p->W, p->H, f0, nf, q, st,
                                                p->esc, p->sph);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace

extern "C" {

int bh_ctx_create(int device, bh_ctx** out) {
  return guarded(nullptr, [&]() -> int {
    if (!out) return fail(nullptr, BH_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) return fail(nullptr, BH_ERR_NO_DEVICE, "no HIP device available");
    if (device < 0 || device >= n) return fail(nullptr, BH_ERR_NO_DEVICE, "device ordinal out of range");
    e = hipSetDevice(device);
    if (e != hipSuccess) return fail(nullptr, BH_ERR_HIP, "hipSetDevice failed");
    bh_ctx* c = new bh_ctx();
    c->device = device;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete c;
      return fail(nullptr, BH_ERR_HIP, "hipStreamCreate failed");
    }
    c->own_stream = true;
    e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
    if (e != hipSuccess) {
      (void)hipStreamDestroy(c->stream);
      delete c;
      return fail(nullptr, BH_ERR_HIP, "hipStreamCreate failed");
    }
    (void)hipEventCreate(&c->ev0);
    (void)hipEventCreate(&c->ev1);
    (void)hipEventCreate(&c->ev2);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
      std::snprintf(c->name, sizeof c->name, "%s (%s)", prop.name, prop.gcnArchName);
      c->n_cu = prop.multiProcessorCount;
      c->clock_khz = prop.clockRate;
    }
    if (hipDeviceGetAttribute(&c->wall_khz, hipDeviceAttributeWallClockRate, device) != hipSuccess)
      c->wall_khz = 0;
    // this device's allocation-free scratch array (device is current); the
    // entry point reports its absence if the lookup fails
    if (bh::noalloc_scratch(&c->noalloc) != hipSuccess) c->noalloc = nullptr;
    *out = c;
    return BH_OK;
  });
}

void bh_ctx_destroy(bh_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto& t : ctx->tex)
    if (t.d) (void)hipFree(t.d);
  for (DevBuf* b : {&ctx->scene, &ctx->rgb, &ctx->hit, &ctx->phi, &ctx->hitpt, &ctx->steps,
                    &ctx->qitems, &ctx->qcount, &ctx->dpart, &ctx->dgrad, &ctx->drgb, &ctx->dhit,
                    &ctx->dgin, &ctx->bscenes, &ctx->bouts, &ctx->boff, &ctx->bctl, &ctx->bband,
                    &ctx->bbandc, &ctx->bend, &ctx->btile, &ctx->btscene, &ctx->probe})
    if (b->p) (void)hipFree(b->p);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->ev2) (void)hipEventDestroy(ctx->ev2);
  if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
  if (ctx->side) (void)hipStreamDestroy(ctx->side);
  delete ctx;
}

const char* bh_last_error(const bh_ctx* ctx) {
  if (ctx) return ctx->err.c_str();
  return g_tls_err.empty() ? g_static_err : g_tls_err.c_str();
}

int bh_device_info(bh_ctx* ctx, char* name, int name_len, int* n_cu, int* clock_khz) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return fail(nullptr, BH_ERR_BAD_HANDLE, "ctx is NULL");
    if (name && name_len > 0) std::snprintf(name, (size_t)name_len, "%s", ctx->name);
    if (n_cu) *n_cu = ctx->n_cu;
    if (clock_khz) *clock_khz = ctx->clock_khz;
    return BH_OK;
  });
}

int bh_ctx_set_stream(bh_ctx* ctx, void* stream) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return fail(nullptr, BH_ERR_BAD_HANDLE, "ctx is NULL");
    (void)hipSetDevice(ctx->device);
    if (ctx->own_stream && ctx->stream) {
      (void)hipStreamSynchronize(ctx->stream);
      (void)hipStreamDestroy(ctx->stream);
    }
    if (stream) {
      ctx->stream = (hipStream_t)stream;
      ctx->own_stream = false;
    } else {
      hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
      if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamCreate");
      ctx->own_stream = true;
    }
    return BH_OK;
  });
}

int bh_synchronize(bh_ctx* ctx) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return fail(nullptr, BH_ERR_BAD_HANDLE, "ctx is NULL");
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return BH_OK;
  });
}

int bh_texture_set(bh_ctx* ctx, int slot, const bh_image* img) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return fail(nullptr, BH_ERR_BAD_HANDLE, "ctx is NULL");
    if (slot < 0 || slot >= BH_MAX_TEXTURES) return fail(ctx, BH_ERR_INVALID_ARG, "texture slot out of range");
    if (!img || !img->data || img->rows <= 0 || img->cols <= 0)
      return fail(ctx, BH_ERR_EMPTY_TEXTURE, "empty texture (Material::SetTexture)");
    if (img->channels != 3 && img->channels != 4)
      return fail(ctx, BH_ERR_INVALID_ARG, "texture channels must be 3 or 4");
    if ((int64_t)img->rows * img->cols > ((int64_t)1 << 28) || img->rows > (1 << 24) ||
        img->cols > (1 << 24))  // sides exact in FP32 texel addressing
      return fail(ctx, BH_ERR_INVALID_ARG, "texture too large");
    (void)hipSetDevice(ctx->device);
    const size_t n = (size_t)img->rows * img->cols;
    std::vector<uint32_t> packed(n);
    const int C = img->channels;
    for (size_t i = 0; i < n; ++i) {
      const uint8_t* p = img->data + i * C;
      packed[i] = (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) |
                  ((C == 4 ? (uint32_t)p[3] : 255u) << 24);
    }
    DeviceTexture& t = ctx->tex[slot];
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (t.d && (size_t)t.rows * t.cols != n) {
      (void)hipFree(t.d);
      t.d = nullptr;
    }
    if (!t.d) HIP_TRY(ctx, hipMalloc(&t.d, n * sizeof(uint32_t)));
    HIP_TRY(ctx, hipMemcpy(t.d, packed.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    t.rows = img->rows;
    t.cols = img->cols;
    return BH_OK;
  });
}

int bh_texture_clear(bh_ctx* ctx, int slot) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return fail(nullptr, BH_ERR_BAD_HANDLE, "ctx is NULL");
    if (slot < 0 || slot >= BH_MAX_TEXTURES) return fail(ctx, BH_ERR_INVALID_ARG, "texture slot out of range");
    DeviceTexture& t = ctx->tex[slot];
    if (t.d) {
      HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
      (void)hipFree(t.d);
    }
    t = DeviceTexture{};
    return BH_OK;
  });
}

int bh_scene_check(bh_ctx* ctx, const bh_scene* scene, const bh_options* opt) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return fail(nullptr, BH_ERR_BAD_HANDLE, "ctx is NULL");
    DevScene d;
    return build_scene(ctx, scene, opt, &d);
  });
}

int bh_render_device(bh_ctx* ctx, const bh_scene* scene, const bh_options* opt,
                     const bh_outputs* out_device) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return fail(nullptr, BH_ERR_BAD_HANDLE, "ctx is NULL");
    if (!out_device) return fail(ctx, BH_ERR_INVALID_ARG, "outputs are NULL");
    DevScene d;
    int rc = build_scene(ctx, scene, opt, &d);
    if (rc) return rc;
    (void)hipSetDevice(ctx->device);
    if ((rc = ensure(ctx, ctx->scene, sizeof(DevScene)))) return rc;
    DevQueue q;
    if ((rc = queue_buffers(ctx, d.W, d.H, &q))) return rc;
    HIP_TRY(ctx, hipMemcpyAsync(ctx->scene.p, &d, sizeof d, hipMemcpyHostToDevice, ctx->stream));
    DevOut o{out_device->rgb, out_device->hit, out_device->phi_end, out_device->hit_point,
             out_device->steps};
    return launch_frame(ctx, d, (DevScene*)ctx->scene.p, o, q, ctx->stream);
  });
}

int bh_render_device_noalloc(bh_ctx* ctx, const bh_scene* scene, const bh_options* opt,
                             const bh_outputs* out_device) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return fail(nullptr, BH_ERR_BAD_HANDLE, "ctx is NULL");
    if (!out_device) return fail(ctx, BH_ERR_INVALID_ARG, "outputs are NULL");
    if (!ctx->noalloc)
      return fail(ctx, BH_ERR_HIP, "the allocation-free scratch array is not available on this device");
    DevScene d;
    int rc = build_scene(ctx, scene, opt, &d);
    if (rc) return rc;
    if (fp32_eligible(d) && bh::queue_items(d.W, d.H) > bh::kNoallocQueueItems)
      return fail(ctx, BH_ERR_INVALID_ARG,
                  "frame too large for bh_render_device_noalloc (its FP64 queue holds " +
                      std::to_string(bh::kNoallocQueueItems) + " items, this frame needs " +
                      std::to_string(bh::queue_items(d.W, d.H)) + ")");
    (void)hipSetDevice(ctx->device);
    const bh::NoallocLayout l = bh::noalloc_layout(ctx->noalloc);
    HIP_TRY(ctx, bh::launch_noalloc_setup(d, l, ctx->stream));
    DevOut o{out_device->rgb, out_device->hit, out_device->phi_end, out_device->hit_point,
             out_device->steps};
    const DevQueue q{l.items, l.count, bh::queue_sub_cap(d.W, d.H)};
    return launch_frame(ctx, d, l.scene, o, q, ctx->stream);
  });
}

int bh_render(bh_ctx* ctx, const bh_scene* scene, const bh_options* opt, const bh_outputs* out) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return fail(nullptr, BH_ERR_BAD_HANDLE, "ctx is NULL");
    if (!out) return fail(ctx, BH_ERR_INVALID_ARG, "outputs are NULL");
    DevScene d;
    int rc = build_scene(ctx, scene, opt, &d);
    if (rc) return rc;
    (void)hipSetDevice(ctx->device);
    const size_t npix = (size_t)d.W * d.H;
    DevOut o{nullptr, nullptr, nullptr, nullptr, nullptr};
    if (out->rgb) { if ((rc = ensure(ctx, ctx->rgb, npix * 3))) return rc; o.rgb = (uint8_t*)ctx->rgb.p; }
    if (out->hit) { if ((rc = ensure(ctx, ctx->hit, npix))) return rc; o.hit = (uint8_t*)ctx->hit.p; }
    if (out->phi_end) { if ((rc = ensure(ctx, ctx->phi, npix * 8))) return rc; o.phi = (double*)ctx->phi.p; }
    if (out->hit_point) { if ((rc = ensure(ctx, ctx->hitpt, npix * 24))) return rc; o.hitpt = (double*)ctx->hitpt.p; }
    if (out->steps) { if ((rc = ensure(ctx, ctx->steps, npix * 4))) return rc; o.steps = (int32_t*)ctx->steps.p; }
    bh_outputs dev{o.rgb, o.hit, o.phi, o.hitpt, o.steps};
    if ((rc = bh_render_device(ctx, scene, opt, &dev))) return rc;
    hipStream_t st = ctx->stream;
    if (out->rgb) HIP_TRY(ctx, hipMemcpyAsync(out->rgb, o.rgb, npix * 3, hipMemcpyDeviceToHost, st));
    if (out->hit) HIP_TRY(ctx, hipMemcpyAsync(out->hit, o.hit, npix, hipMemcpyDeviceToHost, st));
    if (out->phi_end) HIP_TRY(ctx, hipMemcpyAsync(out->phi_end, o.phi, npix * 8, hipMemcpyDeviceToHost, st));
    if (out->hit_point) HIP_TRY(ctx, hipMemcpyAsync(out->hit_point, o.hitpt, npix * 24, hipMemcpyDeviceToHost, st));
    if (out->steps) HIP_TRY(ctx, hipMemcpyAsync(out->steps, o.steps, npix * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipStreamSynchronize(st));
    return BH_OK;
  });
}

int bh_time_render(bh_ctx* ctx, const bh_scene* scene, const bh_options* opt, int iters,
                   double* avg_ms_main, double* avg_ms_total, int64_t* band_rays) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return fail(nullptr, BH_ERR_BAD_HANDLE, "ctx is NULL");
    if (iters <= 0) return fail(ctx, BH_ERR_INVALID_ARG, "iters must be positive");
    DevScene d;
    int rc = build_scene(ctx, scene, opt, &d);
    if (rc) return rc;
    (void)hipSetDevice(ctx->device);
    const size_t npix = (size_t)d.W * d.H;
    if ((rc = ensure(ctx, ctx->rgb, npix * 3))) return rc;
    if ((rc = ensure(ctx, ctx->hit, npix))) return rc;
    if ((rc = ensure(ctx, ctx->scene, sizeof(DevScene)))) return rc;
    DevQueue q;
    if ((rc = queue_buffers(ctx, d.W, d.H, &q))) return rc;
    hipStream_t st = ctx->stream;
    HIP_TRY(ctx, hipMemcpyAsync(ctx->scene.p, &d, sizeof d, hipMemcpyHostToDevice, st));
    DevOut o{(uint8_t*)ctx->rgb.p, (uint8_t*)ctx->hit.p, nullptr, nullptr, nullptr};
    DevScene* ds = (DevScene*)ctx->scene.p;
    const bool fast = fp32_eligible(d);
    // warm-up
    if ((rc = launch_frame(ctx, d, ds, o, q, st))) return rc;
    HIP_TRY(ctx, hipStreamSynchronize(st));
    // main kernel: iters back-to-back launches between one event pair (the
    // queue is cleared afterwards); whole frame: iters march + band launches
    float main_ms = 0, total_ms = 0;
    std::vector<uint32_t> cnt(bh::kQueueCountWords, 0u);
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
    for (int it = 0; it < iters; ++it) {
      if (fast) {
        HIP_TRY(ctx, bh::launch_march_fp32(ds, o, d.W, d.H, q, st, d.escape != 0, d.n_spheres > 0,
                                           d.nstep));
      } else {
        HIP_TRY(ctx, bh::launch_march_fp64(ds, o, d.W, d.H, st));
      }
    }
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
    if (fast) {
      HIP_TRY(ctx, hipMemcpyAsync(cnt.data(), q.count, cnt.size() * 4, hipMemcpyDeviceToHost, st));
      HIP_TRY(ctx, bh::launch_zero_u32(q.count, (int)bh::kQueueCountWords, st));
    }
    HIP_TRY(ctx, hipEventRecord(ctx->ev2, st));
    for (int it = 0; it < iters; ++it)
      if ((rc = launch_frame(ctx, d, ds, o, q, st))) return rc;
    hipEvent_t end;
    HIP_TRY(ctx, hipEventCreate(&end));
    HIP_TRY(ctx, hipEventRecord(end, st));
    HIP_TRY(ctx, hipEventSynchronize(end));
    HIP_TRY(ctx, hipEventElapsedTime(&main_ms, ctx->ev0, ctx->ev1));
    HIP_TRY(ctx, hipEventElapsedTime(&total_ms, ctx->ev2, end));
    (void)hipEventDestroy(end);
    int64_t nband = 0;
    for (uint32_t k = 0; k < bh::kQueueSubs; ++k) nband += cnt[16 * k] + cnt[16 * k + 2];
    if (fast) nband /= iters;
    if (avg_ms_main) *avg_ms_main = main_ms / iters;
    if (avg_ms_total) *avg_ms_total = total_ms / iters;
    if (band_rays) *band_rays = nband;
    return BH_OK;
  });
}

// ---------------------------------------------------------------------------
// camera paths as HIP graphs
// ---------------------------------------------------------------------------
int bh_path_create(bh_ctx* ctx, const bh_scene* scene, const bh_options* opt,
                   const bh_camera* cameras, int n_frames, bh_path** out) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return fail(nullptr, BH_ERR_BAD_HANDLE, "ctx is NULL");
    if (!out || !scene || !cameras || n_frames <= 0)
      return fail(ctx, BH_ERR_INVALID_ARG, "path needs scene, cameras and n_frames > 0");
    *out = nullptr;
    (void)hipSetDevice(ctx->device);
    std::vector<DevScene> frames((size_t)n_frames);
    for (int i = 0; i < n_frames; ++i) {
      bh_scene s = *scene;
      s.camera = cameras[i];
      if (cameras[i].width != cameras[0].width || cameras[i].height != cameras[0].height)
        return fail(ctx, BH_ERR_INVALID_ARG, "path cameras must share width/height");
      int rc = build_scene(ctx, &s, opt, &frames[(size_t)i]);
      if (rc) return rc;
    }
    bh_path* p = new bh_path();
    p->ctx = ctx;
    p->n_frames = n_frames;
    p->W = frames[0].W;
    p->H = frames[0].H;
    const size_t npix = (size_t)p->W * p->H;
    auto cleanup = [&](hipError_t e, const char* what) {
      bh_path_destroy(p);
      return hip_fail(ctx, e, what);
    };
    hipError_t e;
    if ((e = hipMalloc(&p->d_scenes, sizeof(DevScene) * n_frames)) != hipSuccess) return cleanup(e, "hipMalloc scenes");
    if ((e = hipMalloc(&p->d_rgb, npix * 3 * n_frames)) != hipSuccess) return cleanup(e, "hipMalloc rgb");
    if ((e = hipMalloc(&p->d_hit, npix * n_frames)) != hipSuccess) return cleanup(e, "hipMalloc hit");
    // one FP64 queue per frame, so that the FP64 pass of frame i (side
    // stream) overlaps the FP32 march of frame i+1
    p->sub_cap = bh::queue_sub_cap(p->W, p->H);
    const size_t qitems = bh::queue_items(p->W, p->H);
    const size_t qwords = bh::kQueueCountWords;
    p->fused = true;
    for (const DevScene& f : frames) p->fused = p->fused && fp32_eligible(f);
    p->esc = frames[0].escape != 0;
    p->sph = frames[0].n_spheres > 0;
    p->nstep = frames[0].nstep;
    if ((e = hipMalloc(&p->d_qitems, qitems * n_frames * sizeof(uint64_t))) != hipSuccess) return cleanup(e, "hipMalloc queue");
    if ((e = hipMalloc(&p->d_qcount, qwords * 4 * (size_t)n_frames)) != hipSuccess) return cleanup(e, "hipMalloc queue count");
    if ((e = hipMemset(p->d_qcount, 0, qwords * 4 * (size_t)n_frames)) != hipSuccess) return cleanup(e, "hipMemset queue count");
    if ((e = hipMemcpy(p->d_scenes, frames.data(), sizeof(DevScene) * n_frames, hipMemcpyHostToDevice)) != hipSuccess)
      return cleanup(e, "hipMemcpy scenes");
    p->events.resize((size_t)n_frames + 1, nullptr);
    for (auto& ev : p->events)
      if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) return cleanup(e, "hipEventCreate");
    hipStream_t st = ctx->stream, side = ctx->side;
    if ((e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal)) != hipSuccess) return cleanup(e, "capture");
    int rc = BH_OK;
    bool used_side = false;
    if (p->fused) {
      // the march graph here; the FP64 pass is captured into its own graph below
      hipError_t le = path_march(p, st);
      if (le != hipSuccess) rc = hip_fail(ctx, le, "path capture launch");
    }
    for (int i = 0; i < n_frames && rc == BH_OK && !p->fused; ++i) {
      DevOut o{p->d_rgb + npix * 3 * i, p->d_hit + npix * i, nullptr, nullptr, nullptr};
      DevQueue q{p->d_qitems + qitems * i, p->d_qcount + qwords * i, p->sub_cap};
      const DevScene& f = frames[(size_t)i];
      if (fp32_eligible(f)) {
        const bool esc = f.escape != 0;
        const bool sph = f.n_spheres > 0;
        hipError_t le =
            bh::launch_march_fp32(p->d_scenes + i, o, p->W, p->H, q, st, esc, sph, f.nstep);
        if (le == hipSuccess) le = hipEventRecord(p->events[(size_t)i], st);
        if (le == hipSuccess) le = hipStreamWaitEvent(side, p->events[(size_t)i], 0);
        if (le == hipSuccess) le = bh::launch_band(p->d_scenes + i, o, q, side, esc, sph);
        if (le != hipSuccess) rc = hip_fail(ctx, le, "path capture launch");
        used_side = true;
      } else {
        hipError_t le = bh::launch_march_fp64(p->d_scenes + i, o, p->W, p->H, st);
        if (le != hipSuccess) rc = hip_fail(ctx, le, "path capture launch");
      }
    }
    if (rc == BH_OK && used_side) {
      // join the side stream back into the captured stream
      hipError_t le = hipEventRecord(p->events[(size_t)n_frames], side);
      if (le == hipSuccess) le = hipStreamWaitEvent(st, p->events[(size_t)n_frames], 0);
      if (le != hipSuccess) rc = hip_fail(ctx, le, "path capture join");
    }
    if (rc != BH_OK) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(st, &g);
      if (g) (void)hipGraphDestroy(g);
      std::string m = ctx->err;
      bh_path_destroy(p);
      return fail(ctx, rc, m);
    }
    if ((e = hipStreamEndCapture(st, &p->graph)) != hipSuccess) return cleanup(e, "end capture");
    if ((e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0)) != hipSuccess)
      return cleanup(e, "hipGraphInstantiate");
    if (p->fused) {
      if ((e = hipEventCreate(&p->k0)) != hipSuccess || (e = hipEventCreate(&p->k1)) != hipSuccess ||
          (e = hipEventCreate(&p->k2)) != hipSuccess)
        return cleanup(e, "hipEventCreate");
      if ((e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal)) != hipSuccess)
        return cleanup(e, "capture");
      const hipError_t le = path_band(p, st);
      hipGraph_t g = nullptr;
      e = hipStreamEndCapture(st, &g);
      p->graph_band = g;
      if (le != hipSuccess) return cleanup(le, "path capture launch");
      if (e != hipSuccess) return cleanup(e, "end capture");
      if ((e = hipGraphInstantiate(&p->exec_band, p->graph_band, nullptr, nullptr, 0)) != hipSuccess)
        return cleanup(e, "hipGraphInstantiate");
    }
    *out = p;
    return BH_OK;
  });
}

int bh_path_launch(bh_path* p, float* ms) {
  return guarded(nullptr, [&]() -> int {
    if (!p) return fail(nullptr, BH_ERR_BAD_HANDLE, "path is NULL");
    bh_ctx* ctx = p->ctx;
    (void)hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    if (p->exec_band) {
      // the march graph, then the FP64 pass's, with timing events around each
      HIP_TRY(ctx, hipEventRecord(p->k0, st));
      HIP_TRY(ctx, hipGraphLaunch(p->exec, st));
      HIP_TRY(ctx, hipEventRecord(p->k1, st));
      HIP_TRY(ctx, hipGraphLaunch(p->exec_band, st));
      HIP_TRY(ctx, hipEventRecord(p->k2, st));
      HIP_TRY(ctx, hipEventSynchronize(p->k2));
      if (ms) HIP_TRY(ctx, hipEventElapsedTime(ms, p->k0, p->k2));
      return BH_OK;
    }
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
    HIP_TRY(ctx, hipGraphLaunch(p->exec, st));
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
    if (ms) HIP_TRY(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return BH_OK;
  });
}

int bh_path_last_ms(bh_path* p, double* march_ms, double* band_ms) {
  return guarded(nullptr, [&]() -> int {
    if (!p) return fail(nullptr, BH_ERR_BAD_HANDLE, "path is NULL");
    bh_ctx* ctx = p->ctx;
    if (!p->exec_band) return fail(ctx, BH_ERR_UNSUPPORTED, "path frames are not all in the FP32 march");
    (void)hipSetDevice(ctx->device);
    HIP_TRY(ctx, hipEventSynchronize(p->k2));
    float a = 0.f, b = 0.f;
    HIP_TRY(ctx, hipEventElapsedTime(&a, p->k0, p->k1));
    HIP_TRY(ctx, hipEventElapsedTime(&b, p->k1, p->k2));
    if (march_ms) *march_ms = a;
    if (band_ms) *band_ms = b;
    return BH_OK;
  });
}

int bh_fma_issue_rate(bh_ctx* ctx, double* fp32_tflops, double* fp64_tflops) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return fail(nullptr, BH_ERR_BAD_HANDLE, "ctx is NULL");
    if (!fp32_tflops || !fp64_tflops) return fail(ctx, BH_ERR_INVALID_ARG, "outputs are NULL");
    (void)hipSetDevice(ctx->device);
    int rc = ensure(ctx, ctx->probe, 256 * sizeof(double));
    if (rc) return rc;
    const int n_cu = ctx->n_cu > 0 ? ctx->n_cu : 1;
    HIP_TRY(ctx, bh::fma_probe<float>(n_cu, ctx->stream, (float*)ctx->probe.p, fp32_tflops));
    HIP_TRY(ctx, bh::fma_probe<double>(n_cu, ctx->stream, (double*)ctx->probe.p, fp64_tflops));
    return BH_OK;
  });
}

int bh_chord_sincos(bh_ctx* ctx, const float* x_host, float* sc_host, int64_t n) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return fail(nullptr, BH_ERR_BAD_HANDLE, "ctx is NULL");
    if (!x_host || !sc_host) return fail(ctx, BH_ERR_INVALID_ARG, "buffers are NULL");
    if (n <= 0 || n > (int64_t)1 << 24) return fail(ctx, BH_ERR_INVALID_ARG, "n must be in 1..2^24");
    (void)hipSetDevice(ctx->device);
    int rc = ensure(ctx, ctx->probe, (size_t)n * 3 * sizeof(float));
    if (rc) return rc;
    float* x = (float*)ctx->probe.p;
    float* out = x + n;
    hipStream_t st = ctx->stream;
    HIP_TRY(ctx, hipMemcpyAsync(x, x_host, (size_t)n * sizeof(float), hipMemcpyHostToDevice, st));
    HIP_TRY(ctx, bh::launch_chord_sincos(x, out, (int)n, st));
    HIP_TRY(ctx, hipMemcpyAsync(sc_host, out, (size_t)n * 2 * sizeof(float),
                                hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipStreamSynchronize(st));
    return BH_OK;
  });
}

int bh_turning_points(bh_ctx* ctx, const bh_scene* scene, const bh_options* opt,
                      double* l_host) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return fail(nullptr, BH_ERR_BAD_HANDLE, "ctx is NULL");
    if (!l_host) return fail(ctx, BH_ERR_INVALID_ARG, "l_host is NULL");
    DevScene d;
    int rc = build_scene(ctx, scene, opt, &d);
    if (rc) return rc;
    (void)hipSetDevice(ctx->device);
    const size_t npix = (size_t)d.W * d.H;
    if ((rc = ensure(ctx, ctx->scene, sizeof(DevScene)))) return rc;
    if ((rc = ensure(ctx, ctx->probe, npix * sizeof(double)))) return rc;
    hipStream_t st = ctx->stream;
    HIP_TRY(ctx, hipMemcpyAsync(ctx->scene.p, &d, sizeof d, hipMemcpyHostToDevice, st));
    HIP_TRY(ctx, bh::launch_turning((const DevScene*)ctx->scene.p, (double*)ctx->probe.p, d.W,
                                    d.H, st));
    HIP_TRY(ctx, hipMemcpyAsync(l_host, ctx->probe.p, npix * sizeof(double),
                                hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipStreamSynchronize(st));
    return BH_OK;
  });
}

int bh_path_time(bh_path* p, int iters, double* march_ms, double* band_ms) {
  return guarded(nullptr, [&]() -> int {
    if (!p) return fail(nullptr, BH_ERR_BAD_HANDLE, "path is NULL");
    bh_ctx* ctx = p->ctx;
    if (!p->fused) return fail(ctx, BH_ERR_UNSUPPORTED, "path frames are not all in the FP32 march");
    if (iters <= 0) return fail(ctx, BH_ERR_INVALID_ARG, "iters must be positive");
    (void)hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    double tm = 0, tb = 0;
    for (int it = 0; it < iters; ++it) {
      HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
      HIP_TRY(ctx, path_march(p, st));
      HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
      HIP_TRY(ctx, path_band(p, st));
      HIP_TRY(ctx, hipEventRecord(ctx->ev2, st));
      HIP_TRY(ctx, hipEventSynchronize(ctx->ev2));
      float a = 0, b = 0;
      HIP_TRY(ctx, hipEventElapsedTime(&a, ctx->ev0, ctx->ev1));
      HIP_TRY(ctx, hipEventElapsedTime(&b, ctx->ev1, ctx->ev2));
      tm += a;
      tb += b;
    }
    if (march_ms) *march_ms = tm / iters;
    if (band_ms) *band_ms = tb / iters;
    return BH_OK;
  });
}

int bh_path_read(bh_path* p, int frame, uint8_t* rgb, uint8_t* hit) {
  return guarded(nullptr, [&]() -> int {
    if (!p) return fail(nullptr, BH_ERR_BAD_HANDLE, "path is NULL");
    bh_ctx* ctx = p->ctx;
    if (frame < 0 || frame >= p->n_frames) return fail(ctx, BH_ERR_INVALID_ARG, "frame out of range");
    (void)hipSetDevice(ctx->device);
    const size_t npix = (size_t)p->W * p->H;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (rgb) HIP_TRY(ctx, hipMemcpy(rgb, p->d_rgb + npix * 3 * frame, npix * 3, hipMemcpyDeviceToHost));
    if (hit) HIP_TRY(ctx, hipMemcpy(hit, p->d_hit + npix * frame, npix, hipMemcpyDeviceToHost));
    return BH_OK;
  });
}

void bh_path_destroy(bh_path* p) {
  if (!p) return;
  if (p->ctx) {
    (void)hipSetDevice(p->ctx->device);
    (void)hipStreamSynchronize(p->ctx->stream);
  }
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->exec_band) (void)hipGraphExecDestroy(p->exec_band);
  if (p->graph_band) (void)hipGraphDestroy(p->graph_band);
  if (p->d_scenes) (void)hipFree(p->d_scenes);
  if (p->d_rgb) (void)hipFree(p->d_rgb);
  if (p->d_hit) (void)hipFree(p->d_hit);
  if (p->d_qitems) (void)hipFree(p->d_qitems);
  if (p->d_qcount) (void)hipFree(p->d_qcount);
  for (auto ev : p->events)
    if (ev) (void)hipEventDestroy(ev);
  for (hipEvent_t ev : {p->k0, p->k1, p->k2})
    if (ev) (void)hipEventDestroy(ev);
  delete p;
}

/* ---- differentiable render -------------------------------------------------- */
int bh_diff_render_device(bh_ctx* ctx, const bh_diff_params* p, int precision, float* rgb_out,
                          uint8_t* hit_out) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return fail(nullptr, BH_ERR_BAD_HANDLE, "ctx is NULL");
    bh::DiffScene d;
    int rc = build_diff(ctx, p, precision, &d);
    if (rc) return rc;
    (void)hipSetDevice(ctx->device);
    HIP_TRY(ctx, bh::launch_diff_forward(d, precision == BH_PREC_FP64, rgb_out, hit_out, ctx->stream));
    return BH_OK;
  });
}

int bh_diff_adjoint_device(bh_ctx* ctx, const bh_diff_params* p, int precision,
                           const float* grad_rgb, float* rgb_out, uint8_t* hit_out,
                           double* grad_out, uint32_t* overflow_out) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return fail(nullptr, BH_ERR_BAD_HANDLE, "ctx is NULL");
    if (!grad_rgb || !grad_out) return fail(ctx, BH_ERR_INVALID_ARG, "grad_rgb and grad_out are required");
    bh::DiffScene d;
    int rc = build_diff(ctx, p, precision, &d);
    if (rc) return rc;
    (void)hipSetDevice(ctx->device);
    const size_t nb = (size_t)bh::diff_partial_blocks(d.W, d.H);
    if ((rc = ensure(ctx, ctx->dpart, nb * 5 * sizeof(double)))) return rc;
    HIP_TRY(ctx, bh::launch_diff_adjoint(d, precision == BH_PREC_FP64, grad_rgb, rgb_out, hit_out,
                                         (double*)ctx->dpart.p, grad_out, overflow_out, ctx->stream));
    return BH_OK;
  });
}

int bh_diff_time(bh_ctx* ctx, const bh_diff_params* p, int precision, int iters, double* fwd_ms,
                 double* fwd_adj_ms) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return fail(nullptr, BH_ERR_BAD_HANDLE, "ctx is NULL");
    if (iters <= 0) return fail(ctx, BH_ERR_INVALID_ARG, "iters must be positive");
    bh::DiffScene d;
    int rc = build_diff(ctx, p, precision, &d);
    if (rc) return rc;
    (void)hipSetDevice(ctx->device);
    const size_t npix = (size_t)d.W * d.H;
    const size_t nb = (size_t)bh::diff_partial_blocks(d.W, d.H);
    if ((rc = ensure(ctx, ctx->dpart, nb * 5 * sizeof(double)))) return rc;
    if ((rc = ensure(ctx, ctx->dgrad, 8 * sizeof(double)))) return rc;
    if ((rc = ensure(ctx, ctx->drgb, npix * 3 * sizeof(float)))) return rc;
    if ((rc = ensure(ctx, ctx->dhit, npix))) return rc;
    if ((rc = ensure(ctx, ctx->dgin, npix * 3 * sizeof(float)))) return rc;
    hipStream_t st = ctx->stream;
    // dL/drgb: every byte 0x3c, i.e. 0.0115 per channel
    HIP_TRY(ctx, hipMemsetAsync(ctx->dgin.p, 0x3c, npix * 3 * sizeof(float), st));
    float* rgb = (float*)ctx->drgb.p;
    uint8_t* hit = (uint8_t*)ctx->dhit.p;
    const float* gin = (const float*)ctx->dgin.p;
    double* part = (double*)ctx->dpart.p;
    double* grad = (double*)ctx->dgrad.p;
    const bool f64 = precision == BH_PREC_FP64;
    HIP_TRY(ctx, bh::launch_diff_forward(d, f64, rgb, hit, st));
    HIP_TRY(ctx, bh::launch_diff_adjoint(d, f64, gin, rgb, hit, part, grad, nullptr, st));
    HIP_TRY(ctx, hipStreamSynchronize(st));
    float ms_f = 0, ms_a = 0;
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
    for (int i = 0; i < iters; ++i) HIP_TRY(ctx, bh::launch_diff_forward(d, f64, rgb, hit, st));
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
    for (int i = 0; i < iters; ++i)
      HIP_TRY(ctx, bh::launch_diff_adjoint(d, f64, gin, rgb, hit, part, grad, nullptr, st));
    HIP_TRY(ctx, hipEventRecord(ctx->ev2, st));
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev2));
    HIP_TRY(ctx, hipEventElapsedTime(&ms_f, ctx->ev0, ctx->ev1));
    HIP_TRY(ctx, hipEventElapsedTime(&ms_a, ctx->ev1, ctx->ev2));
    if (fwd_ms) *fwd_ms = ms_f / iters;
    if (fwd_adj_ms) *fwd_adj_ms = ms_a / iters;
    return BH_OK;
  });
}

/* ---- batches ------------------------------------------------------------------- */
int bh_render_batch_device(bh_ctx* ctx, int n, const bh_scene* scenes, const bh_options* options,
                           const bh_outputs* outs, const bh_batch_tuning* tune, double* scene_ms,
                           bh_batch_stats* stats) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return fail(nullptr, BH_ERR_BAD_HANDLE, "ctx is NULL");
    const bh_batch_tuning t = normalized_tuning(tune);
    int rc = check_tuning(ctx, t);
    if (rc) return rc;
    BatchPlan p;
    if ((rc = plan_batch(ctx, n, scenes, options, outs, t.mode == BH_BATCH_LANES, &p))) return rc;
    (void)hipSetDevice(ctx->device);
    bh::BatchDev b;
    if ((rc = run_batch(ctx, p, t, &b))) return rc;
    if (!scene_ms && !stats) return BH_OK;
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev2));
    if (stats) {
      float m = 0, tot = 0;
      HIP_TRY(ctx, hipEventElapsedTime(&m, ctx->ev0, ctx->ev1));
      HIP_TRY(ctx, hipEventElapsedTime(&tot, ctx->ev0, ctx->ev2));
      std::vector<uint32_t> cnt(bh::kQueueCountWords);
      HIP_TRY(ctx, hipMemcpy(cnt.data(), b.band.count, cnt.size() * 4, hipMemcpyDeviceToHost));
      int64_t nb = 0;
      for (uint32_t k = 0; k < bh::kQueueSubs; ++k)
        nb += std::min(cnt[16 * k] + cnt[16 * k + 2], b.band.sub_cap);
      stats->march_ms = m;
      stats->band_ms = tot - m;
      stats->total_ms = tot;
      int64_t px = 0;
      for (const auto& d : p.scenes) px += (int64_t)d.W * d.H;
      stats->rays = px;
      stats->band_rays = nb;
      stats->solo_scenes = (int32_t)p.solo.size();
      stats->pad_ = 0;
    }
    if (scene_ms) {
      std::vector<unsigned long long> end(p.fast.size());
      unsigned long long t0 = 0;
      if (!end.empty())
        HIP_TRY(ctx, hipMemcpy(end.data(), b.scene_end, end.size() * 8, hipMemcpyDeviceToHost));
      HIP_TRY(ctx, hipMemcpy(&t0, b.t_start, 8, hipMemcpyDeviceToHost));
      const double khz = ctx->wall_khz > 0 ? ctx->wall_khz : 100000.0;
      float tot = 0;
      HIP_TRY(ctx, hipEventElapsedTime(&tot, ctx->ev0, ctx->ev2));
      for (int i = 0; i < n; ++i) scene_ms[i] = tot;  // solo scenes: end of the call
      for (size_t k = 0; k < p.fast.size(); ++k)
        scene_ms[p.fast[k]] = end[k] >= t0 ? (double)(end[k] - t0) / khz : 0.0;
    }
    return BH_OK;
  });
}

int bh_batch_autotune(bh_ctx* ctx, int n, const bh_scene* scenes, const bh_options* options,
                      const bh_outputs* outs, const bh_batch_tuning* cand, int n_cand, int reps,
                      bh_batch_tuning* best, double* best_ms, double* cand_ms) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return fail(nullptr, BH_ERR_BAD_HANDLE, "ctx is NULL");
    if (!cand || n_cand <= 0 || !best) return fail(ctx, BH_ERR_INVALID_ARG, "need candidates and best");
    if (reps <= 0) reps = 1;
    // one plan per mode (the lane-refill mode renders sphere scenes solo)
    BatchPlan plans[2];
    bool planned[2] = {false, false};
    (void)hipSetDevice(ctx->device);
    double bms = 1e300;
    int bi = -1;
    int rc;
    for (int c = 0; c < n_cand; ++c) {
      const bh_batch_tuning t = normalized_tuning(&cand[c]);
      if ((rc = check_tuning(ctx, t))) return rc;
      const int m = t.mode == BH_BATCH_LANES ? 1 : 0;
      if (!planned[m]) {
        if ((rc = plan_batch(ctx, n, scenes, options, outs, m == 1, &plans[m]))) return rc;
        planned[m] = true;
      }
      const BatchPlan& p = plans[m];
      double cms = 1e300;
      for (int r = 0; r < reps + 1; ++r) {  // first run warms up
        if ((rc = run_batch(ctx, p, t, nullptr))) return rc;
        HIP_TRY(ctx, hipEventSynchronize(ctx->ev2));
        float ms = 0;
        HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev2));
        if (r > 0) cms = std::min(cms, (double)ms);
      }
      if (cand_ms) cand_ms[c] = cms;
      if (cms < bms) {
        bms = cms;
        bi = c;
      }
    }
    *best = normalized_tuning(&cand[bi]);
    if (best_ms) *best_ms = bms;
    return BH_OK;
  });
}

static int build_flat(bh_ctx* ctx, const bh_flat_scene* sc, bh::DevFlat* dp) {
  if (!sc) return fail(ctx, BH_ERR_INVALID_ARG, "scene is NULL");
  const bh_camera& c = sc->camera;
  if (c.width <= 0 || c.height <= 0 || c.width > 32768 || c.height > 32768)
    return fail(ctx, BH_ERR_INVALID_ARG, "camera width/height must be in 1..32768");
  if (!finite3(c.position) || !finite3(c.axis_x) || !finite3(c.axis_y) || !finite3(c.axis_z) ||
      !std::isfinite(c.fov))
    return fail(ctx, BH_ERR_INVALID_ARG, "camera pose/fov must be finite");
  if (sc->n_rects < 0 || sc->n_rects > BH_MAX_RECTS || sc->n_planes < 0 ||
      sc->n_planes > BH_MAX_PLANES)
    return fail(ctx, BH_ERR_INVALID_ARG, "n_rects / n_planes out of range");
  if (sc->steps < 1 || sc->steps > 1000000)
    return fail(ctx, BH_ERR_INVALID_ARG, "steps must be in 1..1e6");
  bh::DevFlat& d = *dp;
  std::memset(&d, 0, sizeof d);
  copy3(d.C, c.position);
  copy3(d.vy, c.axis_y);
  copy3(d.vz, c.axis_z);
  const double f = bh_focal_length(c.width, c.fov);
  for (int i = 0; i < 3; ++i) d.fv[i] = c.axis_x[i] * f;  // Camera::focus_vector
  d.W = c.width;
  d.H = c.height;
  d.n_rects = sc->n_rects;
  d.n_planes = sc->n_planes;
  d.steps = sc->steps;
  for (int i = 0; i < sc->n_rects; ++i) {
    for (int v = 0; v < 5; ++v) {
      if (!finite3(sc->rects[i].vertex[v])) return fail(ctx, BH_ERR_INVALID_ARG, "rectangle vertices must be finite");
      copy3(d.rv[i][v], sc->rects[i].vertex[v]);
    }
    d.rtex[i] = devtex(ctx, sc->rects[i].texture);
    if (!d.rtex[i].data)
      return fail(ctx, BH_ERR_EMPTY_TEXTURE, "rectangle texture slot is empty (Material::SetTexture)");
  }
  for (int i = 0; i < sc->n_planes; ++i) {
    const bh_plane& p = sc->planes[i];
    if (!finite3(p.position) || !finite3(p.axis_x) || !finite3(p.axis_y) || !finite3(p.axis_z) ||
        !std::isfinite(p.pattern_size))
      return fail(ctx, BH_ERR_INVALID_ARG, "plane must be finite");
    // ChessPattern2D(int pattern_size) (pattern.h:20): a whole number; 0 = black
    if (p.pattern_size != 0.0 &&
        !(p.pattern_size >= 1.0 && p.pattern_size <= 536870912.0 &&
          p.pattern_size == std::floor(p.pattern_size)))
      return fail(ctx, BH_ERR_INVALID_ARG, "plane pattern_size must be 0 or a whole number in 1..2^29");
    copy3(d.pp[i], p.position);
    copy3(d.px[i], p.axis_x);
    copy3(d.py[i], p.axis_y);
    copy3(d.pz[i], p.axis_z);
    d.psize[i] = p.pattern_size;
  }
  return BH_OK;
}

int bh_render_flat_device(bh_ctx* ctx, const bh_flat_scene* sc, const bh_outputs* out) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return fail(nullptr, BH_ERR_BAD_HANDLE, "ctx is NULL");
    if (!out) return fail(ctx, BH_ERR_INVALID_ARG, "outputs are NULL");
    bh::DevFlat d;
    int rc = build_flat(ctx, sc, &d);
    if (rc) return rc;
    (void)hipSetDevice(ctx->device);
    DevOut o{out->rgb, out->hit, out->phi_end, out->hit_point, out->steps};
    HIP_TRY(ctx, bh::launch_flat(d, o, ctx->stream));
    return BH_OK;
  });
}

int bh_render_flat(bh_ctx* ctx, const bh_flat_scene* sc, const bh_outputs* out) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return fail(nullptr, BH_ERR_BAD_HANDLE, "ctx is NULL");
    if (!sc || !out) return fail(ctx, BH_ERR_INVALID_ARG, "scene and outputs are required");
    bh::DevFlat d;
    int rc = build_flat(ctx, sc, &d);
    if (rc) return rc;
    (void)hipSetDevice(ctx->device);
    const size_t npix = (size_t)d.W * d.H;
    DevOut o{nullptr, nullptr, nullptr, nullptr, nullptr};
    if (out->rgb) { if ((rc = ensure(ctx, ctx->rgb, npix * 3))) return rc; o.rgb = (uint8_t*)ctx->rgb.p; }
    if (out->hit) { if ((rc = ensure(ctx, ctx->hit, npix))) return rc; o.hit = (uint8_t*)ctx->hit.p; }
    if (out->phi_end) { if ((rc = ensure(ctx, ctx->phi, npix * 8))) return rc; o.phi = (double*)ctx->phi.p; }
    if (out->hit_point) { if ((rc = ensure(ctx, ctx->hitpt, npix * 24))) return rc; o.hitpt = (double*)ctx->hitpt.p; }
    if (out->steps) { if ((rc = ensure(ctx, ctx->steps, npix * 4))) return rc; o.steps = (int32_t*)ctx->steps.p; }
    hipStream_t st = ctx->stream;
    HIP_TRY(ctx, bh::launch_flat(d, o, st));
    if (out->rgb) HIP_TRY(ctx, hipMemcpyAsync(out->rgb, o.rgb, npix * 3, hipMemcpyDeviceToHost, st));
    if (out->hit) HIP_TRY(ctx, hipMemcpyAsync(out->hit, o.hit, npix, hipMemcpyDeviceToHost, st));
    if (out->phi_end) HIP_TRY(ctx, hipMemcpyAsync(out->phi_end, o.phi, npix * 8, hipMemcpyDeviceToHost, st));
    if (out->hit_point) HIP_TRY(ctx, hipMemcpyAsync(out->hit_point, o.hitpt, npix * 24, hipMemcpyDeviceToHost, st));
    if (out->steps) HIP_TRY(ctx, hipMemcpyAsync(out->steps, o.steps, npix * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipStreamSynchronize(st));
    return BH_OK;
  });
}

int bh_time_flat(bh_ctx* ctx, const bh_flat_scene* sc, int iters, double* avg_ms) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return fail(nullptr, BH_ERR_BAD_HANDLE, "ctx is NULL");
    if (iters <= 0) return fail(ctx, BH_ERR_INVALID_ARG, "iters must be positive");
    bh::DevFlat d;
    int rc = build_flat(ctx, sc, &d);
    if (rc) return rc;
    (void)hipSetDevice(ctx->device);
    const size_t npix = (size_t)d.W * d.H;
    if ((rc = ensure(ctx, ctx->rgb, npix * 3))) return rc;
    if ((rc = ensure(ctx, ctx->hit, npix))) return rc;
    DevOut o{(uint8_t*)ctx->rgb.p, (uint8_t*)ctx->hit.p, nullptr, nullptr, nullptr};
    hipStream_t st = ctx->stream;
    HIP_TRY(ctx, bh::launch_flat(d, o, st));  // warm-up
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
    for (int it = 0; it < iters; ++it) HIP_TRY(ctx, bh::launch_flat(d, o, st));
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
    float ms = 0;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (avg_ms) *avg_ms = ms / iters;
    return BH_OK;
  });
}

// ---- frame ring ------------------------------------------------------------
struct bh_ring {
  struct Slot {  // rgb plane then hit plane, one allocation each side
    uint8_t* d_rgb = nullptr;
    uint8_t* d_hit = nullptr;  // d_rgb + 3 npix
    uint8_t* h_rgb = nullptr;  // pinned
    uint8_t* h_hit = nullptr;  // h_rgb + 3 npix
    hipEvent_t start = nullptr, done = nullptr;
    hipEvent_t rendered = nullptr;  // the frame's kernels ended (copy stream waits on it)
    bool busy = false;  // queued or held by the host
  };
  bh_ctx* ctx = nullptr;
  int W = 0, H = 0, next = 0;
  std::vector<Slot> slots;
  std::mutex m;
  std::condition_variable cv;
};

static void ring_free(bh_ring* r) {
  for (auto& s : r->slots) {
    if (s.d_rgb) (void)hipFree(s.d_rgb);
    if (s.h_rgb) (void)hipHostFree(s.h_rgb);
    if (s.start) (void)hipEventDestroy(s.start);
    if (s.done) (void)hipEventDestroy(s.done);
    if (s.rendered) (void)hipEventDestroy(s.rendered);
  }
  delete r;
}

int bh_ring_create(bh_ctx* ctx, int width, int height, int depth, bh_ring** out) {
  return guarded(ctx, [&]() -> int {
    if (!ctx) return fail(nullptr, BH_ERR_BAD_HANDLE, "ctx is NULL");
    if (!out) return fail(ctx, BH_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    if (width <= 0 || height <= 0 || width > 32768 || height > 32768)
      return fail(ctx, BH_ERR_INVALID_ARG, "width/height must be in 1..32768");
    if (depth < 2 || depth > 8) return fail(ctx, BH_ERR_INVALID_ARG, "depth must be in 2..8");
    (void)hipSetDevice(ctx->device);
    bh_ring* r = new bh_ring();
    r->ctx = ctx;
    r->W = width;
    r->H = height;
    r->slots.resize(depth);
    const size_t npix = (size_t)width * height;
    for (auto& s : r->slots) {
      hipError_t e = hipMalloc(&s.d_rgb, npix * 4);
      if (e == hipSuccess) e = hipHostMalloc(&s.h_rgb, npix * 4, hipHostMallocDefault);
      if (e == hipSuccess) {
        s.d_hit = s.d_rgb + npix * 3;
        s.h_hit = s.h_rgb + npix * 3;
      }
      if (e == hipSuccess) e = hipEventCreate(&s.start);
      if (e == hipSuccess) e = hipEventCreate(&s.done);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&s.rendered, hipEventDisableTiming);
      if (e != hipSuccess) {
        ring_free(r);
        return hip_fail(ctx, e, "frame ring allocation");
      }
    }
    *out = r;
    return BH_OK;
  });
}

// Take the next slot, waiting while the host still holds it.
static bh_ring::Slot* ring_acquire(bh_ring* r, int* idx) {
  std::unique_lock<std::mutex> lk(r->m);
  bh_ring::Slot* s = &r->slots[r->next];
  r->cv.wait(lk, [s] { return !s->busy; });
  *idx = r->next;
  r->next = (r->next + 1) % (int)r->slots.size();
  return s;
}

// The frame's copies to pinned memory run on the context's second stream,
// after the frame's kernels, so the next frame's kernels overlap them.  A
// slot's device planes are written again only after the host has released
// the slot, i.e. after its copies completed.
static int ring_finish(bh_ring* r, bh_ring::Slot* s, int idx, int rc, int* slot) {
  bh_ctx* ctx = r->ctx;
  if (rc) return rc;
  const size_t npix = (size_t)r->W * r->H;
  hipStream_t cp = ctx->side;
  HIP_TRY(ctx, hipEventRecord(s->rendered, ctx->stream));
  HIP_TRY(ctx, hipStreamWaitEvent(cp, s->rendered, 0));
  HIP_TRY(ctx, hipMemcpyAsync(s->h_rgb, s->d_rgb, npix * 4, hipMemcpyDeviceToHost, cp));
  HIP_TRY(ctx, hipEventRecord(s->done, cp));
  {
    std::lock_guard<std::mutex> lk(r->m);
    s->busy = true;
  }
  *slot = idx;
  return BH_OK;
}

int bh_ring_render(bh_ring* r, const bh_scene* scene, const bh_options* opt, int* slot) {
  return guarded(nullptr, [&]() -> int {
    if (!r) return fail(nullptr, BH_ERR_BAD_HANDLE, "ring is NULL");
    bh_ctx* ctx = r->ctx;
    if (!scene || !slot) return fail(ctx, BH_ERR_INVALID_ARG, "scene/slot are NULL");
    if (scene->camera.width != r->W || scene->camera.height != r->H)
      return fail(ctx, BH_ERR_INVALID_ARG, "camera size differs from the ring's");
    int idx;
    bh_ring::Slot* s = ring_acquire(r, &idx);
    (void)hipSetDevice(ctx->device);
    HIP_TRY(ctx, hipEventRecord(s->start, ctx->stream));
    bh_outputs o{s->d_rgb, s->d_hit, nullptr, nullptr, nullptr};
    return ring_finish(r, s, idx, bh_render_device(ctx, scene, opt, &o), slot);
  });
}

int bh_ring_render_flat(bh_ring* r, const bh_flat_scene* scene, int* slot) {
  return guarded(nullptr, [&]() -> int {
    if (!r) return fail(nullptr, BH_ERR_BAD_HANDLE, "ring is NULL");
    bh_ctx* ctx = r->ctx;
    if (!scene || !slot) return fail(ctx, BH_ERR_INVALID_ARG, "scene/slot are NULL");
    if (scene->camera.width != r->W || scene->camera.height != r->H)
      return fail(ctx, BH_ERR_INVALID_ARG, "camera size differs from the ring's");
    int idx;
    bh_ring::Slot* s = ring_acquire(r, &idx);
    (void)hipSetDevice(ctx->device);
    HIP_TRY(ctx, hipEventRecord(s->start, ctx->stream));
    bh_outputs o{s->d_rgb, s->d_hit, nullptr, nullptr, nullptr};
    return ring_finish(r, s, idx, bh_render_flat_device(ctx, scene, &o), slot);
  });
}

int bh_ring_wait(bh_ring* r, int slot, const uint8_t** rgb, const uint8_t** hit, float* gpu_ms) {
  return guarded(nullptr, [&]() -> int {
    if (!r) return fail(nullptr, BH_ERR_BAD_HANDLE, "ring is NULL");
    if (slot < 0 || slot >= (int)r->slots.size())
      return fail(nullptr, BH_ERR_INVALID_ARG, "slot out of range");
    bh_ring::Slot& s = r->slots[slot];
    {
      std::lock_guard<std::mutex> lk(r->m);
      if (!s.busy) return fail(nullptr, BH_ERR_INVALID_ARG, "slot holds no frame");
    }
    hipError_t e = hipEventSynchronize(s.done);
    if (e != hipSuccess) return fail(nullptr, BH_ERR_HIP, std::string("hipEventSynchronize: ") + hipGetErrorString(e));
    if (gpu_ms) {
      e = hipEventElapsedTime(gpu_ms, s.start, s.done);
      if (e != hipSuccess) *gpu_ms = -1.0f;
    }
    if (rgb) *rgb = s.h_rgb;
    if (hit) *hit = s.h_hit;
    return BH_OK;
  });
}

int bh_ring_release(bh_ring* r, int slot) {
  return guarded(nullptr, [&]() -> int {
    if (!r) return fail(nullptr, BH_ERR_BAD_HANDLE, "ring is NULL");
    if (slot < 0 || slot >= (int)r->slots.size())
      return fail(nullptr, BH_ERR_INVALID_ARG, "slot out of range");
    {
      std::lock_guard<std::mutex> lk(r->m);
      r->slots[slot].busy = false;
    }
    r->cv.notify_all();
    return BH_OK;
  });
}

void bh_ring_destroy(bh_ring* r) {
  if (!r) return;
  (void)hipSetDevice(r->ctx->device);
  for (auto& s : r->slots)
    if (s.done) (void)hipEventSynchronize(s.done);
  ring_free(r);
}

}  // extern "C"

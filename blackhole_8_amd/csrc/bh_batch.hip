// bh_batch.hip — many independent FP32 scenes (any sizes, masses, cameras)
// marched by one persistent launch from a device ray queue.
//
// The reference renders one scene at a time with a serial x/y loop
// (include/blackhole/blackhole_solution_test.cc:161-298); per ray, this
// kernel runs the same set-up, step and chord tests as the per-frame FP32
// kernel (bh_march.h), so every scene of a batch equals its solo render.
//
// Scheduling, in one of three modes (scenes grouped by class: escape,
// spheres; one kernel instance per class present):
//  grid mode: one ordinary launch with a wave per 8x8 tile of every scene;
//    the hardware dispatcher balances the load, each wave marches its tile
//    with the per-frame kernel's wave-uniform loop (scene constants in
//    scalar registers) and stores its finish time for the per-scene maximum;
//  tile mode: a fixed grid of persistent waves takes ranges of tiles from a
//    device queue and marches them the same way;
//  lane-refill mode: a fixed grid of persistent waves; each lane owns one
//    ray at a time and
//    advances it by its own step counter (a state machine over the reference
//    rule's steps), so lanes of one wave may be at different steps, in
//    different scenes.  After every `refill_steps` steps the wave ballots its
//    idle lanes and, once at least `refill_idle` of its 64 lanes are idle,
//    takes that many ray ids with one atomic add, handing them out by the
//    lane's rank among the idle lanes (mbcnt over the ballot mask).  Live
//    rays stay packed instead of each wave idling until its slowest ray ends,
//    which pays where ray lengths differ widely (escape to r_escape with a
//    large step cap).
// The queue hands out ranges of tiles by guided self-scheduling (one atomic
// per range, ranges shrinking as the queue empties), so the single counter is
// touched a few times per wave rather than once per tile.  Photon-sphere band
// rays are appended (same ballot/prefix-sum scheme) to a queue for the FP64
// pass that follows, and so are the rays whose texel lies within the edge
// guard of a texel edge (the FP64 pass marches them again in the reference's
// order).  A scene's completion time is the latest time a wave finished one
// of its rays in the march (wall clock; one atomic max when a wave moves on
// to another scene); queued rays are finished by the FP64 pass.
#include "bh_internal.h"
#include "bh_march.h"

namespace bh {
namespace {

// Scene owning ray id: the last s with ray_off[s] <= id.
__device__ inline uint32_t scene_of(const BatchDev& b, uint32_t id) {
  uint32_t lo = 0, hi = b.n_scenes;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (b.ray_off[mid] <= id)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Pixel of a scene-local ray id (8x8 tiles); false for tile padding.
__device__ inline bool pixel_of(const DevScene& s, uint32_t local, int* x, int* y) {
  const uint32_t tile = local >> 6, k = local & 63;
  const uint32_t ntx = ((uint32_t)s.W + 7) >> 3;
  *x = (int)((tile % ntx) * 8 + (k & 7));
  *y = (int)((tile / ntx) * 8 + (k >> 3));
  return *x < s.W && *y < s.H;
}

__device__ __forceinline__ uint32_t lane_rank(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ int lane_id() { return (int)lane_rank(~0ull); }

// Per-wave record of the latest finish time of the scene being worked on;
// flushed to the scene's slot (one atomic max) when the wave moves on to
// another scene and when it exits.
struct Finish {
  uint32_t sc;  // wave-uniform; ~0u: none yet
  unsigned long long t;
};

__device__ inline void flush(const BatchDev& b, const Finish& f) {
  if (f.sc != ~0u && lane_id() == 0) atomicMax(b.scene_end + f.sc, f.t);
}

// Note that the lanes with fin set finished a ray of their scene sc now.
__device__ inline void note_finish(const BatchDev& b, Finish& f, bool fin, uint32_t sc) {
  uint64_t m = __ballot(fin);
  if (!m) return;
  const unsigned long long now = (unsigned long long)wall_clock64();
  while (m) {
    const int leader = __builtin_ctzll(m);
    const uint32_t s0 = (uint32_t)__builtin_amdgcn_readlane((int)sc, leader);
    m &= ~__ballot(fin && sc == s0);
    if (s0 != f.sc) {
      flush(b, f);
      f.sc = s0;
    }
    f.t = now;
  }
}

// Guided self-scheduling: take max(chunk, remaining / (4 waves)) units from
// the queue with one atomic add, remaining judged from the head this wave saw
// last.  Returns false once the queue is empty; otherwise [*first, *first +
// *count) are this wave's, all below total.
__device__ inline bool pop_range(const BatchDev& b, uint32_t total, uint32_t waves,
                                 uint32_t* seen, uint32_t* first, uint32_t* count) {
  const uint32_t rem = total > *seen ? total - *seen : 0u;
  uint32_t c = rem / (4u * waves);
  if (c < (uint32_t)b.chunk) c = (uint32_t)b.chunk;
  uint32_t base = 0;
  if (lane_id() == 0) base = atomicAdd(b.head, c);
  base = (uint32_t)__builtin_amdgcn_readfirstlane((int)base);
  *seen = base + c;
  if (base >= total) return false;
  const uint32_t left = total - base;
  *first = base;
  *count = left < c ? left : c;
  return true;
}

struct StepScene {
  double Rh, u0, r_esc;
  float r_in2, r_out2;
  int32_t has_disc, nstep, escape, step_cap;
};

struct RayRun {
  Lane<float> L;
  MarchState<float> st;
  MarchOut<float> o;
  RaySetup r;
  StepScene k;
  uint32_t sc, idx, id;  // scene, pixel index in the scene, ray id
  int n;  // next step
};

// Finish a ray: output planes of its scene.  Returns true when its texel lies
// within the edge guard of a texel edge (the ray goes to the FP64 pass).
__device__ inline bool finish(const BatchDev& b, const RayRun& R) {
  const DevScene& s = b.scenes[R.sc];
  const DevOut out = b.outs[R.sc];
  const float phi0 = phi0_of(out, R.r);
  if (s.escape) return finish_ray<true>(s, out, R.idx, R.r, R.o, phi0);
  return finish_ray<false>(s, out, R.idx, R.r, R.o, phi0);
}

// Advance one ray by one step of the reference rule (the per-frame march's
// loop body, bh_march.h); returns true when the ray has ended.
__device__ inline bool advance(RayRun& R) {
  const StepScene& s = R.k;
  const int nstep = s.nstep, n = R.n;
  R.n = n + 1;
  auto end_ray = [&](int m) {
    R.o.steps = m;
    R.o.a = R.st.a;
    R.o.d = (float)R.st.d;
    return true;
  };
  if (n >= 2 * nstep) {
    // outbound sweep done: end, or (escape) keep stepping outward until the
    // next point would lie at r >= r_escape or the step cap is reached
    if (!s.escape || n > s.step_cap) return end_ray(n - 1);
    const float kt = (float)nstep - 0.1f;
    const float j = (float)(n - nstep);
    const double un = __builtin_fma((double)kt - (double)j, R.o.du64, s.u0);
    if (un * s.r_esc <= 1.0) return end_ray(n - 1);
  }
  float G, ka, kb;
  if (n == nstep) to_outbound<float>(R.L);
  step_coeffs<float>(n, nstep, R.L, &G, &ka, &kb);
  if (march_step<float, StepScene>(s, R.L, R.st, R.r.first_aa, G, ka, kb,
                                   n == 1 && R.r.first_aa != 0.0, R.o.hit))
    return end_ray(n);
  if (n == nstep && R.r.captured) {
    R.o.hit.cls = BH_HIT_HORIZON;
    R.o.hit.ang = R.st.a;
    R.o.hit.hx = (float)s.Rh;
    R.o.hit.hy = 0.f;
    return end_ray(nstep);
  }
  return false;
}

__device__ inline void arm(const DevScene& s, RayRun& R) {
  const double du = (R.r.l - s.u0) * s.inv_nstep;
  setup_lane<float>(s, R.r.l, R.r.ib2, du, R.r.e2x, R.r.e2y, R.r.e2z, R.r.sgn, R.L);
  R.st = MarchState<float>{0.f, 0.0, 0.f};
  R.o.hit.cls = BH_HIT_NONE;
  R.o.hit.ang = 0.0;
  R.o.hit.hx = R.o.hit.hy = 0.f;
  R.o.du64 = du;
  R.o.steps = 0;
  R.o.a = 0.0;
  R.o.d = 0.f;
  R.n = 1;
  R.k = StepScene{s.Rh, s.u0, s.r_esc, s.r_in2, s.r_out2, s.has_disc, s.nstep, s.escape,
                  s.step_cap};
}

// Start ray id on this lane.  Returns true when the lane now holds a live
// ray; otherwise the ray is finished (padding, degenerate) or queued for the
// FP64 band pass (*band).
__device__ inline bool start(const BatchDev& b, uint32_t id, RayRun& R, bool* band, bool* fin) {
  *band = false;
  *fin = false;
  R.sc = scene_of(b, id);
  R.id = id;
  const DevScene& s = b.scenes[R.sc];
  int x, y;
  if (!pixel_of(s, id - b.ray_off[R.sc], &x, &y)) return false;
  R.idx = (uint32_t)y * (uint32_t)s.W + (uint32_t)x;
  setup_ray(s, x, y, false, R.r);
  if (R.r.band && !R.r.degenerate) {
    *band = true;
    return false;
  }
  if (R.r.degenerate) {
    degenerate_out(R.o);
    (void)finish(b, R);  // class 1, black: no texel
    *fin = true;
    return false;
  }
  arm(s, R);
  return true;
}

// Position of slot in sub-queue k: band rays from the front, texel-edge rays
// (back) from the back.
__device__ __forceinline__ size_t queue_pos(const DevQueue& q, uint32_t k, uint32_t slot,
                                            bool back) {
  return (size_t)k * q.sub_cap + (back ? q.sub_cap - 1u - slot : slot);
}

// Append the lanes with flag set, whose ray ids all lie in one tile, to the
// tile's FP64 sub-queue (one atomic per wave): band rays to the front,
// texel-edge rays with their outcome to the back.  Returns true for a lane
// whose ray did not fit; a sub-queue holds every ray id of its tiles, so that
// does not happen.
__device__ inline bool push_band(const BatchDev& b, bool flag, uint32_t id, uint32_t outcome,
                                 bool back) {
  const uint64_t m = __ballot(flag);
  if (!m) return false;
  const int leader = __builtin_ctzll(m);
  const uint32_t k =
      ((uint32_t)__builtin_amdgcn_readlane((int)id, leader) >> 6) % kQueueSubs;
  uint32_t* count = b.band.count + 16 * k + (back ? 2 : 0);
  uint32_t base = 0;
  if (lane_id() == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
  base = (uint32_t)__shfl((int)base, leader);
  if (!flag) return false;
  const uint32_t slot = base + lane_rank(m);
  if (slot < b.band.sub_cap) {
    b.band.items[queue_pos(b.band, k, slot, back)] = (uint64_t)id | ((uint64_t)outcome << 32);
    return false;
  }
  return true;  // overflow: march here
}

// The same for lanes holding rays of different tiles (lane-refill mode): one
// atomic per lane on its own tile's sub-queue.
__device__ inline bool push_band_lane(const BatchDev& b, bool flag, uint32_t id,
                                      uint32_t outcome, bool back) {
  if (!flag) return false;
  const uint32_t k = (id >> 6) % kQueueSubs;
  const uint32_t slot = atomicAdd(b.band.count + 16 * k + (back ? 2 : 0), 1u);
  if (slot >= b.band.sub_cap) return true;
  b.band.items[queue_pos(b.band, k, slot, back)] = (uint64_t)id | ((uint64_t)outcome << 32);
  return false;
}

// Lane-refill mode: every lane its own ray and step counter.  Ray ids come
// from a per-wave reservoir (a range of whole tiles taken by pop_range);
// when the reservoir holds fewer ids than idle lanes, the rest wait for the
// next refill.
__device__ inline void run_lanes(const BatchDev& b, uint32_t waves) {
  RayRun R;
  Finish f{~0u, 0ull};
  bool live = false;
  bool drained = false;  // queue and reservoir empty (wave-uniform)
  uint32_t res = 0, res_end = 0, seen = 0;
  for (;;) {
    const uint64_t idle = __ballot(!live);
    const uint32_t nidle = (uint32_t)__popcll(idle);
    if (!drained && nidle > 0 && (nidle >= (uint32_t)b.refill_idle || nidle == 64)) {
      if (res == res_end) {
        uint32_t t0, cnt;
        if (pop_range(b, b.n_rays >> 6, waves, &seen, &t0, &cnt)) {
          res = t0 << 6;
          res_end = (t0 + cnt) << 6;
        } else {
          drained = true;
        }
      }
      const uint32_t avail = res_end - res;
      bool band = false, fin = false;
      uint32_t id = 0;
      if (!live) {
        const uint32_t rank = lane_rank(idle);
        id = res + rank;
        if (rank < avail) live = start(b, id, R, &band, &fin);
      }
      res += min(nidle, avail);
      note_finish(b, f, fin, R.sc);
      if (push_band_lane(b, band, id, 0u, false)) {
        arm(b.scenes[R.sc], R);  // band queue full: march this ray in FP32 here
        live = true;
      }
    }
    if (!__any(live)) {
      if (drained) break;
      continue;
    }
    bool fin = false, redo = false;
    for (int k = 0; k < b.refill_steps; ++k) {
      if (live && advance(R)) {
        redo = finish(b, R);
        live = false;
        fin = true;
      }
    }
    note_finish(b, f, fin, R.sc);
    // a lane ends at most one ray per refill round, so R is that ray
    push_band_lane(b, redo, R.id, queue_outcome(R.o.hit.cls, R.o.steps), true);
  }
  flush(b, f);
}

// Tile mode: a wave takes a whole 8x8 tile of one scene and marches it with
// the per-frame kernel's wave-uniform loop (scene constants in scalar
// registers).  Returns whether this lane's ray was finished here (not queued
// for the band pass).
template <bool ESCAPE, bool SPH, int NS = 0>
__device__ inline bool march_tile_xy(const BatchDev& b, const DevScene& s, uint32_t sc,
                                     uint32_t id, int x, int y) {
  if (!(x < s.W && y < s.H)) {
    push_band(b, false, id, 0u, false);
    return false;
  }
  const size_t idx = (size_t)y * s.W + x;
  RaySetup r;
  setup_ray(s, x, y, false, r);
  const bool band = r.band && !r.degenerate;
  if (!push_band(b, band, id, 0u, false) && band) return false;
  const float phi0 = phi0_of(b.outs[sc], r);
  MarchOut<float> mo;
  bool sredo = false;
  if (r.degenerate) {
    degenerate_out(mo);
  } else {
    Lane<float> L;
    const double du = (r.l - s.u0) * s.inv_nstep;
    setup_lane<float>(s, r.l, r.ib2, du, r.e2x, r.e2y, r.e2z, r.sgn, L);
    if constexpr (SPH) setup_spheres32(s, r, L);
    march<float, ESCAPE, SPH, NS>(s, L, r.captured, r.first_aa, du, mo);
    if constexpr (SPH) sredo = L.sredo;
  }
  const bool redo = finish_ray<ESCAPE>(s, b.outs[sc], idx, r, mo, phi0) || sredo;
  push_band(b, redo, id, queue_outcome(mo.hit.cls, mo.steps), true);
  return true;
}

template <bool ESCAPE, bool SPH>
__device__ inline bool march_tile(const BatchDev& b, const DevScene& s, uint32_t sc,
                                  uint32_t id) {
  int x, y;
  pixel_of(s, id - b.ray_off[sc], &x, &y);
  return march_tile_xy<ESCAPE, SPH>(b, s, sc, id, x, y);
}

// Tiles [b.tile_lo, b.tile_hi) hold scenes of one class (escape, spheres),
// marched by the kernel instance of that class.
template <bool ESCAPE, bool SPH>
__device__ inline void run_tiles(const BatchDev& b, const DevScene* __restrict__ scenes,
                                 uint32_t waves) {
  const uint32_t n_tiles = b.tile_hi - b.tile_lo;
  Finish f{~0u, 0ull};
  uint32_t seen = 0;
  uint32_t t0, cnt;
  // every successful pop takes at least one tile, so n_tiles + 1 pops bound
  // the loop whatever the queue holds
  for (uint32_t pops = 0; pops <= n_tiles && pop_range(b, n_tiles, waves, &seen, &t0, &cnt);
       ++pops) {
    const uint32_t t_end = min(t0 + cnt, n_tiles);
    for (uint32_t t = t0; t < t_end; ++t) {
      const uint32_t id0 = (b.tile_lo + t) << 6;
      const uint32_t sc = (uint32_t)__builtin_amdgcn_readfirstlane((int)scene_of(b, id0));
      const DevScene& s = scenes[sc];
      const uint32_t id = id0 + (uint32_t)lane_id();
      const bool fin = march_tile<ESCAPE, SPH>(b, s, sc, id);
      note_finish(b, f, fin, sc);
    }
  }
  flush(b, f);
}

// Grid mode: one wave per tile of [b.tile_lo, b.tile_hi), no queue; the
// wave's finish time goes to tile_end[tile] (reduced per scene afterwards).
template <bool ESCAPE, bool SPH, int NS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SPH ? 1 : 8)))
void batch_grid_kernel(BatchDev b, const DevScene* __restrict__ scenes) {
  // the first workgroup is dispatched first: its start is the batch's
  if (blockIdx.x == 0 && threadIdx.x == 0)
    atomicMin(b.t_start, (unsigned long long)wall_clock64());
  const uint32_t tile = b.tile_lo + blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (tile >= b.tile_hi) return;
  const uint32_t id0 = tile << 6;
  // the tile's scene and its position in the scene's grid of 8x8 tiles
  const uint2 ti = b.tile_info[__builtin_amdgcn_readfirstlane((int)tile)];
  const int lane = lane_id();
  const int x = (int)(ti.y & 0xffffu) * 8 + (lane & 7), y = (int)(ti.y >> 16) * 8 + (lane >> 3);
  march_tile_xy<ESCAPE, SPH, NS>(b, scenes[ti.x], ti.x, id0 + (uint32_t)lane, x, y);
  if (lane_id() == 0) b.tile_end[tile] = (unsigned long long)wall_clock64();
}

// Latest tile finish time of each scene (grid mode), and the earliest start.
__global__ __launch_bounds__(256) void batch_grid_finish_kernel(BatchDev b) {
  const uint32_t sc = blockIdx.x;
  const uint32_t t0 = b.ray_off[sc] >> 6, t1 = b.ray_off[sc + 1] >> 6;
  unsigned long long m = 0ull;
  for (uint32_t t = t0 + threadIdx.x; t < t1; t += blockDim.x) m = max(m, b.tile_end[t]);
  __shared__ unsigned long long red[256];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) b.scene_end[sc] = red[0];
}

// The scene table is also passed as a restrict-qualified read-only argument
// so that the tile mode's uniform scene loads go through the scalar cache.
template <bool ESCAPE, bool SPH>
__global__ __launch_bounds__(256) void batch_tile_kernel(BatchDev b,
                                                         const DevScene* __restrict__ scenes) {
  if (lane_id() == 0) atomicMin(b.t_start, (unsigned long long)wall_clock64());
  run_tiles<ESCAPE, SPH>(b, scenes, (gridDim.x * blockDim.x) >> 6);
}

__global__ __launch_bounds__(256) void batch_lane_kernel(BatchDev b) {
  if (lane_id() == 0) atomicMin(b.t_start, (unsigned long long)wall_clock64());
  run_lanes(b, (gridDim.x * blockDim.x) >> 6);
}

// FP64 pass over the queued rays (the per-frame FP64 pass's body with the
// scene of each ray id): workgroup blk takes part blk / kQueueSubs of
// sub-queue blk mod kQueueSubs.
__global__ __launch_bounds__(64) void batch_band_kernel(BatchDev b) {
  const uint32_t k = blockIdx.x % kQueueSubs, part = blockIdx.x / kQueueSubs;
  const uint32_t parts = gridDim.x / kQueueSubs;
  const uint32_t nb = min(b.band.count[16 * k], b.band.sub_cap);
  const uint32_t ne = min(b.band.count[16 * k + 2], b.band.sub_cap - nb);
  const uint64_t* items = b.band.items + (size_t)k * b.band.sub_cap;
  // band rays, then texel-edge rays (from the back)
#pragma unroll 1
  for (uint32_t i = part * blockDim.x + threadIdx.x; i < nb + ne; i += parts * blockDim.x) {
    const uint64_t it = items[i < nb ? i : b.band.sub_cap - 1u - (i - nb)];
    const uint32_t id = (uint32_t)it, outcome = (uint32_t)(it >> 32);
    const uint32_t sc = scene_of(b, id);
    const DevScene& s = b.scenes[sc];
    int x, y;
    pixel_of(s, id - b.ray_off[sc], &x, &y);
    const size_t idx = (size_t)y * s.W + x;
    const bool sph = s.n_spheres > 0;
    if (s.escape) {
      if (sph)
        fp64_ray<true, true>(s, b.outs[sc], idx, x, y, outcome);
      else
        fp64_ray<true, false>(s, b.outs[sc], idx, x, y, outcome);
    } else {
      if (sph)
        fp64_ray<false, true>(s, b.outs[sc], idx, x, y, outcome);
      else
        fp64_ray<false, false>(s, b.outs[sc], idx, x, y, outcome);
    }
  }
}

__global__ void batch_clear_kernel(BatchDev b) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < b.n_scenes) b.scene_end[i] = 0ull;
  if (i < 4) b.head[i] = 0u;
  if (i < kQueueSubs) {
    b.band.count[16 * i] = 0u;
    b.band.count[16 * i + 2] = 0u;
  }
  if (i == 0) b.t_start[0] = ~0ull;
}

// Grid mode: the scene of every tile and the tile's column and row in the
// scene (workgroup s fills scene s's tiles), so that a wave finds its scene
// and pixels with one load rather than a search and a division.
__global__ __launch_bounds__(256) void batch_tile_info_kernel(BatchDev b) {
  const uint32_t s = blockIdx.x;
  const uint32_t t0 = b.ray_off[s] >> 6, t1 = b.ray_off[s + 1] >> 6;
  const uint32_t ntx = ((uint32_t)b.scenes[s].W + 7) >> 3;
  for (uint32_t t = t0 + threadIdx.x; t < t1; t += blockDim.x)
    b.tile_info[t] = make_uint2(s, ((t - t0) % ntx) | (((t - t0) / ntx) << 16));
}

}  // namespace

hipError_t launch_batch_clear(const BatchDev& b, hipStream_t stream) {
  const uint32_t n = b.n_scenes > kQueueSubs ? b.n_scenes : kQueueSubs;
  hipLaunchKernelGGL(batch_clear_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, b);
  if (b.tile_info && b.n_scenes > 0)
    hipLaunchKernelGGL(batch_tile_info_kernel, dim3(b.n_scenes), dim3(256), 0, stream, b);
  return hipGetLastError();
}

hipError_t launch_batch(const BatchDev& b, int cls, int block, int grid, hipStream_t stream) {
  if (b.mode == BH_BATCH_GRID) {
    const uint32_t tiles = b.tile_hi - b.tile_lo;
    const uint32_t per = (uint32_t)block / 64;
    const dim3 g((tiles + per - 1) / per);
    // scenes that all take the reference's step count: the instance compiled
    // for it
    const bool ref = b.nstep == kRefSteps;
    switch (cls) {
      case 0:
        if (ref)
          hipLaunchKernelGGL((batch_grid_kernel<false, false, kRefSteps>), g, dim3(block), 0,
                             stream, b, b.scenes);
        else
          hipLaunchKernelGGL((batch_grid_kernel<false, false, 0>), g, dim3(block), 0, stream, b,
                             b.scenes);
        break;
      case 1:
        if (ref)
          hipLaunchKernelGGL((batch_grid_kernel<true, false, kRefSteps>), g, dim3(block), 0,
                             stream, b, b.scenes);
        else
          hipLaunchKernelGGL((batch_grid_kernel<true, false, 0>), g, dim3(block), 0, stream, b,
                             b.scenes);
        break;
      case 2:
        if (ref)
          hipLaunchKernelGGL((batch_grid_kernel<false, true, kRefSteps>), g, dim3(block), 0,
                             stream, b, b.scenes);
        else
          hipLaunchKernelGGL((batch_grid_kernel<false, true, 0>), g, dim3(block), 0, stream, b,
                             b.scenes);
        break;
      default:
        if (ref)
          hipLaunchKernelGGL((batch_grid_kernel<true, true, kRefSteps>), g, dim3(block), 0,
                             stream, b, b.scenes);
        else
          hipLaunchKernelGGL((batch_grid_kernel<true, true, 0>), g, dim3(block), 0, stream, b,
                             b.scenes);
        break;
    }
    return hipGetLastError();
  }
  if (b.mode == BH_BATCH_LANES) {
    hipLaunchKernelGGL(batch_lane_kernel, dim3(grid), dim3(block), 0, stream, b);
  } else {
    switch (cls) {
      case 0:
        hipLaunchKernelGGL((batch_tile_kernel<false, false>), dim3(grid), dim3(block), 0, stream,
                           b, b.scenes);
        break;
      case 1:
        hipLaunchKernelGGL((batch_tile_kernel<true, false>), dim3(grid), dim3(block), 0, stream,
                           b, b.scenes);
        break;
      case 2:
        hipLaunchKernelGGL((batch_tile_kernel<false, true>), dim3(grid), dim3(block), 0, stream,
                           b, b.scenes);
        break;
      default:
        hipLaunchKernelGGL((batch_tile_kernel<true, true>), dim3(grid), dim3(block), 0, stream,
                           b, b.scenes);
        break;
    }
  }
  return hipGetLastError();
}

hipError_t launch_batch_grid_finish(const BatchDev& b, hipStream_t stream) {
  hipLaunchKernelGGL(batch_grid_finish_kernel, dim3(b.n_scenes), dim3(256), 0, stream, b);
  return hipGetLastError();
}

// 32 one-wave workgroups per sub-queue.
hipError_t launch_batch_band(const BatchDev& b, hipStream_t stream) {
  hipLaunchKernelGGL(batch_band_kernel, dim3(32 * kQueueSubs), dim3(64), 0, stream, b);
  return hipGetLastError();
}

}  // namespace bh

// bh_internal.h — device-side scene layout and kernel launchers shared by the
// C ABI implementation (bh_api.cpp) and the HIP kernels (bh_march.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bh_api.h"

namespace bh {

// Texture in HBM: one packed little-endian RGBA word per texel (channel 0 in
// the low byte), so every lookup is one aligned 4-byte load.
struct DevTex {
  const uint32_t* data;
  int32_t rows, cols;
};

// Everything a ray needs, resolved on the host once per frame.  Lives in
// device memory; every field is wave-uniform and is read with scalar loads.
struct DevScene {
  // camera (camera.h:30-63)
  double C[3];
  double vy[3], vz[3];
  double fv[3];  // axis_x * focal length (Camera::focus_vector)
  int32_t W, H;
  // hole (blackhole_solution.h:24-25,57)
  double B[3];
  double F[3];     // C - B
  double M, bc, Rh;
  double l0;       // SolveG lower bracket cbrt(DBL_EPSILON) (utility.h:17-21)
  double u0;       // 1 / |C - B|
  // disc (Annulus)
  int32_t has_disc, n_spheres;
  double dv0[3], dv1[3], dv2[3], dv4[3];
  double dn[3];
  double r_in, r_out;
  DevTex disc_tex;
  // spheres
  double sc[BH_MAX_SPHERES][3];
  double sr[BH_MAX_SPHERES];
  DevTex stex[BH_MAX_SPHERES];
  // background
  int32_t has_bg, pad0;
  DevTex bg;
  // options
  int32_t nstep, escape, step_cap, filter, root, precision;
  double r_esc, band, edge;

  // ---- per-frame constants of the FP32 march (host-computed) ----
  double fvB[3];          // focus vector relative to B
  double hW, hH;          // W/2 and H/2
  // the pixel point pv = fvB - sx vy - sy vz (sx = W/2 - x, sy = H/2 - y)
  // enters the set-up as dv = F - pv = pxD + sx vy + sy vz,
  // pv x F = pxC - sx pxCy - sy pxCz and dv . F = pxF[0] + sx pxF[1] + sy pxF[2]
  double pxD[3], pxC[3], pxCy[3], pxCz[3], pxF[3];
  double bc2, lo2, hi2;   // b_c^2 and the band's bounds on b^2
  double Uc;              // 1/(3M)
  double r02;             // |C - B|^2
  double cell, inv_cell;  // SolveG grid: (1/(3M) - l0) / 2^20 and its inverse
  double inv_nstep;       // 1 / nstep
  double M2, M3, M6;      // 2M, 3M, 6M
  float Rhf, u0f;         // R_h and u0 in FP32
  float ktf;              // nstep - 0.1 in FP32
  float e1f[3];           // F / |F|
  float nf[3];            // disc normal
  float alpha_f;          // n . e1
  float r_in2, r_out2;    // radius window, squared (r_in2 < 0 when r_in <= 0)
  float Bf[3];
  float tv1[3], ts1[3];   // disc texture square: p1 and s1 = p2 - p1
  float tka, tkb;         // cols / |s1|^2 and rows / (|s1| |s2|)
  float sphf[BH_MAX_SPHERES][4];  // sphere centre relative to B, radius
  // texel-edge guards of the FP32 march: a ray whose texel coordinate lies
  // within these of a texel edge is marched again in FP64 (0: never)
  float edge_a, edge_b;             // disc texel coordinates (texel units)
  float edge_bg;                    // background direction (radians)
  float sph_edge[BH_MAX_SPHERES];   // sphere i's normal direction (radians)
  int32_t fast_ok;        // the FP32 march applies to this scene
  // relative FP32 position error bound: a chord passing a sphere's surface
  // within it (grazing) is marched again in FP64 (0: never)
  float pos_guard;
};

// Output planes (device pointers, any may be null).  Pixel (x, y) of the
// frame lives at index y * W + x of each plane.
struct DevOut {
  uint8_t* rgb;
  uint8_t* hit;
  double* phi;
  double* hitpt;
  int32_t* steps;
};

// FP64 queue: the FP32 march appends the rays it leaves to the FP64 pass —
// the photon-sphere band (|b/b_c - 1| < band) and the rays whose texel lies
// within the edge guard of a texel edge.  An item is the pixel index (low 32
// bits) and, for a texel-edge ray, the outcome of its FP32 march (high 32
// bits: queue_outcome; 0 for a band ray, which the FP32 march left alone).  It is
// split into kQueueSubs sub-queues so that the appends of a frame spread over
// that many counters (one counter serialises every append of the frame):
// sub-queue k takes the pixels of the march's 16x16-pixel workgroups (bx, by)
// with (bx + 37 by) mod kQueueSubs = k (a column or a row of workgroups meets
// every sub-queue, so a line of texel ties is spread over all of them) and
// holds items [k sub_cap, (k + 1) sub_cap): band rays from the front, counted
// by count[16 k], texel-edge rays from the back, counted by count[16 k + 2]
// (the FP64 pass runs the two kinds in separate loops, so that a wave does
// not mix the band rays' full march with the edge rays' short one);
// count[16 k + 1] counts the finished workgroups of the FP64 pass, whose last
// one clears the counts for the next frame.  A pixel is appended at most once
// and sub_cap covers every pixel of the sub-queue's workgroups (at most
// ceil(gx / kQueueSubs) per row of workgroups), so the two ends never meet.
constexpr uint32_t kQueueSubs = 64;
constexpr uint32_t kQueueCountWords = 16 * kQueueSubs;  // count words per queue
struct DevQueue {
  uint64_t* items;
  uint32_t* count;
  uint32_t sub_cap;
};

// Outcome word of a queue item: hit class (BH_HIT_*) and the step it ended
// at (steps < 2^28; the step cap is at most 10^6).
__host__ __device__ inline uint32_t queue_outcome(int cls, int steps) {
  return ((uint32_t)cls << 28) | ((uint32_t)steps & 0x0fffffffu);
}

// Sub-queue of workgroup (bx, by) of a frame.
__host__ __device__ inline uint32_t queue_sub(uint32_t bx, uint32_t by) {
  return (bx + 37u * by) % kQueueSubs;
}
// sub_cap of the queue of a W x H frame: 256 pixels for each of the
// gy ceil(gx / kQueueSubs) workgroups a sub-queue can take.
inline uint32_t queue_sub_cap(int W, int H) {
  const uint64_t gx = (uint64_t)((W + 15) / 16), gy = (uint64_t)((H + 15) / 16);
  return (uint32_t)(gy * ((gx + kQueueSubs - 1) / kQueueSubs) * 256u);
}
inline size_t queue_items(int W, int H) { return (size_t)kQueueSubs * queue_sub_cap(W, H); }

// The allocation-free entry point's file-scope device array (bh_noalloc.hip):
// a scene record, the queue's count words and at most kNoallocQueueItems
// queue items (24 MiB), which take frames up to queue_items(W, H) of that,
// e.g. 2560 x 1080.
constexpr size_t kNoallocQueueItems = size_t(3) << 20;
struct NoallocLayout {
  DevScene* scene;
  uint32_t* count;
  uint64_t* items;
};
// Device address of this device's array (the current device; a host-side
// stand-in of the array's name is no device address).
hipError_t noalloc_scratch(void** ptr);
NoallocLayout noalloc_layout(void* base);
// Scene record into the array, queue counts cleared; on stream.
hipError_t launch_noalloc_setup(const DevScene& s, const NoallocLayout& l, hipStream_t stream);

// FMA issue-rate probe (bh_probe.hip): TFLOP/s of plain FMA chains in
// precision T, 8 waves per SIMD on n_cu compute units; sink: 256 elements.
template <typename T>
hipError_t fma_probe(int n_cu, hipStream_t stream, T* sink, double* tflops);
// The FP32 chord tests' sin and cos of n angles x (bh_probe.hip): out holds
// n (sin, cos) pairs.
hipError_t launch_chord_sincos(const float* x, float* out, int n, hipStream_t stream);

hipError_t launch_march_fp64(const DevScene* scene, const DevOut& out, int W, int H,
                             hipStream_t stream);
// nstep: the scene's step count (the reference's 20 takes the instance
// compiled for it).
hipError_t launch_march_fp32(const DevScene* scene, const DevOut& out, int W, int H,
                             const DevQueue& q, hipStream_t stream, bool escape, bool spheres,
                             int nstep);
hipError_t launch_band(const DevScene* scene, const DevOut& out, const DevQueue& q,
                       hipStream_t stream, bool escape, bool spheres);
// Frames [frame0, frame0 + nf) of a path: frame f's queue is q shifted by f
// queues of queue_items(W, H) items and kQueueCountWords count words.
hipError_t launch_march_frames(const DevScene* scenes, const DevOut& out, int W, int H,
                               int frame0, int nf, const DevQueue& q, hipStream_t stream,
                               bool escape, bool spheres, int nstep);
hipError_t launch_band_frames(const DevScene* scenes, const DevOut& out, int W, int H,
                              int frame0, int nf, const DevQueue& q, hipStream_t stream,
                              bool escape, bool spheres);
hipError_t launch_zero_u32(uint32_t* p, int n, hipStream_t stream);
hipError_t launch_turning(const DevScene* scene, double* l, int W, int H, hipStream_t stream);

// ---------------------------------------------------------------------------
// Batch of independent FP32 scenes marched by one persistent launch.  Ray ids
// enumerate each scene's pixels in 8x8 tiles (row-major tiles, row-major
// within a tile), scenes back to back: scene s owns ids
// [ray_off[s], ray_off[s + 1]), padded to whole tiles.
// ---------------------------------------------------------------------------
struct BatchDev {
  const DevScene* scenes;
  const DevOut* outs;
  const uint32_t* ray_off;   // n_scenes + 1 entries
  uint32_t n_scenes, n_rays;
  uint32_t tile_lo, tile_hi; // tile-mode launch: the tiles of one scene class
  int32_t nstep;             // step count shared by the launch's scenes (0: they differ)
  uint32_t* head;            // queue head of this launch (four slots, cleared together)
  DevQueue band;             // FP64 pass: items hold ray ids (low 32 bits) and
                             // outcomes; sub-queue k takes the tiles k, k +
                             // kQueueSubs, ... (sub_cap = 64 per tile)
  int32_t mode;              // BH_BATCH_GRID, _TILES or _LANES
  int32_t refill_steps;      // steps between refill checks
  int32_t refill_idle;       // refill once at least this many lanes of 64 idle
  int32_t chunk;             // least tiles per queue pop
  unsigned long long* tile_end;   // grid mode: finish time of each tile
  uint2* tile_info;               // grid mode: per tile its scene and (column | row << 16)
                                  // in the scene's tiles (set by the clear launch)
  unsigned long long* scene_end;  // per scene: latest finish time (wall clock)
  unsigned long long* t_start;    // earliest wave start (wall clock)
};

// cls: scene class of a tile-mode launch, escape + 2 spheres
hipError_t launch_batch(const BatchDev& b, int cls, int block, int grid, hipStream_t stream);
hipError_t launch_batch_band(const BatchDev& b, hipStream_t stream);
hipError_t launch_batch_grid_finish(const BatchDev& b, hipStream_t stream);
hipError_t launch_batch_clear(const BatchDev& b, hipStream_t stream);

// ---------------------------------------------------------------------------
// Flat-space scene (RayTracer::Prograde, ray_tracer.h:52-96): straight chords,
// rectangles and chess-patterned infinite planes.  Passed by value.
// ---------------------------------------------------------------------------
struct DevFlat {
  double C[3], fv[3], vy[3], vz[3];  // camera position, focus vector, axes y, z
  int32_t W, H, n_rects, n_planes, steps, pad0;
  double rv[BH_MAX_RECTS][5][3];     // rectangle vertices ([0] pivot, [1..4] p1..p4)
  DevTex rtex[BH_MAX_RECTS];
  double pp[BH_MAX_PLANES][3], px[BH_MAX_PLANES][3], py[BH_MAX_PLANES][3],
      pz[BH_MAX_PLANES][3];
  double psize[BH_MAX_PLANES];
};

hipError_t launch_flat(const DevFlat& s, const DevOut& out, hipStream_t stream);

// ---------------------------------------------------------------------------
// Differentiable hole + disc scene (bh_diff_params), resolved on the host.
// Passed by value as a kernel argument.  Derivatives are with respect to the
// inclination (suffix _i) and the camera distance (suffix _d); the camera
// axes do not depend on the distance.
// ---------------------------------------------------------------------------
struct DiffScene {
  double C[3], C_i[3], C_d[3];        // camera position
  double fx[3], fx_i[3];              // axis_x * focal length (focus vector)
  double ay[3], ay_i[3];              // axis_y
  double az[3], az_i[3];              // axis_z
  double M, bc, Rh, Uc;               // mass, 3 sqrt(3) M, 2M, 1/(3M)
  double r_in, r_out, edge;           // disc radii and soft-edge width
  double half, uscale, vscale;        // texture square half side; cols, rows / (2 half)
  DevTex tex;                         // disc texture (data null: black disc)
  int32_t W, H, nstep, pad0;
};

// Forward: rgb (H*W*3 float) and hit (H*W u8) may be null.
hipError_t launch_diff_forward(const DiffScene& s, bool fp64, float* rgb, uint8_t* hit,
                               hipStream_t stream);
// Forward + adjoint.  partial holds 5 doubles per 256-pixel block (returned by
// diff_partial_blocks); grad_out receives 4 doubles, overflow_out one uint32.
int diff_partial_blocks(int W, int H);
hipError_t launch_diff_adjoint(const DiffScene& s, bool fp64, const float* grad_rgb,
                               float* rgb, uint8_t* hit, double* partial, double* grad_out,
                               uint32_t* overflow_out, hipStream_t stream);

}  // namespace bh

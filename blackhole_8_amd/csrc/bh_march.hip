// bh_march.hip — CDNA4 (gfx950) kernels for the per-pixel photon march.
//
//  march64_kernel  FP64, reference operation order, one lane per pixel
//                  (include/blackhole/blackhole_solution_test.cc:161-298).
//  march_kernel    FP32 event march: FP64 set-up, a 3-FMA + 1-rsqrt step in
//                  the orbit plane, exact object tests only on the steps where
//                  an object can be reached; rays in the photon-sphere band,
//                  and rays whose texel lies within the edge guard of a texel
//                  edge, are queued for band_kernel.
//  band_kernel     the FP64 reference march (march64_kernel's per-pixel code)
//                  of the queued rays.
//  march_frames_kernel / band_frames_kernel
//                  the same two passes over every frame of a camera path in
//                  one launch each.
#include "bh_march.h"

namespace bh {

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

// Append band ray idx to the front of sub-queue k (wave-uniform; the compiler
// turns the lanes' increments into one atomic per wave).  False if it did not
// fit, which the sub-queue sizes rule out.
__device__ __forceinline__ bool queue_push_band(const DevQueue& q, uint32_t k, uint32_t idx) {
  const uint32_t slot = atomicAdd(q.count + 16 * k, 1u);
  if (slot >= q.sub_cap) return false;
  q.items[(size_t)k * q.sub_cap + slot] = (uint64_t)idx;
  return true;
}

// Append texel-edge ray idx and its FP32 outcome to the back of sub-queue k.
__device__ __forceinline__ void queue_push_edge(const DevQueue& q, uint32_t k, uint32_t idx,
                                                uint32_t outcome) {
  const uint32_t slot = atomicAdd(q.count + 16 * k + 2, 1u);
  if (slot < q.sub_cap)
    q.items[(size_t)k * q.sub_cap + (q.sub_cap - 1u - slot)] =
        (uint64_t)idx | ((uint64_t)outcome << 32);
}

// One ray of the FP32 march: pixel (x, y) of frame s, marched by a workgroup
// of the frame whose queue is sub-queue k.
template <bool ESCAPE, bool SPH, int NS>
__device__ __forceinline__ void march_pixel(const DevScene& s, const DevOut& out,
                                            const DevQueue& q, uint32_t k, int x, int y) {
  const size_t idx = (size_t)y * s.W + x;
  RaySetup r;
  setup_ray(s, x, y, false, r);
  // queued for the FP64 march (a full queue would march the ray here)
  if (r.band && !r.degenerate && queue_push_band(q, k, (uint32_t)idx)) return;
  const float phi0 = phi0_of(out, r);
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
  // the texel lies within the FP32 error of a texel edge (or a sphere was
  // grazed): the FP64 pass settles the ray and overwrites what was just stored
  if (finish_ray<ESCAPE>(s, out, idx, r, mo, phi0) || sredo)
    queue_push_edge(q, k, (uint32_t)idx, queue_outcome(mo.hit.cls, mo.steps));
}

// Without spheres the march fits 64 VGPRs: 8 waves per SIMD.
template <bool ESCAPE, bool SPH, int NS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SPH ? 1 : 8)))
void march_kernel(const DevScene* __restrict__ sp, DevOut out, DevQueue q) {
  const DevScene& s = *sp;
  const int x = blockIdx.x * 16 + (threadIdx.x & 15);
  const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= s.W || y >= s.H) return;
  march_pixel<ESCAPE, SPH, NS>(s, out, q, queue_sub(blockIdx.x, blockIdx.y), x, y);
}

// Per-frame planes and queue of frame f of a camera path: planes at offset
// f npix, the queue kQueueSubs sub-queues further per frame.
__device__ __forceinline__ void frame_slice(const DevOut& out, const DevQueue& q, int f,
                                            size_t npix, DevOut* o, DevQueue* qf) {
  o->rgb = out.rgb ? out.rgb + (size_t)f * npix * 3 : nullptr;
  o->hit = out.hit ? out.hit + (size_t)f * npix : nullptr;
  o->phi = out.phi ? out.phi + (size_t)f * npix : nullptr;
  o->hitpt = out.hitpt ? out.hitpt + (size_t)f * npix * 3 : nullptr;
  o->steps = out.steps ? out.steps + (size_t)f * npix : nullptr;
  qf->items = q.items + (size_t)f * kQueueSubs * q.sub_cap;
  qf->count = q.count + (size_t)f * kQueueCountWords;
  qf->sub_cap = q.sub_cap;
}

// Every frame of a camera path in one launch (blockIdx.z = frame - frame0):
// consecutive frames share no ramp-up or drain of the machine.
template <bool ESCAPE, bool SPH, int NS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SPH ? 1 : 8)))
void march_frames_kernel(const DevScene* __restrict__ sp, DevOut out, DevQueue q, size_t npix,
                         int frame0) {
  const int f = frame0 + (int)blockIdx.z;
  const DevScene& s = sp[f];
  const int x = blockIdx.x * 16 + (threadIdx.x & 15);
  const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= s.W || y >= s.H) return;
  DevOut o;
  DevQueue qf;
  frame_slice(out, q, f, npix, &o, &qf);
  march_pixel<ESCAPE, SPH, NS>(s, o, qf, queue_sub(blockIdx.x, blockIdx.y), x, y);
}

// FP64 pass over one frame's queue, by one-wave workgroups of two kinds, so
// that no wave mixes the band rays' full FP64 march with the texel-edge rays'
// short refinement: workgroups blk < kQueueSubs * R take the texel-edge rays
// (the back of the sub-queues; blk takes part blk / kQueueSubs of sub-queue
// blk mod kQueueSubs), the nband workgroups after them the band rays of all
// sub-queues packed together (the front segments, indexed through their
// prefix sum).  The last workgroup to finish clears the frame's counts for
// the next frame; count[1] of sub-queue 0 counts the finished workgroups.
// One round of vector loads over the scene record's 64-byte lines, a lane
// each, before a wave's first ray: a wave of this pass runs nearly alone on
// its CU after the march has filled L2 with framebuffer lines, so the scalar
// loads of the record's constants would each wait on memory in turn.
__device__ __forceinline__ void touch_scene(const DevScene& s, uint32_t lane) {
  constexpr uint32_t kLines = (sizeof(DevScene) + 63) / 64;
  static_assert(kLines <= 64, "one line per lane");
  if (lane < kLines) {
    const uint32_t v = reinterpret_cast<const uint32_t*>(&s)[lane * 16];
    __asm__ volatile("" ::"v"(v) : "memory");
  }
}

template <bool ESCAPE, bool SPH>
__device__ __forceinline__ void band_rays(const DevScene& s, const DevOut& out, const DevQueue& q,
                                          uint32_t blk, uint32_t R, uint32_t nband) {
  __shared__ uint32_t pre[kQueueSubs + 1];
  __shared__ uint32_t last;
  const uint32_t lane = threadIdx.x;
  if (blk < kQueueSubs * R) {
    const uint32_t k = blk % kQueueSubs, part = blk / kQueueSubs;
    const uint32_t* count = q.count + 16 * k;
    const uint64_t* items = q.items + (size_t)k * q.sub_cap;
    const uint32_t nb = min(count[0], q.sub_cap);
    const uint32_t ne = min(count[2], q.sub_cap - nb);
    if (part * 64 < ne) touch_scene(s, lane);
#pragma unroll 1
    for (uint32_t i = part * 64 + lane; i < ne; i += R * 64) {
      const uint64_t it = items[q.sub_cap - 1u - i];
      const uint32_t idx = (uint32_t)it;
      fp64_ray<ESCAPE, SPH>(s, out, idx, (int)(idx % (uint32_t)s.W),
                            (int)(idx / (uint32_t)s.W), (uint32_t)(it >> 32));
    }
  } else {
    // pre[k] = band rays of sub-queues 0 .. k-1
    pre[lane + 1] = min(q.count[16 * lane], q.sub_cap);
    __syncthreads();
    if (lane == 0) {
      pre[0] = 0;
      for (uint32_t k = 1; k <= kQueueSubs; ++k) pre[k] += pre[k - 1];
    }
    __syncthreads();
    const uint32_t total = pre[kQueueSubs];
    if ((blk - kQueueSubs * R) * 64 < total) touch_scene(s, lane);
#pragma unroll 1
    for (uint32_t j = (blk - kQueueSubs * R) * 64 + lane; j < total; j += nband * 64) {
      uint32_t lo = 0, hi = kQueueSubs;  // the k with pre[k] <= j < pre[k + 1]
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (pre[mid] <= j)
          lo = mid;
        else
          hi = mid;
      }
      const uint32_t idx = (uint32_t)q.items[(size_t)lo * q.sub_cap + (j - pre[lo])];
      fp64_ray<ESCAPE, SPH>(s, out, idx, (int)(idx % (uint32_t)s.W),
                            (int)(idx / (uint32_t)s.W), 0u);
    }
  }
  // Every lane's reads of the counts have returned before the barrier, so
  // the reset cannot race with them; the queue is read again only by later
  // launches.  (No device-scope fence here: on this multi-XCD part it writes
  // back and invalidates the XCD's L2, full of framebuffer lines, in every
  // workgroup.)
  __syncthreads();
  if (lane == 0) last = atomicAdd(q.count + 1, 1u) == kQueueSubs * R + nband - 1u ? 1u : 0u;
  __syncthreads();
  if (last) {
    atomicExch(q.count + 16 * lane, 0u);
    atomicExch(q.count + 16 * lane + 2, 0u);
    if (lane == 0) atomicExch(q.count + 1, 0u);
  }
}

// Workgroups of the FP64 pass of one frame: edge workgroups per sub-queue and
// band workgroups.
constexpr uint32_t kEdgeParts1 = 16, kBandWaves1 = 16;     // one frame alone
constexpr uint32_t kEdgePartsN = 2, kBandWavesN = 24;      // each frame of a path

template <bool ESCAPE, bool SPH>
__global__ __launch_bounds__(64) void band_kernel(const DevScene* __restrict__ sp, DevOut out,
                                                  DevQueue q) {
  band_rays<ESCAPE, SPH>(*sp, out, q, blockIdx.x, kEdgeParts1, kBandWaves1);
}

// The queued rays of every frame of a path: blockIdx.y = frame - frame0, the
// gridDim.x workgroups of a row serve that frame's queue.
template <bool ESCAPE, bool SPH>
__global__ __launch_bounds__(64) void band_frames_kernel(const DevScene* __restrict__ sp,
                                                         DevOut out, DevQueue q, size_t npix,
                                                         int frame0) {
  const int f = frame0 + (int)blockIdx.y;
  DevOut o;
  DevQueue qf;
  frame_slice(out, q, f, npix, &o, &qf);
  band_rays<ESCAPE, SPH>(sp[f], o, qf, blockIdx.x, kEdgePartsN, kBandWavesN);
}

__global__ void zero_u32_kernel(uint32_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
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
                             const DevQueue& q, hipStream_t stream, bool escape, bool spheres,
                             int nstep) {
  dim3 grid((W + 15) / 16, (H + 15) / 16);
  const bool ref = nstep == kRefSteps;  // the instance compiled for the reference's nstep
  if (spheres) {
    if (escape && ref)
      hipLaunchKernelGGL((march_kernel<true, true, kRefSteps>), grid, dim3(256), 0, stream,
                         scene, out, q);
    else if (escape)
      hipLaunchKernelGGL((march_kernel<true, true, 0>), grid, dim3(256), 0, stream, scene, out, q);
    else if (ref)
      hipLaunchKernelGGL((march_kernel<false, true, kRefSteps>), grid, dim3(256), 0, stream,
                         scene, out, q);
    else
      hipLaunchKernelGGL((march_kernel<false, true, 0>), grid, dim3(256), 0, stream, scene, out,
                         q);
  } else if (escape) {
    if (ref)
      hipLaunchKernelGGL((march_kernel<true, false, kRefSteps>), grid, dim3(256), 0, stream,
                         scene, out, q);
    else
      hipLaunchKernelGGL((march_kernel<true, false, 0>), grid, dim3(256), 0, stream, scene, out,
                         q);
  } else {
    if (ref)
      hipLaunchKernelGGL((march_kernel<false, false, kRefSteps>), grid, dim3(256), 0, stream,
                         scene, out, q);
    else
      hipLaunchKernelGGL((march_kernel<false, false, 0>), grid, dim3(256), 0, stream, scene, out,
                         q);
  }
  return hipGetLastError();
}

// The turning point of every pixel as the per-ray set-up forms it (0 for a
// captured or degenerate ray).
__global__ __launch_bounds__(256) void turning_kernel(const DevScene* __restrict__ sp,
                                                      double* __restrict__ l) {
  const DevScene& s = *sp;
  const int x = blockIdx.x * 16 + (threadIdx.x & 15);
  const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= s.W || y >= s.H) return;
  RaySetup r;
  setup_ray(s, x, y, false, r);
  l[(size_t)y * s.W + x] = (r.captured || r.degenerate) ? 0.0 : r.l;
}

hipError_t launch_turning(const DevScene* scene, double* l, int W, int H, hipStream_t stream) {
  hipLaunchKernelGGL(turning_kernel, dim3((W + 15) / 16, (H + 15) / 16), dim3(256), 0, stream,
                     scene, l);
  return hipGetLastError();
}

// FP64 pass of one frame; the last workgroup clears the queue, so it is clear
// when the launch completes.
hipError_t launch_band(const DevScene* scene, const DevOut& out, const DevQueue& q,
                       hipStream_t stream, bool escape, bool spheres) {
  const dim3 g(kEdgeParts1 * kQueueSubs + kBandWaves1), b(64);
  if (spheres) {
    if (escape)
      hipLaunchKernelGGL((band_kernel<true, true>), g, b, 0, stream, scene, out, q);
    else
      hipLaunchKernelGGL((band_kernel<false, true>), g, b, 0, stream, scene, out, q);
  } else {
    if (escape)
      hipLaunchKernelGGL((band_kernel<true, false>), g, b, 0, stream, scene, out, q);
    else
      hipLaunchKernelGGL((band_kernel<false, false>), g, b, 0, stream, scene, out, q);
  }
  return hipGetLastError();
}

// Frames [frame0, frame0 + nf) of a camera path, nf <= 65535, all with the
// step count nstep.
hipError_t launch_march_frames(const DevScene* scenes, const DevOut& out, int W, int H,
                               int frame0, int nf, const DevQueue& q, hipStream_t stream,
                               bool escape, bool spheres, int nstep) {
  const dim3 grid((W + 15) / 16, (H + 15) / 16, nf);
  const size_t npix = (size_t)W * H;
  const bool ref = nstep == kRefSteps;
  if (spheres) {
    if (escape && ref)
      hipLaunchKernelGGL((march_frames_kernel<true, true, kRefSteps>), grid, dim3(256), 0,
                         stream, scenes, out, q, npix, frame0);
    else if (escape)
      hipLaunchKernelGGL((march_frames_kernel<true, true, 0>), grid, dim3(256), 0, stream,
                         scenes, out, q, npix, frame0);
    else if (ref)
      hipLaunchKernelGGL((march_frames_kernel<false, true, kRefSteps>), grid, dim3(256), 0,
                         stream, scenes, out, q, npix, frame0);
    else
      hipLaunchKernelGGL((march_frames_kernel<false, true, 0>), grid, dim3(256), 0, stream,
                         scenes, out, q, npix, frame0);
  } else if (escape) {
    if (ref)
      hipLaunchKernelGGL((march_frames_kernel<true, false, kRefSteps>), grid, dim3(256), 0,
                         stream, scenes, out, q, npix, frame0);
    else
      hipLaunchKernelGGL((march_frames_kernel<true, false, 0>), grid, dim3(256), 0, stream,
                         scenes, out, q, npix, frame0);
  } else {
    if (ref)
      hipLaunchKernelGGL((march_frames_kernel<false, false, kRefSteps>), grid, dim3(256), 0,
                         stream, scenes, out, q, npix, frame0);
    else
      hipLaunchKernelGGL((march_frames_kernel<false, false, 0>), grid, dim3(256), 0, stream,
                         scenes, out, q, npix, frame0);
  }
  return hipGetLastError();
}

// FP64 pass of a path: one row of workgroups per frame.
hipError_t launch_band_frames(const DevScene* scenes, const DevOut& out, int W, int H,
                              int frame0, int nf, const DevQueue& q, hipStream_t stream,
                              bool escape, bool spheres) {
  const dim3 g(kEdgePartsN * kQueueSubs + kBandWavesN, nf), b(64);
  const size_t npix = (size_t)W * H;
  if (spheres) {
    if (escape)
      hipLaunchKernelGGL((band_frames_kernel<true, true>), g, b, 0, stream, scenes, out, q, npix,
                         frame0);
    else
      hipLaunchKernelGGL((band_frames_kernel<false, true>), g, b, 0, stream, scenes, out, q, npix,
                         frame0);
  } else {
    if (escape)
      hipLaunchKernelGGL((band_frames_kernel<true, false>), g, b, 0, stream, scenes, out, q, npix,
                         frame0);
    else
      hipLaunchKernelGGL((band_frames_kernel<false, false>), g, b, 0, stream, scenes, out, q,
                         npix, frame0);
  }
  return hipGetLastError();
}

hipError_t launch_zero_u32(uint32_t* p, int n, hipStream_t stream) {
  hipLaunchKernelGGL(zero_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, p, n);
  return hipGetLastError();
}

}  // namespace bh

// bh_march.hip — CDNA4 (gfx950) kernels for the per-pixel photon march.
//
//  march64_kernel  FP64, reference operation order, one lane per pixel
//                  (include/blackhole/blackhole_solution_test.cc:161-298).
//  march_kernel    FP32 event march: FP64 set-up, a 3-FMA + 1-rsqrt step in
//                  the orbit plane, exact object tests only on the steps where
//                  an object can be reached; rays in the photon-sphere band
//                  are queued for band_kernel.
//  band_kernel     the same event march in FP64 for the queued band rays.
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

// One ray of the FP32 march: pixel (x, y) of frame s.
template <bool ESCAPE, bool SPH>
__device__ __forceinline__ void march_pixel(const DevScene& s, const DevOut& out,
                                            const DevQueue& q, int x, int y) {
  const size_t idx = (size_t)y * s.W + x;
  RaySetup r;
  setup_ray(s, x, y, false, r);
  if (r.band && !r.degenerate) {
    // queued for the FP64 march; a full queue marches the ray here in FP32
    const uint32_t slot = atomicAdd(q.count, 1u);
    if (slot < q.capacity) {
      q.items[slot] = (uint32_t)idx;
      return;
    }
  }
  MarchOut mo;
  if (r.degenerate) {
    degenerate_out(mo);
  } else {
    Lane<float> L;
    const double du = (r.l - s.u0) * s.inv_nstep;
    setup_lane<float>(s, r.l, r.ib2, du, r.e2x, r.e2y, r.e2z, r.sgn, L);
    march<float, ESCAPE, SPH>(s, L, r.captured, r.first_aa, du, mo);
  }
  finish_ray<ESCAPE>(s, out, idx, r, mo);
}

template <bool ESCAPE, bool SPH>
__global__ __launch_bounds__(256) void march_kernel(const DevScene* __restrict__ sp, DevOut out,
                                                    DevQueue q) {
  const DevScene& s = *sp;
  const int x = blockIdx.x * 16 + (threadIdx.x & 15);
  const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= s.W || y >= s.H) return;
  march_pixel<ESCAPE, SPH>(s, out, q, x, y);
}

// Per-frame planes and queue slice of frame f of a camera path: planes at
// offset f npix, queue slice f of qstride items and its count pair.
__device__ __forceinline__ void frame_slice(const DevOut& out, const DevQueue& q, int f,
                                            size_t npix, uint32_t qstride, DevOut* o,
                                            DevQueue* qf) {
  o->rgb = out.rgb ? out.rgb + (size_t)f * npix * 3 : nullptr;
  o->hit = out.hit ? out.hit + (size_t)f * npix : nullptr;
  o->phi = out.phi ? out.phi + (size_t)f * npix : nullptr;
  o->hitpt = out.hitpt ? out.hitpt + (size_t)f * npix * 3 : nullptr;
  o->steps = out.steps ? out.steps + (size_t)f * npix : nullptr;
  qf->items = q.items + (size_t)f * qstride;
  qf->count = q.count + 16 * f;
  qf->capacity = q.capacity;
}

// Every frame of a camera path in one launch (blockIdx.z = frame - frame0):
// consecutive frames share no ramp-up or drain of the machine.
template <bool ESCAPE, bool SPH>
__global__ __launch_bounds__(256) void march_frames_kernel(const DevScene* __restrict__ sp,
                                                           DevOut out, DevQueue q, size_t npix,
                                                           uint32_t qstride, int frame0) {
  const int f = frame0 + (int)blockIdx.z;
  const DevScene& s = sp[f];
  const int x = blockIdx.x * 16 + (threadIdx.x & 15);
  const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= s.W || y >= s.H) return;
  DevOut o;
  DevQueue qf;
  frame_slice(out, q, f, npix, qstride, &o, &qf);
  march_pixel<ESCAPE, SPH>(s, o, qf, x, y);
}

// FP64 march of the photon-sphere band rays queued by the FP32 march.  The
// last of the nblk workgroups serving the queue clears it for the next frame.
template <bool ESCAPE, bool SPH>
__device__ __forceinline__ void band_rays(const DevScene& s, const DevOut& out, const DevQueue& q,
                                          uint32_t blk, uint32_t nblk) {
  const uint32_t n = min(q.count[0], q.capacity);
  for (uint32_t i = blk * blockDim.x + threadIdx.x; i < n; i += nblk * blockDim.x) {
    const uint32_t idx = q.items[i];
    const int x = (int)(idx % (uint32_t)s.W), y = (int)(idx / (uint32_t)s.W);
    RaySetup r;
    setup_ray(s, x, y, true, r);
    MarchOut mo;
    Lane<double> L;
    const double du = (r.l - s.u0) * s.inv_nstep;
    setup_lane<double>(s, r.l, r.ib2, du, r.e2x, r.e2y, r.e2z, r.sgn, L);
    march<double, ESCAPE, SPH>(s, L, r.captured, r.first_aa, du, mo);
    finish_ray<ESCAPE>(s, out, idx, r, mo);
  }
  // Every lane's read of count[0] has returned before the barrier, so the
  // reset cannot race with it; the queue is read again only by later
  // launches.  (No device-scope fence here: on this multi-XCD part it writes
  // back and invalidates the XCD's L2, full of framebuffer lines, in every
  // workgroup.)
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t done = atomicAdd(q.count + 1, 1u);
    if (done == nblk - 1) {
      atomicExch(q.count, 0u);
      atomicExch(q.count + 1, 0u);
    }
  }
}

template <bool ESCAPE, bool SPH>
__global__ __launch_bounds__(64) void band_kernel(const DevScene* __restrict__ sp, DevOut out,
                                                  DevQueue q) {
  band_rays<ESCAPE, SPH>(*sp, out, q, blockIdx.x, gridDim.x);
}

// The band rays of every frame of a path: blockIdx.y = frame - frame0, the
// gridDim.x workgroups of a row serve that frame's queue slice.
template <bool ESCAPE, bool SPH>
__global__ __launch_bounds__(64) void band_frames_kernel(const DevScene* __restrict__ sp,
                                                         DevOut out, DevQueue q, size_t npix,
                                                         uint32_t qstride, int frame0) {
  const int f = frame0 + (int)blockIdx.y;
  DevOut o;
  DevQueue qf;
  frame_slice(out, q, f, npix, qstride, &o, &qf);
  band_rays<ESCAPE, SPH>(sp[f], o, qf, blockIdx.x, gridDim.x);
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
                             const DevQueue& q, hipStream_t stream, bool escape, bool spheres) {
  dim3 grid((W + 15) / 16, (H + 15) / 16);
  if (spheres) {
    if (escape)
      hipLaunchKernelGGL((march_kernel<true, true>), grid, dim3(256), 0, stream, scene, out, q);
    else
      hipLaunchKernelGGL((march_kernel<false, true>), grid, dim3(256), 0, stream, scene, out, q);
  } else {
    if (escape)
      hipLaunchKernelGGL((march_kernel<true, false>), grid, dim3(256), 0, stream, scene, out, q);
    else
      hipLaunchKernelGGL((march_kernel<false, false>), grid, dim3(256), 0, stream, scene, out, q);
  }
  return hipGetLastError();
}

// Band grid: 64 workgroups of one wave; every workgroup reaches the queue
// reset, so the queue is clear when the launch completes.
hipError_t launch_band(const DevScene* scene, const DevOut& out, const DevQueue& q,
                       hipStream_t stream, bool escape, bool spheres) {
  if (spheres) {
    if (escape)
      hipLaunchKernelGGL((band_kernel<true, true>), dim3(64), dim3(64), 0, stream, scene, out, q);
    else
      hipLaunchKernelGGL((band_kernel<false, true>), dim3(64), dim3(64), 0, stream, scene, out, q);
  } else {
    if (escape)
      hipLaunchKernelGGL((band_kernel<true, false>), dim3(64), dim3(64), 0, stream, scene, out, q);
    else
      hipLaunchKernelGGL((band_kernel<false, false>), dim3(64), dim3(64), 0, stream, scene, out, q);
  }
  return hipGetLastError();
}

// Frames [frame0, frame0 + nf) of a camera path, nf <= 65535.
hipError_t launch_march_frames(const DevScene* scenes, const DevOut& out, int W, int H,
                               int frame0, int nf, const DevQueue& q, uint32_t qstride,
                               hipStream_t stream, bool escape, bool spheres) {
  const dim3 grid((W + 15) / 16, (H + 15) / 16, nf);
  const size_t npix = (size_t)W * H;
  if (spheres) {
    if (escape)
      hipLaunchKernelGGL((march_frames_kernel<true, true>), grid, dim3(256), 0, stream, scenes,
                         out, q, npix, qstride, frame0);
    else
      hipLaunchKernelGGL((march_frames_kernel<false, true>), grid, dim3(256), 0, stream, scenes,
                         out, q, npix, qstride, frame0);
  } else {
    if (escape)
      hipLaunchKernelGGL((march_frames_kernel<true, false>), grid, dim3(256), 0, stream, scenes,
                         out, q, npix, qstride, frame0);
    else
      hipLaunchKernelGGL((march_frames_kernel<false, false>), grid, dim3(256), 0, stream, scenes,
                         out, q, npix, qstride, frame0);
  }
  return hipGetLastError();
}

// Band grid: 64 one-wave workgroups per frame.
hipError_t launch_band_frames(const DevScene* scenes, const DevOut& out, int W, int H,
                              int frame0, int nf, const DevQueue& q, uint32_t qstride,
                              hipStream_t stream, bool escape, bool spheres) {
  const dim3 grid(64, nf);
  const size_t npix = (size_t)W * H;
  if (spheres) {
    if (escape)
      hipLaunchKernelGGL((band_frames_kernel<true, true>), grid, dim3(64), 0, stream, scenes, out,
                         q, npix, qstride, frame0);
    else
      hipLaunchKernelGGL((band_frames_kernel<false, true>), grid, dim3(64), 0, stream, scenes,
                         out, q, npix, qstride, frame0);
  } else {
    if (escape)
      hipLaunchKernelGGL((band_frames_kernel<true, false>), grid, dim3(64), 0, stream, scenes,
                         out, q, npix, qstride, frame0);
    else
      hipLaunchKernelGGL((band_frames_kernel<false, false>), grid, dim3(64), 0, stream, scenes,
                         out, q, npix, qstride, frame0);
  }
  return hipGetLastError();
}

hipError_t launch_zero_u32(uint32_t* p, int n, hipStream_t stream) {
  hipLaunchKernelGGL(zero_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, p, n);
  return hipGetLastError();
}

}  // namespace bh

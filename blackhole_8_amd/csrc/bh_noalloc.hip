// bh_noalloc.hip — device memory of the allocation-free render entry point
// (bh_render_device_noalloc): one file-scope array per GPU holds the scene
// record and the FP64 queue of a frame, so the entry point needs no hipMalloc.
// Its address on each device is resolved once, when a context is created on
// that device (noalloc_scratch), and only that address is used on the host.
// Every call on a GPU shares that GPU's array: calls that overlap on one GPU
// race.
#include "bh_internal.h"

namespace bh {
namespace {

// Layout: scene record, queue count words, queue items.
constexpr size_t kSceneBytes = (sizeof(DevScene) + 255) / 256 * 256;
constexpr size_t kCountBytes = (kQueueCountWords * sizeof(uint32_t) + 255) / 256 * 256;
constexpr size_t kScratchBytes = kSceneBytes + kCountBytes + kNoallocQueueItems * sizeof(uint64_t);

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Copy the scene record into the array and clear the queue's counts (the
// array keeps the previous call's contents).  A kernel rather than
// hipMemcpyAsync / hipMemsetAsync, so the call may also be captured in a
// graph.
__global__ __launch_bounds__(256) void noalloc_setup_kernel(DevScene s, DevScene* dst,
                                                            uint32_t* count) {
  const uint32_t* src = reinterpret_cast<const uint32_t*>(&s);
  uint32_t* d = reinterpret_cast<uint32_t*>(dst);
  for (uint32_t i = threadIdx.x; i < sizeof(DevScene) / 4; i += blockDim.x) d[i] = src[i];
  for (uint32_t i = threadIdx.x; i < kQueueCountWords; i += blockDim.x) count[i] = 0u;
}

}  // namespace

hipError_t noalloc_scratch(void** ptr) {
  return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch));
}

NoallocLayout noalloc_layout(void* base) {
  unsigned char* b = static_cast<unsigned char*>(base);
  NoallocLayout l;
  l.scene = reinterpret_cast<DevScene*>(b);
  l.count = reinterpret_cast<uint32_t*>(b + kSceneBytes);
  l.items = reinterpret_cast<uint64_t*>(b + kSceneBytes + kCountBytes);
  return l;
}

hipError_t launch_noalloc_setup(const DevScene& s, const NoallocLayout& l, hipStream_t stream) {
  static_assert(sizeof(DevScene) % 4 == 0, "scene record is copied as 32-bit words");
  static_assert(sizeof(DevScene) + 2 * sizeof(void*) <= 4096,
                "the scene record is passed as a kernel argument");
  hipLaunchKernelGGL(noalloc_setup_kernel, dim3(1), dim3(256), 0, stream, s, l.scene, l.count);
  return hipGetLastError();
}

}  // namespace bh

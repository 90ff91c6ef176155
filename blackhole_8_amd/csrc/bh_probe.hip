// bh_probe.hip — issue-rate probe for the measurement harness: the FP32 and
// FP64 fused multiply-add throughput a plain (not packed) VALU instruction
// stream reaches on this device, the ceiling the FP32 march's instruction mix
// is issued against.  Every lane runs eight independent FMA chains, eight
// waves per SIMD, so dependency latency is hidden and the time is issue time.
#include "bh_internal.h"
#include "bh_math.h"

namespace bh {
namespace {

constexpr int kProbeIters = 16384;
constexpr int kProbeChains = 8;

__device__ __forceinline__ float fma_probe_step(float a, float m) {
  return __builtin_fmaf(a, m, 0.5f);
}
__device__ __forceinline__ double fma_probe_step(double a, double m) {
  return __builtin_fma(a, m, 0.5);
}

template <typename T>
__global__ __launch_bounds__(256) void fma_probe_kernel(T* sink, T mul) {
  T a[kProbeChains];
#pragma unroll
  for (int j = 0; j < kProbeChains; ++j) a[j] = (T)threadIdx.x * (T)1e-3 + (T)j;
  for (int i = 0; i < kProbeIters; ++i) {
#pragma unroll
    for (int j = 0; j < kProbeChains; ++j) a[j] = fma_probe_step(a[j], mul);
  }
  T t = 0;
#pragma unroll
  for (int j = 0; j < kProbeChains; ++j) t += a[j];
  // never true for these inputs; keeps the chains live
  if (t == (T)-1) sink[threadIdx.x] = t;
}

}  // namespace

// FMA throughput (TFLOP/s, 2 flops per FMA) of the probe kernel in precision
// T over n_cu compute units, timed with HIP events on stream.
template <typename T>
hipError_t fma_probe(int n_cu, hipStream_t stream, T* sink, double* tflops) {
  const int blocks = n_cu * 8;  // 8 waves per SIMD: 4 SIMDs x 8 / 4 waves per block
  hipEvent_t e0, e1;
  hipError_t e = hipEventCreate(&e0);
  if (e != hipSuccess) return e;
  if ((e = hipEventCreate(&e1)) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return e;
  }
  hipLaunchKernelGGL(fma_probe_kernel<T>, dim3(blocks), dim3(256), 0, stream, sink, (T)0.999);
  (void)hipEventRecord(e0, stream);
  const int reps = 5;
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL(fma_probe_kernel<T>, dim3(blocks), dim3(256), 0, stream, sink, (T)0.999);
  (void)hipEventRecord(e1, stream);
  e = hipEventSynchronize(e1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e == hipSuccess) e = hipGetLastError();
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return e;
  const double fmas = (double)blocks * 256.0 * kProbeIters * kProbeChains * reps;
  *tflops = 2.0 * fmas / (ms * 1e-3) / 1e12;
  return hipSuccess;
}

template hipError_t fma_probe<float>(int, hipStream_t, float*, double*);
template hipError_t fma_probe<double>(int, hipStream_t, double*, double*);

// sin and cos of n FP32 angles as the FP32 march's chord tests form them
// (out: n (sin, cos) pairs).
__global__ __launch_bounds__(256) void chord_sincos_kernel(const float* __restrict__ x,
                                                           float* __restrict__ out, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float s, c;
  sincos_chord<float>(x[i], &s, &c);
  out[2 * i] = s;
  out[2 * i + 1] = c;
}

hipError_t launch_chord_sincos(const float* x, float* out, int n, hipStream_t stream) {
  hipLaunchKernelGGL(chord_sincos_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, x, out, n);
  return hipGetLastError();
}

}  // namespace bh
